// libnbx_objects.so: the reference's FEM input objects for callers without
// libnbots (include/nbx_objects.h).  libnbots_amd.so reads these objects only
// through the accessors of include/nbots_amd.h (NBX_IMPORTS_BEGIN ..
// NBX_IMPORTS_END); this library defines those accessors on objects of its
// own.  It is never linked together with libnbots.
//
// Host code only (no device work): built with the host C++ compiler.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <vector>

#include "../../include/nbx_objects.h"

// ---- triangle mesh (geometric_bot/mesh/mesh2D.c) -----------------------------
struct nb_mesh2D_s {
    std::vector<double> nod;                    // [2 N_nod]
    std::vector<uint32_t> adj;                  // [3 N_elem]
    std::vector<uint32_t> vtx;                  // model vertex -> node
    std::vector<uint32_t> sgm_ptr, sgm_nodes;   // model segment -> node list
};

struct nb_fem_elem_s {
    uint8_t N_nodes, N_gp;
};

struct nb_material_s {
    double E = 0.0, nu = 0.0, density = 0.0;
    double fracture_energy = 0.0, traction_limit = 0.0, compression_limit = 0.0;
    double energy_release_rate = 0.0, damage_length_scale = 0.0;
};

// static_damage2D.c:29-34
struct nb_fem_implicit_s {
    uint32_t N_steps = 0;
    uint32_t N_max_iter = 0;
    uint32_t N_max_iter_without_enhance = 0;
    double residual_tolerance = 0.0;
};

namespace {

struct Atom {
    uint32_t id;
    bool mask[2];
    double val[2];
    void (*fval)(const double* x, double t, double* out);
};

}  // namespace

// Four FIFO lists (Dirichlet/Neumann x point/segment), like the reference's
// NB_QUEUE containers (bcond.c:13).
struct nb_bcond_s {
    uint8_t N_dof = 2;
    std::vector<Atom> lists[4];
};

struct nb_bcond_iter_s {
    const std::vector<Atom>* list = nullptr;
    size_t next = 0;
    const Atom* cur = nullptr;
    uint8_t N_dof = 2;
};

static std::vector<Atom>* bc_list(nb_bcond_t* b, nb_bcond_id t, nb_bcond_where w) {
    return &b->lists[(t == NB_NEUMANN ? 2 : 0) + (w == NB_BC_ON_SEGMENT ? 1 : 0)];
}

static double node_dist(const nb_mesh2D_t* m, uint32_t a, uint32_t b) {
    const double dx = m->nod[2 * (size_t)a] - m->nod[2 * (size_t)b];
    const double dy = m->nod[2 * (size_t)a + 1] - m->nod[2 * (size_t)b + 1];
    return sqrt(dx * dx + dy * dy);
}

extern "C" {

// ---- graph (graph.c:20-118): N_adj and adj share one allocation, as do wij
uint32_t nb_graph_get_memsize(void) { return sizeof(nb_graph_t); }
void nb_graph_init(nb_graph_t* g) { memset(g, 0, sizeof *g); }
void nb_graph_finish(nb_graph_t* g) {
    free(g->N_adj);
    free(g->wi);
    free(g->wij);
}
void nb_graph_clear(nb_graph_t* g) {
    nb_graph_finish(g);
    memset(g, 0, sizeof *g);
}

// ---- mesh ----------------------------------------------------------------------
nb_mesh2D_t* nb_mesh2D_create(nb_mesh2D_type type) {
    if (type != NB_TRIAN) return nullptr;
    return new (std::nothrow) nb_mesh2D_s;
}
void nb_mesh2D_destroy(nb_mesh2D_t* m) { delete m; }
uint32_t nb_mesh2D_get_N_invtx(const nb_mesh2D_t* m) { return (uint32_t)m->vtx.size(); }
uint32_t nb_mesh2D_get_N_insgm(const nb_mesh2D_t* m) {
    return m->sgm_ptr.empty() ? 0 : (uint32_t)m->sgm_ptr.size() - 1;
}
uint32_t nb_mesh2D_get_N_nodes(const nb_mesh2D_t* m) { return (uint32_t)(m->nod.size() / 2); }
uint32_t nb_mesh2D_get_N_elems(const nb_mesh2D_t* m) { return (uint32_t)(m->adj.size() / 3); }
double nb_mesh2D_node_get_x(const nb_mesh2D_t* m, uint32_t id) { return m->nod[2 * (size_t)id]; }
double nb_mesh2D_node_get_y(const nb_mesh2D_t* m, uint32_t id) {
    return m->nod[2 * (size_t)id + 1];
}
uint32_t nb_mesh2D_elem_get_N_adj(const nb_mesh2D_t*, uint32_t) { return 3; }
uint32_t nb_mesh2D_elem_get_adj(const nb_mesh2D_t* m, uint32_t e, uint8_t k) {
    return m->adj[3 * (size_t)e + k];
}
uint32_t nb_mesh2D_get_invtx(const nb_mesh2D_t* m, uint32_t id) { return m->vtx[id]; }
uint32_t nb_mesh2D_insgm_get_N_nodes(const nb_mesh2D_t* m, uint32_t id) {
    return m->sgm_ptr[id + 1] - m->sgm_ptr[id];
}
uint32_t nb_mesh2D_insgm_get_node(const nb_mesh2D_t* m, uint32_t s, uint32_t i) {
    return m->sgm_nodes[m->sgm_ptr[s] + i];
}
// mesh2D.c:650-661: the chord from the segment's first node to its last
double nb_mesh2D_insgm_get_length(const nb_mesh2D_t* m, uint32_t s) {
    const uint32_t n = nb_mesh2D_insgm_get_N_nodes(m, s);
    return node_dist(m, nb_mesh2D_insgm_get_node(m, s, 0), nb_mesh2D_insgm_get_node(m, s, n - 1));
}
// mesh2D.c:663-674: sub-segment i joins the segment's nodes i and i + 1
double nb_mesh2D_insgm_subsgm_get_length(const nb_mesh2D_t* m, uint32_t s, uint32_t i) {
    return node_dist(m, nb_mesh2D_insgm_get_node(m, s, i), nb_mesh2D_insgm_get_node(m, s, i + 1));
}
int nbx_mesh2D_set_trg(nb_mesh2D_t* m, uint32_t N_nod, const double* nod, uint32_t N_elem,
                       const uint32_t* adj, uint32_t N_vtx, const uint32_t* vtx, uint32_t N_sgm,
                       const uint32_t* sgm_ptr, const uint32_t* nod_x_sgm) {
    for (size_t i = 0; i < 3 * (size_t)N_elem; ++i)
        if (adj[i] >= N_nod) return NBX_ERR_ARG;
    for (uint32_t i = 0; i < N_vtx; ++i)
        if (vtx[i] >= N_nod) return NBX_ERR_ARG;
    const uint32_t n_sn = N_sgm ? sgm_ptr[N_sgm] : 0;
    for (uint32_t s = 0; s < N_sgm; ++s)
        if (sgm_ptr[s + 1] < sgm_ptr[s]) return NBX_ERR_ARG;
    for (uint32_t i = 0; i < n_sn; ++i)
        if (nod_x_sgm[i] >= N_nod) return NBX_ERR_ARG;
    m->nod.assign(nod, nod + 2 * (size_t)N_nod);
    m->adj.assign(adj, adj + 3 * (size_t)N_elem);
    m->vtx.assign(vtx, vtx + N_vtx);
    m->sgm_ptr.assign(sgm_ptr, sgm_ptr + N_sgm + 1);
    m->sgm_nodes.assign(nod_x_sgm, nod_x_sgm + n_sn);
    return NBX_OK;
}

// ---- element type (element.c:16-135) ------------------------------------------
nb_fem_elem_t* nb_fem_elem_create(nb_elem_id type) {
    nb_fem_elem_t* e = new (std::nothrow) nb_fem_elem_s;
    if (!e) return nullptr;
    if (type == NB_QUAD_LINEAR) {
        e->N_nodes = 4;
        e->N_gp = 4;
    } else {   // unknown types fall back to linear triangles
        e->N_nodes = 3;
        e->N_gp = 1;
    }
    return e;
}
void nb_fem_elem_destroy(nb_fem_elem_t* e) { delete e; }
uint8_t nb_fem_elem_get_N_gpoints(const nb_fem_elem_t* const e) { return e->N_gp; }
uint8_t nb_fem_elem_get_N_nodes(const nb_fem_elem_t* const e) { return e->N_nodes; }

// ---- material --------------------------------------------------------------------
nb_material_t* nb_material_create(void) { return new (std::nothrow) nb_material_s; }
nb_material_t* nb_material_clone(nb_material_t* m) { return new (std::nothrow) nb_material_s(*m); }
void nb_material_destroy(nb_material_t* m) { delete m; }
void nb_material_set_poisson_module(nb_material_t* m, double v) { m->nu = v; }
void nb_material_set_elasticity_module(nb_material_t* m, double v) { m->E = v; }
void nb_material_set_density(nb_material_t* m, double v) { m->density = v; }
void nb_material_set_fracture_energy(nb_material_t* m, double v) { m->fracture_energy = v; }
void nb_material_set_traction_limit_stress(nb_material_t* m, double v) { m->traction_limit = v; }
void nb_material_set_compression_limit_stress(nb_material_t* m, double v) {
    m->compression_limit = v;
}
void nb_material_set_energy_release_rate(nb_material_t* m, double v) { m->energy_release_rate = v; }
void nb_material_set_damage_length_scale(nb_material_t* m, double v) { m->damage_length_scale = v; }
double nb_material_get_poisson_module(const nb_material_t* const m) { return m->nu; }
double nb_material_get_elasticity_module(const nb_material_t* const m) { return m->E; }
double nb_material_get_density(const nb_material_t* const m) { return m->density; }
double nb_material_get_fracture_energy(const nb_material_t* const m) { return m->fracture_energy; }
double nb_material_get_traction_limit_stress(const nb_material_t* const m) {
    return m->traction_limit;
}
double nb_material_get_compression_limit_stress(const nb_material_t* const m) {
    return m->compression_limit;
}
double nb_material_get_energy_release_rate(const nb_material_t* const m) {
    return m->energy_release_rate;
}
double nb_material_get_damage_length_scale(const nb_material_t* const m) {
    return m->damage_length_scale;
}

// ---- boundary conditions (bcond.c, bcond_iter.c) -----------------------------------
uint16_t nb_bcond_get_memsize(uint8_t) { return (uint16_t)sizeof(nb_bcond_s); }
void nb_bcond_init(void* p, uint8_t N_dof) {
    nb_bcond_s* b = new (p) nb_bcond_s;
    b->N_dof = N_dof;
}
void nb_bcond_finish(void* p) { static_cast<nb_bcond_s*>(p)->~nb_bcond_s(); }
void* nb_bcond_create(uint8_t N_dof) {
    nb_bcond_s* b = new (std::nothrow) nb_bcond_s;
    if (b) b->N_dof = N_dof;
    return b;
}
void nb_bcond_destroy(void* p) { delete static_cast<nb_bcond_s*>(p); }
void nb_bcond_clear(void* p) {
    for (auto& l : static_cast<nb_bcond_s*>(p)->lists) l.clear();
}
uint8_t nb_bcond_get_N_dof(const nb_bcond_t* const b) { return b->N_dof; }

void nb_bcond_push(nb_bcond_t* b, nb_bcond_id t, nb_bcond_where w, uint32_t id, const bool mask[],
                   const double value[]) {
    Atom a{id, {mask[0], b->N_dof > 1 ? mask[1] : false}, {0.0, 0.0}, nullptr};
    for (int j = 0; j < b->N_dof && j < 2; ++j) a.val[j] = value[j];
    bc_list(b, t, w)->push_back(a);
}

void nb_bcond_push_function(nb_bcond_t* b, nb_bcond_id t, nb_bcond_where w, uint32_t id,
                            const bool mask[], void (*fval)(const double*, double, double*)) {
    Atom a{id, {mask[0], b->N_dof > 1 ? mask[1] : false}, {0.0, 0.0}, fval};
    bc_list(b, t, w)->push_back(a);
}

uint16_t nb_bcond_iter_get_memsize(void) { return (uint16_t)sizeof(nb_bcond_iter_s); }
void nb_bcond_iter_init(void* p) { new (p) nb_bcond_iter_s; }
void nb_bcond_iter_finish(void*) {}
void* nb_bcond_iter_create(void) { return new (std::nothrow) nb_bcond_iter_s; }
void nb_bcond_iter_destroy(void* p) { delete static_cast<nb_bcond_iter_s*>(p); }
void nb_bcond_iter_set_conditions(nb_bcond_iter_t* it, const nb_bcond_t* b, nb_bcond_id t,
                                  nb_bcond_where w) {
    it->list = bc_list(const_cast<nb_bcond_t*>(b), t, w);
    it->next = 0;
    it->cur = nullptr;
    it->N_dof = b->N_dof;
}
bool nb_bcond_iter_has_more(const nb_bcond_iter_t* it) {
    return it->list && it->next < it->list->size();
}
void nb_bcond_iter_go_next(nb_bcond_iter_t* it) { it->cur = &(*it->list)[it->next++]; }
uint32_t nb_bcond_iter_get_id(const nb_bcond_iter_t* it) { return it->cur->id; }
bool nb_bcond_iter_get_mask(const nb_bcond_iter_t* it, uint8_t j) {
    return j < 2 ? it->cur->mask[j] : false;
}
bool nb_bcond_iter_val_is_function(const nb_bcond_iter_t* it) { return it->cur->fval != nullptr; }
// bcond_iter.c:116-123: a function is evaluated at (x, t), else the values
void nb_bcond_iter_get_val(const nb_bcond_iter_t* it, uint8_t N_dof, double* x, double t,
                           double val[]) {
    if (it->cur->fval)
        it->cur->fval(x, t, val);
    else
        for (uint8_t j = 0; j < N_dof && j < 2; ++j) val[j] = it->cur->val[j];
}

// ---- implicit-scheme parameters (static_damage2D.c:79-126) -----------------------
nb_fem_implicit_t* nb_fem_implicit_create(void) { return new (std::nothrow) nb_fem_implicit_s; }
void nb_fem_implicit_destroy(nb_fem_implicit_t* p) { delete p; }
void nb_fem_implicit_set_N_steps(nb_fem_implicit_t* p, uint32_t n) { p->N_steps = n; }
void nb_fem_implicit_set_N_max_iter(nb_fem_implicit_t* p, uint32_t n) { p->N_max_iter = n; }
void nb_fem_implicit_set_N_max_iter_without_enhance(nb_fem_implicit_t* p, uint32_t n) {
    p->N_max_iter_without_enhance = n;
}
void nb_fem_implicit_set_residual_tolerance(nb_fem_implicit_t* p, double t) {
    p->residual_tolerance = t;
}
uint32_t nb_fem_implicit_get_N_steps(nb_fem_implicit_t* p) { return p->N_steps; }
uint32_t nb_fem_implicit_get_N_max_iter(nb_fem_implicit_t* p) { return p->N_max_iter; }
uint32_t nb_fem_implicit_get_N_max_iter_without_enhance(nb_fem_implicit_t* p) {
    return p->N_max_iter_without_enhance;
}
double nb_fem_implicit_get_residual_tolerance(nb_fem_implicit_t* p) {
    return p->residual_tolerance;
}

}  // extern "C"

// Problem import on the host: boundary-condition plans in the reference's
// order and arithmetic (set_bconditions.c:52-260), and the reading of the
// reference's own mesh and bcond objects through their public accessors.
//
// This library never defines the mesh, bcond, material or element types:
// nb_fem_compute_2D_Solid_Mechanics receives objects that libnbots built,
// and reads them only through functions that libnbots exports
// (mesh2D.h, bcond.h, bcond_iter.h).  Callers without libnbots get the same
// functions from libnbx_objects.so.
#include <math.h>
#include <string.h>

#include "bcplan.hpp"

namespace nbx {

BcPlan PlanBuilder::finish() const {
    BcPlan p;
    p.n_ptr.push_back(0);
    for (const auto& kv : nadd) {
        p.n_dof.push_back(kv.first);
        for (double a : kv.second) p.n_add.push_back(a);
        p.n_ptr.push_back((uint32_t)p.n_add.size());
    }
    std::map<uint32_t, size_t> where;
    for (size_t i = 0; i < dseq.size(); ++i) {
        const uint32_t dof = dseq[i].first;
        auto it = where.find(dof);
        if (it == where.end()) {
            where[dof] = p.d_dof.size();
            p.d_dof.push_back(dof);
            p.d_rank.push_back((uint32_t)i);
            p.d_vfirst.push_back(dseq[i].second);
            p.d_vlast.push_back(dseq[i].second);
        } else {
            p.d_vlast[it->second] = dseq[i].second;
        }
    }
    return p;
}

namespace {

double seg_len(const double* nod, uint32_t a, uint32_t b) {
    // mesh2D.c:650-674: the distance between two nodes (a segment's length
    // is the chord from its first to its last node)
    const double x1 = nod[2 * a], y1 = nod[2 * a + 1];
    const double x2 = nod[2 * b], y2 = nod[2 * b + 1];
    return sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
}

int bad_id(const char* what, uint32_t id) {
    set_last_error("boundary condition on %s %u does not exist", what, id);
    return NBX_ERR_ARG;
}

int check_nodes(const BcPlan& p, uint32_t nn) {
    for (uint32_t d : p.d_dof)
        if (d >= 2 * nn) return bad_id("node", d / 2);
    for (uint32_t d : p.n_dof)
        if (d >= 2 * nn) return bad_id("node", d / 2);
    return NBX_OK;
}

}  // namespace

// set_neumann_sgm (:63-154), set_neumann_vtx (:173-190),
// set_dirichlet_sgm (:192-238), set_dirichlet_vtx (:240-260).
int plan_from_lists(const double* nod, uint32_t nn, const uint32_t* vtx, uint32_t n_vtx,
                    const uint32_t* sgm_ptr, const uint32_t* sgm_nodes, uint32_t n_sgm,
                    const std::vector<BcAtom>& dv, const std::vector<BcAtom>& nv,
                    const std::vector<BcAtom>& ds, const std::vector<BcAtom>& ns,
                    const std::vector<const double*>& ns_samples, double factor, BcPlan& out) {
    PlanBuilder pb;
    for (size_t c = 0; c < ns.size(); ++c) {
        const BcAtom& a = ns[c];
        if (a.id >= n_sgm) return bad_id("segment", a.id);
        const uint32_t* nodes = sgm_nodes + sgm_ptr[a.id];
        const uint32_t N = sgm_ptr[a.id + 1] - sgm_ptr[a.id];
        if (N < 2) continue;
        const double* fv = ns_samples.empty() ? nullptr : ns_samples[c];
        if (fv) {   // trapezoid integration of sampled tractions (:87-131)
            double val1[2] = {fv[0], fv[1]};
            uint32_t v1 = nodes[0];
            for (uint32_t i = 0; i + 1 < N; ++i) {
                const double sub = seg_len(nod, nodes[i], nodes[i + 1]);
                const uint32_t v2 = nodes[i + 1];
                const double val2[2] = {fv[2 * (i + 1)], fv[2 * (i + 1) + 1]};
                for (int j = 0; j < 2; ++j)
                    if (a.mask[j]) {
                        const double val = 0.5 * (val1[j] + val2[j]) * sub;
                        pb.neumann(v1 * 2 + j, factor * val * 0.5);
                        pb.neumann(v2 * 2 + j, factor * val * 0.5);
                    }
                v1 = v2;
                val1[0] = val2[0];
                val1[1] = val2[1];
            }
        } else {    // constant total force split by length (:133-154)
            const double L = seg_len(nod, nodes[0], nodes[N - 1]);
            for (uint32_t i = 0; i + 1 < N; ++i) {
                const double w = seg_len(nod, nodes[i], nodes[i + 1]) / L;
                const double f = factor * w * 0.5;
                for (int j = 0; j < 2; ++j)
                    if (a.mask[j]) pb.neumann(nodes[i] * 2 + j, f * a.val[j]);
                for (int j = 0; j < 2; ++j)
                    if (a.mask[j]) pb.neumann(nodes[i + 1] * 2 + j, f * a.val[j]);
            }
        }
    }
    for (const BcAtom& a : nv) {
        if (a.id >= n_vtx) return bad_id("vertex", a.id);
        const uint32_t v = vtx[a.id];
        for (int j = 0; j < 2; ++j)
            if (a.mask[j]) pb.neumann(v * 2 + j, factor * a.val[j]);
    }
    auto dirichlet_node = [&](const BcAtom& a, uint32_t v) {
        for (int j = 0; j < 2; ++j)
            if (a.mask[j]) pb.dirichlet(v * 2 + j, factor * a.val[j]);
    };
    for (const BcAtom& a : ds) {
        if (a.id >= n_sgm) return bad_id("segment", a.id);
        for (uint32_t i = sgm_ptr[a.id]; i < sgm_ptr[a.id + 1]; ++i) dirichlet_node(a, sgm_nodes[i]);
    }
    for (const BcAtom& a : dv) {
        if (a.id >= n_vtx) return bad_id("vertex", a.id);
        dirichlet_node(a, vtx[a.id]);
    }
    out = pb.finish();
    return check_nodes(out, nn);
}

int import_mesh(const nb_mesh2D_t* part, MeshArrays& out) {
    const uint32_t nn = nb_mesh2D_get_N_nodes(part);
    const uint32_t ne = nb_mesh2D_get_N_elems(part);
    out.nn = nn;
    out.ne = ne;
    out.nod.resize(2 * (size_t)nn);
    out.conn.resize(3 * (size_t)ne);
    for (uint32_t i = 0; i < nn; ++i) {
        out.nod[2 * (size_t)i] = nb_mesh2D_node_get_x(part, i);
        out.nod[2 * (size_t)i + 1] = nb_mesh2D_node_get_y(part, i);
    }
    for (uint32_t e = 0; e < ne; ++e) {
        if (nb_mesh2D_elem_get_N_adj(part, e) != 3) {
            set_last_error("element %u has %u nodes: the solid-mechanics path takes linear "
                           "triangles", e, nb_mesh2D_elem_get_N_adj(part, e));
            return NBX_ERR_ARG;
        }
        for (uint8_t k = 0; k < 3; ++k) {
            const uint32_t v = nb_mesh2D_elem_get_adj(part, e, k);
            if (v >= nn) {
                set_last_error("triangle %u references node %u >= %u", e, v, nn);
                return NBX_ERR_ARG;
            }
            out.conn[3 * (size_t)e + k] = v;
        }
    }
    return NBX_OK;
}

namespace {

// One category of conditions, visited with the reference's iterator.
struct Iter {
    nb_bcond_iter_t* it;
    Iter(const nb_bcond_t* bc, nb_bcond_id t, nb_bcond_where w)
        : it((nb_bcond_iter_t*)nb_bcond_iter_create()) {
        nb_bcond_iter_set_conditions(it, bc, t, w);
    }
    ~Iter() { nb_bcond_iter_destroy(it); }
    Iter(const Iter&) = delete;
    Iter& operator=(const Iter&) = delete;
};

// set_neumann (:156-171): values at x = (0, 0), each masked dof += f val[j]
void add_neumann(PlanBuilder& pb, const nb_bcond_iter_t* it, double f, uint32_t v) {
    double x0[2] = {0.0, 0.0};
    double val[2] = {0.0, 0.0};
    nb_bcond_iter_get_val(it, 2, x0, 0.0, val);
    for (uint8_t j = 0; j < 2; ++j)
        if (nb_bcond_iter_get_mask(it, j)) pb.neumann(v * 2 + j, f * val[j]);
}

// set_dirichlet (:218-238): values at the node, each masked dof eliminated
void add_dirichlet(PlanBuilder& pb, const nb_mesh2D_t* part, const nb_bcond_iter_t* it,
                   double factor, uint32_t v) {
    double x[2] = {nb_mesh2D_node_get_x(part, v), nb_mesh2D_node_get_y(part, v)};
    double val[2] = {0.0, 0.0};
    nb_bcond_iter_get_val(it, 2, x, 0.0, val);
    for (uint8_t j = 0; j < 2; ++j)
        if (nb_bcond_iter_get_mask(it, j)) pb.dirichlet(v * 2 + j, factor * val[j]);
}

}  // namespace

int import_bconditions(const nb_mesh2D_t* part, const nb_bcond_t* bcond, double factor,
                       BcPlan& out) {
    if (nb_bcond_get_N_dof(bcond) != 2) {
        set_last_error("solid mechanics needs 2 dofs per node (bcond has %u)",
                       (unsigned)nb_bcond_get_N_dof(bcond));
        return NBX_ERR_ARG;
    }
    const uint32_t n_sgm = nb_mesh2D_get_N_insgm(part);
    const uint32_t n_vtx = nb_mesh2D_get_N_invtx(part);
    PlanBuilder pb;
    {   // set_neumann_sgm (:63-85)
        Iter c(bcond, NB_NEUMANN, NB_BC_ON_SEGMENT);
        while (nb_bcond_iter_has_more(c.it)) {
            nb_bcond_iter_go_next(c.it);
            const uint32_t id = nb_bcond_iter_get_id(c.it);
            if (id >= n_sgm) return bad_id("segment", id);
            const uint32_t N = nb_mesh2D_insgm_get_N_nodes(part, id);
            if (N < 2) continue;
            if (nb_bcond_iter_val_is_function(c.it)) {
                // set_neumann_sgm_function (:87-131): trapezoid rule over
                // the sub-segments, tractions sampled at the nodes
                uint32_t v1 = nb_mesh2D_insgm_get_node(part, id, 0);
                double x[2] = {nb_mesh2D_node_get_x(part, v1), nb_mesh2D_node_get_y(part, v1)};
                double val1[2] = {0.0, 0.0}, val2[2] = {0.0, 0.0};
                nb_bcond_iter_get_val(c.it, 2, x, 0.0, val1);
                for (uint32_t i = 0; i + 1 < N; ++i) {
                    const double sub = nb_mesh2D_insgm_subsgm_get_length(part, id, i);
                    const uint32_t v2 = nb_mesh2D_insgm_get_node(part, id, i + 1);
                    x[0] = nb_mesh2D_node_get_x(part, v2);
                    x[1] = nb_mesh2D_node_get_y(part, v2);
                    nb_bcond_iter_get_val(c.it, 2, x, 0.0, val2);
                    for (uint8_t j = 0; j < 2; ++j)
                        if (nb_bcond_iter_get_mask(c.it, j)) {
                            const double val = 0.5 * (val1[j] + val2[j]) * sub;
                            pb.neumann(v1 * 2 + j, factor * val * 0.5);
                            pb.neumann(v2 * 2 + j, factor * val * 0.5);
                        }
                    v1 = v2;
                    val1[0] = val2[0];
                    val1[1] = val2[1];
                }
            } else {
                // set_neumann_sgm_integrated (:133-154): the value is a total
                // force, split by sub-segment length over the chord
                const double L = nb_mesh2D_insgm_get_length(part, id);
                for (uint32_t i = 0; i + 1 < N; ++i) {
                    const double w = nb_mesh2D_insgm_subsgm_get_length(part, id, i) / L;
                    const uint32_t v1 = nb_mesh2D_insgm_get_node(part, id, i);
                    const uint32_t v2 = nb_mesh2D_insgm_get_node(part, id, i + 1);
                    add_neumann(pb, c.it, factor * w * 0.5, v1);
                    add_neumann(pb, c.it, factor * w * 0.5, v2);
                }
            }
        }
    }
    {   // set_neumann_vtx (:173-190)
        Iter c(bcond, NB_NEUMANN, NB_BC_ON_POINT);
        while (nb_bcond_iter_has_more(c.it)) {
            nb_bcond_iter_go_next(c.it);
            const uint32_t id = nb_bcond_iter_get_id(c.it);
            if (id >= n_vtx) return bad_id("vertex", id);
            add_neumann(pb, c.it, factor, nb_mesh2D_get_invtx(part, id));
        }
    }
    {   // set_dirichlet_sgm (:192-216)
        Iter c(bcond, NB_DIRICHLET, NB_BC_ON_SEGMENT);
        while (nb_bcond_iter_has_more(c.it)) {
            nb_bcond_iter_go_next(c.it);
            const uint32_t id = nb_bcond_iter_get_id(c.it);
            if (id >= n_sgm) return bad_id("segment", id);
            const uint32_t N = nb_mesh2D_insgm_get_N_nodes(part, id);
            for (uint32_t i = 0; i < N; ++i)
                add_dirichlet(pb, part, c.it, factor, nb_mesh2D_insgm_get_node(part, id, i));
        }
    }
    {   // set_dirichlet_vtx (:240-260)
        Iter c(bcond, NB_DIRICHLET, NB_BC_ON_POINT);
        while (nb_bcond_iter_has_more(c.it)) {
            nb_bcond_iter_go_next(c.it);
            const uint32_t id = nb_bcond_iter_get_id(c.it);
            if (id >= n_vtx) return bad_id("vertex", id);
            add_dirichlet(pb, part, c.it, factor, nb_mesh2D_get_invtx(part, id));
        }
    }
    out = pb.finish();
    return check_nodes(out, nb_mesh2D_get_N_nodes(part));
}

}  // namespace nbx

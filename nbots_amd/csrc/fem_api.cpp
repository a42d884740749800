// C-ABI of the device FEM system (nbx_fem_*) and the drop-in FEM driver
// nb_fem_compute_2D_Solid_Mechanics with the objects it takes (triangle
// mesh, material, element type, boundary conditions).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <new>
#include <vector>

#include "../../include/nbots_amd.h"
#include "fem.hpp"
#include "opts.hpp"
#include "simp.hpp"

struct nbx_fem_s {
    nbx::FemSys sys;
};

// ---- drop-in objects -------------------------------------------------------
struct nb_mesh2D_s {
    std::vector<double> nod;
    std::vector<uint32_t> adj, vtx, sgm_ptr, sgm_nodes;
};

struct nb_fem_elem_s {
    nb_elem_id type;
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


struct BcAtom {
    uint32_t id;
    bool mask[2];
    double val[2];
    void (*fval)(const double* x, double t, double* out);
};

}  // namespace

// FIFO lists, like the reference's NB_QUEUE containers (bcond.c:13).
struct nb_bcond_s {
    uint8_t N_dof = 2;
    std::vector<BcAtom> dv, nv, ds, ns;   // Dirichlet/Neumann x vertex/segment
};

namespace {

double seg_len(const double* nod, uint32_t a, uint32_t b) {
    // geometric_bot/mesh/mesh2D/mesh2D.c:650-674
    const double x1 = nod[2 * a], y1 = nod[2 * a + 1];
    const double x2 = nod[2 * b], y2 = nod[2 * b + 1];
    return sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
}

// Collects the Neumann additions per dof and the Dirichlet eliminations in
// the order of pde_bot/finite_element/solid_mechanics/set_bconditions.c.
struct PlanBuilder {
    std::map<uint32_t, std::vector<double>> nadd;
    std::vector<std::pair<uint32_t, double>> dseq;

    void neumann(uint32_t dof, double v) { nadd[dof].push_back(v); }
    void dirichlet(uint32_t dof, double v) { dseq.emplace_back(dof, v); }

    nbx::BcPlan finish() const {
        nbx::BcPlan p;
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
};

// set_neumann_sgm (:63-154), set_neumann_vtx (:173-190),
// set_dirichlet_sgm (:192-238), set_dirichlet_vtx (:240-260).
int plan_from_lists(const double* nod, uint32_t nn, const uint32_t* vtx, uint32_t n_vtx,
                    const uint32_t* sgm_ptr, const uint32_t* sgm_nodes, uint32_t n_sgm,
                    const std::vector<BcAtom>& dv, const std::vector<BcAtom>& nv,
                    const std::vector<BcAtom>& ds, const std::vector<BcAtom>& ns,
                    const std::vector<const double*>& ns_samples, double factor,
                    nbx::BcPlan& out) {
    PlanBuilder pb;
    auto bad = [&](const char* what, uint32_t id) {
        nbx::set_last_error("boundary condition on %s %u does not exist", what, id);
        return nbx::NBX_ERR_ARG;
    };
    for (size_t c = 0; c < ns.size(); ++c) {
        const BcAtom& a = ns[c];
        if (a.id >= n_sgm) return bad("segment", a.id);
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
        if (a.id >= n_vtx) return bad("vertex", a.id);
        const uint32_t v = vtx[a.id];
        double val[2] = {a.val[0], a.val[1]};
        if (a.fval) {   // set_neumann evaluates at x = (0, 0) (:161-163)
            const double x0[2] = {0.0, 0.0};
            a.fval(x0, 0.0, val);
        }
        for (int j = 0; j < 2; ++j)
            if (a.mask[j]) pb.neumann(v * 2 + j, factor * val[j]);
    }
    auto dirichlet_node = [&](const BcAtom& a, uint32_t v) {
        double val[2] = {a.val[0], a.val[1]};
        if (a.fval) {   // set_dirichlet evaluates at the node (:224-229)
            const double x[2] = {nod[2 * v], nod[2 * v + 1]};
            a.fval(x, 0.0, val);
        }
        for (int j = 0; j < 2; ++j)
            if (a.mask[j]) pb.dirichlet(v * 2 + j, factor * val[j]);
    };
    for (const BcAtom& a : ds) {
        if (a.id >= n_sgm) return bad("segment", a.id);
        for (uint32_t i = sgm_ptr[a.id]; i < sgm_ptr[a.id + 1]; ++i) dirichlet_node(a, sgm_nodes[i]);
    }
    for (const BcAtom& a : dv) {
        if (a.id >= n_vtx) return bad("vertex", a.id);
        dirichlet_node(a, vtx[a.id]);
    }
    out = pb.finish();
    for (uint32_t d : out.d_dof)
        if (d >= 2 * nn) return bad("node", d / 2);
    for (uint32_t d : out.n_dof)
        if (d >= 2 * nn) return bad("node", d / 2);
    return nbx::NBX_OK;
}

std::vector<BcAtom> atoms(uint32_t n, const uint32_t* id, const uint8_t* mask, const double* val) {
    std::vector<BcAtom> v(n);
    for (uint32_t c = 0; c < n; ++c) {
        v[c].id = id[c];
        v[c].mask[0] = mask[2 * c] != 0;
        v[c].mask[1] = mask[2 * c + 1] != 0;
        v[c].val[0] = val[2 * c];
        v[c].val[1] = val[2 * c + 1];
        v[c].fval = nullptr;
    }
    return v;
}

int fem_solve(nbx::FemSys& S, int mode, double tol, uint32_t max_iter, const nbx::MixedOpts& mo,
              nbx_solve_info_t* info) {
    nbx::SolveStats st;
    int status = 0;
    memset(info, 0, sizeof *info);
    NBX_TRY(nbx::mat_prepare_sym(S.K, S.s));
    info->layout = S.K.has_sym ? 1 : 0;
    if (mode == NBX_SOLVE_REFERENCE) {
        uint32_t it = 0;
        double tr = 0.0;
        status = nbx::pcg_jacobi(S.K, S.F, S.u, max_iter, tol, &it, &tr, S.ws, S.s, &st);
        if (status < 0) return -status;
        info->iters = it;
        info->fp64_iters = it;
        info->tol_reached = tr;
    } else if (mode == NBX_SOLVE_FP64) {
        status = nbx::solve_fp64(S.K, S.F, S.u, tol, mo.true_residual, max_iter, S.ws, S.s, &st);
        if (status > 1) return status;
        info->iters = st.iters;
        info->fp64_iters = st.fp64_iters;
        info->rel_res = st.rel_res;
    } else if (mode == NBX_SOLVE_MIXED) {
        NBX_TRY(nbx::mat_prepare_mixed(S.K, S.s, mo));
        status = nbx::solve_mixed(S.K, S.F, S.u, tol, mo, max_iter,
                                  S.ws, S.s, &st);
        if (status > 1) return status;
        info->layout |= st.layout & 6;   // the preconditioner's fp16 copy
        info->iters = st.iters;
        info->outer = st.outer;
        info->fp64_iters = st.fp64_iters;
        info->rel_res = st.rel_res;
        info->fallback = st.fallback;
    } else {
        nbx::set_last_error("unknown solve mode %d", mode);
        return nbx::NBX_ERR_ARG;
    }
    info->ms = st.ms;
    info->ms_spmv = st.ms_spmv_avg;
    info->status = status;
    return status;
}

}  // namespace

extern "C" {

// ---- device-resident FEM system ---------------------------------------------
int nbx_fem_create(nbx_fem_t** out, uint32_t N_nod, const double* nod, uint32_t N_elem,
                   const uint32_t* conn) {
    *out = nullptr;
    if (N_nod == 0) {
        nbx::set_last_error("empty mesh");
        return nbx::NBX_ERR_ARG;
    }
    if ((uint64_t)N_nod * 2 >= 0xffffffffull || (uint64_t)N_elem * 3 >= 0xffffffffull) {
        nbx::set_last_error("mesh too large for 32-bit indices");
        return nbx::NBX_ERR_ARG;
    }
    nbx_fem_t* f = new nbx_fem_t;
    int rc = nbx::fem_create(f->sys, N_nod, nod, N_elem, conn);
    if (rc) {
        nbx::fem_free(f->sys);
        delete f;
        return rc;
    }
    *out = f;
    return nbx::NBX_OK;
}

void nbx_fem_destroy(nbx_fem_t* f) {
    if (!f) return;
    nbx::fem_free(f->sys);
    delete f;
}

int nbx_fem_sizes(const nbx_fem_t* f, uint32_t* N, uint64_t* nnz, uint64_t* blocks) {
    if (N) *N = 2 * f->sys.nn;
    if (nnz) *nnz = f->sys.nnz;
    if (blocks) *blocks = f->sys.K.s.n_blocks;
    return nbx::NBX_OK;
}

int nbx_fem_assemble(nbx_fem_t* f, double E, double nu, double thickness, const double* elem_E,
                     const uint8_t* enabled, int self_weight, double density,
                     const double* gravity) {
    NBX_TRY(nbx::fem_assemble(f->sys, E, nu, thickness, elem_E, enabled, self_weight, density,
                              gravity));
    const double g0[2] = {0.0, 0.0};
    return nbx::fem_body_force(f->sys, self_weight, density, gravity ? gravity : g0);
}

int nbx_fem_set_bconditions(nbx_fem_t* f, const uint32_t* vtx, const uint32_t* sgm_ptr,
                            const uint32_t* sgm_nodes, uint32_t n_dv, const uint32_t* dv_id,
                            const uint8_t* dv_mask, const double* dv_val, uint32_t n_nv,
                            const uint32_t* nv_id, const uint8_t* nv_mask, const double* nv_val,
                            uint32_t n_ds, const uint32_t* ds_id, const uint8_t* ds_mask,
                            const double* ds_val, uint32_t n_ns, const uint32_t* ns_id,
                            const uint8_t* ns_mask, const double* ns_val, const int32_t* ns_fptr,
                            const double* ns_fval, double factor) {
    nbx::FemSys& S = f->sys;
    std::vector<double> nod(2 * (size_t)S.nn);
    NBX_HIP(nbx::copy_on(nod.data(), S.nod, nod.size() * 8, hipMemcpyDeviceToHost, S.s));
    std::vector<const double*> samples;
    if (ns_fptr) {
        samples.resize(n_ns, nullptr);
        for (uint32_t c = 0; c < n_ns; ++c)
            if (ns_fptr[c] >= 0) samples[c] = ns_fval + 2 * (size_t)ns_fptr[c];
    }
    // model sizes: the largest ids the lists may use are checked against
    // the arrays the caller passes, which carry no length; trust ids below
    // the number of segments implied by sgm_ptr being monotone is not
    // checkable, so only node indices are validated.
    nbx::BcPlan plan;
    NBX_TRY(plan_from_lists(nod.data(), S.nn, vtx, 0xffffffffu, sgm_ptr, sgm_nodes, 0xffffffffu,
                            atoms(n_dv, dv_id, dv_mask, dv_val), atoms(n_nv, nv_id, nv_mask, nv_val),
                            atoms(n_ds, ds_id, ds_mask, ds_val), atoms(n_ns, ns_id, ns_mask, ns_val),
                            samples, factor, plan));
    return nbx::fem_apply_bc(S, plan);
}

static int fem_solve_entry(nbx_fem_t* f, int mode, double tol, uint32_t max_iter,
                           const nbx::MixedOpts& mo, const double* u_init, nbx_solve_info_t* info) {
    nbx::FemSys& S = f->sys;
    nbx_solve_info_t tmp;
    if (!info) info = &tmp;
    if (u_init)
        NBX_HIP(hipMemcpyAsync(S.u, u_init, 2 * (size_t)S.nn * 8, hipMemcpyHostToDevice, S.s));
    else
        NBX_HIP(hipMemsetAsync(S.u, 0, 2 * (size_t)S.nn * 8, S.s));
    return fem_solve(S, mode, tol, max_iter, mo, info);
}

static void opts_from(const nbx::MixedOpts& d, nbx_solve_opts_t* o) {
    o->mixed_variant = (int32_t)d.variant;
    o->cheb_degree = d.cheb_degree;
    o->cheb_ratio = d.cheb_ratio;
    o->inner_rtol = d.inner_rtol;
    o->reliable_delta = d.delta;
    o->precond_bits = d.precond_bits;
    o->true_residual = d.true_residual ? 1 : 0;
}

void nbx_solve_opts_default(nbx_solve_opts_t* o) { opts_from(nbx::MixedOpts(), o); }

void nbx_simp_opts_default(nbx_solve_opts_t* o) { opts_from(nbx::simp_mixed_defaults(), o); }

int nbx_fem_solve(nbx_fem_t* f, int mode, double tol, uint32_t max_iter, double inner_rtol,
                  const double* u_init, nbx_solve_info_t* info) {
    return fem_solve_entry(f, mode, tol, max_iter, nbx::mixed_opts(inner_rtol), u_init, info);
}

int nbx_fem_solve_ex(nbx_fem_t* f, int mode, double tol, uint32_t max_iter,
                     const nbx_solve_opts_t* opts, const double* u_init, nbx_solve_info_t* info) {
    nbx::MixedOpts mo;
    NBX_TRY(nbx::mixed_opts(opts, &mo));
    return fem_solve_entry(f, mode, tol, max_iter, mo, u_init, info);
}

int nbx_fem_get_displacement(const nbx_fem_t* f, double* u) {
    NBX_HIP(nbx::copy_on(u, f->sys.u, 2 * (size_t)f->sys.nn * 8, hipMemcpyDeviceToHost, f->sys.s));
    return nbx::NBX_OK;
}

int nbx_fem_true_residual(nbx_fem_t* f, int compensated, double* rel_res) {
    nbx::FemSys& S = f->sys;
    NBX_TRY(nbx::ws_alloc(S.ws, 2 * S.nn));
    double r = 0.0, fn = 0.0;
    if (compensated)
        NBX_TRY(nbx::fem_residual_dot2(S, &r));
    else
        NBX_TRY(nbx::fem_residual(S, S.ws.v64[4], &r));
    int rc = nbx::NBX_OK;
    fn = sqrt(nbx::device_norm2(2 * S.nn, S.F, S.ws, S.s, &rc));
    if (rc) return rc;
    *rel_res = fn > 0.0 ? r / fn : r;
    return nbx::NBX_OK;
}

int nbx_fem_get_system(const nbx_fem_t* f, uint32_t* rowptr, uint32_t* col, double* val, double* F) {
    if (rowptr && col && val) NBX_TRY(nbx::fem_download_csr(f->sys, rowptr, col, val));
    if (F) NBX_HIP(nbx::copy_on(F, f->sys.F, 2 * (size_t)f->sys.nn * 8, hipMemcpyDeviceToHost, f->sys.s));
    return nbx::NBX_OK;
}

int nbx_fem_strain_stress(const nbx_fem_t* f, double* strain, double* stress) {
    nbx::FemSys& S = const_cast<nbx::FemSys&>(f->sys);
    double *ds = nullptr, *dt = nullptr;
    const size_t bytes = 3 * (size_t)S.ne * 8 + 8;
    NBX_HIP(hipMalloc(&ds, bytes));
    NBX_HIP(hipMalloc(&dt, bytes));
    int rc = nbx::fem_strain_stress(S, ds, dt);
    if (!rc && strain && nbx::copy_on(strain, ds, bytes - 8, hipMemcpyDeviceToHost, S.s) != hipSuccess)
        rc = nbx::NBX_ERR_HIP;
    if (!rc && stress && nbx::copy_on(stress, dt, bytes - 8, hipMemcpyDeviceToHost, S.s) != hipSuccess)
        rc = nbx::NBX_ERR_HIP;
    (void)hipFree(ds);
    (void)hipFree(dt);
    return rc;
}

int nbx_fem_compliance(nbx_fem_t* f, const double* rho, double penal, double E0, double Emin,
                       double* compliance, double* dc, double* ce) {
    nbx::FemSys& S = f->sys;
    const size_t bytes = (size_t)S.ne * 8 + 8;
    if (rho) {
        if (!S.rho) NBX_HIP(hipMalloc(&S.rho, bytes));
        NBX_HIP(hipMemcpyAsync(S.rho, rho, (size_t)S.ne * 8, hipMemcpyHostToDevice, S.s));
    }
    if (!S.dc) NBX_HIP(hipMalloc(&S.dc, bytes));
    if (!S.ce) NBX_HIP(hipMalloc(&S.ce, bytes));
    double C = 0.0;
    NBX_TRY(nbx::fem_compliance(S, rho ? S.rho : nullptr, penal, E0, Emin, S.ce, S.dc, &C));
    if (compliance) *compliance = C;
    if (dc) NBX_HIP(nbx::copy_on(dc, S.dc, (size_t)S.ne * 8, hipMemcpyDeviceToHost, S.s));
    if (ce) NBX_HIP(nbx::copy_on(ce, S.ce, (size_t)S.ne * 8, hipMemcpyDeviceToHost, S.s));
    return nbx::NBX_OK;
}

int nbx_fem_adjoint(nbx_fem_t* f, const double* rhs, double* lambda, int mode, double rtol,
                    uint32_t max_iter, nbx_solve_info_t* info) {
    // K lambda = rhs with lambda = 0 on the Dirichlet dofs: swap the system's
    // F and u for the adjoint pair, solve, swap back.
    nbx::FemSys& S = f->sys;
    const size_t bytes = 2 * (size_t)S.nn * 8;
    if (!S.aux) NBX_HIP(hipMalloc(&S.aux, bytes + 8));
    if (!S.aux2) NBX_HIP(hipMalloc(&S.aux2, bytes + 8));
    NBX_HIP(hipMemcpyAsync(S.aux, rhs, bytes, hipMemcpyHostToDevice, S.s));
    NBX_HIP(hipMemsetAsync(S.aux2, 0, bytes, S.s));
    NBX_TRY(nbx::fem_zero_constrained(S, S.aux));
    std::swap(S.F, S.aux);
    std::swap(S.u, S.aux2);
    nbx_solve_info_t tmp;
    int rc = fem_solve(S, mode, rtol, max_iter, nbx::MixedOpts(), info ? info : &tmp);
    std::swap(S.F, S.aux);
    std::swap(S.u, S.aux2);
    if (rc > 1) return rc;
    NBX_HIP(nbx::copy_on(lambda, S.aux2, bytes, hipMemcpyDeviceToHost, S.s));
    return rc;
}

int nbx_fem_elem_cross(nbx_fem_t* f, const double* a, const double* b, double* out) {
    nbx::FemSys& S = f->sys;
    const size_t bytes = 2 * (size_t)S.nn * 8;
    if (!S.aux) NBX_HIP(hipMalloc(&S.aux, bytes + 8));
    if (!S.aux2) NBX_HIP(hipMalloc(&S.aux2, bytes + 8));
    if (!S.ce) NBX_HIP(hipMalloc(&S.ce, (size_t)S.ne * 8 + 8));
    double* da = S.u;
    double* db = S.u;
    if (a) {
        NBX_HIP(hipMemcpyAsync(S.aux, a, bytes, hipMemcpyHostToDevice, S.s));
        da = S.aux;
    }
    if (b) {
        NBX_HIP(hipMemcpyAsync(S.aux2, b, bytes, hipMemcpyHostToDevice, S.s));
        db = S.aux2;
    }
    NBX_TRY(nbx::fem_elem_cross(S, da, db, S.ce));
    NBX_HIP(hipMemcpyAsync(out, S.ce, (size_t)S.ne * 8, hipMemcpyDeviceToHost, S.s));
    NBX_HIP(hipStreamSynchronize(S.s));
    return nbx::NBX_OK;
}

int nbx_fem_spmv_bench(nbx_fem_t* f, int precision, uint32_t reps, double* ms_per_spmv) {
    nbx::FemSys& S = f->sys;
    if (precision == 32 || precision == 16) NBX_TRY(nbx::mat_make_scaled(S.K, S.s, precision));
    if (precision == 65) NBX_TRY(nbx::mat_prepare_sym(S.K, S.s));
    return nbx::spmv_bench(S.K, precision, reps, S.ws, S.s, ms_per_spmv);
}

// ---- SIMP loop ----------------------------------------------------------------
struct nbx_simp_s {
    nbx_fem_t* fem;
    nbx::Simp o;
};

int nbx_simp_create(nbx_simp_t** out, nbx_fem_t* f, double rmin, double volfrac, double penal,
                    double E0, double Emin, double move, const double* x0) {
    if (!out || !f) {
        nbx::set_last_error("nbx_simp_create: null argument");
        return nbx::NBX_ERR_ARG;
    }
    *out = nullptr;
    nbx_simp_t* h = new (std::nothrow) nbx_simp_s;
    if (!h) {
        nbx::set_last_error("out of host memory");
        return nbx::NBX_ERR_ARG;
    }
    h->fem = f;
    const int rc = nbx::simp_create(h->o, f->sys, rmin, volfrac, penal, E0, Emin, move, x0);
    if (rc) {
        nbx::simp_free(h->o);
        delete h;
        return rc;
    }
    *out = h;
    return nbx::NBX_OK;
}

void nbx_simp_destroy(nbx_simp_t* h) {
    if (!h) return;
    nbx::simp_free(h->o);
    delete h;
}

int nbx_simp_iterate(nbx_simp_t* h, int mode, double tol, uint32_t max_iter,
                     const nbx_solve_opts_t* opts, double* compliance, double* change,
                     nbx_solve_info_t* info) {
    nbx::FemSys& S = h->fem->sys;
    nbx_solve_info_t tmp;
    if (!info) info = &tmp;
    nbx::MixedOpts mo;
    NBX_TRY(nbx::mixed_opts(opts, &mo, nbx::simp_mixed_defaults()));
    // K(E(xPhys)) with the system's boundary conditions; the solve starts
    // from the previous displacement
    NBX_TRY(nbx::simp_moduli(h->o, S));
    NBX_TRY(nbx::fem_assemble_device(S));
    NBX_TRY(nbx::fem_reapply_bc(S));
    const int rc = fem_solve(S, mode, tol, max_iter ? max_iter : 20 * 2 * S.nn, mo, info);
    if (rc > 1) return rc;
    double C = 0.0;
    NBX_TRY(nbx::fem_compliance(S, h->o.xphys, h->o.penal, h->o.E0, h->o.Emin, S.ce, S.dc, &C));
    if (compliance) *compliance = C;
    NBX_TRY(nbx::simp_update(h->o, S, change));
    return rc;
}

int nbx_simp_get(const nbx_simp_t* h, double* x, double* xphys, double* dc, uint32_t* oc_steps) {
    const nbx::FemSys& S = h->fem->sys;
    const size_t bytes = (size_t)S.ne * 8;
    if (x) NBX_HIP(nbx::copy_on(x, h->o.x, bytes, hipMemcpyDeviceToHost, h->fem->sys.s));
    if (xphys) NBX_HIP(nbx::copy_on(xphys, h->o.xphys, bytes, hipMemcpyDeviceToHost, h->fem->sys.s));
    if (dc) NBX_HIP(nbx::copy_on(dc, h->o.dcf, bytes, hipMemcpyDeviceToHost, h->fem->sys.s));
    if (oc_steps) *oc_steps = h->o.last_oc_steps;
    return nbx::NBX_OK;
}

int nbx_simp_filter(const nbx_simp_t* h, uint64_t* n_entries, uint32_t* Hptr, uint32_t* Hidx,
                    double* Hval, double* Hs) {
    const nbx::Simp& o = h->o;
    if (n_entries) *n_entries = o.h_idx.size();
    if (Hptr) std::copy(o.h_ptr.begin(), o.h_ptr.end(), Hptr);
    if (Hidx) std::copy(o.h_idx.begin(), o.h_idx.end(), Hidx);
    if (Hval) std::copy(o.h_val.begin(), o.h_val.end(), Hval);
    if (Hs) std::copy(o.h_hs.begin(), o.h_hs.end(), Hs);
    return nbx::NBX_OK;
}

// ---- drop-in objects ----------------------------------------------------------
nb_mesh2D_t* nb_mesh2D_create(nb_mesh2D_type type) {
    if (type != NB_TRIAN) {
        nbx::set_last_error("only triangle meshes (NB_TRIAN) are supported");
        return nullptr;
    }
    return new nb_mesh2D_s;
}
void nb_mesh2D_destroy(nb_mesh2D_t* m) { delete m; }
uint32_t nb_mesh2D_get_N_invtx(const nb_mesh2D_t* m) { return (uint32_t)m->vtx.size(); }
uint32_t nb_mesh2D_get_N_insgm(const nb_mesh2D_t* m) {
    return m->sgm_ptr.empty() ? 0 : (uint32_t)m->sgm_ptr.size() - 1;
}
uint32_t nb_mesh2D_get_N_nodes(const nb_mesh2D_t* m) { return (uint32_t)(m->nod.size() / 2); }
uint32_t nb_mesh2D_get_N_elems(const nb_mesh2D_t* m) { return (uint32_t)(m->adj.size() / 3); }
double nb_mesh2D_node_get_x(const nb_mesh2D_t* m, uint32_t id) { return m->nod[2 * id]; }
double nb_mesh2D_node_get_y(const nb_mesh2D_t* m, uint32_t id) { return m->nod[2 * id + 1]; }
uint32_t nb_mesh2D_elem_get_adj(const nb_mesh2D_t* m, uint32_t e, uint8_t k) { return m->adj[3 * e + k]; }
uint32_t nb_mesh2D_get_invtx(const nb_mesh2D_t* m, uint32_t id) { return m->vtx[id]; }
uint32_t nb_mesh2D_insgm_get_N_nodes(const nb_mesh2D_t* m, uint32_t id) {
    return m->sgm_ptr[id + 1] - m->sgm_ptr[id];
}
uint32_t nb_mesh2D_insgm_get_node(const nb_mesh2D_t* m, uint32_t s, uint32_t i) {
    return m->sgm_nodes[m->sgm_ptr[s] + i];
}
int nbx_mesh2D_set_trg(nb_mesh2D_t* m, uint32_t N_nod, const double* nod, uint32_t N_elem,
                       const uint32_t* adj, uint32_t N_vtx, const uint32_t* vtx, uint32_t N_sgm,
                       const uint32_t* sgm_ptr, const uint32_t* nod_x_sgm) {
    m->nod.assign(nod, nod + 2 * (size_t)N_nod);
    m->adj.assign(adj, adj + 3 * (size_t)N_elem);
    m->vtx.assign(vtx, vtx + N_vtx);
    m->sgm_ptr.assign(sgm_ptr, sgm_ptr + N_sgm + 1);
    m->sgm_nodes.assign(nod_x_sgm, nod_x_sgm + (N_sgm ? sgm_ptr[N_sgm] : 0));
    for (uint32_t a : m->adj)
        if (a >= N_nod) {
            nbx::set_last_error("triangle references node %u >= %u", a, N_nod);
            return nbx::NBX_ERR_ARG;
        }
    return nbx::NBX_OK;
}

nb_fem_elem_t* nb_fem_elem_create(nb_elem_id type) {
    // element.c:16-33 (unknown types fall back to linear triangles)
    nb_fem_elem_t* e = new nb_fem_elem_s;
    e->type = (type == NB_QUAD_LINEAR) ? NB_QUAD_LINEAR : NB_TRG_LINEAR;
    return e;
}
void nb_fem_elem_destroy(nb_fem_elem_t* e) { delete e; }
uint8_t nb_fem_elem_get_N_gpoints(const nb_fem_elem_t* const e) { return e->type == NB_TRG_LINEAR ? 1 : 4; }
uint8_t nb_fem_elem_get_N_nodes(const nb_fem_elem_t* const e) { return e->type == NB_TRG_LINEAR ? 3 : 4; }

nb_material_t* nb_material_create(void) { return new nb_material_s; }
void nb_material_destroy(nb_material_t* m) { delete m; }
void nb_material_set_poisson_module(nb_material_t* m, double v) { m->nu = v; }
void nb_material_set_elasticity_module(nb_material_t* m, double v) { m->E = v; }
void nb_material_set_density(nb_material_t* m, double v) { m->density = v; }
double nb_material_get_poisson_module(const nb_material_t* const m) { return m->nu; }
double nb_material_get_elasticity_module(const nb_material_t* const m) { return m->E; }
double nb_material_get_density(const nb_material_t* const m) { return m->density; }
nb_material_t* nb_material_clone(nb_material_t* m) { return new nb_material_s(*m); }
void nb_material_set_fracture_energy(nb_material_t* m, double v) { m->fracture_energy = v; }
void nb_material_set_traction_limit_stress(nb_material_t* m, double v) { m->traction_limit = v; }
void nb_material_set_compression_limit_stress(nb_material_t* m, double v) {
    m->compression_limit = v;
}
void nb_material_set_energy_release_rate(nb_material_t* m, double v) { m->energy_release_rate = v; }
void nb_material_set_damage_length_scale(nb_material_t* m, double v) { m->damage_length_scale = v; }
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

// ---- implicit-integration parameters (static_damage2D.c:79-126) -------------
nb_fem_implicit_t* nb_fem_implicit_create(void) { return new nb_fem_implicit_s; }
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

uint16_t nb_bcond_get_memsize(uint8_t) { return (uint16_t)sizeof(nb_bcond_s); }
void nb_bcond_init(void* p, uint8_t N_dof) {
    nb_bcond_s* b = new (p) nb_bcond_s;
    b->N_dof = N_dof;
}
void nb_bcond_finish(void* p) { static_cast<nb_bcond_s*>(p)->~nb_bcond_s(); }
void* nb_bcond_create(uint8_t N_dof) {
    nb_bcond_s* b = new nb_bcond_s;
    b->N_dof = N_dof;
    return b;
}
void nb_bcond_destroy(void* p) { delete static_cast<nb_bcond_s*>(p); }
void nb_bcond_clear(void* p) {
    nb_bcond_s* b = static_cast<nb_bcond_s*>(p);
    b->dv.clear();
    b->nv.clear();
    b->ds.clear();
    b->ns.clear();
}
uint8_t nb_bcond_get_N_dof(const nb_bcond_t* const b) { return b->N_dof; }

static std::vector<BcAtom>* bc_list(nb_bcond_t* b, nb_bcond_id t, nb_bcond_where w) {
    if (t == NB_DIRICHLET) return w == NB_BC_ON_POINT ? &b->dv : &b->ds;
    return w == NB_BC_ON_POINT ? &b->nv : &b->ns;
}

void nb_bcond_push(nb_bcond_t* b, nb_bcond_id t, nb_bcond_where w, uint32_t id, const bool mask[],
                   const double value[]) {
    BcAtom a{id, {mask[0], b->N_dof > 1 ? mask[1] : false}, {0.0, 0.0}, nullptr};
    for (int j = 0; j < b->N_dof && j < 2; ++j) a.val[j] = value[j];
    bc_list(b, t, w)->push_back(a);
}

void nb_bcond_push_function(nb_bcond_t* b, nb_bcond_id t, nb_bcond_where w, uint32_t id,
                            const bool mask[], void (*fval)(const double*, double, double*)) {
    BcAtom a{id, {mask[0], b->N_dof > 1 ? mask[1] : false}, {0.0, 0.0}, fval};
    bc_list(b, t, w)->push_back(a);
}

// static_elasticity2D.c:31-97 on the GPU: pattern, assembly, BCs, Jacobi-PCG
// with x = 0, max_iter = N, tol = 1e-8 (absolute), strain.
}  // extern "C"

// The boundary conditions of a bcond object on a mesh, values scaled by
// factor (nb_fem_set_bconditions, set_bconditions.c:52-61); function-valued
// Neumann segments are sampled at their nodes at time 0.
static int plan_for(const nb_mesh2D_t* part, const nb_bcond_t* bcond, double factor,
                    nbx::BcPlan& plan) {
    std::vector<std::vector<double>> buf(bcond->ns.size());
    std::vector<const double*> samples(bcond->ns.size(), nullptr);
    for (size_t c = 0; c < bcond->ns.size(); ++c) {
        const BcAtom& a = bcond->ns[c];
        if (!a.fval || a.id >= nb_mesh2D_get_N_insgm(part)) continue;
        const uint32_t N = nb_mesh2D_insgm_get_N_nodes(part, a.id);
        buf[c].resize(2 * (size_t)N);
        for (uint32_t i = 0; i < N; ++i) {
            const uint32_t v = nb_mesh2D_insgm_get_node(part, a.id, i);
            const double x[2] = {part->nod[2 * v], part->nod[2 * v + 1]};
            a.fval(x, 0.0, &buf[c][2 * i]);
        }
        samples[c] = buf[c].data();
    }
    return plan_from_lists(part->nod.data(), nb_mesh2D_get_N_nodes(part), part->vtx.data(),
                           (uint32_t)part->vtx.size(), part->sgm_ptr.data(),
                           part->sgm_nodes.data(), nb_mesh2D_get_N_insgm(part), bcond->dv,
                           bcond->nv, bcond->ds, bcond->ns, samples, factor, plan);
}

extern "C" {

int nb_fem_compute_2D_Solid_Mechanics(const nb_mesh2D_t* const part,
                                      const nb_fem_elem_t* const elemtype,
                                      const nb_material_t* const material,
                                      const nb_bcond_t* const bcond, bool enable_self_weight,
                                      double gravity[2], nb_analysis2D_t,
                                      nb_analysis2D_params* params2D, const bool* elements_enabled,
                                      double* displacement, double* strain) {
    if (elemtype->type != NB_TRG_LINEAR) {
        nbx::set_last_error("only linear triangles are implemented on the GPU path");
        fprintf(stderr, "nb_fem_compute_2D_Solid_Mechanics: %s\n", nbx::last_error());
        return -1;
    }
    if (bcond->N_dof != 2) {
        nbx::set_last_error("solid mechanics needs 2 dofs per node");
        return -1;
    }
    const uint32_t nn = nb_mesh2D_get_N_nodes(part), ne = nb_mesh2D_get_N_elems(part);
    nbx_fem_t* f = nullptr;
    int rc = nbx_fem_create(&f, nn, part->nod.data(), ne, part->adj.data());
    if (rc) {
        fprintf(stderr, "nb_fem_compute_2D_Solid_Mechanics: %s\n", nbx::last_error());
        return -1;
    }
    std::vector<uint8_t> en;
    if (elements_enabled) {
        en.resize(ne);
        for (uint32_t e = 0; e < ne; ++e) en[e] = elements_enabled[e] ? 1 : 0;
    }
    int status = 0;
    rc = nbx_fem_assemble(f, material->E, material->nu, params2D->thickness, nullptr,
                          elements_enabled ? en.data() : nullptr, enable_self_weight ? 1 : 0,
                          material->density, gravity);
    if (rc == nbx::NBX_ERR_DISTORTED) {
        status = 1;   // static_elasticity2D.c:62-65
        goto DONE;
    }
    if (rc) goto FAIL;
    {
        nbx::BcPlan plan;
        rc = plan_for(part, bcond, 1.0, plan);
        if (rc) goto FAIL;
        rc = nbx::fem_apply_bc(f->sys, plan);
        if (rc) goto FAIL;
    }
    {
        nbx_solve_info_t info;
        rc = nbx_fem_solve(f, NBX_SOLVE_REFERENCE, 1e-8, 2 * nn, 0.0, nullptr, &info);
        if (rc > 1) goto FAIL;   // 0 or 1: both count as success (:91-96)
        rc = nbx_fem_get_displacement(f, displacement);
        if (rc) goto FAIL;
        rc = nbx_fem_strain_stress(f, strain, nullptr);
        if (rc) goto FAIL;
    }
DONE:
    nbx_fem_destroy(f);
    return status;
FAIL:
    fprintf(stderr, "nb_fem_compute_2D_Solid_Mechanics: GPU failure: %s\n", nbx::last_error());
    nbx_fem_destroy(f);
    return -1;
}

// Quasi-static damage (static_damage2D.c:239-472) on the device, with the
// conjugate-gradient path: for each load step n the boundary values are
// scaled by (n+1)/N_steps, and the iteration assembles K with D (1 - d),
// computes F - K u, stops on (||F - K u|| < tol after the first iteration)
// or N_max_iter or N_max_iter_without_enhance iterations without a new best
// residual, else solves K du = F - K u with Jacobi-PCG (absolute tol 1e-8,
// N iterations, from the previous du; du = 0 at the start of a step),
// u += du, and updates strain and damage.  The log has the reference's lines.
// The direct-solver option runs the same PCG (its factorisations are not on
// this path).
int nbx_fem_damage_solve(const nb_mesh2D_t* part, const nb_fem_elem_t* elemtype,
                         const nb_material_t* material, const nb_bcond_t* bcond,
                         int enable_self_weight, const double* gravity,
                         const nb_analysis2D_params* params2D, nb_fem_implicit_t* params,
                         const char* logfile, double* displacement, double* damage) {
    if (elemtype->type != NB_TRG_LINEAR || bcond->N_dof != 2) {
        nbx::set_last_error("the damage solver needs linear triangles with 2 dofs per node");
        return nbx::NBX_ERR_ARG;
    }
    const uint32_t nn = nb_mesh2D_get_N_nodes(part), ne = nb_mesh2D_get_N_elems(part);
    const uint32_t N = 2 * nn;
    FILE* log = logfile ? fopen(logfile, "a") : nullptr;
    auto logf = [&](const char* fmt, auto... a) {
        if (log) {
            fprintf(log, fmt, a...);
            fflush(log);
        }
    };
    logf("FEM: Damage Model\n");
    nbx_fem_t* f = nullptr;
    int rc = nbx_fem_create(&f, nn, part->nod.data(), ne, part->adj.data());
    double *dmg = nullptr, *r = nullptr, *r_prev = nullptr, *strain = nullptr, *res = nullptr,
           *du = nullptr;
    const double E = material->E, nu = material->nu;
    const double Gf = material->fracture_energy, ft = material->traction_limit;
    const double g0[2] = {0.0, 0.0};
    const double* grav = gravity ? gravity : g0;
    if (rc) goto END;
    {
        nbx::FemSys& S = f->sys;
        const size_t eb = (size_t)(ne ? ne : 1) * 8, vb = (size_t)N * 8 + 8;
        if (hipMalloc(&dmg, eb) != hipSuccess || hipMalloc(&r, eb) != hipSuccess ||
            hipMalloc(&r_prev, eb) != hipSuccess || hipMalloc(&strain, 3 * eb) != hipSuccess ||
            hipMalloc(&res, vb) != hipSuccess || hipMalloc(&du, vb) != hipSuccess) {
            nbx::set_last_error("out of device memory");
            rc = nbx::NBX_ERR_HIP;
            goto END;
        }
        {
            std::vector<double> r0(ne, ft / sqrt(E));   // :275-276
            if (nbx::copy_on(r, r0.data(), (size_t)ne * 8, hipMemcpyHostToDevice, S.s) != hipSuccess ||
                hipMemsetAsync(dmg, 0, eb, S.s) != hipSuccess ||
                hipMemsetAsync(S.u, 0, vb - 8, S.s) != hipSuccess) {
                rc = nbx::NBX_ERR_HIP;
                goto END;
            }
        }
        S.P = nbx::ElemParams();
        S.P.E = E;
        S.P.nu = nu;
        S.P.thickness = params2D->thickness;
        S.P.dmg = dmg;
        for (uint32_t n = 0; n < params->N_steps; ++n) {
            logf("  [ Load step %i]\n", n + 1);
            if (hipMemcpyAsync(r_prev, r, (size_t)ne * 8, hipMemcpyDeviceToDevice, S.s) !=
                    hipSuccess ||
                hipMemsetAsync(du, 0, vb - 8, S.s) != hipSuccess) {
                rc = nbx::NBX_ERR_HIP;
                goto END;
            }
            uint32_t it = 0, wo = 0;
            double best = 0.0;
            const double factor = (n + 1.0) / (double)params->N_steps;
            for (;;) {
                if ((rc = nbx::fem_assemble_device(S))) goto END;
                if ((rc = nbx::fem_body_force(S, enable_self_weight, material->density, grav)))
                    goto END;
                nbx::BcPlan plan;
                if ((rc = plan_for(part, bcond, factor, plan))) goto END;
                if ((rc = nbx::fem_apply_bc(S, plan))) goto END;
                double nrm = 0.0;
                if ((rc = nbx::fem_residual(S, res, &nrm))) goto END;
                if (it == 0) {
                    best = nrm;
                } else if (nrm < best) {
                    best = nrm;
                    wo = 0;
                } else {
                    wo += 1;
                }
                if ((nrm < params->residual_tolerance && it > 0) || it >= params->N_max_iter ||
                    wo >= params->N_max_iter_without_enhance) {
                    logf("    [%i: Residual: %e ]\n", it, nrm);
                    break;
                }
                uint32_t iters = 0;
                double err = 0.0;
                if ((rc = nbx::mat_prepare_sym(S.K, S.s))) goto END;
                const int st = nbx::pcg_jacobi(S.K, res, du, N, 1e-8, &iters, &err, S.ws, S.s,
                                               nullptr);
                if (st < 0) {
                    rc = -st;
                    goto END;
                }
                if (st != 0)   // the reference then tries Gauss-Seidel (:417-427)
                    logf("    [Conjugate Gradient (p. Jacobi) did not converge: %u iterations, "
                         "error %e ]\n", iters, err);
                logf("    [%i: Residual: %e]\n", it, nrm);
                if ((rc = nbx::fem_add(S, S.u, du))) goto END;
                if ((rc = nbx::fem_damage_update(S, E, nu, Gf, ft, r_prev, r, dmg, strain)))
                    goto END;
                ++it;
            }
        }
        if (displacement &&
            nbx::copy_on(displacement, S.u, (size_t)N * 8, hipMemcpyDeviceToHost, S.s) != hipSuccess)
            rc = nbx::NBX_ERR_HIP;
        if (!rc && damage && nbx::copy_on(damage, dmg, (size_t)ne * 8, hipMemcpyDeviceToHost, S.s) != hipSuccess)
            rc = nbx::NBX_ERR_HIP;
    }
END:
    if (log) fclose(log);
    void* ps[] = {dmg, r, r_prev, strain, res, du};
    for (void* q : ps)
        if (q) (void)hipFree(q);
    if (f) {
        f->sys.P.dmg = nullptr;
        nbx_fem_destroy(f);
    }
    return rc;
}

void nb_fem_compute_2D_Non_Linear_Solid_Mechanics(
    const nb_mesh2D_t* const part, const nb_fem_elem_t* const elemtype,
    const nb_material_t* const material, const nb_bcond_t* const bcond, bool enable_self_weight,
    double gravity[2], bool, nb_analysis2D_t, nb_analysis2D_params* params2D,
    nb_fem_implicit_t* params, const char* logfile) {
    const int rc = nbx_fem_damage_solve(part, elemtype, material, bcond, enable_self_weight ? 1 : 0,
                                        gravity, params2D, params, logfile, nullptr, nullptr);
    if (rc)
        fprintf(stderr, "nb_fem_compute_2D_Non_Linear_Solid_Mechanics: %s\n", nbx::last_error());
}

void nb_fem_compute_stress_from_strain(uint32_t N_elements, const nb_fem_elem_t* const,
                                       const nb_material_t* const material, nb_analysis2D_t,
                                       double* strain, const bool* elements_enabled,
                                       double* stress) {
    std::vector<uint8_t> en;
    if (elements_enabled) {
        en.resize(N_elements);
        for (uint32_t e = 0; e < N_elements; ++e) en[e] = elements_enabled[e] ? 1 : 0;
    }
    if (nbx::stress_from_strain(N_elements, material->E, material->nu,
                                elements_enabled ? en.data() : nullptr, strain, stress))
        fprintf(stderr, "nb_fem_compute_stress_from_strain: GPU failure: %s\n", nbx::last_error());
}

}  // extern "C"

// C-ABI of the device FEM system (nbx_fem_*) and the drop-in FEM drivers
// nb_fem_compute_2D_Solid_Mechanics and nb_fem_compute_2D_Non_Linear_Solid_Mechanics.
// The drivers read the caller's mesh, element type, material, boundary
// conditions and implicit-scheme parameters only through the reference's
// public accessors (bcplan.hpp): those objects belong to libnbots.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "../../include/nbots_amd.h"
#include "bcplan.hpp"
#include "fem.hpp"
#include "gauss_seidel.hpp"
#include "opts.hpp"
#include "simp.hpp"

struct nbx_fem_s {
    nbx::FemSys sys;
};

namespace {

using nbx::BcAtom;

std::vector<BcAtom> atoms(uint32_t n, const uint32_t* id, const uint8_t* mask, const double* val) {
    std::vector<BcAtom> v(n);
    for (uint32_t c = 0; c < n; ++c) {
        v[c].id = id[c];
        v[c].mask[0] = mask[2 * c] != 0;
        v[c].mask[1] = mask[2 * c + 1] != 0;
        v[c].val[0] = val[2 * c];
        v[c].val[1] = val[2 * c + 1];
    }
    return v;
}

int fem_solve(nbx::FemSys& S, int mode, double tol, uint32_t max_iter, const nbx::MixedOpts& mo,
              nbx_solve_info_t* info) {
    nbx::SolveStats st;
    int status = 0;
    memset(info, 0, sizeof *info);
    S.ws.replay = mo.replay;
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
        if (mo.variant == nbx::AMG_FCG)
            NBX_TRY(nbx::amg_prepare(S));
        else
            NBX_TRY(nbx::mat_prepare_mixed(S.K, S.s, mo));
        status = nbx::solve_mixed(S.K, S.F, S.u, tol, mo, max_iter,
                                  S.ws, S.s, &st);
        if (status > 1) return status;
        info->layout |= st.layout & 30;   // the preconditioner's fp16 copy, or AMG
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
    NBX_TRY(nbx::plan_from_lists(nod.data(), S.nn, vtx, 0xffffffffu, sgm_ptr, sgm_nodes, 0xffffffffu,
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
    o->replay = d.replay ? 1 : 0;
    o->reserved = 0;
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
    // max_iter 0: 20 N iterations, or 5,000 outer iterations of the
    // multigrid-preconditioned solve (which takes tens; a cap in N would let a
    // solve that has stopped progressing run for minutes)
    const bool amg = mode == NBX_SOLVE_MIXED && mo.variant == nbx::AMG_FCG;
    const uint32_t cap = max_iter ? max_iter : amg ? 5000u : 20 * 2 * S.nn;
    const int rc = fem_solve(S, mode, tol, cap, mo, info);
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

}  // extern "C"

// The element type the drivers take: linear triangles, one Gauss point
// (element.c:54-73), read through the element's accessors.
static int check_t3(const nb_fem_elem_t* elemtype) {
    if (nb_fem_elem_get_N_nodes(elemtype) != 3 || nb_fem_elem_get_N_gpoints(elemtype) != 1) {
        nbx::set_last_error("only linear triangles (NB_TRG_LINEAR) are implemented on the GPU "
                            "path");
        return nbx::NBX_ERR_ARG;
    }
    return nbx::NBX_OK;
}

extern "C" {

// static_elasticity2D.c:31-97 on the GPU: pattern, assembly, BCs, Jacobi-PCG
// with x = 0, max_iter = N, tol = 1e-8 (absolute), strain.  The mesh and the
// boundary conditions are read on the host first (import_mesh,
// import_bconditions); an invalid object gives -1 before any device work.
int nb_fem_compute_2D_Solid_Mechanics(const nb_mesh2D_t* const part,
                                      const nb_fem_elem_t* const elemtype,
                                      const nb_material_t* const material,
                                      const nb_bcond_t* const bcond, bool enable_self_weight,
                                      double gravity[2], nb_analysis2D_t,
                                      nb_analysis2D_params* params2D, const bool* elements_enabled,
                                      double* displacement, double* strain) {
    nbx::MeshArrays mesh;
    nbx::BcPlan plan;
    int rc = check_t3(elemtype);
    if (!rc) rc = nbx::import_mesh(part, mesh);
    if (!rc) rc = nbx::import_bconditions(part, bcond, 1.0, plan);
    if (rc) {
        fprintf(stderr, "nb_fem_compute_2D_Solid_Mechanics: %s\n", nbx::last_error());
        return -1;
    }
    const uint32_t nn = mesh.nn, ne = mesh.ne;
    nbx_fem_t* f = nullptr;
    rc = nbx_fem_create(&f, nn, mesh.nod.data(), ne, mesh.conn.data());
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
    rc = nbx_fem_assemble(f, nb_material_get_elasticity_module(material),
                          nb_material_get_poisson_module(material), params2D->thickness, nullptr,
                          elements_enabled ? en.data() : nullptr, enable_self_weight ? 1 : 0,
                          nb_material_get_density(material), gravity);
    if (rc == nbx::NBX_ERR_DISTORTED) {
        status = 1;   // static_elasticity2D.c:62-65
        goto DONE;
    }
    if (rc) goto FAIL;
    rc = nbx::fem_apply_bc(f->sys, plan);
    if (rc) goto FAIL;
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
// u += du, and updates strain and damage; when the PCG does not converge,
// Gauss-Seidel continues from its du (gauss_seidel.hip).  The log has the
// reference's lines.  The direct-solver option runs the same PCG (its
// factorisations are not on this path).
int nbx_fem_damage_solve(const nb_mesh2D_t* part, const nb_fem_elem_t* elemtype,
                         const nb_material_t* material, const nb_bcond_t* bcond,
                         int enable_self_weight, const double* gravity,
                         const nb_analysis2D_params* params2D, nb_fem_implicit_t* params,
                         const char* logfile, double* displacement, double* damage) {
    nbx::MeshArrays mesh;
    NBX_TRY(check_t3(elemtype));
    NBX_TRY(nbx::import_mesh(part, mesh));
    const uint32_t n_steps = nb_fem_implicit_get_N_steps(params);
    const uint32_t max_iter = nb_fem_implicit_get_N_max_iter(params);
    const uint32_t max_wo = nb_fem_implicit_get_N_max_iter_without_enhance(params);
    const double res_tol = nb_fem_implicit_get_residual_tolerance(params);
    // the conditions scaled by the load factor of each step (n + 1) / N_steps
    std::vector<nbx::BcPlan> plans(n_steps);
    for (uint32_t n = 0; n < n_steps; ++n)
        NBX_TRY(nbx::import_bconditions(part, bcond, (n + 1.0) / (double)n_steps, plans[n]));
    const uint32_t nn = mesh.nn, ne = mesh.ne;
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
    int rc = nbx_fem_create(&f, nn, mesh.nod.data(), ne, mesh.conn.data());
    double *dmg = nullptr, *r = nullptr, *r_prev = nullptr, *strain = nullptr, *res = nullptr,
           *du = nullptr;
    nbx::GsSystem gs;                           // the Gauss-Seidel fallback's copy of K
    std::vector<uint32_t> csr_rp, csr_col;
    std::vector<double> csr_val;
    const double E = nb_material_get_elasticity_module(material);
    const double nu = nb_material_get_poisson_module(material);
    const double Gf = nb_material_get_fracture_energy(material);
    const double ft = nb_material_get_traction_limit_stress(material);
    const double density = nb_material_get_density(material);
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
        for (uint32_t n = 0; n < n_steps; ++n) {
            logf("  [ Load step %i]\n", n + 1);
            if (hipMemcpyAsync(r_prev, r, (size_t)ne * 8, hipMemcpyDeviceToDevice, S.s) !=
                    hipSuccess ||
                hipMemsetAsync(du, 0, vb - 8, S.s) != hipSuccess) {
                rc = nbx::NBX_ERR_HIP;
                goto END;
            }
            uint32_t it = 0, wo = 0;
            double best = 0.0;
            for (;;) {
                if ((rc = nbx::fem_assemble_device(S))) goto END;
                if ((rc = nbx::fem_body_force(S, enable_self_weight, density, grav))) goto END;
                if ((rc = nbx::fem_apply_bc(S, plans[n]))) goto END;
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
                if ((nrm < res_tol && it > 0) || it >= max_iter || wo >= max_wo) {
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
                if (st != 0) {
                    // :414-427: Gauss-Seidel from the CG's du, 60 N sweeps,
                    // tolerance 1e-8; its failure is logged
                    if (csr_rp.empty()) {
                        csr_rp.resize(N + 1);
                        csr_col.resize(S.nnz);
                        csr_val.resize(S.nnz);
                    }
                    if ((rc = nbx::fem_download_csr(S, csr_rp.data(), csr_col.data(),
                                                    csr_val.data())))
                        goto END;
                    if ((rc = nbx::gs_prepare(gs, N, csr_rp.data(), csr_col.data(),
                                              csr_val.data(), S.s)))
                        goto END;
                    uint32_t gi = 0;
                    double ge = 0.0;
                    const int gst = nbx::gauss_seidel(gs, res, du, 60 * N, 1e-8, &gi, &ge, S.s);
                    if (gst > 1) {
                        rc = gst;
                        goto END;
                    }
                    if (gst == 1) {
                        logf("    [Conjugate Gradient (p. Jacobi) and Gauss-Seidel fails.]\n");
                        logf("      [Gauss-Seidel iterations: %u]\n", gi);
                        logf("      [Gauss-Seidel error: %e ]\n", ge);
                    }
                }
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
    nbx::gs_free(gs);
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
    if (nbx::stress_from_strain(N_elements, nb_material_get_elasticity_module(material),
                                nb_material_get_poisson_module(material),
                                elements_enabled ? en.data() : nullptr, strain, stress))
        fprintf(stderr, "nb_fem_compute_stress_from_strain: GPU failure: %s\n", nbx::last_error());
}

}  // extern "C"

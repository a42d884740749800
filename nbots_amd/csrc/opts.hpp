// Solver options of the C-ABI (nbx_solve_opts_t, inner_rtol) mapped to the
// solver's MixedOpts.
#pragma once

#include "../../include/nbots_amd.h"
#include "internal.hpp"

namespace nbx {

// inner_rtol > 0: FP32 CG sweeps to that tolerance in FP64 refinement;
// < 0: FP32 CG with FP64 reliable updates; 0: the default variant.
inline MixedOpts mixed_opts(double inner_rtol) {
    MixedOpts o;
    if (inner_rtol > 0.0) {
        o.variant = CG_RESTART;
        o.inner_rtol = inner_rtol;
    } else if (inner_rtol < 0.0) {
        o.variant = CG_RELIABLE;
    }
    return o;
}

// Fields at 0 keep the defaults in `base`.  NBX_ERR_ARG for an unknown
// variant or a precond_bits other than 0, 16, 32 or 64.
inline int mixed_opts(const nbx_solve_opts_t* p, MixedOpts* o, const MixedOpts& base = MixedOpts()) {
    *o = base;
    if (!p) return NBX_OK;
    if (p->mixed_variant < NBX_MIXED_CG_RESTART || p->mixed_variant > NBX_MIXED_AMG_FCG) {
        set_last_error("unknown mixed-precision variant %d", (int)p->mixed_variant);
        return NBX_ERR_ARG;
    }
    if (p->precond_bits != 0 && p->precond_bits != 16 && p->precond_bits != 32 &&
        p->precond_bits != 64) {
        set_last_error("precond_bits %d (0, 16, 32 or 64)", (int)p->precond_bits);
        return NBX_ERR_ARG;
    }
    o->variant = (MixedVariant)p->mixed_variant;
    if (p->cheb_degree) o->cheb_degree = p->cheb_degree;
    if (p->cheb_ratio > 1.0) o->cheb_ratio = p->cheb_ratio;
    if (p->inner_rtol > 0.0) o->inner_rtol = p->inner_rtol;
    if (p->reliable_delta > 0.0) o->delta = p->reliable_delta;
    if (p->precond_bits) o->precond_bits = p->precond_bits;
    o->true_residual = p->true_residual != 0;
    o->replay = p->replay != 0;
    return NBX_OK;
}

// Defaults of the SIMP loop's mixed solves: the multigrid-preconditioned
// flexible CG (1.05M-DOF MBB beam, 50 iterations: 45 ms per iteration
// against 231 with the Chebyshev preconditioner).  Its systems carry void
// elements (E = Emin = 1e-9 E0) that spread the spectrum, which a
// polynomial pays for in products and a multigrid hierarchy much less.
// The Chebyshev fields keep the plate defaults (degree 40 on [lmax/3e4,
// lmax], also the best point swept on the beam: 261 ms per iteration over
// 25; degree 32: 265, degree 48 on [lmax/1e5, lmax]: 263, degree 24: 271)
// for a caller that selects that variant.
inline MixedOpts simp_mixed_defaults() {
    MixedOpts o;
    o.variant = AMG_FCG;
    o.cheb_degree = 40;
    o.cheb_ratio = 3e4;
    return o;
}

}  // namespace nbx

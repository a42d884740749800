// Conjugate-gradient solvers for gfx950, device resident.
//
// * Jacobi-preconditioned CG with the reference's recurrences and stop test
//   (solver_bot/sparse/solvers/cg_precond_jacobi.c:13-90).  Three kernels per
//   iteration: SpMV + p.w, the g update + g.q and g.g, and the x and
//   direction update (x += alpha p reads the p that the direction update
//   reads anyway, so the vectors cross HBM 12 times per iteration, not 13).  Reducing kernels write one partial per workgroup; a one-
//   workgroup finalize kernel sums them in index order and computes alpha,
//   beta and the stop decision on the GPU (runs are reproducible, and no
//   device-scope fence or atomic sits in the loop); the host only polls a
//   done flag between batches of iterations.
// * Mixed precision: CG on the Jacobi-scaled FP32 copy D^-1/2 K D^-1/2 with
//   FP32 vectors and compensated FP32 dot products, and FP64 reliable updates
//   (Sleijpen & van der Vorst 1996): when the iterated residual has dropped
//   by delta, the FP32 correction is folded into the FP64 solution u, the
//   true residual b - K u is recomputed with the FP64 matrix, and the FP32
//   recurrence continues from it without restarting the Krylov directions.
//   A restarted variant (inner tolerance per sweep) is kept for comparison.
//   CG on the symmetrically Jacobi-scaled system is Jacobi-PCG in other
//   variables.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <type_traits>
#include <vector>

#include "rows.hpp"

namespace nbx {

// ---------------------------------------------------------------------------
// FP64 Jacobi PCG, reference semantics
// ---------------------------------------------------------------------------
// init: g = A x - b; q = g / Aii; p = -q; gg = g.g; rz = g.q
//   (cg_precond_jacobi.c:33-43)
template <class Op>
__global__ __launch_bounds__(BLOCK) void k_pcg_init(Op op, const typename Op::V* __restrict__ b,
                                                    const typename Op::V* __restrict__ x,
                                                    typename Op::V* __restrict__ g,
                                                    typename Op::V* __restrict__ p,
                                                    const typename Op::V* __restrict__ diag,
                                                    PcgState* st, double* partials, uint32_t n) {
    using V = typename Op::V;
    double acc[2] = {0.0, 0.0};
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) {
        const V gv = subu(op.row((uint32_t)u, x), b[u]);
        const V q = divu(gv, diag[u]);
        g[u] = gv;
        p[u] = negu(q);
        acc[0] += dotu(gv, gv);
        acc[1] += dotu(gv, q);
    }
    __shared__ double red[2][BLOCK / WAVE];
    block_sum<2>(acc, red);
    put_partials<2>(acc, partials);
}

__global__ __launch_bounds__(FIN) void k_fin_init(PcgState* st, const double* partials, unsigned np) {
    double tot[2];
    fin_sum<2>(partials, np, tot);
    if (threadIdx.x == 0) {
        const double gg = tot[0];
        st->rz = tot[1];
        st->gg_cur = gg;
        st->gg_prev = gg;
        st->gg0 = gg;
        st->k = 0;
        if (!(gg > st->tol2 && 0u < st->max_iter)) {
            st->done = 1;
            st->tol_reached = sqrt(gg);
            st->status = gg > st->tol2 ? 1u : 0u;
        }
    }
}

// w = A p; pw = p.w; alpha = rz / pw   (cg_precond_jacobi.c:50-59)
template <class Op>
__global__ __launch_bounds__(BLOCK) void k_pcg_A(Op op, const typename Op::V* __restrict__ p,
                                                 typename Op::V* __restrict__ w, PcgState* st,
                                                 double* partials, uint32_t n) {
    if (st->done) return;
    double acc[1] = {0.0};
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) {
        const typename Op::V wv = op.row((uint32_t)u, p);
        w[u] = wv;
        acc[0] += dotu(p[u], wv);
    }
    __shared__ double red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    put_partials<1>(acc, partials);
}

// The finalize kernels read the state they need together with the partials
// and test `done` afterwards: one memory round trip instead of three in a
// row (state, partials, state).  The partials buffer holds 4 x MAX_GRID
// values, so the reads stay in it whatever the state says; a finished
// solve ignores the sums.
__global__ __launch_bounds__(FIN) void k_fin_A(PcgState* st, const double* partials, unsigned np) {
    const uint32_t done = st->done, j = st->k;   // this is iteration j
    const double rz = st->rz;
    double tot[1];
    fin_sum<1>(partials, np, tot);
    if (done) return;
    if (threadIdx.x == 0) {
        st->pw = tot[0];
        st->alpha = rz / tot[0];
        st->alpha_hist[j & 3u] = st->alpha;
    }
}

__global__ __launch_bounds__(FIN) void k_fin_B(PcgState* st, const double* partials, unsigned np) {
    const uint32_t done = st->done, k = st->k + 1, max_iter = st->max_iter;
    const double rz = st->rz, gg_prev = st->gg_cur, tol2 = st->tol2;
    double tot[2];
    fin_sum<2>(partials, np, tot);
    if (done) return;
    if (threadIdx.x == 0) {
        st->beta = tot[0] / rz;
        st->rz = tot[0];
        st->gg_prev = gg_prev;
        st->gg_cur = tot[1];
        st->k = k;
        if (!(gg_prev > tol2 && k < max_iter)) {
            st->done = 1;
            st->tol_reached = sqrt(gg_prev);
            st->status = gg_prev > tol2 ? 1u : 0u;
        }
    }
}

// g += alpha w; q = g / Aii; rz' = g.q; gg' = g.g, then the lagged stop test
// (cg_precond_jacobi.c:63-70, :75); x is left to the direction kernel.
template <class V, bool NT>
__global__ __launch_bounds__(BLOCK) void k_pcg_B(V* __restrict__ g, const V* __restrict__ w,
                                                  const V* __restrict__ diag, PcgState* st,
                                                  double* partials, uint32_t n) {
    if (st->done) return;
    const double alpha = st->alpha;
    double acc[2] = {0.0, 0.0};
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) {
        const V gv = axpyu(ldv<NT>(&g[u]), alpha, ldv<NT>(&w[u]));
        stv<NT>(&g[u], gv);
        const V q = divu(gv, ldv<NT>(&diag[u]));
        acc[0] += dotu(gv, q);
        acc[1] += dotu(gv, gv);
    }
    __shared__ double red[2][BLOCK / WAVE];
    block_sum<2>(acc, red);
    put_partials<2>(acc, partials);
}

// The four direction buffers of the FP64 PCG: p_j lives in p[j % 4].
template <class V>
struct PBufs {
    V* p[4];
};

// p_{j+1} = -q + beta p_j  (cg_precond_jacobi.c:72-74) into the next of four
// direction buffers, and x += alpha p (:62) for four iterations at a time:
// iteration j reads p_j from buffer j % 4 and writes p_{j+1} into buffer
// (j + 1) % 4, which still holds p_{j-3}.  Iterations j % 4 = 0, 1, 2 leave
// x alone; iteration j % 4 = 3 (FLUSH) applies
// x = (((x + a_{j-3} p_{j-3}) + a_{j-2} p_{j-2}) + a_{j-1} p_{j-1}) + a_j p_j,
// the same roundings in the same order as one update per iteration, so x is
// bit-identical, and x is read and written every fourth iteration only.
// Once the stop test has fired this kernel no longer runs, and k_pcg_tail
// applies the updates still pending.
template <class V, bool NT, bool FLUSH>
__global__ __launch_bounds__(BLOCK) void k_pcg_C(V* __restrict__ x, PBufs<V> P, uint32_t j,
                                                  const V* __restrict__ g,
                                                  const V* __restrict__ diag, PcgState* st,
                                                  uint32_t n) {
    if (st->done) return;
    const double alpha = st->alpha, beta = st->beta;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if constexpr (FLUSH) {
        a0 = st->alpha_hist[0];
        a1 = st->alpha_hist[1];
        a2 = st->alpha_hist[2];
    }
    const V* __restrict__ p = P.p[j & 3u];
    V* __restrict__ p_next = P.p[(j + 1) & 3u];
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) {
        const V pu = ldv<NT>(&p[u]);
        if constexpr (FLUSH) {
            // j % 4 = 3: p_{j-3} is in p_next (read before p_{j+1} replaces
            // it), p_{j-2} in P.p[1], p_{j-1} in P.p[2]
            V xv = ldv<NT>(&x[u]);
            xv = axpyu(xv, a0, ldv<NT>(&p_next[u]));
            xv = axpyu(xv, a1, ldv<NT>(&P.p[1][u]));
            xv = axpyu(xv, a2, ldv<NT>(&P.p[2][u]));
            stv<NT>(&x[u], axpyu(xv, alpha, pu));
        }
        const V q = divu(ldv<NT>(&g[u]), ldv<NT>(&diag[u]));
        p_next[u] = pupd(q, beta, pu);   // gathered by the next SpMV: cached store
    }
}

// The x updates still pending when the loop has stopped (cg_precond_jacobi.c:
// 62): iteration J = k - 1 did not reach k_pcg_C, and the updates of the
// iterations 4 floor(J / 4) .. J are pending, applied in ascending order.
// Nothing if no iteration ran.
template <class V>
__global__ __launch_bounds__(BLOCK) void k_pcg_tail(V* __restrict__ x, PBufs<V> P,
                                                    const PcgState* st, uint32_t n) {
    const uint32_t k = st->k;
    if (k == 0) return;
    const uint32_t last = (k - 1) & 3u;   // buffers 0 .. last hold p of the pending ones
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) {
        V xv = x[u];
        for (uint32_t i = 0; i <= last; ++i) xv = axpyu(xv, st->alpha_hist[i], P.p[i][u]);
        x[u] = xv;
    }
}

// ---------------------------------------------------------------------------
// mixed precision
// ---------------------------------------------------------------------------
// True residual in FP64 and its scaled FP32 image:
//   r = b - K u; r32 = fl32(dis * r); rr64 = r.r; rr32 = r32.r32
// first: p32 = r32, y32 = 0 (a new CG); otherwise beta = rr32 / rz so that
// the direction update continues the old Krylov sequence.
__global__ __launch_bounds__(BLOCK) void k_mix_residual(Sell2 K, const double2* __restrict__ b,
                                                        const double2* __restrict__ u,
                                                        const double2* __restrict__ dis,
                                                        float2* __restrict__ r32,
                                                        float2* __restrict__ p32,
                                                        float2* __restrict__ y32, PcgState* st,
                                                        double* partials, uint32_t n, int first) {
    if (st->done) return;
    double acc[2] = {0.0, 0.0};
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const double2 ku = row64(K, (uint32_t)i, u);
        const double2 bi = b[i];
        const double2 r = make_double2(bi.x - ku.x, bi.y - ku.y);
        const double2 d = dis[i];
        const float2 rs = make_float2((float)(d.x * r.x), (float)(d.y * r.y));
        r32[i] = rs;
        if (first) {
            p32[i] = rs;
            y32[i] = make_float2(0.f, 0.f);
        }
        acc[0] += r.x * r.x + r.y * r.y;
        acc[1] += (double)rs.x * rs.x + (double)rs.y * rs.y;
    }
    __shared__ double red[2][BLOCK / WAVE];
    block_sum<2>(acc, red);
    put_partials<2>(acc, partials);
}

__global__ __launch_bounds__(FIN) void k_fin_residual(PcgState* st, const double* partials,
                                                      unsigned np, int first) {
    if (st->done) return;
    double tot[2];
    fin_sum<2>(partials, np, tot);
    if (threadIdx.x == 0) {
        const double rr64 = tot[0], rr32 = tot[1];
        st->tol_reached = sqrt(rr64);
        st->gg_cur = rr64;
        if (!(rr64 > st->tol2)) {
            st->done = 1;
            st->status = 0;
        } else if (st->k >= st->max_iter || !isfinite(rr64)) {
            st->done = 1;
            st->status = 1;
        }
        if (!first) st->beta = rr32 / st->rz;
        st->rz = rr32;
        st->maxrr = rr32;
        // the iterated residual at which the true one is predicted to meet
        // the tolerance (ratio of the two norms taken from this update)
        st->conv_rr = rr64 > 0.0 ? 0.5 * st->tol2 * rr32 / rr64 : 0.0;
        st->hold = 0;
        st->updates += 1;
    }
}

// w = Kt p; alpha = rz / (p.w)
__global__ __launch_bounds__(BLOCK) void k_mix_A(Sell2 K, const float2* __restrict__ p,
                                                 float2* __restrict__ w, PcgState* st,
                                                 float* partials, uint32_t n) {
    if (st->done | st->hold) return;
    f32c acc[1] = {{0.f, 0.f}};
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const float2 wv = row32(K, (uint32_t)i, p);
        w[i] = wv;
        const float2 pv = p[i];
        acc[0] = cadd(acc[0], pv.x * wv.x);
        acc[0] = cadd(acc[0], pv.y * wv.y);
    }
    __shared__ f32c red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    put_partials<1>(acc, partials);
}

__global__ __launch_bounds__(FIN) void k_fin_mixA(PcgState* st, const float* partials, unsigned np) {
    if (st->done | st->hold) return;
    f32c tot[1];
    fin_sum<1>(partials, np, tot);
    if (threadIdx.x == 0) {
        const double pw = (double)tot[0].s + (double)tot[0].c;
        st->pw = pw;
        st->alpha = st->rz / pw;
        if (!(pw > 0.0) || !isfinite(pw)) {
            st->breakdown = 1;
            st->hold = 1;
        }
    }
}

// y += alpha p; r -= alpha w; rz' = r.r; then continue (beta) or pause for
// a reliable update / the end of a restarted sweep.
__global__ __launch_bounds__(BLOCK) void k_mix_B(float2* __restrict__ y, float2* __restrict__ r,
                                                 const float2* __restrict__ p,
                                                 const float2* __restrict__ w, PcgState* st,
                                                 float* partials, uint32_t n) {
    if (st->done | st->hold) return;
    const float alpha = (float)st->alpha;
    f32c acc[1] = {{0.f, 0.f}};
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const float2 pv = p[i], wv = w[i];
        float2 yv = y[i], rv = r[i];
        yv.x = yv.x + alpha * pv.x;
        yv.y = yv.y + alpha * pv.y;
        rv.x = rv.x - alpha * wv.x;
        rv.y = rv.y - alpha * wv.y;
        y[i] = yv;
        r[i] = rv;
        acc[0] = cadd(acc[0], rv.x * rv.x);
        acc[0] = cadd(acc[0], rv.y * rv.y);
    }
    __shared__ f32c red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    put_partials<1>(acc, partials);
}

__global__ __launch_bounds__(FIN) void k_fin_mixB(PcgState* st, const float* partials, unsigned np,
                                                  double delta2, double inner_stop) {
    if (st->done | st->hold) return;
    f32c tot[1];
    fin_sum<1>(partials, np, tot);
    if (threadIdx.x == 0) {
        const double rz = (double)tot[0].s + (double)tot[0].c;
        st->k += 1;
        const bool upd = inner_stop > 0.0 ? !(rz > inner_stop * st->maxrr)   // restarted sweeps
                                          : (rz < delta2 * st->maxrr || rz < st->conv_rr);
        if (upd || st->k >= st->max_iter || !isfinite(rz)) {
            st->hold = 1;
            if (!isfinite(rz)) st->breakdown = 1;
        } else {
            st->beta = rz / st->rz;
            st->rz = rz;
            if (inner_stop <= 0.0 && rz > st->maxrr) st->maxrr = rz;
        }
    }
}

// p = r + beta p
__global__ __launch_bounds__(BLOCK) void k_mix_C(float2* __restrict__ p, const float2* __restrict__ r,
                                                 PcgState* st, uint32_t n) {
    if (st->done | st->hold) return;
    const float beta = (float)st->beta;
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const float2 rv = r[i], pv = p[i];
        p[i] = make_float2(rv.x + beta * pv.x, rv.y + beta * pv.y);
    }
}

// u += dis * y (FP64); y = 0
__global__ __launch_bounds__(BLOCK) void k_mix_update(double2* __restrict__ u,
                                                      const double2* __restrict__ dis,
                                                      float2* __restrict__ y, PcgState* st,
                                                      uint32_t n) {
    if (st->done) return;
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const double2 d = dis[i];
        const float2 yv = y[i];
        double2 uv = u[i];
        uv.x = uv.x + d.x * (double)yv.x;
        uv.y = uv.y + d.y * (double)yv.y;
        u[i] = uv;
        y[i] = make_float2(0.f, 0.f);
    }
}

// ---------------------------------------------------------------------------
// mixed precision: FP64 flexible CG, FP32 Chebyshev-Jacobi preconditioner
// ---------------------------------------------------------------------------
// z = D^-1/2 c(Kt) D^-1/2 r, where c(Kt) s is m steps of the Chebyshev
// iteration for Kt y = s from y = 0 on [lmin, lmax] (Saad, Iterative Methods
// for Sparse Linear Systems, 2nd ed., Alg. 12.1), Kt = D^-1/2 K D^-1/2 in
// FP32.  The outer iteration is CG with the Polak-Ribiere beta (Notay 2000,
// flexible CG), all in FP64 with the FP64 matrix, so FP32 rounding only
// weakens the preconditioner and never the attainable residual.

// first: r = b - K x; otherwise x += alpha p, r -= alpha w.  Then the
// preconditioner's start, P = float2 (scaled, FP32): s = fl32(dis r),
// res = s, d = z = s / theta; P = double2 (unscaled, FP64): res = r,
// d = z = D^-1 r / theta.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_fcg_update(Sell2 K, const double2* __restrict__ b,
                                                      double2* __restrict__ x,
                                                      double2* __restrict__ r,
                                                      const double2* __restrict__ p,
                                                      const double2* __restrict__ w,
                                                      const double2* __restrict__ dis,
                                                      P* __restrict__ res,
                                                      P* __restrict__ d,
                                                      P* __restrict__ z, PcgState* st,
                                                      double* partials, uint32_t n, int first,
                                                      double inv_theta) {
    if (st->done) return;
    const double alpha = st->alpha;
    double acc[1] = {0.0};
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        double2 rv;
        if (first) {
            const double2 ku = row64(K, (uint32_t)i, x);
            const double2 bi = b[i];
            rv = make_double2(bi.x - ku.x, bi.y - ku.y);
        } else {
            x[i] = axpyu(x[i], alpha, p[i]);
            const double2 ro = r[i], wv = w[i];
            rv = make_double2(ro.x - alpha * wv.x, ro.y - alpha * wv.y);
        }
        r[i] = rv;
        const double2 dd = dis[i];
        if constexpr (std::is_same<P, float2>::value) {
            const float2 sv = make_float2((float)(dd.x * rv.x), (float)(dd.y * rv.y));
            const float it = (float)inv_theta;
            res[i] = sv;
            const float2 dv = make_float2(sv.x * it, sv.y * it);
            d[i] = dv;
            z[i] = dv;
        } else {
            res[i] = rv;
            const double2 dv = make_double2(dd.x * dd.x * rv.x * inv_theta,
                                            dd.y * dd.y * rv.y * inv_theta);
            d[i] = dv;
            z[i] = dv;
        }
        acc[0] += rv.x * rv.x + rv.y * rv.y;
    }
    __shared__ double red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    put_partials<1>(acc, partials);
}

__global__ __launch_bounds__(FIN) void k_fin_fcg_update(PcgState* st, const double* partials,
                                                        unsigned np, int first) {
    const uint32_t done = st->done, k = first ? st->k : st->k + 1, max_iter = st->max_iter;
    const double tol2 = st->tol2;
    double tot[1];
    fin_sum<1>(partials, np, tot);
    if (done) return;
    if (threadIdx.x == 0) {
        const double gg = tot[0];
        st->k = k;
        if (first) st->gg0 = gg;
        st->gg_cur = gg;
        st->tol_reached = sqrt(gg);
        if (!(gg > tol2)) {
            st->done = 1;
            st->status = 0;
        } else if (k >= max_iter || !isfinite(gg)) {
            st->done = 1;
            st->status = 1;
        }
    }
}

// one Chebyshev step: res -= Kt d; d' = c1 d + c2 res; z += d'
template <int BITS>
__global__ __launch_bounds__(BLOCK) void k_cheb_step(Sell2 K, const float2* __restrict__ din,
                                                     float2* __restrict__ dout,
                                                     float2* __restrict__ res,
                                                     float2* __restrict__ z, PcgState* st,
                                                     uint32_t n, float c1, float c2) {
    if (st->done) return;
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const float2 t = row_scaled<BITS>(K, (uint32_t)i, din);
        float2 rs = res[i];
        rs.x = rs.x - t.x;
        rs.y = rs.y - t.y;
        const float2 di = din[i];
        const float2 dn = make_float2(c1 * di.x + c2 * rs.x, c1 * di.y + c2 * rs.y);
        float2 zv = z[i];
        zv.x = zv.x + dn.x;
        zv.y = zv.y + dn.y;
        res[i] = rs;
        dout[i] = dn;
        z[i] = zv;
    }
}

// The step on the fp16 copy, with the row sum and the recurrence in FP64
// (vectors stored in FP32).  The FP32 sum loses most of the small result of
// a cancelling stiffness row; with it the outer iteration count jumps up
// and down with the degree (4.2M DOF, degree 28-40: 1006, 940, 1118, 996,
// 711), with FP64 sums it falls steadily (1006, 943, 882, 787, 709) at the
// same cost per step.  NT: non-temporal matrix loads.
// ZM says how z = d_0 + d_1 + ... takes this step's direction:
//   Z_SKIP  not yet (the next step adds it with its own),
//   Z_PAIR  z += d_in + d_out (the previous step's and this one's),
//   Z_ONE   z += d_out.
// Adding the directions in pairs reads and writes z on half the steps
// (16 B per node less on every other step) and rounds z once per pair.
// M: the fp16 matrix, Sell2 (full copy) or Off16 (without diagonal blocks).
enum { Z_SKIP = 0, Z_PAIR = 1, Z_ONE = 2 };
template <bool NT, int ZM, class M>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_cheb_step64(M K, const float2* __restrict__ din,
                                                       float2* __restrict__ dout,
                                                       float2* __restrict__ res,
                                                       float2* __restrict__ z, PcgState* st,
                                                       uint32_t n, double c1, double c2) {
    if (st->done) return;
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const float2 di = din[i];
        const double2 t = cheb_row<NT>(K, (uint32_t)i, din, di);
        const float2 r0 = res[i];
        const double rx = (double)r0.x - t.x, ry = (double)r0.y - t.y;
        const double dx = c1 * (double)di.x + c2 * rx, dy = c1 * (double)di.y + c2 * ry;
        res[i] = make_float2((float)rx, (float)ry);
        dout[i] = make_float2((float)dx, (float)dy);
        if constexpr (ZM != Z_SKIP) {
            const float2 zv = z[i];
            const double ax = ZM == Z_PAIR ? (double)di.x + dx : dx;
            const double ay = ZM == Z_PAIR ? (double)di.y + dy : dy;
            z[i] = make_float2((float)((double)zv.x + ax), (float)((double)zv.y + ay));
        }
    }
}

// Step k of 1..last (z = d_0 on entry): even steps add d_{k-1} + d_k, odd
// steps wait for the next one, except the last, which adds d_k alone.
template <bool NT, class M>
static void launch_cheb_step64(uint32_t k, uint32_t last, unsigned G, hipStream_t s, const M& K,
                               const float2* din, float2* dout, float2* res, float2* z,
                               PcgState* st, uint32_t n, double c1, double c2) {
    if (k % 2 == 0)
        hipLaunchKernelGGL((k_cheb_step64<NT, Z_PAIR, M>), dim3(G), dim3(BLOCK), 0, s, K, din, dout,
                           res, z, st, n, c1, c2);
    else if (k == last)
        hipLaunchKernelGGL((k_cheb_step64<NT, Z_ONE, M>), dim3(G), dim3(BLOCK), 0, s, K, din, dout,
                           res, z, st, n, c1, c2);
    else
        hipLaunchKernelGGL((k_cheb_step64<NT, Z_SKIP, M>), dim3(G), dim3(BLOCK), 0, s, K, din, dout,
                           res, z, st, n, c1, c2);
}

// One V-cycle of the multigrid preconditioner whose finest level is
// smoothed on the fp16 scaled copy K~ = D^-1/2 K D^-1/2, in scaled variables
// (z~ = D^1/2 z, res~ = D^-1/2 res): Chebyshev smoothing in K~ is Chebyshev
// smoothing in D^-1 K.  On entry the flexible-CG update has written
// res = r~, din = r~ / theta, z = din (the first smoothing step from 0).
//   pre-smoothing: steps 1 .. degree - 1;
//   res = r~ - K~ z (a step with c1 = c2 = 0 completes the residual);
//   c = P~ M_c^-1 P~^T res, the coarse levels' V-cycle (P~ = D^1/2 P);
//   post-smoothing from z + c: its first step takes res -= K~ c,
//   d = res / theta, z += c + d; then steps 1 .. degree - 1.
template <bool NT, class M>
static void amg16_cycle(const M& K, const AmgFine& f, AmgDev* amg, float2* res, float2* da,
                        float2* db, float2* z, PcgState* st, uint32_t nu, unsigned G,
                        hipStream_t s) {
    const uint32_t last = f.degree - 1;
    auto smooth = [&]() {
        float2 *din = da, *dout = db;
        for (uint32_t k = 1; k <= last; ++k) {
            launch_cheb_step64<NT>(k, last, G, s, K, din, dout, res, z, st, nu, f.c1[k], f.c2[k]);
            std::swap(din, dout);
        }
        return din;
    };
    const float2* dl = smooth();
    hipLaunchKernelGGL((k_cheb_step64<NT, Z_SKIP, M>), dim3(G), dim3(BLOCK), 0, s, K, dl,
                       dl == da ? db : da, res, z, st, nu, 0.0, 0.0);
    (void)amg_coarse_correction(amg, (const float*)res, f.c, st, s);
    hipLaunchKernelGGL((k_cheb_step64<NT, Z_PAIR, M>), dim3(G), dim3(BLOCK), 0, s, K,
                       (const float2*)f.c, da, res, z, st, nu, 0.0, f.inv_theta);
    (void)smooth();
}

// One step of the all-FP64 preconditioner (precond_bits = 64): the same
// Chebyshev recurrence on the FP64 matrix itself (the symmetric copy the
// FP64 products read, or the full one), vectors and arithmetic in FP64, in
// unscaled variables: res -= K d; d' = c1 d + c2 D^-1 res; z += d'.  With
// d = D^-1/2 d~ and res = D^1/2 res~ this is the scaled step on
// D^-1/2 K D^-1/2 exactly; z then holds D^-1/2 z~, the preconditioned
// residual.  ZM as in k_cheb_step64.
template <class Op, int ZM>
__global__ __launch_bounds__(BLOCK) void k_cheb_step_f64(Op op, const double2* __restrict__ din,
                                                         double2* __restrict__ dout,
                                                         double2* __restrict__ res,
                                                         double2* __restrict__ z,
                                                         const double2* __restrict__ dis,
                                                         PcgState* st, uint32_t n, double c1,
                                                         double c2) {
    if (st->done) return;
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const double2 t = op.row((uint32_t)i, din);
        const double2 r0 = res[i], di = din[i], dd = dis[i];
        const double rx = r0.x - t.x, ry = r0.y - t.y;
        const double dx = c1 * di.x + c2 * (dd.x * dd.x * rx);
        const double dy = c1 * di.y + c2 * (dd.y * dd.y * ry);
        res[i] = make_double2(rx, ry);
        dout[i] = make_double2(dx, dy);
        if constexpr (ZM != Z_SKIP) {
            const double2 zv = z[i];
            const double ax = ZM == Z_PAIR ? di.x + dx : dx;
            const double ay = ZM == Z_PAIR ? di.y + dy : dy;
            z[i] = make_double2(zv.x + ax, zv.y + ay);
        }
    }
}

template <class Op>
static void launch_cheb_step_f64(uint32_t k, uint32_t last, unsigned G, hipStream_t s, const Op& op,
                                 const double2* din, double2* dout, double2* res, double2* z,
                                 const double2* dis, PcgState* st, uint32_t n, double c1,
                                 double c2) {
    if (k % 2 == 0)
        hipLaunchKernelGGL((k_cheb_step_f64<Op, Z_PAIR>), dim3(G), dim3(BLOCK), 0, s, op, din, dout,
                           res, z, dis, st, n, c1, c2);
    else if (k == last)
        hipLaunchKernelGGL((k_cheb_step_f64<Op, Z_ONE>), dim3(G), dim3(BLOCK), 0, s, op, din, dout,
                           res, z, dis, st, n, c1, c2);
    else
        hipLaunchKernelGGL((k_cheb_step_f64<Op, Z_SKIP>), dim3(G), dim3(BLOCK), 0, s, op, din, dout,
                           res, z, dis, st, n, c1, c2);
}

// the preconditioned residual z64: dis z (scaled FP32 z) or z itself
// (unscaled FP64 z)
__device__ __forceinline__ double2 z_unscaled(const float2* z, const double2* dis, size_t i) {
    const double2 dd = dis[i];
    const float2 zv = z[i];
    return make_double2(dd.x * (double)zv.x, dd.y * (double)zv.y);
}
__device__ __forceinline__ double2 z_unscaled(const double2* z, const double2*, size_t i) {
    return z[i];
}

// partials z64.w (Polak-Ribiere numerator) and r.z64
template <class P>
__global__ __launch_bounds__(BLOCK) void k_fcg_dots(const P* __restrict__ z,
                                                    const double2* __restrict__ dis,
                                                    const double2* __restrict__ w,
                                                    const double2* __restrict__ r, PcgState* st,
                                                    double* partials, uint32_t n, int first) {
    if (st->done) return;
    double acc[2] = {0.0, 0.0};
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const double2 zz = z_unscaled(z, dis, i);
        if (!first) acc[0] += dotu(zz, w[i]);
        acc[1] += dotu(r[i], zz);
    }
    __shared__ double red[2][BLOCK / WAVE];
    block_sum<2>(acc, red);
    put_partials<2>(acc, partials);
}

__global__ __launch_bounds__(FIN) void k_fin_fcg_dots(PcgState* st, const double* partials,
                                                      unsigned np, int first) {
    const uint32_t done = st->done;
    const double alpha = st->alpha, rz_old = st->rz;
    double tot[2];
    fin_sum<2>(partials, np, tot);
    if (done) return;
    if (threadIdx.x == 0) {
        const double zw = tot[0], rz = tot[1];
        double beta = first ? 0.0 : -alpha * zw / rz_old;
        if (!(rz > 0.0) || !isfinite(rz) || !isfinite(beta)) {
            // the preconditioned residual lost positivity: restart the
            // direction (p = z) and count the event
            beta = 0.0;
            st->breakdown += 1;
        }
        st->beta = beta;
        st->rz = rz;
    }
}

// p = z64 + beta p  (p = z64 on the first step)
template <class P>
__global__ __launch_bounds__(BLOCK) void k_fcg_p(const P* __restrict__ z,
                                                 const double2* __restrict__ dis,
                                                 double2* __restrict__ p, PcgState* st,
                                                 uint32_t n, int first) {
    if (st->done) return;
    const double beta = st->beta;
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const double2 zz = z_unscaled(z, dis, i);
        p[i] = first ? zz : axpyu(zz, beta, p[i]);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int poll_state(Workspace& w, hipStream_t s) {
    NBX_HIP(hipMemcpyAsync(w.h_st, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    return NBX_OK;
}

static int sampled_ms(Workspace& w, uint32_t n, double* avg) {
    double tot = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        float t = 0.f;
        NBX_HIP(hipEventElapsedTime(&t, w.sev[2 * i], w.sev[2 * i + 1]));
        tot += t;
    }
    *avg = n ? tot / n : 0.0;
    return NBX_OK;
}

// A fixed sequence of kernel launches captured once and replayed with
// hipGraphLaunch.  The solvers' repeated steps launch the same kernels with
// the same arguments every time (their state lives on the device), so one
// replay issues a step's kernels with one host call instead of one per
// kernel: the same kernels in the same order as launched one by one, so the
// same results.  The sequences hold kernel launches only.
struct Replay {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Replay() = default;
    Replay(const Replay&) = delete;
    Replay& operator=(const Replay&) = delete;
    ~Replay() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

template <class F>
static int capture(hipStream_t s, F&& launches, Replay& r) {
    NBX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    launches();
    const hipError_t le = hipGetLastError();
    const hipError_t ee = hipStreamEndCapture(s, &r.graph);
    if (le != hipSuccess || ee != hipSuccess) {
        set_last_error("stream capture of the solver's steps: %s",
                       hipGetErrorString(le != hipSuccess ? le : ee));
        return NBX_ERR_HIP;
    }
    NBX_HIP(hipGraphInstantiate(&r.exec, r.graph, nullptr, nullptr, 0));
    return NBX_OK;
}

// diag: the Jacobi diagonal, or all ones for unpreconditioned CG (q = g/1
// is g exactly, so the same kernels give conjugate_gradient.c:13-77).
template <class Op>
static int pcg_run(const Op& op, uint32_t nu, const double* diag_in, const typename Op::V* b,
                   typename Op::V* x, uint32_t max_iter, double tol, bool stream, Workspace& w,
                   hipStream_t s, SolveStats* stats) {
    using V = typename Op::V;
    V* g = (V*)w.v64[0];
    V* p = (V*)w.v64[1];
    V* wv = (V*)w.v64[2];
    const PBufs<V> pbuf{{p, (V*)w.v64[6], (V*)w.v64[7], (V*)w.v64[8]}};   // p_j in [j % 4]
    const V* diag = (const V*)diag_in;
    PcgState h{};
    h.tol2 = tol * tol;
    h.max_iter = max_iter;
    NBX_HIP(hipMemcpyAsync(w.st, &h, sizeof h, hipMemcpyHostToDevice, s));
    const unsigned G = red_grid(nu);
    // a workspace without events (the no-allocation entry point) is not timed
    const bool timed = w.ev[0] != nullptr;
    if (timed) NBX_HIP(hipEventRecord(w.ev[0], s));
    hipLaunchKernelGGL(k_pcg_init<Op>, dim3(G), dim3(BLOCK), 0, s, op, b, x, g, p, diag, w.st,
                       w.partials, nu);
    hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G);
    NBX_HIP(hipGetLastError());
    uint32_t nsamp = 0;   // sampled SpMV launches so far
    // Iteration j's kernels (after the stop they return at once).  They
    // depend on j only through j % 4: the direction buffers rotate and every
    // fourth iteration flushes the x updates.
    auto iteration = [&](uint32_t j, bool samp) {
        j &= 3u;
        const V* pj = pbuf.p[j];
        const bool flush = j == 3u;
        if (samp) (void)hipEventRecord(w.sev[2 * nsamp], s);
        hipLaunchKernelGGL(k_pcg_A<Op>, dim3(G), dim3(BLOCK), 0, s, op, pj, wv, w.st, w.partials, nu);
        if (samp) (void)hipEventRecord(w.sev[2 * nsamp++ + 1], s);
        hipLaunchKernelGGL(k_fin_A, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G);
        if (stream) {
            hipLaunchKernelGGL((k_pcg_B<V, true>), dim3(G), dim3(BLOCK), 0, s, g, wv, diag, w.st,
                               w.partials, nu);
            hipLaunchKernelGGL(k_fin_B, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G);
            if (flush)
                hipLaunchKernelGGL((k_pcg_C<V, true, true>), dim3(G), dim3(BLOCK), 0, s, x, pbuf, j,
                                   g, diag, w.st, nu);
            else
                hipLaunchKernelGGL((k_pcg_C<V, true, false>), dim3(G), dim3(BLOCK), 0, s, x, pbuf, j,
                                   g, diag, w.st, nu);
        } else {
            hipLaunchKernelGGL((k_pcg_B<V, false>), dim3(G), dim3(BLOCK), 0, s, g, wv, diag, w.st,
                               w.partials, nu);
            hipLaunchKernelGGL(k_fin_B, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G);
            if (flush)
                hipLaunchKernelGGL((k_pcg_C<V, false, true>), dim3(G), dim3(BLOCK), 0, s, x, pbuf, j,
                                   g, diag, w.st, nu);
            else
                hipLaunchKernelGGL((k_pcg_C<V, false, false>), dim3(G), dim3(BLOCK), 0, s, x, pbuf,
                                   j, g, diag, w.st, nu);
        }
    };
    // Four iterations (j % 4 = 0..3) captured once and replayed: batches
    // hold multiples of four iterations and start at j % 4 = 0.
    Replay four;
    if (w.replay && s != nullptr)
        NBX_TRY(capture(s, [&] {
            for (uint32_t q = 0; q < 4; ++q) iteration(q, false);
        }, four));
    // 64-bit, so that max_iter near UINT32_MAX ("no limit") cannot wrap the
    // guard below
    uint64_t launched = 0;
    uint32_t batch = 4;
    for (;;) {
        for (uint32_t i = 0; i < batch; i += 4) {
            // iterations j0 .. j0 + 3; a sampled one (j0 % SAMPLE_EVERY = 0)
            // is launched kernel by kernel between its events
            const uint64_t j0 = launched + i;
            const bool samp = timed && j0 % Workspace::SAMPLE_EVERY == 0 &&
                              nsamp < (uint32_t)Workspace::NSAMP;
            if (four.exec && !samp) {
                NBX_HIP(hipGraphLaunch(four.exec, s));
            } else {
                for (uint32_t q = 0; q < 4; ++q) iteration(q, samp && q == 0);
            }
        }
        NBX_HIP(hipGetLastError());
        launched += batch;
        NBX_TRY(poll_state(w, s));
        if (w.h_st->done) break;
        if (launched > (uint64_t)max_iter + 8u) {
            set_last_error("PCG did not report completion after %llu iterations",
                           (unsigned long long)launched);
            return NBX_ERR_BREAKDOWN;
        }
        batch = batch < 64 ? batch * 2 : 64;
    }
    hipLaunchKernelGGL(k_pcg_tail<V>, dim3(G), dim3(BLOCK), 0, s, x, pbuf, w.st, nu);
    NBX_HIP(hipGetLastError());
    if (!timed) {
        NBX_HIP(hipStreamSynchronize(s));
        if (stats) {
            stats->iters = stats->fp64_iters = w.h_st->k;
            stats->res0 = sqrt(w.h_st->gg0);
        }
        return NBX_OK;
    }
    NBX_HIP(hipEventRecord(w.ev[1], s));
    NBX_HIP(hipEventSynchronize(w.ev[1]));
    if (stats) {
        float ms = 0.f;
        NBX_HIP(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
        stats->ms = ms;
        stats->iters = w.h_st->k;
        stats->fp64_iters = w.h_st->k;
        stats->res0 = sqrt(w.h_st->gg0);
        // samples of iterations that did run (iteration j * SAMPLE_EVERY < k)
        const uint32_t ran = (w.h_st->k + Workspace::SAMPLE_EVERY - 1) / Workspace::SAMPLE_EVERY;
        NBX_TRY(sampled_ms(w, nsamp < ran ? nsamp : ran, &stats->ms_spmv_avg));
    }
    return NBX_OK;
}

__global__ void k_fill_one(double* __restrict__ v, uint32_t n) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) v[i] = 1.0;
}

int pcg_jacobi(const DevMatrix& m, const double* b, double* x, uint32_t max_iter, double tol,
               uint32_t* niter, double* tol_reached, Workspace& w, hipStream_t s,
               SolveStats* stats, bool unpreconditioned) {
    if (m.n == 0) {
        // empty system: no iteration, ||g|| = 0 (cg_precond_jacobi.c:44-89,
        // conjugate_gradient.c:38-76)
        if (niter) *niter = 0;
        if (tol_reached) *tol_reached = 0.0;
        return 0;
    }
    int rc = ws_alloc(w, m.n);
    if (rc) return -rc;
    const double* diag = m.diag;
    if (unpreconditioned) {
        // v64[5] is free during a PCG run (pcg_run uses v64[0..2] and the
        // direction buffers v64[6..8])
        hipLaunchKernelGGL(k_fill_one, dim3(grid_for(m.n)), dim3(BLOCK), 0, s, w.v64[5], m.n);
        if (hipGetLastError() != hipSuccess) return -NBX_ERR_HIP;
        diag = w.v64[5];
    }
    // Stream the update kernels' vectors past the caches when the
    // iteration's working set (matrix + five vectors) exceeds the 256 MiB
    // Infinity Cache; below that, caching them is faster.
    const double mat_bytes = !m.blocked ? 12.0 * (double)m.c.nnz
                             : m.has_sym ? 24.0 * (double)m.y.n_nodes + 36.0 * (double)m.y.n_upper +
                                               8.0 * (double)m.y.n_lower
                                         : 36.0 * (double)m.s.n_blocks;
    const bool stream = mat_bytes + 40.0 * (double)m.n > (double)INFINITY_CACHE_BYTES;
    if (m.blocked && m.has_sym && stream)
        rc = pcg_run(OpSym{m.y}, m.s.n_nodes, diag, (const double2*)b, (double2*)x, max_iter, tol,
                     stream, w, s, stats);
    else if (m.blocked && m.has_sym)
        rc = pcg_run(OpSymC{m.y}, m.s.n_nodes, diag, (const double2*)b, (double2*)x, max_iter, tol,
                     stream, w, s, stats);
    else if (m.blocked)
        rc = pcg_run(OpS2{m.s}, m.s.n_nodes, diag, (const double2*)b, (double2*)x, max_iter, tol,
                     stream, w, s, stats);
    else
        rc = pcg_run(OpCsr{m.c}, m.c.n, diag, b, x, max_iter, tol, stream, w, s, stats);
    if (rc) return -rc;
    if (stats) stats->layout = m.blocked && m.has_sym ? 1 : 0;
    if (niter) *niter = w.h_st->k;
    if (tol_reached) *tol_reached = w.h_st->tol_reached;
    return (int)w.h_st->status;
}

int pcg_jacobi_csr_view(const DevCsr& c, const double* diag, const double* b, double* x,
                        uint32_t max_iter, double tol, uint32_t* niter, double* tol_reached,
                        Workspace& w, hipStream_t s) {
    if (c.n == 0) {
        if (niter) *niter = 0;
        if (tol_reached) *tol_reached = 0.0;
        return 0;
    }
    const int rc = pcg_run(OpCsr{c}, c.n, diag, b, x, max_iter, tol, false, w, s, nullptr);
    if (rc) return -rc;
    if (niter) *niter = w.h_st->k;
    if (tol_reached) *tol_reached = w.h_st->tol_reached;
    return (int)w.h_st->status;
}

// Relative-tolerance solves stop on the true residual.  When the
// recursion's stop test fires, the solver restarts from its iterate; the
// restart's first step computes ||b - A x|| afresh, and a restart that
// finds it within the tolerance ends the solve with no further iteration.
// On ill-conditioned K the recursively updated residual drifts from the
// true one (at 4.2M DOF the reference algorithm's 1e-8 leaves 3.5e-8).
// A restart that finds the residual less than 10 % below the previous
// restart's has met the attainable accuracy: the solve then reports
// NBX_NOT_CONVERGED with that residual.
constexpr int SOLVE_PASSES = 8;

static bool stagnated(int pass, double res0, double prev_res0) {
    return pass > 0 && res0 >= 0.9 * prev_res0;
}

// The relative-tolerance solves refuse a right-hand side with a non-finite
// entry, and report a non-finite final residual (a non-finite matrix value
// or initial guess, or an iteration that broke down) as a breakdown rather
// than as convergence: their stop tests compare with `>`, which a NaN fails.
static int finite_rhs(double bb) {
    if (std::isfinite(bb)) return NBX_OK;
    set_last_error("the right-hand side has a non-finite entry (||b||^2 = %g)", bb);
    return NBX_ERR_BREAKDOWN;
}
static int finite_result(double rel_res) {
    if (std::isfinite(rel_res)) return NBX_OK;
    set_last_error("non-finite residual (%g): a matrix value or the initial guess is not "
                   "finite, or the iteration broke down", rel_res);
    return NBX_ERR_BREAKDOWN;
}

static void add_pass(SolveStats& tot, const SolveStats& p, int pass) {
    tot.ms += p.ms;
    tot.iters += p.iters;
    tot.outer += p.outer;
    tot.fp64_iters += p.fp64_iters;
    tot.fallback += p.fallback;
    tot.layout = p.layout;
    tot.rel_res = p.rel_res;
    if (pass == 0) tot.ms_spmv_avg = p.ms_spmv_avg;
}

int solve_fp64(const DevMatrix& m, const double* b, double* x, double rtol, bool true_residual,
               uint32_t max_iter, Workspace& w, hipStream_t s, SolveStats* stats) {
    int rc = ws_alloc(w, m.n);
    if (rc) return rc;
    const double bb = device_norm2(m.n, b, w, s, &rc);
    if (rc) return rc;
    NBX_TRY(finite_rhs(bb));
    SolveStats tot{};
    int status = NBX_NOT_CONVERGED;
    uint32_t used = 0;
    double prev = 0.0;
    const int passes = true_residual ? SOLVE_PASSES : 1;
    for (int pass = 0; pass < passes; ++pass) {
        SolveStats ps{};
        uint32_t it = 0;
        double tr = 0.0;
        const int st = pcg_jacobi(m, b, x, max_iter - used, rtol * sqrt(bb), &it, &tr, w, s, &ps);
        if (st < 0) return -st;
        used += it;
        ps.rel_res = bb > 0 ? tr / sqrt(bb) : 0.0;
        add_pass(tot, ps, pass);
        if (st != 0) break;          // out of iterations
        if (it == 0 || !true_residual) {   // the fresh residual meets the tolerance,
            status = NBX_OK;               // or the recursion's test is the stop
            break;
        }
        if (stagnated(pass, ps.res0, prev)) {
            tot.rel_res = bb > 0 ? ps.res0 / sqrt(bb) : 0.0;
            break;
        }
        prev = ps.res0;
    }
    if (stats) *stats = tot;
    NBX_TRY(finite_result(tot.rel_res));
    return status;
}

static int solve_mixed_cg(const DevMatrix& m, const double* b, double* x, double rtol,
                          const MixedOpts& opt, uint32_t max_iter, Workspace& w, hipStream_t s,
                          SolveStats* stats) {
    int rc = NBX_OK;
    const uint32_t nu = m.s.n_nodes;
    const double bb = device_norm2(m.n, b, w, s, &rc);
    if (rc) return rc;
    NBX_TRY(finite_rhs(bb));
    PcgState h{};
    h.tol2 = rtol * rtol * bb;
    h.max_iter = max_iter;
    NBX_HIP(hipMemcpyAsync(w.st, &h, sizeof h, hipMemcpyHostToDevice, s));
    float2* r32 = (float2*)w.v32[0];
    float2* p32 = (float2*)w.v32[1];
    float2* w32 = (float2*)w.v32[2];
    float2* y32 = (float2*)w.v32[3];
    const double2* dis = (const double2*)m.dis;
    const unsigned G = red_grid(nu);
    const double delta2 = opt.delta * opt.delta;
    const bool restart = opt.variant == CG_RESTART;
    const double inner_stop = restart ? opt.inner_rtol * opt.inner_rtol : 0.0;
    uint32_t nsamp = 0, samp_valid = 0, updates = 0;
    double prev_res = -1.0;
    int stalled = 0;
    NBX_HIP(hipEventRecord(w.ev[0], s));
    hipLaunchKernelGGL(k_mix_residual, dim3(G), dim3(BLOCK), 0, s, m.s, (const double2*)b,
                       (const double2*)x, dis, r32, p32, y32, w.st, w.partials, nu, 1);
    hipLaunchKernelGGL(k_fin_residual, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G, 1);
    NBX_HIP(hipGetLastError());
    NBX_TRY(poll_state(w, s));
    prev_res = w.h_st->tol_reached;
    uint32_t batch = 8;
    bool first_batch = true;
    while (!w.h_st->done) {
        for (uint32_t i = 0; i < batch; ++i) {
            const bool samp = first_batch && nsamp < (uint32_t)Workspace::NSAMP;
            if (samp) NBX_HIP(hipEventRecord(w.sev[2 * nsamp], s));
            hipLaunchKernelGGL(k_mix_A, dim3(G), dim3(BLOCK), 0, s, m.s, p32, w32, w.st,
                               w.partials32, nu);
            if (samp) NBX_HIP(hipEventRecord(w.sev[2 * nsamp++ + 1], s));
            hipLaunchKernelGGL(k_fin_mixA, dim3(1), dim3(FIN), 0, s, w.st, w.partials32, G);
            hipLaunchKernelGGL(k_mix_B, dim3(G), dim3(BLOCK), 0, s, y32, r32, p32, w32, w.st,
                               w.partials32, nu);
            hipLaunchKernelGGL(k_fin_mixB, dim3(1), dim3(FIN), 0, s, w.st, w.partials32, G, delta2,
                               inner_stop);
            hipLaunchKernelGGL(k_mix_C, dim3(G), dim3(BLOCK), 0, s, p32, r32, w.st, nu);
        }
        NBX_HIP(hipGetLastError());
        NBX_TRY(poll_state(w, s));
        if (first_batch) samp_valid = w.h_st->k;
        first_batch = false;
        if (w.h_st->done) break;
        if (w.h_st->hold) {
            if (w.h_st->breakdown) {
                stalled = 1;
                break;
            }
            // reliable update (or the end of a restarted sweep): fold y into
            // u and recompute the true residual with the FP64 matrix
            const int first = restart ? 1 : 0;
            hipLaunchKernelGGL(k_mix_update, dim3(G), dim3(BLOCK), 0, s, (double2*)x, dis, y32, w.st,
                               nu);
            hipLaunchKernelGGL(k_mix_residual, dim3(G), dim3(BLOCK), 0, s, m.s, (const double2*)b,
                               (const double2*)x, dis, r32, p32, y32, w.st, w.partials, nu, first);
            hipLaunchKernelGGL(k_fin_residual, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G, first);
            if (!first)
                hipLaunchKernelGGL(k_mix_C, dim3(G), dim3(BLOCK), 0, s, p32, r32, w.st, nu);
            NBX_HIP(hipGetLastError());
            ++updates;
            NBX_TRY(poll_state(w, s));
            const double res = w.h_st->tol_reached;
            if (!w.h_st->done && restart && res > 0.5 * prev_res) {
                stalled = 1;   // a sweep that no longer reduces the FP64 residual
                break;
            }
            prev_res = res;
            batch = 8;
            continue;
        }
        batch = batch < 64 ? batch * 2 : 64;
    }
    NBX_HIP(hipEventRecord(w.ev[1], s));
    NBX_HIP(hipEventSynchronize(w.ev[1]));
    float ms = 0.f;
    NBX_HIP(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
    const uint32_t inner = w.h_st->k;
    int status = w.h_st->status ? NBX_NOT_CONVERGED : NBX_OK;
    double rel = bb > 0 ? w.h_st->tol_reached / sqrt(bb) : 0.0;
    double spmv_avg = 0.0;
    NBX_TRY(sampled_ms(w, nsamp < samp_valid ? nsamp : samp_valid, &spmv_avg));
    SolveStats fb{};
    if (stalled) {
        // the FP32 iteration broke down or stopped reducing the residual:
        // FP64 PCG finishes from the current iterate
        rc = solve_fp64(m, b, x, rtol, opt.true_residual, max_iter, w, s, &fb);
        if (rc != NBX_OK && rc != NBX_NOT_CONVERGED) return rc;
        status = rc;
        rel = fb.rel_res;
        ms += (float)fb.ms;
    }
    if (stats) {
        stats->ms_spmv_avg = spmv_avg;
        stats->iters = inner;
        stats->outer = updates;
        stats->fp64_iters = fb.fp64_iters;
        stats->rel_res = rel;
        stats->ms = ms;
        stats->fallback = stalled;
    }
    NBX_TRY(finite_result(rel));
    return status;
}


// FP64 flexible CG with the FP32 Chebyshev-Jacobi preconditioner.
static int fcg_cheb_pass(const DevMatrix& m, const double* b, double* x, double rtol,
                         const MixedOpts& opt, uint32_t max_iter, uint32_t hint, Workspace& w,
                         hipStream_t s, SolveStats* stats) {
    const uint32_t nu = m.s.n_nodes;
    int rc = NBX_OK;
    const double bb = device_norm2(m.n, b, w, s, &rc);
    if (rc) return rc;
    PcgState h{};
    h.tol2 = rtol * rtol * bb;
    h.max_iter = max_iter;
    NBX_HIP(hipMemcpyAsync(w.st, &h, sizeof h, hipMemcpyHostToDevice, s));
    // Chebyshev coefficients on [lmax/ratio, lmax] (Saad, Alg. 12.1)
    const uint32_t deg = opt.cheb_degree < 1 ? 1 : opt.cheb_degree;
    const double lmax = m.lambda_max, lmin = lmax / opt.cheb_ratio;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
    std::vector<float> c1(deg), c2(deg);
    std::vector<double> c1d(deg), c2d(deg);
    double rho = 1.0 / sigma;
    for (uint32_t k = 1; k < deg; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        c1d[k] = rn * rho;
        c2d[k] = 2.0 * rn / delta;
        c1[k] = (float)c1d[k];
        c2[k] = (float)c2d[k];
        rho = rn;
    }
    const double inv_theta = 1.0 / theta;
    double2* r = (double2*)w.v64[0];
    double2* p = (double2*)w.v64[1];
    double2* wv = (double2*)w.v64[2];
    float2* res = (float2*)w.v32[0];
    float2* da = (float2*)w.v32[1];
    float2* db = (float2*)w.v32[2];
    float2* z = (float2*)w.v32[3];
    // the all-FP64 preconditioner's vectors (free during a flexible CG)
    double2* res64 = (double2*)w.v64[5];
    double2* da64 = (double2*)w.v64[6];
    double2* db64 = (double2*)w.v64[7];
    double2* z64 = (double2*)w.v64[8];
    const bool p64 = opt.precond_bits == 64;
    const double2* dis = (const double2*)m.dis;
    const double2* bv = (const double2*)b;
    double2* xv = (double2*)x;
    const unsigned G = red_grid(nu);
    OpS2 op{m.s};
    OpSym ops{m.y};
    // The fp16 preconditioner's working set (matrix + four FP32 vectors) is
    // read by each of the degree - 1 steps in a row: when it fits in the
    // Infinity Cache, cached loads keep it there between steps (at 4.2M DOF,
    // 243 MB: 1.48 -> 1.32 s per solve); otherwise the matrix streams.  The
    // FP64 preconditioner reads the FP64 matrix the same way, cached when
    // its working set (matrix + five FP64 vectors) fits.
    const bool off16 = opt.precond_bits == 16 && m.has_off16;
    const bool packed = off16 && m.off_packed;
    const double sym_mat = 24.0 * (double)m.y.n_nodes + 36.0 * (double)m.y.n_upper +
                           8.0 * (double)m.y.n_lower;
    const double cheb_mat = p64     ? (m.has_sym ? sym_mat : 36.0 * (double)m.s.n_blocks)
                            : packed ? 8.0 * (double)m.o.n_blocks + 4.0 * (double)m.o.n_pairs +
                                           4.0 * (double)nu
                            : off16 ? 12.0 * (double)m.o.n_blocks + 4.0 * (double)nu
                                    : 12.0 * (double)m.s.n_blocks;
    const bool cheb_cached =
        cheb_mat + (p64 ? 80.0 : 33.0) * (double)nu <= (double)INFINITY_CACHE_BYTES;
    const bool amg = opt.variant == AMG_FCG;
    const double2* ones = amg ? (const double2*)amg_ones(m.amg) : nullptr;
    const AmgFine fine = amg ? amg_fine(m.amg) : AmgFine{};
    // the fp16 copy's working set, as above, for the multigrid's finest level
    const double fine_mat = m.off_packed ? 8.0 * (double)m.o.n_blocks +
                                               4.0 * (double)m.o.n_pairs + 4.0 * (double)nu
                                         : 12.0 * (double)m.o.n_blocks + 4.0 * (double)nu;
    const bool fine_cached = fine_mat + 41.0 * (double)nu <= (double)INFINITY_CACHE_BYTES;
    auto precondition_amg16 = [&](int first) {
        hipLaunchKernelGGL(k_fcg_update<float2>, dim3(G), dim3(BLOCK), 0, s, m.s, bv, xv, r, p, wv,
                           dis, res, da, z, w.st, w.partials, nu, first, fine.inv_theta);
        hipLaunchKernelGGL(k_fin_fcg_update, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G, first);
        auto cycle = [&](const auto& K) {
            if (fine_cached)
                amg16_cycle<false>(K, fine, m.amg, res, da, db, z, w.st, nu, G, s);
            else
                amg16_cycle<true>(K, fine, m.amg, res, da, db, z, w.st, nu, G, s);
        };
        if (m.off_packed && m.o.width)
            cycle(Off16U{m.o});
        else if (m.off_packed)
            cycle(Off16P{m.o});
        else
            cycle(m.o);
        hipLaunchKernelGGL(k_fcg_dots<float2>, dim3(G), dim3(BLOCK), 0, s, z, dis, wv, r, w.st,
                           w.partials, nu, first);
        hipLaunchKernelGGL(k_fin_fcg_dots, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G, first);
        hipLaunchKernelGGL(k_fcg_p<float2>, dim3(G), dim3(BLOCK), 0, s, z, dis, p, w.st, nu, first);
    };
    auto precondition_amg = [&](int first) {
        // a one-level hierarchy: the update with a unit scaling writes r in
        // FP32 (res), the dense solve z = C^-1 r, the dot products and
        // direction read z unscaled
        hipLaunchKernelGGL(k_fcg_update<float2>, dim3(G), dim3(BLOCK), 0, s, m.s, bv, xv, r, p, wv,
                           ones, res, da, z, w.st, w.partials, nu, first, 1.0);
        hipLaunchKernelGGL(k_fin_fcg_update, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G, first);
        (void)amg_vcycle(m.amg, (const float*)res, (float*)z, w.st, s);
        hipLaunchKernelGGL(k_fcg_dots<float2>, dim3(G), dim3(BLOCK), 0, s, z, ones, wv, r, w.st,
                           w.partials, nu, first);
        hipLaunchKernelGGL(k_fin_fcg_dots, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G, first);
        hipLaunchKernelGGL(k_fcg_p<float2>, dim3(G), dim3(BLOCK), 0, s, z, ones, p, w.st, nu,
                           first);
    };
    auto precondition = [&](int first) {
        if (amg && fine.scaled16) {
            precondition_amg16(first);
            return;
        }
        if (amg) {
            precondition_amg(first);
            return;
        }
        if (p64) {
            hipLaunchKernelGGL(k_fcg_update<double2>, dim3(G), dim3(BLOCK), 0, s, m.s, bv, xv, r, p,
                               wv, dis, res64, da64, z64, w.st, w.partials, nu, first, inv_theta);
        } else {
            hipLaunchKernelGGL(k_fcg_update<float2>, dim3(G), dim3(BLOCK), 0, s, m.s, bv, xv, r, p,
                               wv, dis, res, da, z, w.st, w.partials, nu, first, inv_theta);
        }
        hipLaunchKernelGGL(k_fin_fcg_update, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G, first);
        float2* din = da;
        float2* dout = db;
        double2* din64 = da64;
        double2* dout64 = db64;
        for (uint32_t k = 1; k < deg; ++k) {
            auto step = [&](const auto& K) {
                if (cheb_cached)
                    launch_cheb_step64<false>(k, deg - 1, G, s, K, din, dout, res, z, w.st, nu,
                                              c1d[k], c2d[k]);
                else
                    launch_cheb_step64<true>(k, deg - 1, G, s, K, din, dout, res, z, w.st, nu,
                                             c1d[k], c2d[k]);
            };
            if (p64 && m.has_sym && cheb_cached)
                launch_cheb_step_f64(k, deg - 1, G, s, OpSymC{m.y}, din64, dout64, res64, z64, dis,
                                     w.st, nu, c1d[k], c2d[k]);
            else if (p64 && m.has_sym)
                launch_cheb_step_f64(k, deg - 1, G, s, ops, din64, dout64, res64, z64, dis, w.st,
                                     nu, c1d[k], c2d[k]);
            else if (p64)
                launch_cheb_step_f64(k, deg - 1, G, s, op, din64, dout64, res64, z64, dis, w.st,
                                     nu, c1d[k], c2d[k]);
            else if (packed && m.o.width)
                step(Off16U{m.o});
            else if (packed)
                step(Off16P{m.o});
            else if (off16)
                step(m.o);
            else if (opt.precond_bits == 16)
                step(m.s);
            else
                hipLaunchKernelGGL(k_cheb_step<32>, dim3(G), dim3(BLOCK), 0, s, m.s, din, dout, res, z,
                                   w.st, nu, c1[k], c2[k]);
            std::swap(din, dout);
            std::swap(din64, dout64);
        }
        if (p64) {
            hipLaunchKernelGGL(k_fcg_dots<double2>, dim3(G), dim3(BLOCK), 0, s, z64, dis, wv, r,
                               w.st, w.partials, nu, first);
        } else {
            hipLaunchKernelGGL(k_fcg_dots<float2>, dim3(G), dim3(BLOCK), 0, s, z, dis, wv, r, w.st,
                               w.partials, nu, first);
        }
        hipLaunchKernelGGL(k_fin_fcg_dots, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G, first);
        if (p64)
            hipLaunchKernelGGL(k_fcg_p<double2>, dim3(G), dim3(BLOCK), 0, s, z64, dis, p, w.st, nu,
                               first);
        else
            hipLaunchKernelGGL(k_fcg_p<float2>, dim3(G), dim3(BLOCK), 0, s, z, dis, p, w.st, nu,
                               first);
    };
    NBX_HIP(hipEventRecord(w.ev[0], s));
    precondition(1);
    NBX_HIP(hipGetLastError());
    NBX_TRY(poll_state(w, s));
    // Iterations launched between polls of the device state.  Kernels
    // launched after the stop test fired return at once but still cost a
    // launch each (about 2.5 us; some 50 per outer iteration with the
    // multigrid preconditioner), so the batches end near the expected count:
    // with a hint (the iterations of a similar earlier solve) the first
    // batch runs to 7/8 of it and the next ones take 4, 4, 8, 16, 32; without
    // one they take 4, 8, 16, 32.
    uint32_t batch = 4, nsamp = 0;
    int fours = 0;   // batches of 4 still to come before the doubling
    if (hint >= 16) {
        batch = hint - hint / 8;
        if (batch > max_iter) batch = max_iter > 4 ? max_iter : 4u;
        fours = 2;
    }
    uint64_t launched = 0;   // 64-bit: max_iter near UINT32_MAX cannot wrap the guard
    constexpr uint32_t every = 8;   // sample one outer iteration in eight
    // An outer iteration: the FP64 product, then the preconditioner and the
    // flexible-CG updates; the same kernels and arguments every time.
    auto outer = [&](bool samp) {
        if (samp) (void)hipEventRecord(w.sev[2 * nsamp], s);
        if (m.has_sym)
            hipLaunchKernelGGL(k_pcg_A<OpSym>, dim3(G), dim3(BLOCK), 0, s, ops,
                               (const double2*)p, wv, w.st, w.partials, nu);
        else
            hipLaunchKernelGGL(k_pcg_A<OpS2>, dim3(G), dim3(BLOCK), 0, s, op,
                               (const double2*)p, wv, w.st, w.partials, nu);
        if (samp) (void)hipEventRecord(w.sev[2 * nsamp++ + 1], s);
        hipLaunchKernelGGL(k_fin_A, dim3(1), dim3(FIN), 0, s, w.st, w.partials, G);
        precondition(0);
    };
    Replay one;
    if (w.replay && s != nullptr && !w.h_st->done)
        NBX_TRY(capture(s, [&] { outer(false); }, one));
    while (!w.h_st->done) {
        for (uint32_t i = 0; i < batch; ++i) {
            const bool samp = (launched + i) % every == 0 && nsamp < (uint32_t)Workspace::NSAMP;
            if (one.exec && !samp)
                NBX_HIP(hipGraphLaunch(one.exec, s));
            else
                outer(samp);
        }
        NBX_HIP(hipGetLastError());
        launched += batch;
        NBX_TRY(poll_state(w, s));
        if (launched > (uint64_t)max_iter + 8u && !w.h_st->done) {
            set_last_error("flexible CG did not report completion");
            return NBX_ERR_BREAKDOWN;
        }
        if (fours > 0) {
            batch = 4;
            --fours;
        } else {
            batch = batch < 32 ? batch * 2 : 32;
        }
    }
    NBX_HIP(hipEventRecord(w.ev[1], s));
    NBX_HIP(hipEventSynchronize(w.ev[1]));
    if (stats) {
        float ms = 0.f;
        NBX_HIP(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
        stats->ms = ms;
        stats->iters = w.h_st->k;
        // SpMVs in the preconditioner (Chebyshev), V-cycles (AMG)
        stats->outer = amg ? w.h_st->k : w.h_st->k * (deg - 1);
        stats->fp64_iters = w.h_st->k;
        stats->rel_res = bb > 0 ? w.h_st->tol_reached / sqrt(bb) : 0.0;
        stats->fallback = (int)w.h_st->breakdown;
        // bits 2, 4, 8: the fp16 copy (packed columns, uniform width) that
        // the Chebyshev preconditioner or the multigrid's finest level read
        const bool o16 = amg ? fine.scaled16 : off16;
        const bool pk = amg ? fine.scaled16 && m.off_packed : packed;
        stats->layout = (m.has_sym ? 1 : 0) | (o16 ? 2 : 0) | (pk ? 4 : 0) |
                        (pk && m.o.width ? 8 : 0) | (amg ? 16 : 0);
        stats->res0 = sqrt(w.h_st->gg0);
        const uint32_t ran = (w.h_st->k + every - 1) / every;
        NBX_TRY(sampled_ms(w, nsamp < ran ? nsamp : ran, &stats->ms_spmv_avg));
    }
    return w.h_st->status ? NBX_NOT_CONVERGED : NBX_OK;
}

// Flexible CG passes until a restart's fresh residual meets the tolerance.
static int solve_fcg_cheb(const DevMatrix& m, const double* b, double* x, double rtol,
                          const MixedOpts& opt, uint32_t max_iter, Workspace& w, hipStream_t s,
                          SolveStats* stats) {
    int rc = NBX_OK;
    const double bn = sqrt(device_norm2(m.n, b, w, s, &rc));
    if (rc) return rc;
    NBX_TRY(finite_rhs(bn * bn));
    SolveStats tot{};
    int status = NBX_NOT_CONVERGED;
    uint32_t used = 0;
    double prev = 0.0;
    const int passes = opt.true_residual ? SOLVE_PASSES : 1;
    for (int pass = 0; pass < passes; ++pass) {
        SolveStats ps{};
        rc = fcg_cheb_pass(m, b, x, rtol, opt, max_iter - used, pass == 0 ? w.fcg_hint : 0u, w,
                           s, &ps);
        if (pass == 0) w.fcg_hint = rc == NBX_OK ? ps.iters : 0u;
        if (rc != NBX_OK && rc != NBX_NOT_CONVERGED) return rc;
        used += ps.iters;
        add_pass(tot, ps, pass);
        if (rc == NBX_NOT_CONVERGED) break;
        if (ps.iters == 0 || !opt.true_residual) {
            status = NBX_OK;
            break;
        }
        if (stagnated(pass, ps.res0, prev)) {
            tot.rel_res = bn > 0 ? ps.res0 / bn : 0.0;
            break;
        }
        prev = ps.res0;
    }
    if (stats) *stats = tot;
    NBX_TRY(finite_result(tot.rel_res));
    return status;
}

int solve_mixed(const DevMatrix& m, const double* b, double* x, double rtol, const MixedOpts& opt,
                uint32_t max_iter, Workspace& w, hipStream_t s, SolveStats* stats) {
    const bool cheb = opt.variant == CHEB_FCG;
    const bool fp16 = cheb && opt.precond_bits == 16;
    const bool fp64 = cheb && opt.precond_bits == 64;
    if (opt.variant == AMG_FCG) {
        if (!m.blocked || !m.amg || !m.amg_current) {
            set_last_error("solve_mixed: no current AMG hierarchy (the AMG preconditioner is "
                           "built for nbx_fem systems, whose node coordinates it needs)");
            return NBX_ERR_ARG;
        }
        int rc = ws_alloc(w, m.n);
        if (rc) return rc;
        return solve_fcg_cheb(m, b, x, rtol, opt, max_iter, w, s, stats);
    }
    if (!m.blocked || !m.has_scale ||
        !(fp64 || (fp16 ? m.has_off16 || m.has_val16 : m.has_val32))) {
        set_last_error("solve_mixed: matrix has no scaled %d-bit copy", fp16 ? 16 : 32);
        return NBX_ERR_ARG;
    }
    int rc = ws_alloc(w, m.n);
    if (rc) return rc;
    if (opt.variant == CHEB_FCG) return solve_fcg_cheb(m, b, x, rtol, opt, max_iter, w, s, stats);
    return solve_mixed_cg(m, b, x, rtol, opt, max_iter, w, s, stats);
}

}  // namespace nbx

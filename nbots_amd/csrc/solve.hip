// Sparse matrix-vector products and conjugate-gradient solvers for gfx950.
//
// * SpMV on the 2x2-block sliced-ELL format (one thread per node row, two
//   scalar rows per thread) or on scalar CSR (one thread per row).  Each
//   scalar row is summed in ascending column order, exactly like
//   nb_sparse_multiply_vector (solver_bot/sparse/sparse.c:405-414), so the
//   products are bit-identical to the reference's.
// * Jacobi-preconditioned CG with the reference's recurrences and stop test
//   (solver_bot/sparse/solvers/cg_precond_jacobi.c:13-90), device resident:
//   alpha/beta and the stop decision are computed on the GPU by the last
//   workgroup of each reduction; the host only polls a done flag.
// * Mixed precision: FP32 CG on the Jacobi-scaled FP32 matrix copy
//   (D^-1/2 K D^-1/2, computed in FP64 and rounded once) inside FP64
//   iterative refinement (residual r = b - K u and update u += D^-1/2 y in
//   FP64).  CG on the symmetrically Jacobi-scaled system is Jacobi-PCG in
//   other variables.
//
// The path is bound by HBM bandwidth (about 0.17 flop/byte in FP64); the
// kernels stream the matrix once per product, keep x gathers in L2 via the
// XCD-contiguous workgroup mapping, and fuse the vector updates with the
// dot products.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "internal.hpp"

namespace nbx {

static char g_err[512];
void set_last_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
const char* last_error() { return g_err; }

// ---------------------------------------------------------------------------
// per-unit vector arithmetic (unit = one node with 2 dofs, or one scalar row)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dotu(double a, double b) { return a * b; }
__device__ __forceinline__ double dotu(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ double axpyu(double x, double a, double p) { return x + a * p; }
__device__ __forceinline__ double2 axpyu(double2 x, double a, double2 p) {
    return make_double2(x.x + a * p.x, x.y + a * p.y);
}
__device__ __forceinline__ double divu(double g, double d) { return g / d; }
__device__ __forceinline__ double2 divu(double2 g, double2 d) {
    return make_double2(g.x / d.x, g.y / d.y);
}
// p = -q + beta * p  (cg_precond_jacobi.c:72-74)
__device__ __forceinline__ double pupd(double q, double b, double p) { return -q + b * p; }
__device__ __forceinline__ double2 pupd(double2 q, double b, double2 p) {
    return make_double2(-q.x + b * p.x, -q.y + b * p.y);
}
__device__ __forceinline__ double negu(double a) { return -a; }
__device__ __forceinline__ double2 negu(double2 a) { return make_double2(-a.x, -a.y); }
__device__ __forceinline__ double subu(double a, double b) { return a - b; }
__device__ __forceinline__ double2 subu(double2 a, double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
}

// ---------------------------------------------------------------------------
// matrix row operators
// ---------------------------------------------------------------------------
struct OpS2 {
    Sell2 A;
    using V = double2;
    __device__ __forceinline__ double2 row(uint32_t v, const double2* __restrict__ x) const {
        const uint32_t s = v >> 6, l = v & 63;
        size_t idx = (size_t)A.slice_ptr[s] + l;
        const int len = A.rowlen[v];
        const double2* __restrict__ val = reinterpret_cast<const double2*>(A.val);
        double s0 = 0.0, s1 = 0.0;
        for (int k = 0; k < len; ++k, idx += SLICE) {
            const uint32_t w = A.bcol[idx];
            const double2 a01 = val[2 * idx];
            const double2 a23 = val[2 * idx + 1];
            const double2 xv = x[w];
            s0 = s0 + a01.x * xv.x;
            s0 = s0 + a01.y * xv.y;
            s1 = s1 + a23.x * xv.x;
            s1 = s1 + a23.y * xv.y;
        }
        return make_double2(s0, s1);
    }
};

struct OpCsr {
    DevCsr A;
    using V = double;
    __device__ __forceinline__ double row(uint32_t r, const double* __restrict__ x) const {
        double s = 0.0;
        const uint32_t e = A.rowptr[r + 1];
        for (uint32_t j = A.rowptr[r]; j < e; ++j) s = s + A.val[j] * x[A.col[j]];
        return s;
    }
};

// Contiguous chunk of units owned by this workgroup (XCD-aware).
struct Chunk {
    size_t beg, end;
};
__device__ __forceinline__ Chunk my_chunk(uint32_t n) {
    const unsigned G = gridDim.x;
    const unsigned lb = xcd_remap(blockIdx.x, G);
    size_t per = ((size_t)n + (size_t)G * BLOCK - 1) / ((size_t)G * BLOCK) * BLOCK;
    size_t b = (size_t)lb * per;
    size_t e = b + per;
    if (b > n) b = n;
    if (e > n) e = n;
    return {b, e};
}

// ---------------------------------------------------------------------------
// SpMV
// ---------------------------------------------------------------------------
template <class Op>
__global__ __launch_bounds__(BLOCK) void k_spmv(Op op, const typename Op::V* __restrict__ x,
                                                typename Op::V* __restrict__ y, uint32_t n) {
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) y[u] = op.row((uint32_t)u, x);
}

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
                                                    PcgState* st, double* partials,
                                                    unsigned* ticket, uint32_t n) {
    using V = typename Op::V;
    double acc[2] = {0.0, 0.0};
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) {
        V gv = subu(op.row((uint32_t)u, x), b[u]);
        V q = divu(gv, diag[u]);
        g[u] = gv;
        p[u] = negu(q);
        acc[0] += dotu(gv, gv);
        acc[1] += dotu(gv, q);
    }
    __shared__ double red[2][BLOCK / WAVE];
    block_sum<2>(acc, red);
    double tot[2];
    if (grid_sum_last<2>(acc, partials, ticket, tot)) {
        const double gg = tot[0];
        st->rz = tot[1];
        st->gg_cur = gg;
        st->gg_prev = gg;
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
                                                 double* partials, unsigned* ticket, uint32_t n) {
    if (st->done) return;
    double acc[1] = {0.0};
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) {
        typename Op::V wv = op.row((uint32_t)u, p);
        w[u] = wv;
        acc[0] += dotu(p[u], wv);
    }
    __shared__ double red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    double tot[1];
    if (grid_sum_last<1>(acc, partials, ticket, tot)) {
        st->pw = tot[0];
        st->alpha = st->rz / tot[0];
    }
}

// x += alpha p; g += alpha w; q = g / Aii; rz' = g.q; gg' = g.g; the lagged
// stop test of cg_precond_jacobi.c:45 (cg_precond_jacobi.c:62-70, :75)
template <class V>
__global__ __launch_bounds__(BLOCK) void k_pcg_B(V* __restrict__ x, V* __restrict__ g,
                                                 const V* __restrict__ p,
                                                 const V* __restrict__ w,
                                                 const V* __restrict__ diag, PcgState* st,
                                                 double* partials, unsigned* ticket, uint32_t n) {
    if (st->done) return;
    const double alpha = st->alpha;
    double acc[2] = {0.0, 0.0};
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) {
        x[u] = axpyu(x[u], alpha, p[u]);
        V gv = axpyu(g[u], alpha, w[u]);
        g[u] = gv;
        V q = divu(gv, diag[u]);
        acc[0] += dotu(gv, q);
        acc[1] += dotu(gv, gv);
    }
    __shared__ double red[2][BLOCK / WAVE];
    block_sum<2>(acc, red);
    double tot[2];
    if (grid_sum_last<2>(acc, partials, ticket, tot)) {
        st->beta = tot[0] / st->rz;
        st->rz = tot[0];
        const double gg_prev = st->gg_cur;
        st->gg_prev = gg_prev;
        st->gg_cur = tot[1];
        const uint32_t k = st->k + 1;
        st->k = k;
        if (!(gg_prev > st->tol2 && k < st->max_iter)) {
            st->done = 1;
            st->tol_reached = sqrt(gg_prev);
            st->status = gg_prev > st->tol2 ? 1u : 0u;
        }
    }
}

// p = -q + beta p   (cg_precond_jacobi.c:72-74)
template <class V>
__global__ __launch_bounds__(BLOCK) void k_pcg_C(V* __restrict__ p, const V* __restrict__ g,
                                                 const V* __restrict__ diag, PcgState* st,
                                                 uint32_t n) {
    if (st->done) return;
    const double beta = st->beta;
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) {
        V q = divu(g[u], diag[u]);
        p[u] = pupd(q, beta, p[u]);
    }
}

// ---------------------------------------------------------------------------
// small utility kernels
// ---------------------------------------------------------------------------
__global__ void k_norm2(const double* __restrict__ v, size_t n, double* partials,
                        unsigned* ticket, double* out) {
    double acc[1] = {0.0};
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK)
        acc[0] += v[i] * v[i];
    __shared__ double red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    double tot[1];
    if (grid_sum_last<1>(acc, partials, ticket, tot)) *out = tot[0];
}

__global__ void k_diag_sell2(Sell2 A, double* __restrict__ diag) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    size_t idx = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    double d0 = 0.0, d1 = 0.0;
    for (int k = 0; k < A.rowlen[v]; ++k, idx += SLICE) {
        if (A.bcol[idx] == v) {
            d0 = A.val[4 * idx];
            d1 = A.val[4 * idx + 3];
        }
    }
    diag[2 * v] = d0;
    diag[2 * v + 1] = d1;
}

// nb_sparse_get(A, i, i): 0.0 when (i,i) is not stored (sparse.c:164-171)
__global__ void k_diag_csr(DevCsr A, double* __restrict__ diag) {
    const uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r >= A.n) return;
    int64_t lo = A.rowptr[r], hi = (int64_t)A.rowptr[r + 1] - 1;
    double d = 0.0;
    while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;
        uint32_t cm = A.col[mid];
        if (cm == r) {
            d = A.val[mid];
            break;
        }
        if (cm < r) lo = mid + 1;
        else hi = mid - 1;
    }
    diag[r] = d;
}

// dis = 1/sqrt(diag); val32 = fl32(a_ij * dis_i * dis_j), computed in FP64.
__global__ void k_dis(const double* __restrict__ diag, double* __restrict__ dis, uint32_t n,
                      unsigned* bad) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double d = diag[i];
    if (!(d > 0.0)) {
        atomicAdd(bad, 1u);
        dis[i] = 1.0;
    } else {
        dis[i] = 1.0 / sqrt(d);
    }
}

__global__ void k_scale32(Sell2 A, const double* __restrict__ dis) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    size_t idx = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const int len = A.rowlen[v];
    const double r0 = dis[2 * v], r1 = dis[2 * v + 1];
    for (int k = 0; k < len; ++k, idx += SLICE) {
        const uint32_t w = A.bcol[idx];
        const double c0 = dis[2 * w], c1 = dis[2 * w + 1];
        const double* a = A.val + 4 * idx;
        float4 o;
        o.x = (float)(a[0] * r0 * c0);
        o.y = (float)(a[1] * r0 * c1);
        o.z = (float)(a[2] * r1 * c0);
        o.w = (float)(a[3] * r1 * c1);
        reinterpret_cast<float4*>(A.val32)[idx] = o;
    }
}

// ---------------------------------------------------------------------------
// mixed precision: FP32 scaled CG inside FP64 iterative refinement
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 row32(const Sell2& A, uint32_t v, const float2* __restrict__ x) {
    const uint32_t s = v >> 6, l = v & 63;
    size_t idx = (size_t)A.slice_ptr[s] + l;
    const int len = A.rowlen[v];
    const float4* __restrict__ val = reinterpret_cast<const float4*>(A.val32);
    float s0 = 0.f, s1 = 0.f;
    for (int k = 0; k < len; ++k, idx += SLICE) {
        const uint32_t w = A.bcol[idx];
        const float4 a = val[idx];
        const float2 xv = x[w];
        s0 = s0 + a.x * xv.x;
        s0 = s0 + a.y * xv.y;
        s1 = s1 + a.z * xv.x;
        s1 = s1 + a.w * xv.y;
    }
    return make_float2(s0, s1);
}

__global__ __launch_bounds__(BLOCK) void k_spmv32(Sell2 A, const float2* __restrict__ x,
                                                  float2* __restrict__ y, uint32_t n) {
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) y[i] = row32(A, (uint32_t)i, x);
}

// Outer step: r = b - K u (FP64), ||r||^2; r32 = fl32(dis * r); y32 = 0;
// p32 = r32; inner rz = r32.r32.  st[0] = outer state, st[1] = inner.
__global__ __launch_bounds__(BLOCK) void k_ir_residual(OpS2 op, const double2* __restrict__ b,
                                                       const double2* __restrict__ u,
                                                       const double2* __restrict__ dis,
                                                       float2* __restrict__ r32,
                                                       float2* __restrict__ p32,
                                                       float2* __restrict__ y32, PcgState* st,
                                                       double* partials, unsigned* ticket,
                                                       uint32_t n, double inner_rtol2) {
    if (st[0].done) return;
    double acc[2] = {0.0, 0.0};
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const double2 au = op.row((uint32_t)i, u);
        const double2 bi = b[i];
        const double2 r = make_double2(bi.x - au.x, bi.y - au.y);
        const double2 d = dis[i];
        const float2 rs = make_float2((float)(d.x * r.x), (float)(d.y * r.y));
        r32[i] = rs;
        p32[i] = rs;
        y32[i] = make_float2(0.f, 0.f);
        acc[0] += r.x * r.x + r.y * r.y;
        acc[1] += (double)rs.x * rs.x + (double)rs.y * rs.y;
    }
    __shared__ double red[2][BLOCK / WAVE];
    block_sum<2>(acc, red);
    double tot[2];
    if (grid_sum_last<2>(acc, partials, ticket, tot)) {
        PcgState& o = st[0];
        PcgState& in = st[1];
        o.gg_prev = o.gg_cur;
        o.gg_cur = tot[0];
        o.tol_reached = sqrt(tot[0]);
        if (!(tot[0] > o.tol2)) {
            o.done = 1;
            o.status = 0;
        } else if (o.k >= o.max_iter) {
            o.done = 1;
            o.status = 1;
        }
        o.k += 1;
        in.rz = tot[1];
        in.tol2 = inner_rtol2 * tot[1];
        in.k = 0;
        in.done = (o.done || !(tot[1] > 0.0)) ? 1u : 0u;
        in.breakdown = 0;
    }
}

// inner: w = Kt p; pw; alpha = rz / pw
__global__ __launch_bounds__(BLOCK) void k_in_A(Sell2 A, const float2* __restrict__ p,
                                                float2* __restrict__ w, PcgState* st,
                                                float* partials, unsigned* ticket, uint32_t n) {
    if (st[1].done) return;
    f32c acc[1] = {{0.f, 0.f}};
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const float2 wv = row32(A, (uint32_t)i, p);
        w[i] = wv;
        const float2 pv = p[i];
        acc[0] = cadd(acc[0], pv.x * wv.x);
        acc[0] = cadd(acc[0], pv.y * wv.y);
    }
    __shared__ f32c red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    f32c tot[1];
    if (grid_sum_last<1>(acc, partials, ticket, tot)) {
        const double pw = (double)tot[0].s + (double)tot[0].c;
        PcgState& in = st[1];
        in.pw = pw;
        if (!(pw > 0.0) || !isfinite(pw)) {
            in.breakdown = 1;
            in.done = 1;
        }
        in.alpha = in.rz / pw;
    }
}

// inner: y += alpha p; r -= alpha w; rz' = r.r; beta = rz'/rz; stop test
__global__ __launch_bounds__(BLOCK) void k_in_B(float2* __restrict__ y, float2* __restrict__ r,
                                                const float2* __restrict__ p,
                                                const float2* __restrict__ w, PcgState* st,
                                                float* partials, unsigned* ticket, uint32_t n) {
    if (st[1].done) return;
    const float alpha = (float)st[1].alpha;
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
    f32c tot[1];
    if (grid_sum_last<1>(acc, partials, ticket, tot)) {
        PcgState& in = st[1];
        const double rz = (double)tot[0].s + (double)tot[0].c;
        in.beta = rz / in.rz;
        in.rz = rz;
        in.k += 1;
        if (!(rz > in.tol2) || in.k >= in.max_iter) in.done = 1;
        if (!isfinite(rz)) {
            in.breakdown = 1;
            in.done = 1;
        }
    }
}

// inner: p = r + beta p
__global__ __launch_bounds__(BLOCK) void k_in_C(float2* __restrict__ p, const float2* __restrict__ r,
                                                PcgState* st, uint32_t n) {
    if (st[1].done) return;
    const float beta = (float)st[1].beta;
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const float2 rv = r[i], pv = p[i];
        p[i] = make_float2(rv.x + beta * pv.x, rv.y + beta * pv.y);
    }
}

// outer: u += dis * y  (FP64)
__global__ __launch_bounds__(BLOCK) void k_ir_update(double2* __restrict__ u,
                                                     const double2* __restrict__ dis,
                                                     const float2* __restrict__ y, PcgState* st,
                                                     uint32_t n) {
    if (st[0].done) return;
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const double2 d = dis[i];
        const float2 yv = y[i];
        double2 uv = u[i];
        uv.x = uv.x + d.x * (double)yv.x;
        uv.y = uv.y + d.y * (double)yv.y;
        u[i] = uv;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static unsigned red_grid(uint32_t n_units) {
    unsigned g = grid_for(n_units);
    return g > (unsigned)MAX_GRID ? (unsigned)MAX_GRID : g;
}
static unsigned map_grid(uint32_t n_units) {
    unsigned g = grid_for(n_units);
    return g > 65535u ? 65535u : g;
}

int ws_alloc(Workspace& w, uint32_t n) {
    if (w.n >= n && w.st) return NBX_OK;
    ws_free(w);
    const size_t nn = (size_t)(n ? n : 1) + 2;
    for (auto& p : w.v64) NBX_HIP(hipMalloc(&p, nn * sizeof(double)));
    for (auto& p : w.v32) NBX_HIP(hipMalloc(&p, nn * sizeof(float)));
    NBX_HIP(hipMalloc(&w.partials, 4 * MAX_GRID * sizeof(double)));
    NBX_HIP(hipMalloc(&w.partials32, 4 * MAX_GRID * sizeof(float)));
    NBX_HIP(hipMalloc(&w.ticket, 8 * sizeof(unsigned)));
    NBX_HIP(hipMemset(w.ticket, 0, 8 * sizeof(unsigned)));
    NBX_HIP(hipMalloc(&w.st, 2 * sizeof(PcgState)));
    NBX_HIP(hipHostMalloc(&w.h_st, 2 * sizeof(PcgState), hipHostMallocDefault));
    for (auto& e : w.ev) NBX_HIP(hipEventCreate(&e));
    for (auto& e : w.sev) NBX_HIP(hipEventCreate(&e));
    w.n = n;
    return NBX_OK;
}

void ws_free(Workspace& w) {
    for (auto& p : w.v64) {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    for (auto& p : w.v32) {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    if (w.partials) (void)hipFree(w.partials);
    if (w.partials32) (void)hipFree(w.partials32);
    if (w.ticket) (void)hipFree(w.ticket);
    if (w.st) (void)hipFree(w.st);
    if (w.h_st) (void)hipHostFree(w.h_st);
    for (auto& e : w.ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    for (auto& e : w.sev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    w.partials = nullptr;
    w.partials32 = nullptr;
    w.ticket = nullptr;
    w.st = nullptr;
    w.h_st = nullptr;
    w.n = 0;
}

void mat_free(DevMatrix& m) {
    void* ps[] = {m.s.slice_ptr, m.s.rowlen, m.s.bcol, m.s.val, m.s.val32,
                  m.c.rowptr,    m.c.col,    m.c.val,  m.diag,  m.dis};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    m = DevMatrix();
}

// Is the CSR a 2-dof node matrix (rows 2v, 2v+1 share their columns, which
// come in pairs 2w, 2w+1)?
static bool csr_blockable(uint32_t n, const uint32_t* rp, const uint32_t* col) {
    if (n == 0 || (n & 1u)) return false;
    for (uint32_t v = 0; v < n / 2; ++v) {
        const uint32_t a = rp[2 * v], b = rp[2 * v + 1], e = rp[2 * v + 2];
        const uint32_t la = b - a, lb = e - b;
        if (la != lb || (la & 1u) || la / 2 > 255u) return false;
        for (uint32_t k = 0; k < la; k += 2) {
            const uint32_t c0 = col[a + k], c1 = col[a + k + 1];
            if ((c0 & 1u) || c1 != c0 + 1 || col[b + k] != c0 || col[b + k + 1] != c1) return false;
        }
    }
    return true;
}

static void sell2_values_host(uint32_t nn, const uint32_t* rp, const double* val,
                              const std::vector<uint32_t>& sptr, std::vector<double>& hv) {
    for (uint32_t v = 0; v < nn; ++v) {
        const uint32_t a = rp[2 * v], b = rp[2 * v + 1], len = (b - a) / 2;
        size_t idx = (size_t)sptr[v >> 6] + (v & 63);
        for (uint32_t k = 0; k < len; ++k, idx += SLICE) {
            hv[4 * idx + 0] = val[a + 2 * k];
            hv[4 * idx + 1] = val[a + 2 * k + 1];
            hv[4 * idx + 2] = val[b + 2 * k];
            hv[4 * idx + 3] = val[b + 2 * k + 1];
        }
    }
}

int mat_upload_csr(DevMatrix& m, uint32_t n, const uint32_t* rp, const uint32_t* col,
                   const double* val, bool values_only, hipStream_t s) {
    if (!values_only) {
        mat_free(m);
        m.n = n;
        m.blocked = csr_blockable(n, rp, col);
        NBX_HIP(hipMalloc(&m.diag, (size_t)(n ? n : 1) * sizeof(double)));
    }
    if (m.blocked) {
        const uint32_t nn = n / 2;
        const uint32_t ns = (nn + SLICE - 1) / SLICE;
        std::vector<uint32_t> sptr(ns + 1, 0);
        std::vector<uint8_t> len(nn);
        uint32_t maxlen = 0;
        for (uint32_t v = 0; v < nn; ++v) {
            len[v] = (uint8_t)((rp[2 * v + 1] - rp[2 * v]) / 2);
            if (len[v] > maxlen) maxlen = len[v];
        }
        for (uint32_t sl = 0; sl < ns; ++sl) {
            uint32_t wdt = 0;
            for (uint32_t v = sl * SLICE; v < nn && v < (sl + 1) * SLICE; ++v)
                if (len[v] > wdt) wdt = len[v];
            sptr[sl + 1] = sptr[sl] + wdt * SLICE;
        }
        const size_t nb = sptr[ns];
        std::vector<double> hv(4 * nb + 4, 0.0);
        sell2_values_host(nn, rp, val, sptr, hv);
        if (!values_only) {
            std::vector<uint32_t> hc(nb + 1);
            for (uint32_t v = 0; v < nn; ++v) {
                const uint32_t a = rp[2 * v];
                size_t idx = (size_t)sptr[v >> 6] + (v & 63);
                const uint32_t wdt = (sptr[(v >> 6) + 1] - sptr[v >> 6]) / SLICE;
                for (uint32_t k = 0; k < wdt; ++k, idx += SLICE)
                    hc[idx] = k < len[v] ? col[a + 2 * k] / 2 : v;
            }
            Sell2& S = m.s;
            S.n_nodes = nn;
            S.n_slices = ns;
            S.n_blocks = nb;
            S.n_valid = rp[n] / 4;
            S.max_rowlen = maxlen;
            NBX_HIP(hipMalloc(&S.slice_ptr, (ns + 1) * sizeof(uint32_t)));
            NBX_HIP(hipMalloc(&S.rowlen, (nn + 1) * sizeof(uint8_t)));
            NBX_HIP(hipMalloc(&S.bcol, (nb + 1) * sizeof(uint32_t)));
            NBX_HIP(hipMalloc(&S.val, (4 * nb + 4) * sizeof(double)));
            NBX_HIP(hipMemcpyAsync(S.slice_ptr, sptr.data(), (ns + 1) * sizeof(uint32_t),
                                   hipMemcpyHostToDevice, s));
            NBX_HIP(hipMemcpyAsync(S.rowlen, len.data(), nn, hipMemcpyHostToDevice, s));
            NBX_HIP(hipMemcpyAsync(S.bcol, hc.data(), (nb + 1) * sizeof(uint32_t),
                                   hipMemcpyHostToDevice, s));
            NBX_HIP(hipMemcpyAsync(S.val, hv.data(), (4 * nb + 4) * sizeof(double),
                                   hipMemcpyHostToDevice, s));
            NBX_HIP(hipStreamSynchronize(s));
        } else {
            NBX_HIP(hipMemcpyAsync(m.s.val, hv.data(), (4 * nb + 4) * sizeof(double),
                                   hipMemcpyHostToDevice, s));
            NBX_HIP(hipStreamSynchronize(s));
        }
        m.has_val32 = false;
    } else {
        DevCsr& Cm = m.c;
        const size_t nnz = n ? rp[n] : 0;
        if (!values_only) {
            Cm.n = n;
            Cm.nnz = nnz;
            NBX_HIP(hipMalloc(&Cm.rowptr, (size_t)(n + 1) * sizeof(uint32_t)));
            NBX_HIP(hipMalloc(&Cm.col, (nnz + 1) * sizeof(uint32_t)));
            NBX_HIP(hipMalloc(&Cm.val, (nnz + 1) * sizeof(double)));
            NBX_HIP(hipMemcpyAsync(Cm.rowptr, rp, (size_t)(n + 1) * sizeof(uint32_t),
                                   hipMemcpyHostToDevice, s));
            if (nnz)
                NBX_HIP(hipMemcpyAsync(Cm.col, col, nnz * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        }
        if (nnz) NBX_HIP(hipMemcpyAsync(Cm.val, val, nnz * sizeof(double), hipMemcpyHostToDevice, s));
        NBX_HIP(hipStreamSynchronize(s));
    }
    return mat_extract_diag(m, s);
}

int mat_extract_diag(DevMatrix& m, hipStream_t s) {
    if (m.n == 0) return NBX_OK;
    if (m.blocked)
        hipLaunchKernelGGL(k_diag_sell2, dim3(grid_for(m.s.n_nodes)), dim3(BLOCK), 0, s, m.s, m.diag);
    else
        hipLaunchKernelGGL(k_diag_csr, dim3(grid_for(m.c.n)), dim3(BLOCK), 0, s, m.c, m.diag);
    NBX_HIP(hipGetLastError());
    m.has_val32 = false;
    return NBX_OK;
}

int mat_make_scaled32(DevMatrix& m, hipStream_t s) {
    if (!m.blocked) {
        set_last_error("mixed precision needs a 2-dof node matrix");
        return NBX_ERR_ARG;
    }
    if (!m.dis) NBX_HIP(hipMalloc(&m.dis, (size_t)m.n * sizeof(double)));
    if (!m.s.val32) NBX_HIP(hipMalloc(&m.s.val32, (4 * m.s.n_blocks + 4) * sizeof(float)));
    unsigned* bad = nullptr;
    NBX_HIP(hipMalloc(&bad, sizeof(unsigned)));
    NBX_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_dis, dim3(grid_for(m.n)), dim3(BLOCK), 0, s, m.diag, m.dis, m.n, bad);
    hipLaunchKernelGGL(k_scale32, dim3(grid_for(m.s.n_nodes)), dim3(BLOCK), 0, s, m.s, m.dis);
    NBX_HIP(hipGetLastError());
    unsigned hbad = 0;
    NBX_HIP(hipMemcpyAsync(&hbad, bad, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    (void)hipFree(bad);
    if (hbad) {
        set_last_error("matrix diagonal has %u non-positive entries", hbad);
        return NBX_ERR_ARG;
    }
    m.has_val32 = true;
    return NBX_OK;
}

int spmv(const DevMatrix& m, const double* x, double* y, hipStream_t s) {
    if (m.n == 0) return NBX_OK;
    if (m.blocked) {
        OpS2 op{m.s};
        hipLaunchKernelGGL(k_spmv<OpS2>, dim3(map_grid(m.s.n_nodes)), dim3(BLOCK), 0, s, op,
                           (const double2*)x, (double2*)y, m.s.n_nodes);
    } else {
        OpCsr op{m.c};
        hipLaunchKernelGGL(k_spmv<OpCsr>, dim3(map_grid(m.c.n)), dim3(BLOCK), 0, s, op, x, y, m.c.n);
    }
    NBX_HIP(hipGetLastError());
    return NBX_OK;
}

// Average duration of `reps` back-to-back SpMV launches (ms), FP64 or the
// scaled FP32 copy, on x = 1.
int spmv_bench(const DevMatrix& m, int prec, uint32_t reps, Workspace& w, hipStream_t s,
               double* ms_avg) {
    if (!m.blocked || m.n == 0) {
        set_last_error("spmv_bench needs a 2-dof node matrix");
        return NBX_ERR_ARG;
    }
    NBX_TRY(ws_alloc(w, m.n));
    const uint32_t nu = m.s.n_nodes;
    std::vector<double> ones(m.n, 1.0);
    std::vector<float> onesf(m.n, 1.0f);
    NBX_HIP(hipMemcpyAsync(w.v64[0], ones.data(), m.n * 8, hipMemcpyHostToDevice, s));
    NBX_HIP(hipMemcpyAsync(w.v32[0], onesf.data(), m.n * 4, hipMemcpyHostToDevice, s));
    const unsigned G = map_grid(nu);
    auto launch = [&]() {
        if (prec == 32)
            hipLaunchKernelGGL(k_spmv32, dim3(G), dim3(BLOCK), 0, s, m.s, (const float2*)w.v32[0],
                               (float2*)w.v32[1], nu);
        else
            hipLaunchKernelGGL(k_spmv<OpS2>, dim3(G), dim3(BLOCK), 0, s, OpS2{m.s},
                               (const double2*)w.v64[0], (double2*)w.v64[1], nu);
    };
    launch();   // warm-up
    NBX_HIP(hipEventRecord(w.ev[2], s));
    for (uint32_t r = 0; r < reps; ++r) launch();
    NBX_HIP(hipEventRecord(w.ev[3], s));
    NBX_HIP(hipEventSynchronize(w.ev[3]));
    NBX_HIP(hipGetLastError());
    float ms = 0.f;
    NBX_HIP(hipEventElapsedTime(&ms, w.ev[2], w.ev[3]));
    *ms_avg = reps ? ms / reps : 0.0;
    return NBX_OK;
}

double device_norm2(uint32_t n, const double* v, Workspace& w, hipStream_t s, int* err) {
    *err = NBX_OK;
    double* out = w.v64[5] + w.n;   // spare slot past the vector end
    unsigned g = red_grid(n);
    hipLaunchKernelGGL(k_norm2, dim3(g), dim3(BLOCK), 0, s, v, (size_t)n, w.partials, w.ticket, out);
    double h = 0.0;
    if (hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        *err = NBX_ERR_HIP;
        return 0.0;
    }
    return h;
}

static int poll_state(Workspace& w, hipStream_t s, int slot = 0) {
    NBX_HIP(hipMemcpyAsync(w.h_st, w.st, 2 * sizeof(PcgState), hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    (void)slot;
    return NBX_OK;
}

template <class Op>
static int pcg_run(const Op& op, uint32_t nu, const DevMatrix& m, const typename Op::V* b,
                   typename Op::V* x, uint32_t max_iter, double tol, Workspace& w,
                   hipStream_t s, SolveStats* stats) {
    using V = typename Op::V;
    V* g = (V*)w.v64[0];
    V* p = (V*)w.v64[1];
    V* wv = (V*)w.v64[2];
    const V* diag = (const V*)m.diag;
    PcgState h{};
    h.tol2 = tol * tol;
    h.max_iter = max_iter;
    NBX_HIP(hipMemcpyAsync(w.st, &h, sizeof h, hipMemcpyHostToDevice, s));
    const unsigned G = red_grid(nu);
    NBX_HIP(hipEventRecord(w.ev[0], s));
    hipLaunchKernelGGL(k_pcg_init<Op>, dim3(G), dim3(BLOCK), 0, s, op, b, x, g, p, diag, w.st,
                       w.partials, w.ticket, nu);
    NBX_HIP(hipGetLastError());
    uint32_t launched = 0, batch = 4, nsamp = 0;
    for (;;) {
        for (uint32_t i = 0; i < batch; ++i) {
            const bool samp = launched == 0 && nsamp < (uint32_t)Workspace::NSAMP;
            if (samp) NBX_HIP(hipEventRecord(w.sev[2 * nsamp], s));
            hipLaunchKernelGGL(k_pcg_A<Op>, dim3(G), dim3(BLOCK), 0, s, op, p, wv, w.st, w.partials,
                               w.ticket, nu);
            if (samp) NBX_HIP(hipEventRecord(w.sev[2 * nsamp++ + 1], s));
            hipLaunchKernelGGL(k_pcg_B<V>, dim3(G), dim3(BLOCK), 0, s, x, g, p, wv, diag, w.st,
                               w.partials, w.ticket, nu);
            hipLaunchKernelGGL(k_pcg_C<V>, dim3(G), dim3(BLOCK), 0, s, p, g, diag, w.st, nu);
        }
        NBX_HIP(hipGetLastError());
        launched += batch;
        NBX_TRY(poll_state(w, s));
        if (w.h_st[0].done) break;
        if (launched > max_iter + 8u) {
            set_last_error("PCG did not report completion after %u iterations", launched);
            return NBX_ERR_BREAKDOWN;
        }
        batch = batch < 64 ? batch * 2 : 64;
    }
    NBX_HIP(hipEventRecord(w.ev[1], s));
    NBX_HIP(hipEventSynchronize(w.ev[1]));
    if (stats) {
        float ms = 0.f;
        NBX_HIP(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
        stats->ms = ms;
        stats->iters = w.h_st[0].k;
        stats->fp64_iters = w.h_st[0].k;
        // only samples of iterations that ran (k > sample index) count
        double tot = 0.0;
        uint32_t cnt = 0;
        for (uint32_t i = 0; i < nsamp && i < w.h_st[0].k; ++i) {
            float t = 0.f;
            NBX_HIP(hipEventElapsedTime(&t, w.sev[2 * i], w.sev[2 * i + 1]));
            tot += t;
            ++cnt;
        }
        stats->ms_spmv_avg = cnt ? tot / cnt : 0.0;
    }
    return NBX_OK;
}

int pcg_jacobi(const DevMatrix& m, const double* b, double* x, uint32_t max_iter, double tol,
               uint32_t* niter, double* tol_reached, Workspace& w, hipStream_t s,
               SolveStats* stats) {
    if (m.n == 0) {
        // empty system: no iteration, ||g|| = 0 (cg_precond_jacobi.c:44-89)
        if (niter) *niter = 0;
        if (tol_reached) *tol_reached = 0.0;
        return 0;
    }
    int rc = ws_alloc(w, m.n);
    if (rc) return -rc;
    if (m.blocked)
        rc = pcg_run(OpS2{m.s}, m.s.n_nodes, m, (const double2*)b, (double2*)x, max_iter, tol, w,
                     s, stats);
    else
        rc = pcg_run(OpCsr{m.c}, m.c.n, m, b, x, max_iter, tol, w, s, stats);
    if (rc) return -rc;
    if (niter) *niter = w.h_st[0].k;
    if (tol_reached) *tol_reached = w.h_st[0].tol_reached;
    return (int)w.h_st[0].status;
}

int solve_fp64(const DevMatrix& m, const double* b, double* x, double rtol, uint32_t max_iter,
               Workspace& w, hipStream_t s, SolveStats* stats) {
    int rc = ws_alloc(w, m.n);
    if (rc) return rc;
    const double bb = device_norm2(m.n, b, w, s, &rc);
    if (rc) return rc;
    uint32_t it = 0;
    double tr = 0.0;
    int st = pcg_jacobi(m, b, x, max_iter, rtol * sqrt(bb), &it, &tr, w, s, stats);
    if (st < 0) return -st;
    if (stats) stats->rel_res = bb > 0 ? tr / sqrt(bb) : 0.0;
    return st == 0 ? NBX_OK : NBX_NOT_CONVERGED;
}

int solve_mixed(const DevMatrix& m, const double* b, double* x, double rtol, double inner_rtol,
                uint32_t max_iter, Workspace& w, hipStream_t s, SolveStats* stats) {
    if (!m.blocked || !m.has_val32) {
        set_last_error("solve_mixed: matrix has no scaled FP32 copy");
        return NBX_ERR_ARG;
    }
    int rc = ws_alloc(w, m.n);
    if (rc) return rc;
    const uint32_t nu = m.s.n_nodes;
    const double bb = device_norm2(m.n, b, w, s, &rc);
    if (rc) return rc;
    PcgState h[2] = {};
    const uint32_t max_outer = 100;
    h[0].tol2 = rtol * rtol * bb;
    h[0].max_iter = max_outer;
    h[1].max_iter = max_iter;
    NBX_HIP(hipMemcpyAsync(w.st, h, sizeof h, hipMemcpyHostToDevice, s));
    float2* r32 = (float2*)w.v32[0];
    float2* p32 = (float2*)w.v32[1];
    float2* w32 = (float2*)w.v32[2];
    float2* y32 = (float2*)w.v32[3];
    const unsigned G = red_grid(nu);
    OpS2 op{m.s};
    uint32_t inner_total = 0, outer = 0, nsamp = 0, samp_valid = 0;
    double prev_res = -1.0;
    int stalled = 0;
    NBX_HIP(hipEventRecord(w.ev[0], s));
    for (;;) {
        hipLaunchKernelGGL(k_ir_residual, dim3(G), dim3(BLOCK), 0, s, op, (const double2*)b,
                           (const double2*)x, (const double2*)m.dis, r32, p32, y32, w.st,
                           w.partials, w.ticket, nu, inner_rtol * inner_rtol);
        NBX_HIP(hipGetLastError());
        NBX_TRY(poll_state(w, s));
        if (w.h_st[0].done) break;
        const double res = w.h_st[0].tol_reached;
        if (prev_res > 0.0 && res > 0.5 * prev_res) {
            // an FP32 sweep that no longer reduces the FP64 residual: finish
            // with FP64 PCG from the current iterate
            stalled = 1;
            break;
        }
        prev_res = res;
        uint32_t batch = 8;
        bool first_batch = true;
        for (;;) {
            for (uint32_t i = 0; i < batch; ++i) {
                const bool samp = outer == 0 && first_batch && nsamp < (uint32_t)Workspace::NSAMP;
                if (samp) NBX_HIP(hipEventRecord(w.sev[2 * nsamp], s));
                hipLaunchKernelGGL(k_in_A, dim3(G), dim3(BLOCK), 0, s, m.s, p32, w32, w.st,
                                   w.partials32, w.ticket, nu);
                if (samp) NBX_HIP(hipEventRecord(w.sev[2 * nsamp++ + 1], s));
                hipLaunchKernelGGL(k_in_B, dim3(G), dim3(BLOCK), 0, s, y32, r32, p32, w32, w.st,
                                   w.partials32, w.ticket, nu);
                hipLaunchKernelGGL(k_in_C, dim3(G), dim3(BLOCK), 0, s, p32, r32, w.st, nu);
            }
            NBX_HIP(hipGetLastError());
            NBX_TRY(poll_state(w, s));
            if (outer == 0 && first_batch) samp_valid = w.h_st[1].k;
            first_batch = false;
            if (w.h_st[1].done) break;
            batch = batch < 64 ? batch * 2 : 64;
        }
        inner_total += w.h_st[1].k;
        hipLaunchKernelGGL(k_ir_update, dim3(G), dim3(BLOCK), 0, s, (double2*)x,
                           (const double2*)m.dis, y32, w.st, nu);
        NBX_HIP(hipGetLastError());
        ++outer;
        if (w.h_st[1].breakdown) {
            stalled = 1;
            break;
        }
    }
    NBX_HIP(hipEventRecord(w.ev[1], s));
    NBX_HIP(hipEventSynchronize(w.ev[1]));
    float ms = 0.f;
    NBX_HIP(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
    SolveStats fb{};
    int status = w.h_st[0].status ? NBX_NOT_CONVERGED : NBX_OK;
    double rel = bb > 0 ? w.h_st[0].tol_reached / sqrt(bb) : 0.0;
    if (stalled) {
        rc = solve_fp64(m, b, x, rtol, max_iter, w, s, &fb);
        if (rc != NBX_OK && rc != NBX_NOT_CONVERGED) return rc;
        status = rc;
        rel = fb.rel_res;
        ms += (float)fb.ms;
    }
    if (stats) {
        double tot = 0.0;
        uint32_t cnt = 0;
        for (uint32_t i = 0; i < nsamp && i < samp_valid; ++i) {
            float t = 0.f;
            NBX_HIP(hipEventElapsedTime(&t, w.sev[2 * i], w.sev[2 * i + 1]));
            tot += t;
            ++cnt;
        }
        stats->ms_spmv_avg = cnt ? tot / cnt : 0.0;
        stats->iters = inner_total;
        stats->outer = outer;
        stats->fp64_iters = fb.fp64_iters;
        stats->rel_res = rel;
        stats->ms = ms;
        stats->fallback = stalled;
    }
    return status;
}

}  // namespace nbx

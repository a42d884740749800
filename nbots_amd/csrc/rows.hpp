// Device primitives shared by the SpMV and CG kernels: row products of the
// 2x2-block sliced-ELL and scalar-CSR formats, per-unit vector arithmetic,
// and the XCD-aware work split.
#pragma once

#include "internal.hpp"

namespace nbx {

typedef double dv2 __attribute__((ext_vector_type(2)));
typedef float fv4 __attribute__((ext_vector_type(4)));
typedef _Float16 hv4 __attribute__((ext_vector_type(4)));

// The matrix stream is read once per product: non-temporal loads keep it
// from evicting the gathered vector out of L2.
template <class T>
__device__ __forceinline__ T nt(const T* p) {
    return __builtin_nontemporal_load(p);
}

// ---- per-unit vector arithmetic (unit = node with 2 dofs, or scalar row) --
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

// ---- 2x2-block sliced-ELL rows --------------------------------------------
// Node row v: the two scalar rows 2v, 2v+1, each summed in ascending column
// order (block by block, column 2w then 2w+1), as nb_sparse_multiply_vector
// sums a row (solver_bot/sparse/sparse.c:405-414).  Blocks are taken eight
// at a time with every independent load (column, values) issued before the
// dependent gathers of x, so a wavefront keeps many requests in flight.
constexpr int UNROLL = 8;

__device__ __forceinline__ double2 row64(const Sell2& A, uint32_t v, const double2* __restrict__ x) {
    const size_t base = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const int len = A.rowlen[v];
    const dv2* __restrict__ pa = reinterpret_cast<const dv2*>(A.val);
    const dv2* __restrict__ pb = reinterpret_cast<const dv2*>(A.val + 2 * A.n_blocks);
    double s0 = 0.0, s1 = 0.0;
    for (int k0 = 0; k0 < len; k0 += UNROLL) {
        uint32_t w[UNROLL];
        dv2 a[UNROLL], b[UNROLL];
        double2 xv[UNROLL];
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
            if (k0 + k < len) {
                const size_t i = base + (size_t)(k0 + k) * SLICE;
                w[k] = nt(&A.bcol[i]);
                a[k] = nt(&pa[i]);
                b[k] = nt(&pb[i]);
            }
        }
#pragma unroll
        for (int k = 0; k < UNROLL; ++k)
            if (k0 + k < len) xv[k] = x[w[k]];
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
            if (k0 + k < len) {
                s0 = s0 + a[k].x * xv[k].x;
                s0 = s0 + a[k].y * xv[k].y;
                s1 = s1 + b[k].x * xv[k].x;
                s1 = s1 + b[k].y * xv[k].y;
            }
        }
    }
    return make_double2(s0, s1);
}

// FP32 copy (Jacobi-scaled values, float2 vectors).
__device__ __forceinline__ float2 row32(const Sell2& A, uint32_t v, const float2* __restrict__ x) {
    const size_t base = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const int len = A.rowlen[v];
    const fv4* __restrict__ val = reinterpret_cast<const fv4*>(A.val32);
    float s0 = 0.f, s1 = 0.f;
    for (int k0 = 0; k0 < len; k0 += UNROLL) {
        uint32_t w[UNROLL];
        fv4 a[UNROLL];
        float2 xv[UNROLL];
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
            if (k0 + k < len) {
                const size_t i = base + (size_t)(k0 + k) * SLICE;
                w[k] = nt(&A.bcol[i]);
                a[k] = nt(&val[i]);
            }
        }
#pragma unroll
        for (int k = 0; k < UNROLL; ++k)
            if (k0 + k < len) xv[k] = x[w[k]];
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
            if (k0 + k < len) {
                s0 = s0 + a[k].x * xv[k].x;
                s0 = s0 + a[k].y * xv[k].y;
                s1 = s1 + a[k].z * xv[k].x;
                s1 = s1 + a[k].w * xv[k].y;
            }
        }
    }
    return make_float2(s0, s1);
}

// fp16 copy (Jacobi-scaled values widened to FP32, float2 vectors): 12 B
// per block instead of 20 B.
__device__ __forceinline__ float2 row16(const Sell2& A, uint32_t v, const float2* __restrict__ x) {
    const size_t base = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const int len = A.rowlen[v];
    const hv4* __restrict__ val = reinterpret_cast<const hv4*>(A.val16);
    float s0 = 0.f, s1 = 0.f;
    for (int k0 = 0; k0 < len; k0 += UNROLL) {
        uint32_t w[UNROLL];
        hv4 a[UNROLL];
        float2 xv[UNROLL];
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
            if (k0 + k < len) {
                const size_t i = base + (size_t)(k0 + k) * SLICE;
                w[k] = nt(&A.bcol[i]);
                a[k] = nt(&val[i]);
            }
        }
#pragma unroll
        for (int k = 0; k < UNROLL; ++k)
            if (k0 + k < len) xv[k] = x[w[k]];
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
            if (k0 + k < len) {
                s0 = s0 + (float)a[k].x * xv[k].x;
                s0 = s0 + (float)a[k].y * xv[k].y;
                s1 = s1 + (float)a[k].z * xv[k].x;
                s1 = s1 + (float)a[k].w * xv[k].y;
            }
        }
    }
    return make_float2(s0, s1);
}

// y = A x with the scaled copy of BITS-bit values
template <int BITS>
__device__ __forceinline__ float2 row_scaled(const Sell2& A, uint32_t v, const float2* __restrict__ x) {
    if constexpr (BITS == 16)
        return row16(A, v, x);
    else
        return row32(A, v, x);
}

struct OpS2 {
    Sell2 A;
    using V = double2;
    __device__ __forceinline__ double2 row(uint32_t v, const double2* __restrict__ x) const {
        return row64(A, v, x);
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

// ---- work split -----------------------------------------------------------
// Workgroup b (XCD-remapped) owns one contiguous chunk of units, a multiple
// of 256 = four 64-row slices.
struct Chunk {
    size_t beg, end;
};
__device__ __forceinline__ Chunk my_chunk(uint32_t n) {
    const unsigned G = gridDim.x;
    const unsigned lb = xcd_remap(blockIdx.x, G);
    const size_t per = ((size_t)n + (size_t)G * BLOCK - 1) / ((size_t)G * BLOCK) * BLOCK;
    size_t b = (size_t)lb * per;
    size_t e = b + per;
    if (b > n) b = n;
    if (e > n) e = n;
    return {b, e};
}

inline unsigned red_grid(uint32_t n_units) {
    const unsigned g = grid_for(n_units);
    return g > (unsigned)MAX_GRID ? (unsigned)MAX_GRID : g;
}

}  // namespace nbx

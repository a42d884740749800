// Device primitives shared by the SpMV and CG kernels: row products of the
// 2x2-block sliced-ELL and scalar-CSR formats, per-unit vector arithmetic,
// and the XCD-aware work split.
#pragma once

#include "internal.hpp"

namespace nbx {

typedef double dv2 __attribute__((ext_vector_type(2)));
typedef float fv4 __attribute__((ext_vector_type(4)));
typedef float fv2 __attribute__((ext_vector_type(2)));
typedef _Float16 hv4 __attribute__((ext_vector_type(4)));
typedef _Float16 hv2 __attribute__((ext_vector_type(2)));
typedef uint32_t uv2 __attribute__((ext_vector_type(2)));

// The matrix stream is read once per product: non-temporal loads keep it
// from evicting the gathered vector out of L2.
template <class T>
__device__ __forceinline__ T nt(const T* p) {
    return __builtin_nontemporal_load(p);
}

// Vector streams that the next kernels do not gather (PCG updates, the
// Chebyshev recurrences) use non-temporal loads and stores as well: they
// then leave the L2 and the Infinity Cache to the matrix and to the
// gathered vector.
__device__ __forceinline__ double ldn(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double2 ldn(const double2* p) {
    const dv2 v = __builtin_nontemporal_load(reinterpret_cast<const dv2*>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void stn(double* p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void stn(double2* p, double2 v) {
    dv2 t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<dv2*>(p));
}
__device__ __forceinline__ float2 ldn(const float2* p) {
    const fv2 v = __builtin_nontemporal_load(reinterpret_cast<const fv2*>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void stn(float2* p, float2 v) {
    fv2 t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<fv2*>(p));
}

template <bool NT, class V>
__device__ __forceinline__ V ldv(const V* p) {
    if constexpr (NT)
        return ldn(p);
    else
        return *p;
}
template <bool NT, class V>
__device__ __forceinline__ void stv(V* p, V v) {
    if constexpr (NT)
        stn(p, v);
    else
        *p = v;
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
// blocks per batch of a row without its diagonal block (6 covers an interior
// row of a T3 mesh; 8 would not fit in the registers of 8 waves per SIMD)
constexpr int UNROLL_OD = 6;

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

// Symmetric layout: the lower blocks (w < v, ascending), each the transpose
// of the stored upper block (w, v), the diagonal block (three stored values,
// read with x_v in whole-wave coalesced loads issued first), then the upper
// blocks (w > v, ascending); the same products in the same order as row64.
// The transposed blocks belong to rows that nearby workgroups of the same
// XCD stream at about the same time, so their loads are served by L2.
// Lower and upper blocks are taken UL and UU at a time; few registers per
// row keep many waves resident, which hides the dependent loads better than
// longer batches.  An interior row of a T3 mesh has 3 lower and 3 upper
// blocks: the streamed product takes them 3 and 3 (70 VGPRs, 7 waves per
// SIMD), the cached one 4 and 4 (84 VGPRs, 5 waves), which measured faster
// where the matrix is read from cache.  Issuing the diagonal loads after the
// lower blocks instead saves registers but exposes their latency.
// NT: non-temporal loads of the matrix stream (when the working set exceeds
// the Infinity Cache); otherwise ordinary loads, so that the matrix stays
// cached from one product to the next.
template <bool NT, class T>
__device__ __forceinline__ T ldm(const T* p) {
    if constexpr (NT)
        return nt(p);
    else
        return *p;
}
template <int UL, int UU, bool NT = true>
__device__ __forceinline__ double2 row64s(const Sym2& Y, uint32_t v, const double2* __restrict__ x) {
    const dv2* __restrict__ pa = reinterpret_cast<const dv2*>(Y.u_val);
    const dv2* __restrict__ pb = reinterpret_cast<const dv2*>(Y.u_val + 2 * Y.n_upper);
    const dv2 d01 = ldm<NT>(reinterpret_cast<const dv2*>(Y.d_val) + v);
    const double d11 = ldm<NT>(Y.d_val + 2 * (size_t)Y.n_nodes + v);
    const double2 xd = x[v];
    double s0 = 0.0, s1 = 0.0;
    {
        const uv2* __restrict__ li = reinterpret_cast<const uv2*>(Y.l_idx);
        const size_t base = (size_t)Y.l_slice_ptr[v >> 6] + (v & 63);
        const int len = Y.l_rowlen[v];
        for (int k0 = 0; k0 < len; k0 += UL) {
            uv2 e[UL];
            dv2 a[UL], b[UL];
            double2 xv[UL];
#pragma unroll
            for (int k = 0; k < UL; ++k)
                if (k0 + k < len) e[k] = ldm<NT>(&li[base + (size_t)(k0 + k) * SLICE]);
#pragma unroll
            for (int k = 0; k < UL; ++k) {
                if (k0 + k < len) {
                    a[k] = pa[e[k].y];
                    b[k] = pb[e[k].y];
                    xv[k] = x[e[k].x];
                }
            }
            // block (v, w) = [[a.x, b.x], [a.y, b.y]]
#pragma unroll
            for (int k = 0; k < UL; ++k) {
                if (k0 + k < len) {
                    s0 = s0 + a[k].x * xv[k].x;
                    s0 = s0 + b[k].x * xv[k].y;
                    s1 = s1 + a[k].y * xv[k].x;
                    s1 = s1 + b[k].y * xv[k].y;
                }
            }
        }
    }
    // block (v, v) = [[d01.x, d01.y], [d01.y, d11]]
    s0 = s0 + d01.x * xd.x;
    s0 = s0 + d01.y * xd.y;
    s1 = s1 + d01.y * xd.x;
    s1 = s1 + d11 * xd.y;
    {
        const size_t base = (size_t)Y.u_slice_ptr[v >> 6] + (v & 63);
        const int len = Y.u_rowlen[v];
        for (int k0 = 0; k0 < len; k0 += UU) {
            uint32_t w[UU];
            dv2 a[UU], b[UU];
            double2 xv[UU];
#pragma unroll
            for (int k = 0; k < UU; ++k) {
                if (k0 + k < len) {
                    const size_t i = base + (size_t)(k0 + k) * SLICE;
                    w[k] = ldm<NT>(&Y.u_bcol[i]);
                    a[k] = ldm<NT>(&pa[i]);
                    b[k] = ldm<NT>(&pb[i]);
                }
            }
#pragma unroll
            for (int k = 0; k < UU; ++k)
                if (k0 + k < len) xv[k] = x[w[k]];
#pragma unroll
            for (int k = 0; k < UU; ++k) {
                if (k0 + k < len) {
                    s0 = s0 + a[k].x * xv[k].x;
                    s0 = s0 + a[k].y * xv[k].y;
                    s1 = s1 + b[k].x * xv[k].x;
                    s1 = s1 + b[k].y * xv[k].y;
                }
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

// The same row with FP64 accumulation: the fp16 values and FP32 vector
// entries are exact in FP64, so each product is exact and only the sum
// rounds; stiffness rows cancel heavily, and an FP32 sum loses most of a
// small result.
template <bool NT = true>
__device__ __forceinline__ double2 row16d(const Sell2& A, uint32_t v, const float2* __restrict__ x) {
    const size_t base = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const int len = A.rowlen[v];
    const hv4* __restrict__ val = reinterpret_cast<const hv4*>(A.val16);
    double s0 = 0.0, s1 = 0.0;
    for (int k0 = 0; k0 < len; k0 += UNROLL) {
        uint32_t w[UNROLL];
        hv4 a[UNROLL];
        float2 xv[UNROLL];
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
            if (k0 + k < len) {
                const size_t i = base + (size_t)(k0 + k) * SLICE;
                if constexpr (NT) {
                    w[k] = nt(&A.bcol[i]);
                    a[k] = nt(&val[i]);
                } else {
                    w[k] = A.bcol[i];
                    a[k] = val[i];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < UNROLL; ++k)
            if (k0 + k < len) xv[k] = x[w[k]];
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
            if (k0 + k < len) {
                s0 = s0 + (double)a[k].x * (double)xv[k].x;
                s0 = s0 + (double)a[k].y * (double)xv[k].y;
                s1 = s1 + (double)a[k].z * (double)xv[k].x;
                s1 = s1 + (double)a[k].w * (double)xv[k].y;
            }
        }
    }
    return make_double2(s0, s1);
}

// The same row on the copy without diagonal blocks: the diagonal block
// [[1, od.x], [od.y, 1]] times the row's own entry xo (which the caller has
// loaded), then the off-diagonal blocks.
// P: read the columns as packed 16-bit deltas (A.pcol), one 32-bit load per
// two blocks, instead of A.bcol.
// U: the uniform-width layout (A.width > 0), whose rows run over the full
// width, padding included (value 0, column v: adds 0).
template <bool NT, bool P, bool U = false>
__device__ __forceinline__ double2 row16od(const Off16& A, uint32_t v, const float2* __restrict__ x,
                                           float2 xo) {
    const size_t base = U ? (size_t)(v >> 6) * SLICE * A.width + (v & 63)
                          : (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const int len = U ? (int)A.width : (int)A.rowlen[v];
    const hv4* __restrict__ val = reinterpret_cast<const hv4*>(A.val);
    hv2 od;
    if constexpr (NT)
        od = nt(reinterpret_cast<const hv2*>(A.od) + v);
    else
        od = reinterpret_cast<const hv2*>(A.od)[v];
    double s0 = (double)xo.x + (double)od.x * (double)xo.y;
    double s1 = (double)od.y * (double)xo.x + (double)xo.y;
    for (int k0 = 0; k0 < len; k0 += UNROLL_OD) {
        uint32_t w[UNROLL_OD];
        hv4 a[UNROLL_OD];
        float2 xv[UNROLL_OD];
        if constexpr (P) {
            const size_t pb = (U ? (size_t)(v >> 6) * SLICE * ((A.width + 1) / 2)
                                 : (size_t)A.pslice_ptr[v >> 6]) +
                              (v & 63) + (size_t)(k0 / 2) * SLICE;
            uint32_t pw[UNROLL_OD / 2];
#pragma unroll
            for (int j = 0; j < UNROLL_OD / 2; ++j)
                if (k0 + 2 * j < len) pw[j] = NT ? nt(&A.pcol[pb + (size_t)j * SLICE])
                                                 : A.pcol[pb + (size_t)j * SLICE];
#pragma unroll
            for (int k = 0; k < UNROLL_OD; ++k)
                if (k0 + k < len)
                    w[k] = v + (uint32_t)(int32_t)(int16_t)(uint16_t)(pw[k / 2] >> (16 * (k & 1)));
        }
#pragma unroll
        for (int k = 0; k < UNROLL_OD; ++k) {
            if (k0 + k < len) {
                const size_t i = base + (size_t)(k0 + k) * SLICE;
                if constexpr (NT) {
                    if constexpr (!P) w[k] = nt(&A.bcol[i]);
                    a[k] = nt(&val[i]);
                } else {
                    if constexpr (!P) w[k] = A.bcol[i];
                    a[k] = val[i];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < UNROLL_OD; ++k)
            if (k0 + k < len) xv[k] = x[w[k]];
#pragma unroll
        for (int k = 0; k < UNROLL_OD; ++k) {
            if (k0 + k < len) {
                s0 = s0 + (double)a[k].x * (double)xv[k].x;
                s0 = s0 + (double)a[k].y * (double)xv[k].y;
                s1 = s1 + (double)a[k].z * (double)xv[k].x;
                s1 = s1 + (double)a[k].w * (double)xv[k].y;
            }
        }
    }
    return make_double2(s0, s1);
}

// the Chebyshev step's row on either fp16 layout
template <bool NT>
__device__ __forceinline__ double2 cheb_row(const Sell2& A, uint32_t v, const float2* __restrict__ x,
                                            float2) {
    return row16d<NT>(A, v, x);
}
template <bool NT>
__device__ __forceinline__ double2 cheb_row(const Off16& A, uint32_t v, const float2* __restrict__ x,
                                            float2 xo) {
    return row16od<NT, false>(A, v, x, xo);
}
// the copy without diagonal blocks read through its packed column deltas
struct Off16P {
    Off16 O;
};
template <bool NT>
__device__ __forceinline__ double2 cheb_row(const Off16P& A, uint32_t v, const float2* __restrict__ x,
                                            float2 xo) {
    return row16od<NT, true>(A.O, v, x, xo);
}
// ... in the uniform-width layout
struct Off16U {
    Off16 O;
};
template <bool NT>
__device__ __forceinline__ double2 cheb_row(const Off16U& A, uint32_t v, const float2* __restrict__ x,
                                            float2 xo) {
    return row16od<NT, true, true>(A.O, v, x, xo);
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

struct OpSym {
    Sym2 Y;
    using V = double2;
    __device__ __forceinline__ double2 row(uint32_t v, const double2* __restrict__ x) const {
        return row64s<3, 3>(Y, v, x);
    }
};
// the same with cached matrix loads (working set within the Infinity Cache)
struct OpSymC {
    Sym2 Y;
    using V = double2;
    __device__ __forceinline__ double2 row(uint32_t v, const double2* __restrict__ x) const {
        return row64s<4, 4, false>(Y, v, x);
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

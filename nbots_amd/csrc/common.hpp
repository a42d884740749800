// Shared definitions for the nbots MI355X (gfx950) solve path.
//
// Every translation unit of libnbots_amd.so is compiled with
// -ffp-contract=off: the reference's arithmetic has no fused multiply-add
// (its default build targets plain x86-64), and the assembly, SpMV, strain
// and BC kernels reproduce its results bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace nbx {

// ---- error handling --------------------------------------------------------
// New entry points never exit(); they return one of these codes.
enum Status : int {
    NBX_OK = 0,
    NBX_NOT_CONVERGED = 1,
    NBX_ERR_DISTORTED = 2,      // an element with detJ < 0
    NBX_ERR_HIP = 3,            // a HIP runtime call failed
    NBX_ERR_ARG = 4,            // invalid argument / size
    NBX_ERR_PATTERN = 5,        // a row has more blocks than the format holds
    NBX_ERR_BREAKDOWN = 6,      // non-finite or non-positive curvature in CG
};

void set_last_error(const char* fmt, ...);
const char* last_error();

#define NBX_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            ::nbx::set_last_error("%s:%d %s -> %s", __FILE__, __LINE__, #call, \
                                  hipGetErrorString(e_));                      \
            return ::nbx::NBX_ERR_HIP;                                         \
        }                                                                      \
    } while (0)

#define NBX_TRY(call)                 \
    do {                              \
        int rc_ = (call);             \
        if (rc_ != 0) return rc_;     \
    } while (0)

// ---- launch geometry -------------------------------------------------------
constexpr int WAVE = 64;          // CDNA wavefront
constexpr int BLOCK = 256;        // 4 waves per workgroup
constexpr int SLICE = 64;         // sliced-ELL slice height = one wavefront
constexpr int NUM_XCD = 8;        // MI355X: 8 XCDs, 32 CUs each
constexpr int MAX_GRID = 65535;   // workgroups of a reduction kernel (partials)
constexpr size_t INFINITY_CACHE_BYTES = 256ull << 20;   // MI355X last-level cache

// A copy ordered on a handle's stream, complete on return (a plain hipMemcpy
// would also wait for every other handle's work).
inline hipError_t copy_on(void* dst, const void* src, size_t n, hipMemcpyKind kind, hipStream_t s) {
    const hipError_t e = hipMemcpyAsync(dst, src, n, kind, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}

inline unsigned grid_for(size_t n, int per_block = BLOCK) {
    size_t g = (n + per_block - 1) / per_block;
    return (unsigned)(g == 0 ? 1 : g);
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs
// on XCD b % 8).  Remapping b so that each XCD walks one contiguous eighth
// of the rows keeps the neighbours' x entries of a sparse row in that XCD's
// own L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nb) {
    unsigned per = (nb + NUM_XCD - 1) / NUM_XCD;
    unsigned xcd = b % NUM_XCD, pos = b / NUM_XCD;
    unsigned full = nb % NUM_XCD;             // XCDs with `per` blocks
    if (full == 0) return xcd * per + pos;
    // the first `full` XCDs own `per` blocks, the rest `per-1`
    if (xcd < full) return xcd * per + pos;
    return full * per + (xcd - full) * (per - 1) + pos;
}

// ---- 2x2-block sliced ELL (node rows of a 2-dof matrix) -------------------
// Node row v holds rowlen[v] blocks; block k of row v is stored at
//   slice_ptr[v / 64] + k * 64 + (v % 64)
// so that the 64 lanes of a wavefront read block k of 64 consecutive rows as
// one contiguous run.  Block columns within a row ascend, exactly as the
// scalar rows 2v, 2v+1 of the reference's CSR (columns 2w, 2w+1 per block).
// FP64 values sit in two planes so that each wavefront load is one
// contiguous run: plane a holds {a(2v,2w), a(2v,2w+1)} of every stored
// block, plane b = val + 2*n_blocks holds {a(2v+1,2w), a(2v+1,2w+1)}.
// The FP32 copy keeps the four values of a block together (16 bytes).
struct Sell2 {
    uint32_t n_nodes = 0;
    uint32_t n_slices = 0;
    size_t n_blocks = 0;           // stored blocks, padding included
    size_t n_valid = 0;            // blocks in the pattern (nnz / 4)
    uint32_t max_rowlen = 0;
    uint32_t* slice_ptr = nullptr; // [n_slices + 1]
    uint8_t* rowlen = nullptr;     // [n_nodes]
    uint32_t* bcol = nullptr;      // [n_blocks]
    double* val = nullptr;         // [4 * n_blocks]
    float* val32 = nullptr;        // [4 * n_blocks], Jacobi-scaled FP32 copy
    _Float16* val16 = nullptr;     // [4 * n_blocks], Jacobi-scaled fp16 copy
};

// The fp16 Jacobi-scaled copy D^-1/2 K D^-1/2 without its diagonal blocks,
// for the Chebyshev preconditioner.  Each diagonal entry of the scaled
// matrix rounds to 1.0 in fp16 (the build checks it), so a diagonal block
// is [[1, od.x], [od.y, 1]] and costs 4 B per node instead of 12 B of
// column and values in the ELL; the off-diagonal blocks keep the Sell2 form.
struct Off16 {
    uint32_t n_nodes = 0;
    uint32_t n_slices = 0;
    size_t n_blocks = 0;             // stored off-diagonal blocks, padding included
    uint32_t* slice_ptr = nullptr;   // [n_slices + 1]
    uint8_t* rowlen = nullptr;       // [n_nodes]
    uint32_t* bcol = nullptr;        // [n_blocks]
    _Float16* val = nullptr;         // [4 * n_blocks], 8 B per block as Sell2::val16
    _Float16* od = nullptr;          // [2 * n_nodes]: (K01, K10) scaled, per node
    // the same columns as row-relative 16-bit deltas, two blocks per word:
    // pair j of row v (blocks 2j, 2j+1) at pslice_ptr[v/64] + v%64 + 64 j,
    // low half (w_2j - v), high half (w_2j+1 - v); built when every delta
    // fits in int16
    size_t n_pairs = 0;              // stored words, padding included
    uint32_t* pslice_ptr = nullptr;  // [n_slices + 1]
    uint32_t* pcol = nullptr;        // [n_pairs]
    // > 0: every slice is `width` blocks (and (width + 1) / 2 words) wide,
    // so row v's entries start at (v / 64) 64 width + v % 64 and
    // (v / 64) 64 ((width + 1) / 2) + v % 64; padding has value 0, column v
    uint32_t width = 0;
};

// Symmetric 2x2-block sliced ELL, for a node matrix whose blocks satisfy
// K(w, v) = K(v, w)^T bit for bit (what the T3 assembly produces).  Row v
// stores its diagonal block (v, v) as three values (it is symmetric itself),
// its blocks (v, w > v) in the Sell2 layout (upper part, ascending w), and
// its blocks (v, w < v) as pairs (w, position of the upper block (w, v))
// (lower part, ascending w), read transposed.  A row product visits the
// lower blocks, the diagonal block, then the upper blocks, i.e. every column
// in ascending order with the same products as the full layout, so results
// are identical; the matrix stream shrinks from 36 B per block to 24 B per
// diagonal, 36 B per upper and 8 B per lower block.
struct Sym2 {
    uint32_t n_nodes = 0;
    uint32_t n_slices = 0;
    size_t n_upper = 0;              // stored upper blocks, padding included
    size_t n_lower = 0;              // stored lower entries, padding included
    double* d_val = nullptr;         // diagonal blocks: [2 * n_nodes] (K00, K01) per
                                     // node, then [n_nodes] K11 (K10 == K01)
    uint32_t* u_slice_ptr = nullptr; // [n_slices + 1]
    uint8_t* u_rowlen = nullptr;     // [n_nodes]
    uint32_t* u_bcol = nullptr;      // [n_upper], block columns w > v
    double* u_val = nullptr;         // [4 * n_upper], two row planes as Sell2::val
    uint32_t* l_slice_ptr = nullptr; // [n_slices + 1]
    uint8_t* l_rowlen = nullptr;     // [n_nodes]
    uint32_t* l_idx = nullptr;       // [2 * n_lower]: (w, upper position of (w, v))
};

// index of entry (i, j) of stored block idx in Sell2::val
__host__ __device__ __forceinline__ size_t ent(size_t n_blocks, size_t idx, int i, int j) {
    return (i ? 2 * n_blocks : 0) + 2 * idx + j;
}

// Generic scalar CSR on the device (any vars_per_node).
struct DevCsr {
    uint32_t n = 0;
    size_t nnz = 0;
    uint32_t* rowptr = nullptr;
    uint32_t* col = nullptr;
    double* val = nullptr;
};

// ---- compensated FP32 accumulation (TwoSum, Knuth) -------------------------
struct f32c {
    float s, c;
};
__device__ __forceinline__ f32c two_sum(float a, float b) {
    float s = a + b;
    float bb = s - a;
    float e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
__device__ __forceinline__ f32c cadd(f32c x, float v) {
    f32c t = two_sum(x.s, v);
    return {t.s, x.c + t.c};
}
__device__ __forceinline__ f32c cadd(f32c x, f32c y) {
    f32c t = two_sum(x.s, y.s);
    return {t.s, (x.c + y.c) + t.c};
}

// ---- wavefront / workgroup reductions in a fixed order --------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_down(v, off, WAVE);
    return v;
}
__device__ __forceinline__ f32c wave_sum(f32c v) {
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        f32c o{__shfl_down(v.s, off, WAVE), __shfl_down(v.c, off, WAVE)};
        v = cadd(v, o);
    }
    return v;
}

// Sum of v over the workgroup; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*smem)[BLOCK / WAVE]) {
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = wave_sum(v[i]);
        if (lane == 0) smem[i][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int w = 0; w < BLOCK / WAVE; ++w) s += smem[i][w];
            v[i] = s;
        }
    }
}
template <int NV>
__device__ __forceinline__ void block_sum(f32c (&v)[NV], f32c (*smem)[BLOCK / WAVE]) {
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        f32c s = wave_sum(v[i]);
        if (lane == 0) smem[i][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            f32c s{0.f, 0.f};
            for (int w = 0; w < BLOCK / WAVE; ++w) s = cadd(s, smem[i][w]);
            v[i] = s;
        }
    }
}

// Grid reductions without fences or atomics: a producing kernel writes one
// partial per workgroup (thread 0, after block_sum); a one-workgroup
// finalize kernel launched next on the same stream sums the partials in
// index order.  Kernel boundaries order the writes, so no device-scope
// fence is needed (on a multi-XCD part such a fence writes back and
// invalidates L2, which costs more than the SpMV itself).
template <int NV>
__device__ __forceinline__ void put_partials(const double (&v)[NV], double* partials) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) partials[(size_t)i * gridDim.x + blockIdx.x] = v[i];
    }
}
template <int NV>
__device__ __forceinline__ void put_partials(const f32c (&v)[NV], float* partials) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            partials[(size_t)(2 * i) * gridDim.x + blockIdx.x] = v[i].s;
            partials[(size_t)(2 * i + 1) * gridDim.x + blockIdx.x] = v[i].c;
        }
    }
}

constexpr int FIN = 1024;   // threads of a finalize kernel
constexpr int FIN_LOADS = 8; // partials a finalize thread loads before it adds

// Sum of np partials per value, in a fixed order; valid in thread 0.  Each
// thread adds partials t, t + FIN, t + 2 FIN, ... in that order; the loads of
// a group of FIN_LOADS are all issued before the first add, so one memory
// latency covers the group.
template <int NV>
__device__ __forceinline__ void fin_sum(const double* partials, unsigned np, double (&tot)[NV]) {
    __shared__ double red[NV][FIN / WAVE];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double* pi = partials + (size_t)i * np;
        double s = 0.0;
        for (unsigned b0 = threadIdx.x; b0 < np; b0 += FIN * FIN_LOADS) {
            double v[FIN_LOADS];
#pragma unroll
            for (int j = 0; j < FIN_LOADS; ++j) {
                const unsigned b = b0 + (unsigned)j * FIN;
                v[j] = b < np ? pi[b] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < FIN_LOADS; ++j)
                if (b0 + (unsigned)j * FIN < np) s += v[j];
        }
        s = wave_sum(s);
        if (lane == 0) red[i][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int w = 0; w < FIN / WAVE; ++w) s += red[i][w];
            tot[i] = s;
        }
    }
}
template <int NV>
__device__ __forceinline__ void fin_sum(const float* partials, unsigned np, f32c (&tot)[NV]) {
    __shared__ f32c red[NV][FIN / WAVE];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const float* ps = partials + (size_t)(2 * i) * np;
        const float* pc = partials + (size_t)(2 * i + 1) * np;
        f32c s{0.f, 0.f};
        for (unsigned b0 = threadIdx.x; b0 < np; b0 += FIN * FIN_LOADS) {
            f32c v[FIN_LOADS];
#pragma unroll
            for (int j = 0; j < FIN_LOADS; ++j) {
                const unsigned b = b0 + (unsigned)j * FIN;
                v[j] = b < np ? f32c{ps[b], pc[b]} : f32c{0.f, 0.f};
            }
#pragma unroll
            for (int j = 0; j < FIN_LOADS; ++j)
                if (b0 + (unsigned)j * FIN < np) s = cadd(s, v[j]);
        }
        s = wave_sum(s);
        if (lane == 0) red[i][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            f32c s{0.f, 0.f};
            for (int w = 0; w < FIN / WAVE; ++w) s = cadd(s, red[i][w]);
            tot[i] = s;
        }
    }
}

}  // namespace nbx

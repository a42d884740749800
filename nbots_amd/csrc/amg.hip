// Device side of the smoothed-aggregation preconditioner (amg.hpp): the
// hierarchy's arrays in HBM, the numeric phase that follows the matrix
// values (fine values from the 2x2-block matrix, diagonal inverses, a power
// iteration for lambda_max(D^-1 A), the smoothed prolongator, W = A P and
// A_c = P^T W, the coarsest inverse), and the FP32 V-cycle.
//
// Every kernel sums in the pattern's order, so runs are reproducible.  The
// V-cycle kernels return at once when the solve has stopped (st->done), like
// the flexible CG's own kernels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <type_traits>
#include <vector>

#include "amg.hpp"
#include "fem.hpp"
#include "rows.hpp"

namespace nbx {

// Block sliced ELL for the V-cycle's FP32 products: block rows of R
// scalar rows (one node or aggregate), blocks of R x C values, 64 block
// rows per slice; block k of block row I at slot slice_ptr[I/64] + 64 k +
// I%64, its value (r, c) in plane r C + c (val[(r C + c) slots + slot]), so
// a wavefront's loads of one block position are coalesced.  The levels'
// operators (3 x 3 blocks below the finest level), P (2 x 3 on the finest
// level, 3 x 3 below) and P^T (3 x 2, 3 x 3) take this form: all R rows of
// a block row share one pattern of C-column groups.
struct BellDev {
    uint32_t rows = 0;
    int lanes = 1;             // lanes per block row in the V-cycle's products
    size_t slots = 0;
    uint32_t* slice_ptr = nullptr;
    uint16_t* len = nullptr;
    uint32_t* bcol = nullptr;
    float* val = nullptr;
};

struct AmgDevLevel {
    uint32_t n = 0, nc = 0, bs = 2;
    size_t a_nnz = 0, p_nnz = 0, w_nnz = 0;
    uint32_t *a_ptr = nullptr, *a_col = nullptr;
    double* a64 = nullptr;
    double* dinv64 = nullptr;
    float* dinv32 = nullptr;
    double lmax = 0.0;
    double* t = nullptr;
    uint32_t* agg = nullptr;   // [n / bs] each node's aggregate, or AMG_NONE
    uint32_t *p_ptr = nullptr, *p_col = nullptr;
    double* p64 = nullptr;
    uint32_t *w_ptr = nullptr, *w_col = nullptr;
    double* w64 = nullptr;
    uint32_t *pt_ptr = nullptr, *pt_row = nullptr, *pt_pos = nullptr;
    BellDev ab, pb, tb;        // the operator (below the finest level), P, P^T
    // V-cycle vectors: right-hand side and solution of this level (the
    // finest level's come from the caller), residual, two directions
    float *b = nullptr, *x = nullptr, *r = nullptr, *d0 = nullptr, *d1 = nullptr;
    // Chebyshev smoothing on [lambda/30, 1.1 lambda]
    float inv_theta = 0.f, c1[AMG_MAX_DEGREE] = {}, c2[AMG_MAX_DEGREE] = {};
    int degree = AMG_DEGREE;   // smoothing steps before and after (amg_level_degree)
};

struct AmgDev {
    std::vector<AmgDevLevel> lv;
    float* cinv = nullptr;          // coarsest inverse, n x n
    double* ones = nullptr;         // [finest n + 2] the flexible CG's unit scaling
    double* pow_x = nullptr;        // power iteration: [finest n + 2] x, y
    double* pow_y = nullptr;
    double* scal = nullptr;         // [2] device scalars
    double* partials = nullptr;     // [MAX_GRID] reduction partials
    std::vector<double> host_inv;
    size_t bytes = 0;
    // more than one level: the finest is smoothed on the matrix's fp16
    // scaled copy, in scaled variables, and its P values are stored as
    // D^1/2 P (FP32)
    bool scaled16 = false;
};

namespace {

template <class T, class Al>
int up(T** d, const std::vector<T, Al>& h, size_t& bytes) {
    const size_t nb = (h.size() + 1) * sizeof(T);
    NBX_HIP(hipMalloc(d, nb));
    bytes += nb;
    if (!h.empty()) NBX_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return NBX_OK;
}
template <class T>
int alloc(T** d, size_t n, size_t& bytes) {
    const size_t nb = (n + 2) * sizeof(T);
    NBX_HIP(hipMalloc(d, nb));
    bytes += nb;
    return NBX_OK;
}

// ---- numeric phase ----------------------------------------------------------

// the slice holding slot sl of a sliced layout: the last s < slices with
// slice_ptr[s] <= sl
__device__ __forceinline__ uint32_t slice_of(const uint32_t* __restrict__ slice_ptr,
                                             uint32_t slices, size_t sl) {
    uint32_t lo = 0, hi = slices;
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (slice_ptr[mid] <= sl) lo = mid; else hi = mid;
    }
    return lo;
}

// The finest level's CSR values from the 2x2-block layout, one thread per
// slot: block k of node v holds entries (i, j) of rows 2v + i at positions
// a_ptr[2v + i] + 2k + j (the CSR the host built from the same layout)
__global__ __launch_bounds__(BLOCK) void k_amg_fine_vals(Sell2 K, const uint32_t* __restrict__ a_ptr,
                                                         double* __restrict__ a) {
    const size_t sl = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (sl >= K.n_blocks) return;
    const uint32_t s = slice_of(K.slice_ptr, K.n_slices, sl);
    const size_t off = sl - K.slice_ptr[s];
    const uint32_t v = s * SLICE + (uint32_t)(off % SLICE);
    const uint32_t k = (uint32_t)(off / SLICE);
    if (v >= K.n_nodes || k >= K.rowlen[v]) return;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            a[a_ptr[2 * (size_t)v + i] + 2 * (size_t)k + j] = K.val[ent(K.n_blocks, sl, i, j)];
}
// bit 2 i + j of a slot: entry (i, j) of its block is nonzero
__global__ __launch_bounds__(BLOCK) void k_block_nzbits(Sell2 K, uint8_t* __restrict__ bits) {
    const size_t sl = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (sl >= K.n_blocks) return;
    uint32_t b = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            if (K.val[ent(K.n_blocks, sl, i, j)] != 0.0) b |= 1u << (2 * i + j);
    bits[sl] = (uint8_t)b;
}

// D^-1 (0 where the diagonal is not positive: an unknown with no coarse or
// smoothing part)
__global__ __launch_bounds__(BLOCK) void k_amg_dinv(const uint32_t* __restrict__ ptr,
                                                    const uint32_t* __restrict__ col,
                                                    const double* __restrict__ a,
                                                    double* __restrict__ d64,
                                                    float* __restrict__ d32, uint32_t n) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (uint32_t e = ptr[i]; e < ptr[i + 1]; ++e)
        if (col[e] == i) d = a[e];
    const double v = d > 0.0 ? 1.0 / d : 0.0;
    d64[i] = v;
    d32[i] = (float)v;
}

// power iteration on D^-1 A: x_0 = +-1 from a hash of the index (amg_sign),
// normalised by the first step
__global__ __launch_bounds__(BLOCK) void k_amg_pow_init(double* __restrict__ x, uint32_t n) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    uint32_t h = i * 2654435761u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    x[i] = (h & 1u) ? 1.0 : -1.0;
}
// x.x partials
__global__ __launch_bounds__(BLOCK) void k_amg_sq(const double* __restrict__ x, double* partials,
                                                  uint32_t n) {
    double acc[1] = {0.0};
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) acc[0] = x[i] * x[i];
    __shared__ double red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    put_partials<1>(acc, partials);
}
// y = D^-1 A (x * s), s = *scale (1 / ||x|| of the previous step); y.y
// partials.  Eight lanes per row, rows strided over the grid (at most
// MAX_GRID workgroups, one partial each).
constexpr int POW_LANES = 8;
__global__ __launch_bounds__(BLOCK) void k_amg_pow_step(const uint32_t* __restrict__ ptr,
                                                        const uint32_t* __restrict__ col,
                                                        const double* __restrict__ a,
                                                        const double* __restrict__ dinv,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ y,
                                                        const double* __restrict__ scale,
                                                        double* partials, uint32_t n) {
    const double sc = *scale;
    double acc[1] = {0.0};
    const size_t gid = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t lane = (uint32_t)(gid % POW_LANES);
    const size_t stride = (size_t)gridDim.x * BLOCK / POW_LANES;
    const size_t rows = ((size_t)n + stride - 1) / stride * stride;   // every group runs alike
    for (size_t r = gid / POW_LANES; r < rows; r += stride) {
        const bool valid = r < n;
        double s = 0.0;
        if (valid)
            for (uint32_t e = ptr[r] + lane; e < ptr[r + 1]; e += POW_LANES) s += a[e] * x[col[e]];
#pragma unroll
        for (int off = POW_LANES / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, POW_LANES);
        if (valid && lane == 0) {
            const double v = dinv[r] * (s * sc);
            y[r] = v;
            acc[0] += v * v;
        }
    }
    __shared__ double red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    put_partials<1>(acc, partials);
}
// scal[0] = 1 / sqrt(sum), scal[1] = sqrt(sum)
__global__ __launch_bounds__(FIN) void k_amg_pow_fin(const double* partials, unsigned np,
                                                     double* scal) {
    double tot[1];
    fin_sum<1>(partials, np, tot);
    if (threadIdx.x == 0) {
        const double nr = sqrt(tot[0]);
        scal[0] = nr > 0.0 ? 1.0 / nr : 0.0;
        scal[1] = nr;
    }
}

// P = T - omega D^-1 A T: SMOOTH_SLOTS threads per node, each taking every
// SMOOTH_SLOTS-th aggregate column group of the node's P rows (its R rows
// share one A and one P pattern; a group is the three columns of one
// aggregate) and summing, in entry order, each row's A entries whose column
// node belongs to that aggregate
constexpr int SMOOTH_SLOTS = 8;
template <int R>
__global__ __launch_bounds__(BLOCK) void k_amg_smooth_p(
    const uint32_t* __restrict__ a_ptr, const uint32_t* __restrict__ a_col,
    const double* __restrict__ a, const double* __restrict__ dinv, const double* __restrict__ t,
    const uint32_t* __restrict__ agg, const uint32_t* __restrict__ p_ptr,
    const uint32_t* __restrict__ p_col, double* __restrict__ p, double omega, uint32_t nodes) {
    const size_t tid = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t I = (uint32_t)(tid / SMOOTH_SLOTS);
    if (I >= nodes) return;
    const uint32_t i0 = R * I;
    const uint32_t ng = (p_ptr[i0 + 1] - p_ptr[i0]) / 3;
    double od[R];
    uint32_t ao[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        od[r] = omega * dinv[i0 + r];
        ao[r] = a_ptr[i0 + r] - a_ptr[i0];   // row r's entry e + ao[r] pairs with e
    }
    for (uint32_t gi = (uint32_t)(tid % SMOOTH_SLOTS); gi < ng; gi += SMOOTH_SLOTS) {
        const uint32_t qo = 3 * gi;          // the group's offset in each row
        const uint32_t ga = p_col[p_ptr[i0] + qo] / 3;   // its aggregate
        const bool own = agg[I] == ga;
        double acc[R][3];
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int m = 0; m < 3; ++m) acc[r][m] = 0.0;
            if (own)
#pragma unroll
                for (int m = 0; m < 3; ++m) acc[r][m] += t[3 * (size_t)(i0 + r) + m];
        }
        for (uint32_t e = a_ptr[i0]; e < a_ptr[i0 + 1]; ++e) {
            if (agg[a_col[e] / R] != ga) continue;
            const uint32_t j = a_col[e];
            double tj[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) tj[m] = t[3 * (size_t)j + m];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double coef = od[r] * a[e + ao[r]];
#pragma unroll
                for (int m = 0; m < 3; ++m) acc[r][m] -= coef * tj[m];
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int m = 0; m < 3; ++m) p[p_ptr[i0 + r] + qo + m] = acc[r][m];
    }
}

// Block values from a CSR whose R rows per block row share one pattern of
// C-column groups: block k of block row I holds the rows' k-th groups
// (scaled by 1 / dis of the row when dis is given: D^1/2 P on the finest
// level in the scaled variables).
template <int R, int C>
__global__ __launch_bounds__(BLOCK) void k_bell_fill(BellDev B, const uint32_t* __restrict__ ptr,
                                                     const double* __restrict__ v,
                                                     const double* __restrict__ dis) {
    const uint32_t I = blockIdx.x * BLOCK + threadIdx.x;
    if (I >= B.rows) return;
    const size_t base = (size_t)B.slice_ptr[I >> 6] + (I & 63);
    const int len = B.len[I];
    for (int r = 0; r < R; ++r) {
        const size_t row = (size_t)R * I + r;
        const uint32_t off = ptr[row];
        const double sc = dis ? 1.0 / dis[row] : 1.0;
        for (int k = 0; k < len; ++k)
            for (int c = 0; c < C; ++c)
                B.val[(size_t)(r * C + c) * B.slots + base + (size_t)k * SLICE] =
                    (float)(v[off + (size_t)C * k + c] * sc);
    }
}
// P^T's blocks from P's values through the P^T lists: block k of aggregate
// a is node J = bcol; its value (c, r) is P(R J + r, 3 a + c), the
// (R k + r)-th entry of column 3 a + c's list.  One thread per slot of the
// layout (padding slots return).
template <int R>
__global__ __launch_bounds__(BLOCK) void k_bell_fill_t(BellDev B, uint32_t slices,
                                                       const uint32_t* __restrict__ pt_ptr,
                                                       const uint32_t* __restrict__ pt_pos,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ dis) {
    const size_t sl = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (sl >= B.slots) return;
    // the slice holding slot sl: the last s with slice_ptr[s] <= sl
    uint32_t lo = 0, hi = slices;
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (B.slice_ptr[mid] <= sl) lo = mid; else hi = mid;
    }
    const size_t off = sl - B.slice_ptr[lo];
    const uint32_t a = lo * SLICE + (uint32_t)(off % SLICE);
    const uint32_t k = (uint32_t)(off / SLICE);
    if (a >= B.rows || k >= B.len[a]) return;
    const uint32_t J = B.bcol[sl];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < R; ++r) {
            const uint32_t q = pt_pos[pt_ptr[3 * (size_t)a + c] + (size_t)R * k + r];
            const double sc = dis ? 1.0 / dis[(size_t)R * J + r] : 1.0;
            B.val[(size_t)(c * R + r) * B.slots + sl] = (float)(p[q] * sc);
        }
}

// W = A P and A_c = P^T W, one thread per output block.  The host built
// every pattern in groups: A's rows of one node (R rows: R = 2 on the finest
// level, 3 below) share its pattern of R-column groups, P's, W's and A_c's
// rows share patterns of 3-column groups (aggregates), all ascending; W's
// row holds every aggregate of every P row its A row reaches, A_c's row
// every aggregate of every W row of a node holding it.  An output block sums
// its block products in a fixed order (the neighbour blocks, or the nodes
// holding the aggregate, ascending): no atomics, and runs are reproducible.

// the block row whose blocks start at or before block t: the last i with
// ptr[stride i] <= per t (ptr ascending, rows in [0, rows))
__device__ __forceinline__ uint32_t block_row(const uint32_t* __restrict__ ptr, uint32_t stride,
                                              size_t per, size_t t, uint32_t rows) {
    uint32_t lo = 0, hi = rows;   // invariant: ptr[stride lo] <= per t < ptr[stride hi]
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if ((size_t)ptr[(size_t)stride * mid] <= per * t) lo = mid; else hi = mid;
    }
    return lo;
}
// position of group b3 (a multiple of 3) in the ascending groups
// col[b], col[b + 3], ..., col[b + 3 (len - 1)], or len
__device__ __forceinline__ uint32_t find_group(const uint32_t* __restrict__ col, uint32_t b,
                                               uint32_t len, uint32_t b3) {
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (col[b + 3 * mid] < b3) lo = mid + 1; else hi = mid;
    }
    return lo < len && col[b + 3 * lo] == b3 ? lo : len;
}

// W(I, b) = sum over neighbour blocks J of A(I, J) P(J, b).  LANES
// consecutive lanes take an output block (each sums every LANES-th
// neighbour, then a butterfly adds the lanes' sums): more lanes on the small
// levels, whose few output blocks each sum many products.
template <int R, int LANES>
__global__ __launch_bounds__(BLOCK) void k_amg_w(const uint32_t* __restrict__ a_ptr,
                                                 const uint32_t* __restrict__ a_col,
                                                 const double* __restrict__ a,
                                                 const uint32_t* __restrict__ p_ptr,
                                                 const uint32_t* __restrict__ p_col,
                                                 const double* __restrict__ p,
                                                 const uint32_t* __restrict__ w_ptr,
                                                 const uint32_t* __restrict__ w_col,
                                                 double* __restrict__ w, uint32_t nodes,
                                                 size_t blocks) {
    const size_t tid = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    const size_t t = tid / LANES;
    const uint32_t lane = (uint32_t)(tid % LANES);
    if (t >= blocks) return;   // (a whole group: blocks end at a group boundary)
    const uint32_t I = block_row(w_ptr, R, 3 * R, t, nodes);
    const uint32_t j = (uint32_t)(t - w_ptr[(size_t)R * I] / (3 * R));
    const uint32_t b3 = w_col[w_ptr[(size_t)R * I] + 3 * j];
    uint32_t ar[R];
#pragma unroll
    for (int r = 0; r < R; ++r) ar[r] = a_ptr[(size_t)R * I + r];
    const uint32_t al = (a_ptr[(size_t)R * I + 1] - ar[0]) / R;
    double acc[R][3] = {};
    for (uint32_t k = lane; k < al; k += LANES) {
        const uint32_t J = a_col[ar[0] + R * k] / R;
        const uint32_t p0 = p_ptr[(size_t)R * J];
        const uint32_t kk = find_group(p_col, p0, (p_ptr[(size_t)R * J + 1] - p0) / 3, b3);
        if (kk == (p_ptr[(size_t)R * J + 1] - p0) / 3) continue;
        double pk[R][3];
#pragma unroll
        for (int c = 0; c < R; ++c)
#pragma unroll
            for (int m = 0; m < 3; ++m) pk[c][m] = p[p_ptr[(size_t)R * J + c] + 3 * kk + m];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                double s = 0.0;
#pragma unroll
                for (int c = 0; c < R; ++c) s += a[ar[r] + R * k + c] * pk[c][m];
                acc[r][m] += s;
            }
    }
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1)
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int m = 0; m < 3; ++m) acc[r][m] += __shfl_xor(acc[r][m], off, LANES);
    if (lane != 0) return;
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int m = 0; m < 3; ++m) w[w_ptr[(size_t)R * I + r] + 3 * j + m] = acc[r][m];
}

// A_c(a, b) = sum over the nodes J holding aggregate a of P(J, a)^T W(J, b),
// LANES lanes per output block as in k_amg_w
template <int R, int LANES>
__global__ __launch_bounds__(BLOCK) void k_amg_galerkin(
    const uint32_t* __restrict__ pt_ptr, const uint32_t* __restrict__ pt_row,
    const uint32_t* __restrict__ pt_pos, const double* __restrict__ p,
    const uint32_t* __restrict__ w_ptr, const uint32_t* __restrict__ w_col,
    const double* __restrict__ w, const uint32_t* __restrict__ c_ptr,
    const uint32_t* __restrict__ c_col, double* __restrict__ ac, uint32_t aggs, size_t blocks) {
    const size_t tid = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    const size_t t = tid / LANES;
    const uint32_t lane = (uint32_t)(tid % LANES);
    if (t >= blocks) return;   // (a whole group)
    const uint32_t A = block_row(c_ptr, 3, 9, t, aggs);
    const uint32_t j = (uint32_t)(t - c_ptr[3 * (size_t)A] / 9);
    const uint32_t b3 = c_col[c_ptr[3 * (size_t)A] + 3 * j];
    uint32_t tr[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) tr[m] = pt_ptr[3 * (size_t)A + m];
    const uint32_t hl = (pt_ptr[3 * (size_t)A + 1] - tr[0]) / R;
    double acc[3][3] = {};
    for (uint32_t h = lane; h < hl; h += LANES) {
        const uint32_t J = pt_row[tr[0] + R * h] / R;
        const uint32_t w0 = w_ptr[(size_t)R * J];
        const uint32_t wl = (w_ptr[(size_t)R * J + 1] - w0) / 3;
        const uint32_t q = find_group(w_col, w0, wl, b3);
        if (q == wl) continue;
        double wq[R][3], ph[R][3];
#pragma unroll
        for (int c = 0; c < R; ++c)
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                wq[c][n] = w[w_ptr[(size_t)R * J + c] + 3 * q + n];
                ph[c][n] = p[pt_pos[tr[n] + R * h + c]];
            }
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                double s = 0.0;
#pragma unroll
                for (int c = 0; c < R; ++c) s += ph[c][m] * wq[c][n];
                acc[m][n] += s;
            }
    }
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1)
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = 0; n < 3; ++n) acc[m][n] += __shfl_xor(acc[m][n], off, LANES);
    if (lane != 0) return;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n) ac[c_ptr[3 * (size_t)A + m] + 3 * j + n] = acc[m][n];
}

// ---- V-cycle (FP32) ----------------------------------------------------------
// A group of LANES consecutive lanes takes a block row: each lane sums every
// LANES-th block, the group adds the lanes' sums by a butterfly, and the
// group's first lane does the row's update.  The order of the additions is
// fixed by the pattern, so runs are reproducible.

template <int LANES>
__device__ __forceinline__ float group_sum(float s) {
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LANES);
    return s;
}

// ---- block sliced ELL products (FP32) ---------------------------------------

template <int R, int C>
struct Bell {
    size_t slots;
    const uint32_t* __restrict__ slice_ptr;
    const uint16_t* __restrict__ len;
    const uint32_t* __restrict__ bcol;
    const float* __restrict__ val;
};
template <int R, int C>
Bell<R, C> bell(const BellDev& B) {
    return Bell<R, C>{B.slots, B.slice_ptr, B.len, B.bcol, B.val};
}

// y = (block row I) x, x indexed by C per block column, by a group of LANES
// lanes: lane l takes blocks l, l + LANES, ..., two at a time (their columns
// and values loaded before the dependent gathers), and the group adds the
// lanes' sums by a butterfly; every lane gets the result.
template <int R, int C, int LANES>
__device__ __forceinline__ void bell_row(const Bell<R, C>& B, uint32_t I, uint32_t lane,
                                         const float* __restrict__ x, float (&y)[R]) {
#pragma unroll
    for (int r = 0; r < R; ++r) y[r] = 0.f;
    const size_t base = (size_t)B.slice_ptr[I >> 6] + (I & 63);
    const int len = B.len[I];
    for (int k0 = (int)lane; k0 < len; k0 += 2 * LANES) {
        uint32_t J[2];
        float a[2][R * C], xv[2][C];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (k0 + u * LANES < len) {
                const size_t sl = base + (size_t)(k0 + u * LANES) * SLICE;
                J[u] = B.bcol[sl];
#pragma unroll
                for (int m = 0; m < R * C; ++m) a[u][m] = B.val[(size_t)m * B.slots + sl];
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
            if (k0 + u * LANES < len)
#pragma unroll
                for (int c = 0; c < C; ++c) xv[u][c] = x[(size_t)C * J[u] + c];
#pragma unroll
        for (int u = 0; u < 2; ++u)
            if (k0 + u * LANES < len)
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int c = 0; c < C; ++c) y[r] += a[u][r * C + c] * xv[u][c];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) y[r] = group_sum<LANES>(y[r]);
}

// the group's block row and lane (a group's lanes share its validity, so a
// group leaves as a whole)
#define BELL_ROW(LANES, rows)                                                       \
    const size_t gid_ = (size_t)blockIdx.x * BLOCK + threadIdx.x;                  \
    const uint32_t I = (uint32_t)(gid_ / LANES), lane = (uint32_t)(gid_ % LANES); \
    if (I >= (rows)) return

// Chebyshev smoothing below the finest level, a group per aggregate (3
// rows).  The start (r = D^-1 b, d = r / theta, x = d) rides in the
// restriction that writes b; the post-smoothing start takes A x.
template <int LANES>
__global__ __launch_bounds__(BLOCK) void k_bsm_init(Bell<3, 3> A, const float* __restrict__ dinv,
                                                    const float* __restrict__ b,
                                                    const float* __restrict__ x,
                                                    float* __restrict__ r, float* __restrict__ d,
                                                    float inv_theta, const PcgState* st,
                                                    uint32_t rows) {
    if (st->done) return;
    BELL_ROW(LANES, rows);
    float ax[3];
    bell_row<3, 3, LANES>(A, I, lane, x, ax);
    if (lane != 0) return;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const size_t i = 3 * (size_t)I + m;
        const float ri = dinv[i] * (b[i] - ax[m]);
        r[i] = ri;
        d[i] = ri * inv_theta;
    }
}

// r -= D^-1 A d_in; d_out = c1 d_in + c2 r; x += d_out (and d_in when ADD_IN)
template <bool ADD_IN, int LANES>
__global__ __launch_bounds__(BLOCK) void k_bsm_step(Bell<3, 3> A, const float* __restrict__ dinv,
                                                    const float* __restrict__ din,
                                                    float* __restrict__ dout,
                                                    float* __restrict__ r, float* __restrict__ x,
                                                    float c1, float c2, const PcgState* st,
                                                    uint32_t rows) {
    if (st->done) return;
    BELL_ROW(LANES, rows);
    float ad[3];
    bell_row<3, 3, LANES>(A, I, lane, din, ad);
    if (lane != 0) return;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const size_t i = 3 * (size_t)I + m;
        const float di = din[i];
        const float ri = r[i] - dinv[i] * ad[m];
        const float dn = c1 * di + c2 * ri;
        r[i] = ri;
        dout[i] = dn;
        x[i] += ADD_IN ? di + dn : dn;
    }
}

template <int LANES>
__global__ __launch_bounds__(BLOCK) void k_bresidual(Bell<3, 3> A, const float* __restrict__ b,
                                                     const float* __restrict__ x,
                                                     float* __restrict__ r, const PcgState* st,
                                                     uint32_t rows) {
    if (st->done) return;
    BELL_ROW(LANES, rows);
    float ax[3];
    bell_row<3, 3, LANES>(A, I, lane, x, ax);
    if (lane != 0) return;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const size_t i = 3 * (size_t)I + m;
        r[i] = b[i] - ax[m];
    }
}

// b_c = P^T r, a group per aggregate; INIT: also the next level's
// smoothing start from x = 0 (r_c = D_c^-1 b_c, d_c = r_c / theta_c, x_c = d_c)
template <int R, bool INIT, int LANES>
__global__ __launch_bounds__(BLOCK) void k_brestrict(Bell<3, R> T, const float* __restrict__ r,
                                                     float* __restrict__ bc,
                                                     const float* __restrict__ dinv_c,
                                                     float* __restrict__ xc,
                                                     float* __restrict__ rc,
                                                     float* __restrict__ dc, float inv_theta_c,
                                                     const PcgState* st, uint32_t aggs) {
    if (st->done) return;
    BELL_ROW(LANES, aggs);
    float y[3];
    bell_row<3, R, LANES>(T, I, lane, r, y);
    if (lane != 0) return;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const size_t i = 3 * (size_t)I + m;
        bc[i] = y[m];
        if constexpr (INIT) {
            const float ri = dinv_c[i] * y[m];
            const float di = ri * inv_theta_c;
            rc[i] = ri;
            dc[i] = di;
            xc[i] = di;
        }
    }
}

// x += P x_c (ADD), or x = P x_c; a group per node
template <int R, bool ADD, int LANES>
__global__ __launch_bounds__(BLOCK) void k_bprolong(Bell<R, 3> P, const float* __restrict__ xc,
                                                    float* __restrict__ x, const PcgState* st,
                                                    uint32_t rows) {
    if (st->done) return;
    BELL_ROW(LANES, rows);
    float y[R];
    bell_row<R, 3, LANES>(P, I, lane, xc, y);
    if (lane != 0) return;
#pragma unroll
    for (int m = 0; m < R; ++m) {
        const size_t i = (size_t)R * I + m;
        x[i] = ADD ? x[i] + y[m] : y[m];
    }
}

// x = C^-1 b on the coarsest level (n <= AMG_COARSEST), one workgroup
constexpr int DENSE_BLOCK = 256;
__global__ __launch_bounds__(DENSE_BLOCK) void k_amg_dense(const float* __restrict__ cinv,
                                                           const float* __restrict__ b,
                                                           float* __restrict__ x,
                                                           const PcgState* st, uint32_t n) {
    if (st->done) return;
    __shared__ float sb[AMG_COARSEST];
    for (uint32_t j = threadIdx.x; j < n; j += DENSE_BLOCK) sb[j] = b[j];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += DENSE_BLOCK) {
        float s = 0.f;
        for (uint32_t j = 0; j < n; ++j) s += cinv[(size_t)i * n + j] * sb[j];
        x[i] = s;
    }
}

}  // namespace


void amg_free(AmgDev* A) {
    if (!A) return;
    for (AmgDevLevel& L : A->lv) {
        void* ps[] = {L.a_ptr, L.a_col, L.a64, L.dinv64, L.dinv32, L.t, L.agg, L.p_ptr,
                      L.p_col, L.p64, L.w_ptr, L.w_col, L.w64, L.pt_ptr,
                      L.pt_row, L.pt_pos, L.b, L.x, L.r, L.d0, L.d1};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        for (BellDev* B : {&L.ab, &L.pb, &L.tb}) {
            void* qs[] = {B->slice_ptr, B->len, B->bcol, B->val};
            for (void* q : qs)
                if (q) (void)hipFree(q);
        }
    }
    void* ps[] = {A->cinv, A->ones, A->pow_x, A->pow_y, A->scal, A->partials};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete A;
}

// The block sliced-ELL pattern of `rows` block rows: len_of(I) blocks in
// row I, col_of(I, k) the k-th block's column; uploaded with room for R x C
// FP32 values per slot.  NBX_ERR_ARG when a row has more blocks than the
// format holds (65,535) or the layout more than 2^32 slots.
template <int R, int C, class Len, class Col>
static int bell_upload(uint32_t rows, Len len_of, Col col_of, BellDev& B, size_t& bytes) {
    const uint32_t ns = (rows + SLICE - 1) / SLICE;
    std::vector<uint32_t> sp(ns + 1, 0);
    std::vector<uint16_t> len(rows, 0);
    for (uint32_t sl = 0; sl < ns; ++sl) {
        uint32_t w = 0;
        for (uint32_t I = sl * SLICE; I < std::min<uint32_t>(rows, (sl + 1) * SLICE); ++I) {
            const uint32_t l = len_of(I);
            if (l > 0xffffu) {
                set_last_error("AMG: a block row of %u blocks", l);
                return NBX_ERR_ARG;
            }
            len[I] = (uint16_t)l;
            w = std::max(w, l);
        }
        if ((uint64_t)sp[sl] + (uint64_t)w * SLICE > 0xffffffffull) {
            set_last_error("AMG: block layout beyond 2^32 slots");
            return NBX_ERR_ARG;
        }
        sp[sl + 1] = sp[sl] + w * SLICE;
    }
    std::vector<uint32_t> bcol(sp[ns], 0);
    for (uint32_t I = 0; I < rows; ++I)
        for (uint32_t k = 0; k < len[I]; ++k)
            bcol[sp[I / SLICE] + (size_t)k * SLICE + I % SLICE] = col_of(I, k);
    B.rows = rows;
    B.slots = sp[ns];
    // about three blocks per lane (1, 2, 4, 8 or 16 lanes)
    size_t used = 0;
    for (uint32_t I = 0; I < rows; ++I) used += len[I];
    const double avg = rows ? (double)used / rows : 0.0;
    B.lanes = 1;
    while (B.lanes < 16 && avg / (2 * B.lanes) >= 2.5) B.lanes *= 2;
    NBX_TRY(up(&B.slice_ptr, sp, bytes));
    NBX_TRY(up(&B.len, len, bytes));
    NBX_TRY(up(&B.bcol, bcol, bytes));
    NBX_TRY(alloc(&B.val, B.slots * R * C, bytes));
    return NBX_OK;
}

// Whether the R rows of every block row share one pattern of complete
// C-column groups (what the block layout stores).
static bool grouped(uint32_t rows, int R, int C, const HostVec<uint32_t>& ptr,
                    const HostVec<uint32_t>& col) {
    for (uint32_t I = 0; I < rows; ++I) {
        const uint32_t b0 = ptr[(size_t)R * I], l = ptr[(size_t)R * I + 1] - b0;
        if (l % C) return false;
        for (int r = 1; r < R; ++r) {
            const uint32_t br = ptr[(size_t)R * I + r];
            if (ptr[(size_t)R * I + r + 1] - br != l) return false;
            for (uint32_t q = 0; q < l; ++q)
                if (col[br + q] != col[b0 + q]) return false;
        }
        for (uint32_t q = 0; q < l; q += C)
            for (int c = 1; c < C; ++c)
                if (col[b0 + q + c] != col[b0 + q] + c || col[b0 + q] % C) return false;
    }
    return true;
}

// Symbolic hierarchy of the system's current K (pattern, Dirichlet rows,
// node coordinates), uploaded; replaces any previous one.
static int amg_symbolic(FemSys& S) {
    amg_free(S.K.amg);
    S.K.amg = nullptr;
    S.K.amg_current = false;
    const uint32_t n = 2 * S.nn;
    const size_t nnz = S.K.s.n_valid * 4;
    if ((size_t)n > (size_t)MAX_GRID * BLOCK || nnz > 0xffffffffull) {
        set_last_error("AMG: %u unknowns (at most %u) or %zu nonzeros", n,
                       (unsigned)(MAX_GRID * BLOCK), nnz);
        return NBX_ERR_ARG;
    }
    // the finest pattern and its nonzero entries, from the block layout's
    // columns and one flag byte per block (the values stay on the device)
    const Sell2& K = S.K.s;
    std::vector<uint32_t> sp(K.n_slices + 1);
    std::vector<uint8_t> len(S.nn + 1);
    HostVec<uint32_t> bc(K.n_blocks + 1);
    HostVec<uint8_t> bits(K.n_blocks + 1);
    std::vector<double> xy(2 * (size_t)S.nn);
    {
        uint8_t* dbits = nullptr;
        NBX_HIP(hipMalloc(&dbits, K.n_blocks + 1));
        hipLaunchKernelGGL(k_block_nzbits, dim3(grid_for(K.n_blocks)), dim3(BLOCK), 0, S.s, K,
                           dbits);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(bits.data(), dbits, K.n_blocks, hipMemcpyDeviceToHost, S.s);
        if (e == hipSuccess) e = hipStreamSynchronize(S.s);
        (void)hipFree(dbits);
        NBX_HIP(e);
    }
    NBX_HIP(hipMemcpyAsync(sp.data(), K.slice_ptr, (K.n_slices + 1) * 4, hipMemcpyDeviceToHost,
                           S.s));
    if (S.nn) NBX_HIP(hipMemcpyAsync(len.data(), K.rowlen, S.nn, hipMemcpyDeviceToHost, S.s));
    NBX_HIP(hipMemcpyAsync(bc.data(), K.bcol, K.n_blocks * 4, hipMemcpyDeviceToHost, S.s));
    NBX_HIP(copy_on(xy.data(), S.nod, xy.size() * sizeof(double), hipMemcpyDeviceToHost, S.s));
    HostVec<uint32_t> rp, col;
    HostVec<uint8_t> nzf;
    amg_expand_blocks(S.nn, sp.data(), len.data(), bc.data(), bits.data(), rp, col, nzf);
    if (rp.size() != (size_t)n + 1 || col.size() != nnz) {
        set_last_error("AMG: the block layout holds %zu entries, the pattern %zu", col.size(), nnz);
        return NBX_ERR_ARG;
    }
    AmgHost H;
    const int hb = amg_build_host(n, rp.data(), col.data(), nzf.data(), xy.data(), H);
    if (hb == -2) {
        set_last_error("AMG: host memory ran out building the hierarchy");
        return NBX_ERR_ARG;
    }
    if (hb) {
        set_last_error("AMG: the matrix is not a 2-dof node pattern");
        return NBX_ERR_ARG;
    }
    if (H.lv.back().n > AMG_COARSEST) {
        set_last_error("AMG: coarsening stopped at %u unknowns (at most %u are solved directly)",
                       H.lv.back().n, AMG_COARSEST);
        return NBX_ERR_ARG;
    }
    AmgDev* A = new AmgDev();
    S.K.amg = A;
    A->lv.resize(H.lv.size());
    const size_t nl = H.lv.size();
    for (size_t l = 0; l < nl; ++l) {
        const AmgHostLevel& h = H.lv[l];
        const uint32_t bs = l == 0 ? 2 : 3;
        const bool ok = h.n % bs == 0 && grouped(h.n / bs, (int)bs, (int)bs, h.a_ptr, h.a_col) &&
                        (l + 1 == nl || (h.nc % 3 == 0 &&
                                          grouped(h.n / bs, (int)bs, 3, h.p_ptr, h.p_col) &&
                                          grouped(h.n / bs, (int)bs, 3, h.w_ptr, h.w_col)));
        if (!ok) {
            set_last_error("AMG: level %zu is not in %ux%u / %ux3 node blocks", l, bs, bs, bs);
            return NBX_ERR_ARG;
        }
    }
    for (size_t l = 0; l < nl; ++l) {
        const AmgHostLevel& h = H.lv[l];
        AmgDevLevel& L = A->lv[l];
        L.n = h.n;
        L.nc = h.nc;
        L.bs = l == 0 ? 2 : 3;
        L.degree = amg_level_degree(l);
        L.a_nnz = h.a_col.size();
        NBX_TRY(up(&L.a_ptr, h.a_ptr, A->bytes));
        NBX_TRY(up(&L.a_col, h.a_col, A->bytes));
        NBX_TRY(alloc(&L.a64, L.a_nnz, A->bytes));
        if (l > 0 && l + 1 < nl) {
            const uint32_t* ap = h.a_ptr.data();
            const uint32_t* ac = h.a_col.data();
            auto alen = [&](uint32_t I) { return (ap[3 * (size_t)I + 1] - ap[3 * (size_t)I]) / 3; };
            auto acol = [&](uint32_t I, uint32_t k) { return ac[ap[3 * (size_t)I] + 3 * k] / 3; };
            NBX_TRY((bell_upload<3, 3>(h.n / 3, alen, acol, L.ab, A->bytes)));
        }
        NBX_TRY(alloc(&L.dinv64, L.n, A->bytes));
        NBX_TRY(alloc(&L.dinv32, L.n, A->bytes));
        NBX_TRY(alloc(&L.r, L.n, A->bytes));
        NBX_TRY(alloc(&L.d0, L.n, A->bytes));
        NBX_TRY(alloc(&L.d1, L.n, A->bytes));
        if (l > 0) {
            NBX_TRY(alloc(&L.b, L.n, A->bytes));
            NBX_TRY(alloc(&L.x, L.n, A->bytes));
        }
        if (l + 1 < H.lv.size()) {
            L.p_nnz = h.p_col.size();
            L.w_nnz = h.w_col.size();
            const uint32_t R = L.bs;
            const uint32_t* pp = h.p_ptr.data();
            const uint32_t* pc = h.p_col.data();
            const uint32_t* tp = h.pt_ptr.data();
            const uint32_t* tr = h.pt_row.data();
            auto plen = [&](uint32_t I) { return (pp[R * (size_t)I + 1] - pp[R * (size_t)I]) / 3; };
            auto pcol = [&](uint32_t I, uint32_t k) { return pc[pp[R * (size_t)I] + 3 * k] / 3; };
            auto tlen = [&](uint32_t a) { return (tp[3 * (size_t)a + 1] - tp[3 * (size_t)a]) / R; };
            auto tcol = [&](uint32_t a, uint32_t k) { return tr[tp[3 * (size_t)a] + R * k] / R; };
            if (R == 2) {
                NBX_TRY((bell_upload<2, 3>(L.n / 2, plen, pcol, L.pb, A->bytes)));
                NBX_TRY((bell_upload<3, 2>(L.nc / 3, tlen, tcol, L.tb, A->bytes)));
            } else {
                NBX_TRY((bell_upload<3, 3>(L.n / 3, plen, pcol, L.pb, A->bytes)));
                NBX_TRY((bell_upload<3, 3>(L.nc / 3, tlen, tcol, L.tb, A->bytes)));
            }
            NBX_TRY(up(&L.t, h.t, A->bytes));
            NBX_TRY(up(&L.agg, h.agg, A->bytes));
            NBX_TRY(up(&L.p_ptr, h.p_ptr, A->bytes));
            NBX_TRY(up(&L.p_col, h.p_col, A->bytes));
            NBX_TRY(alloc(&L.p64, L.p_nnz, A->bytes));
            NBX_TRY(up(&L.w_ptr, h.w_ptr, A->bytes));
            NBX_TRY(up(&L.w_col, h.w_col, A->bytes));
            NBX_TRY(alloc(&L.w64, L.w_nnz, A->bytes));
            NBX_TRY(up(&L.pt_ptr, h.pt_ptr, A->bytes));
            NBX_TRY(up(&L.pt_row, h.pt_row, A->bytes));
            NBX_TRY(up(&L.pt_pos, h.pt_pos, A->bytes));
        }
    }
    const uint32_t nco = A->lv.back().n;
    NBX_TRY(alloc(&A->cinv, (size_t)nco * nco, A->bytes));
    NBX_TRY(alloc(&A->ones, n, A->bytes));
    NBX_TRY(alloc(&A->pow_x, n, A->bytes));
    NBX_TRY(alloc(&A->pow_y, n, A->bytes));
    NBX_TRY(alloc(&A->scal, 2, A->bytes));
    NBX_TRY(alloc(&A->partials, (size_t)MAX_GRID, A->bytes));
    std::vector<double> ones(n + 2, 1.0);
    NBX_HIP(hipMemcpy(A->ones, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice));
    A->host_inv.assign((size_t)nco * nco, 0.0);
    return NBX_OK;
}

// f(integral_constant LANES) for a block layout's lanes per row
template <class F>
static void with_bl(int lanes, F&& f) {
    switch (lanes) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        default: f(std::integral_constant<int, 16>{}); break;
    }
}
// f(integral_constant R) for a level's rows per node
template <class F>
static void with_bs(uint32_t bs, F&& f) {
    if (bs == 2)
        f(std::integral_constant<int, 2>{});
    else
        f(std::integral_constant<int, 3>{});
}

// f(integral_constant LANES) for the numeric phase's products: lanes per
// output block, more on the small levels (0 blocks: one lane)
template <class F>
static void with_out_lanes(size_t blocks, F&& f) {
    if (blocks == 0)
        f(std::integral_constant<int, 1>{});
    else if (blocks >= (size_t)1 << 14)
        f(std::integral_constant<int, 4>{});
    else
        f(std::integral_constant<int, 16>{});
}

// Numeric phase for the current values of K.
static int amg_numeric(FemSys& S) {
    AmgDev* A = S.K.amg;
    const hipStream_t s = S.s;
    AmgDevLevel& L0 = A->lv[0];
    hipLaunchKernelGGL(k_amg_fine_vals, dim3(grid_for(S.K.s.n_blocks)), dim3(BLOCK), 0, s, S.K.s,
                       L0.a_ptr, L0.a64);
    NBX_HIP(hipGetLastError());
    for (size_t l = 0; l < A->lv.size(); ++l) {
        AmgDevLevel& L = A->lv[l];
        const unsigned g = grid_for(L.n);
        hipLaunchKernelGGL(k_amg_dinv, dim3(g), dim3(BLOCK), 0, s, L.a_ptr, L.a_col, L.a64,
                           L.dinv64, L.dinv32, L.n);
        // (the finest level's smoothing reads the fp16 scaled copy; the
        // coarser levels' operators are filled in block form by the level
        // above)
        // lambda_max(D^-1 A): 15 power steps from a fixed start (the
        // coarsest level, solved directly, needs none)
        if (l + 1 < A->lv.size()) {
        hipLaunchKernelGGL(k_amg_pow_init, dim3(g), dim3(BLOCK), 0, s, A->pow_x, L.n);
        hipLaunchKernelGGL(k_amg_sq, dim3(g), dim3(BLOCK), 0, s, A->pow_x, A->partials, L.n);
        hipLaunchKernelGGL(k_amg_pow_fin, dim3(1), dim3(FIN), 0, s, A->partials, g, A->scal);
        double *px = A->pow_x, *py = A->pow_y;
        const unsigned gp = std::min<size_t>(grid_for((size_t)L.n * POW_LANES), (size_t)MAX_GRID);
        for (int it = 0; it < 15; ++it) {
            hipLaunchKernelGGL(k_amg_pow_step, dim3(gp), dim3(BLOCK), 0, s, L.a_ptr, L.a_col, L.a64,
                               L.dinv64, px, py, A->scal, A->partials, L.n);
            hipLaunchKernelGGL(k_amg_pow_fin, dim3(1), dim3(FIN), 0, s, A->partials, gp, A->scal);
            std::swap(px, py);
        }
        NBX_HIP(hipGetLastError());
        double sc[2];
        NBX_HIP(copy_on(sc, A->scal, sizeof sc, hipMemcpyDeviceToHost, s));
        L.lmax = sc[1];
        if (!(L.lmax > 0.0) || !std::isfinite(L.lmax)) {
            set_last_error("AMG: lambda_max(D^-1 A) of level %zu is %g", l, L.lmax);
            return NBX_ERR_BREAKDOWN;
        }
        {
            const double hi = 1.1 * L.lmax, lo = L.lmax / 30.0;
            const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
            double rho = 1.0 / sigma;
            L.inv_theta = (float)(1.0 / theta);
            for (int k = 1; k < L.degree; ++k) {
                const double rn = 1.0 / (2.0 * sigma - rho);
                L.c1[k] = (float)(rn * rho);
                L.c2[k] = (float)(2.0 * rn / delta);
                rho = rn;
            }
        }
        }
        if (l + 1 == A->lv.size()) {
            // coarsest: its inverse, on the host (at most AMG_COARSEST^2 entries)
            std::vector<uint32_t> ptr(L.n + 1), col(L.a_nnz);
            std::vector<double> a(L.a_nnz);
            NBX_HIP(hipMemcpyAsync(ptr.data(), L.a_ptr, ptr.size() * 4, hipMemcpyDeviceToHost, s));
            NBX_HIP(hipMemcpyAsync(col.data(), L.a_col, col.size() * 4, hipMemcpyDeviceToHost, s));
            NBX_HIP(copy_on(a.data(), L.a64, a.size() * 8, hipMemcpyDeviceToHost, s));
            std::vector<double>& inv = A->host_inv;
            std::fill(inv.begin(), inv.end(), 0.0);
            for (uint32_t i = 0; i < L.n; ++i)
                for (uint32_t e = ptr[i]; e < ptr[i + 1]; ++e) inv[(size_t)i * L.n + col[e]] = a[e];
            // symmetrise (P^T A P rounds slightly unsymmetrically)
            for (uint32_t i = 0; i < L.n; ++i)
                for (uint32_t j = 0; j < i; ++j) {
                    const double m = 0.5 * (inv[(size_t)i * L.n + j] + inv[(size_t)j * L.n + i]);
                    inv[(size_t)i * L.n + j] = inv[(size_t)j * L.n + i] = m;
                }
            amg_dense_inverse(L.n, inv.data());
            std::vector<float> f(inv.size());
            for (size_t k = 0; k < inv.size(); ++k) f[k] = (float)inv[k];
            NBX_HIP(hipMemcpyAsync(A->cinv, f.data(), f.size() * 4, hipMemcpyHostToDevice, s));
            NBX_HIP(hipStreamSynchronize(s));
            break;
        }
        AmgDevLevel& C = A->lv[l + 1];
        const double omega = 4.0 / (3.0 * L.lmax);
        with_bs(L.bs, [&](auto rr) {
            constexpr int R = decltype(rr)::value;
            hipLaunchKernelGGL(k_amg_smooth_p<R>,
                               dim3(grid_for((size_t)(L.n / R) * SMOOTH_SLOTS)), dim3(BLOCK), 0,
                               s, L.a_ptr, L.a_col, L.a64, L.dinv64, L.t, L.agg, L.p_ptr,
                               L.p_col, L.p64, omega, L.n / R);
        });
        // P and P^T in block form (D^1/2 P on the finest level when it is
        // smoothed in the scaled variables)
        const double* pdis = l == 0 && A->scaled16 ? S.K.dis : nullptr;
        if (L.bs == 2) {
            hipLaunchKernelGGL((k_bell_fill<2, 3>), dim3(grid_for(L.pb.rows)), dim3(BLOCK), 0, s,
                               L.pb, L.p_ptr, L.p64, pdis);
            hipLaunchKernelGGL((k_bell_fill_t<2>), dim3(grid_for(L.tb.slots)), dim3(BLOCK), 0, s,
                               L.tb, (L.tb.rows + SLICE - 1) / SLICE, L.pt_ptr, L.pt_pos, L.p64,
                               pdis);
        } else {
            hipLaunchKernelGGL((k_bell_fill<3, 3>), dim3(grid_for(L.pb.rows)), dim3(BLOCK), 0, s,
                               L.pb, L.p_ptr, L.p64, pdis);
            hipLaunchKernelGGL((k_bell_fill_t<3>), dim3(grid_for(L.tb.slots)), dim3(BLOCK), 0, s,
                               L.tb, (L.tb.rows + SLICE - 1) / SLICE, L.pt_ptr, L.pt_pos, L.p64,
                               pdis);
        }
        const uint32_t nodes = L.n / L.bs, aggs = L.nc / 3;
        const size_t wblocks = L.w_nnz / (3 * L.bs), cblocks = C.a_nnz / 9;
        with_bs(L.bs, [&](auto rr) {
            constexpr int R = decltype(rr)::value;
            // (the finest level's W blocks sum about nine products each: one
            // lane per block measured fastest there)
            with_out_lanes(R == 2 ? 0 : wblocks, [&](auto ln) {
                constexpr int LN = decltype(ln)::value;
                hipLaunchKernelGGL((k_amg_w<R, LN>), dim3(grid_for(wblocks * LN)), dim3(BLOCK), 0,
                                   s, L.a_ptr, L.a_col, L.a64, L.p_ptr, L.p_col, L.p64, L.w_ptr,
                                   L.w_col, L.w64, nodes, wblocks);
            });
            with_out_lanes(cblocks, [&](auto ln) {
                constexpr int LN = decltype(ln)::value;
                hipLaunchKernelGGL((k_amg_galerkin<R, LN>), dim3(grid_for(cblocks * LN)),
                                   dim3(BLOCK), 0, s, L.pt_ptr, L.pt_row, L.pt_pos, L.p64, L.w_ptr,
                                   L.w_col, L.w64, C.a_ptr, C.a_col, C.a64, aggs, cblocks);
            });
        });
        if (C.ab.rows)
            hipLaunchKernelGGL((k_bell_fill<3, 3>), dim3(grid_for(C.ab.rows)), dim3(BLOCK), 0, s,
                               C.ab, C.a_ptr, C.a64, (const double*)nullptr);
        NBX_HIP(hipGetLastError());
    }
    return NBX_OK;
}

int amg_prepare(FemSys& S) {
    if (!S.K.amg || !S.amg_symbolic_current) {
        // (the host arrays of a large system can exhaust host memory: an
        // error, not an exception through the C interface)
        int rc;
        try {
            rc = amg_symbolic(S);
        } catch (const std::bad_alloc&) {
            set_last_error("AMG: host memory ran out building the hierarchy");
            rc = NBX_ERR_ARG;
        }
        if (rc != NBX_OK) return rc;
        S.amg_symbolic_current = true;
    }
    if (!S.K.amg_current) {
        // the finest level's smoothing reads the fp16 Jacobi-scaled copy
        // (a FEM matrix always has one: its scaled entries lie in [-1, 1])
        AmgDev* A = S.K.amg;
        A->scaled16 = false;
        if (A->lv.size() > 1) {
            NBX_TRY(mat_make_off16(S.K, S.s));
            if (!S.K.has_off16 || !S.K.has_scale) {
                set_last_error("AMG: the matrix has no fp16 Jacobi-scaled copy for the finest "
                               "level's smoothing");
                return NBX_ERR_ARG;
            }
            A->scaled16 = true;
        }
        NBX_TRY(amg_numeric(S));
        S.K.amg_current = true;
    }
    return NBX_OK;
}

AmgFine amg_fine(const AmgDev* A) {
    AmgFine f;
    if (!A || !A->scaled16 || A->lv.size() < 2) return f;
    const AmgDevLevel& L = A->lv[0];
    f.scaled16 = true;
    f.degree = AMG_FINE_DEGREE;
    const double hi = 1.1 * L.lmax, lo = L.lmax / 30.0;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    f.inv_theta = 1.0 / theta;
    for (int k = 1; k < AMG_FINE_DEGREE; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        f.c1[k] = rn * rho;
        f.c2[k] = 2.0 * rn / delta;
        rho = rn;
    }
    f.c = L.d0;
    return f;
}

const double* amg_ones(const AmgDev* A) { return A->ones; }
uint32_t amg_levels(const AmgDev* A) { return (uint32_t)A->lv.size(); }


// b_{l+1} = P_l^T r_l, with the next level's smoothing start unless that
// level is the coarsest
static void restrict_to(AmgDev* A, size_t l, const float* r, const PcgState* st, hipStream_t s) {
    AmgDevLevel& V = A->lv[l];
    AmgDevLevel& W = A->lv[l + 1];
    const bool init = l + 2 < A->lv.size();
    with_bs(V.bs, [&](auto R) {
        constexpr int RR = decltype(R)::value;
        with_bl(V.tb.lanes, [&](auto ln) {
            constexpr int LN = decltype(ln)::value;
            const unsigned g = grid_for((size_t)V.tb.rows * LN);
            if (init)
                hipLaunchKernelGGL((k_brestrict<RR, true, LN>), dim3(g), dim3(BLOCK), 0, s,
                                   bell<3, RR>(V.tb), r, W.b, W.dinv32, W.x, W.r, W.d0,
                                   W.inv_theta, st, V.tb.rows);
            else
                hipLaunchKernelGGL((k_brestrict<RR, false, LN>), dim3(g), dim3(BLOCK), 0, s,
                                   bell<3, RR>(V.tb), r, W.b, W.dinv32, W.x, W.r, W.d0,
                                   W.inv_theta, st, V.tb.rows);
        });
    });
}

// x_l (+)= P_l x_{l+1}
static void prolong_from(AmgDev* A, size_t l, float* x, bool add, const PcgState* st,
                         hipStream_t s) {
    AmgDevLevel& V = A->lv[l];
    const float* xc = A->lv[l + 1].x;
    with_bs(V.bs, [&](auto R) {
        constexpr int RR = decltype(R)::value;
        with_bl(V.pb.lanes, [&](auto ln) {
            constexpr int LN = decltype(ln)::value;
            const unsigned g = grid_for((size_t)V.pb.rows * LN);
            if (add)
                hipLaunchKernelGGL((k_bprolong<RR, true, LN>), dim3(g), dim3(BLOCK), 0, s,
                                   bell<RR, 3>(V.pb), xc, x, st, V.pb.rows);
            else
                hipLaunchKernelGGL((k_bprolong<RR, false, LN>), dim3(g), dim3(BLOCK), 0, s,
                                   bell<RR, 3>(V.pb), xc, x, st, V.pb.rows);
        });
    });
}

// The V-cycle from level l0 >= 1 down, x = M^-1 b on that level (b, x: that
// level's arrays; the finest level is smoothed by the flexible CG's own
// kernels on the fp16 copy).  The restriction has already started each
// level's pre-smoothing (r, d0, x).  A one-level hierarchy is its dense
// solve alone (l0 = 0).
static void cycle_from(AmgDev* A, size_t l0, const float* b, float* x, const PcgState* st,
                       hipStream_t s) {
    const size_t L = A->lv.size();
    for (size_t l = l0; l + 1 < L; ++l) {
        AmgDevLevel& V = A->lv[l];
        const float* bl = l == l0 ? b : V.b;
        float* xl = l == l0 ? x : V.x;
        float *din = V.d0, *dout = V.d1;
        with_bl(V.ab.lanes, [&](auto ln) {
            constexpr int LN = decltype(ln)::value;
            const unsigned g = grid_for((size_t)V.ab.rows * LN);
            for (int k = 1; k < V.degree; ++k) {
                hipLaunchKernelGGL((k_bsm_step<false, LN>), dim3(g), dim3(BLOCK), 0, s,
                                   bell<3, 3>(V.ab), V.dinv32, din, dout, V.r, xl, V.c1[k],
                                   V.c2[k], st, V.ab.rows);
                std::swap(din, dout);
            }
            hipLaunchKernelGGL(k_bresidual<LN>, dim3(g), dim3(BLOCK), 0, s, bell<3, 3>(V.ab), bl,
                               xl, V.r, st, V.ab.rows);
        });
        restrict_to(A, l, V.r, st, s);
    }
    {
        AmgDevLevel& C = A->lv[L - 1];
        hipLaunchKernelGGL(k_amg_dense, dim3(1), dim3(DENSE_BLOCK), 0, s, A->cinv,
                           L == l0 + 1 ? b : C.b, L == l0 + 1 ? x : C.x, st, C.n);
    }
    for (size_t l = L - 1; l-- > l0;) {
        AmgDevLevel& V = A->lv[l];
        const float* bl = l == l0 ? b : V.b;
        float* xl = l == l0 ? x : V.x;
        prolong_from(A, l, xl, true, st, s);
        float *din = V.d0, *dout = V.d1;
        with_bl(V.ab.lanes, [&](auto ln) {
            constexpr int LN = decltype(ln)::value;
            const unsigned g = grid_for((size_t)V.ab.rows * LN);
            hipLaunchKernelGGL(k_bsm_init<LN>, dim3(g), dim3(BLOCK), 0, s, bell<3, 3>(V.ab),
                               V.dinv32, bl, xl, V.r, V.d0, V.inv_theta, st, V.ab.rows);
            for (int k = 1; k < V.degree; ++k) {
                if (k == 1)
                    hipLaunchKernelGGL((k_bsm_step<true, LN>), dim3(g), dim3(BLOCK), 0, s,
                                       bell<3, 3>(V.ab), V.dinv32, din, dout, V.r, xl, V.c1[k],
                                       V.c2[k], st, V.ab.rows);
                else
                    hipLaunchKernelGGL((k_bsm_step<false, LN>), dim3(g), dim3(BLOCK), 0, s,
                                       bell<3, 3>(V.ab), V.dinv32, din, dout, V.r, xl, V.c1[k],
                                       V.c2[k], st, V.ab.rows);
                std::swap(din, dout);
            }
        });
    }
}

// z = C^-1 r for a one-level hierarchy (the finest level is the coarsest)
int amg_vcycle(AmgDev* A, const float* b, float* x, const PcgState* st, hipStream_t s) {
    if (A->lv.size() != 1) {
        set_last_error("AMG: the V-cycle of a multi-level hierarchy runs on the fp16 copy");
        return NBX_ERR_ARG;
    }
    cycle_from(A, 0, b, x, st, s);
    NBX_HIP(hipGetLastError());
    return NBX_OK;
}

int amg_coarse_correction(AmgDev* A, const float* r, float* c, const PcgState* st,
                          hipStream_t s) {
    restrict_to(A, 0, r, st, s);
    cycle_from(A, 1, A->lv[1].b, A->lv[1].x, st, s);
    prolong_from(A, 0, c, false, st, s);
    NBX_HIP(hipGetLastError());
    return NBX_OK;
}

}  // namespace nbx

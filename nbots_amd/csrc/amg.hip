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
#include <type_traits>
#include <vector>

#include "amg.hpp"
#include "fem.hpp"
#include "rows.hpp"

namespace nbx {

struct AmgDevLevel {
    uint32_t n = 0, nc = 0;
    size_t a_nnz = 0, p_nnz = 0, w_nnz = 0;
    uint32_t *a_ptr = nullptr, *a_col = nullptr;
    double* a64 = nullptr;
    float* a32 = nullptr;
    double* dinv64 = nullptr;
    float* dinv32 = nullptr;
    double lmax = 0.0;
    double* t = nullptr;
    uint32_t* self = nullptr;
    uint32_t *p_ptr = nullptr, *p_col = nullptr;
    double* p64 = nullptr;
    float* p32 = nullptr;
    uint32_t* ap_map = nullptr;
    uint32_t *w_ptr = nullptr, *w_col = nullptr;
    double* w64 = nullptr;
    uint32_t *pt_ptr = nullptr, *pt_row = nullptr, *pt_pos = nullptr;
    // V-cycle vectors: right-hand side and solution of this level (the
    // finest level's come from the caller), residual, two directions
    float *b = nullptr, *x = nullptr, *r = nullptr, *d0 = nullptr, *d1 = nullptr;
    // Chebyshev smoothing on [lambda/30, 1.1 lambda]
    float inv_theta = 0.f, c1[AMG_DEGREE] = {}, c2[AMG_DEGREE] = {};
    int la = 8, lp = 8, lr = 16;   // lanes per row: A, P, P^T lists
};

struct AmgDev {
    std::vector<AmgDevLevel> lv;
    uint32_t* fine_map = nullptr;   // [nnz of the finest level] index into the 2x2-block values
    float* cinv = nullptr;          // coarsest inverse, n x n
    double* ones = nullptr;         // [finest n + 2] the flexible CG's unit scaling
    double* pow_x = nullptr;        // power iteration: [finest n + 2] x, y
    double* pow_y = nullptr;
    double* scal = nullptr;         // [2] device scalars
    double* partials = nullptr;     // [MAX_GRID] reduction partials
    std::vector<double> host_inv;
    size_t bytes = 0;
    // the finest level is smoothed on the matrix's fp16 scaled copy, in
    // scaled variables; its P values are stored as D^1/2 P (FP32)
    bool scaled16 = false;
};

namespace {

template <class T>
int up(T** d, const std::vector<T>& h, size_t& bytes) {
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

__global__ __launch_bounds__(BLOCK) void k_amg_fine_vals(const double* __restrict__ kval,
                                                         const uint32_t* __restrict__ map,
                                                         double* __restrict__ a, size_t nnz) {
    const size_t e = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e < nnz) a[e] = kval[map[e]];
}

__global__ __launch_bounds__(BLOCK) void k_amg_to32(const double* __restrict__ in,
                                                    float* __restrict__ out, size_t n) {
    const size_t e = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e < n) out[e] = (float)in[e];
}

// the finest P for the scaled variables: row i times D_ii^1/2 = 1 / dis_i
__global__ __launch_bounds__(BLOCK) void k_amg_p_scaled(const uint32_t* __restrict__ p_ptr,
                                                        const double* __restrict__ p,
                                                        const double* __restrict__ dis,
                                                        float* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double sq = 1.0 / dis[i];
    for (uint32_t q = p_ptr[i]; q < p_ptr[i + 1]; ++q) out[q] = (float)(p[q] * sq);
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

// P = T - omega D^-1 A T, row i
__global__ __launch_bounds__(BLOCK) void k_amg_smooth_p(
    const uint32_t* __restrict__ a_ptr, const uint32_t* __restrict__ a_col,
    const double* __restrict__ a, const double* __restrict__ dinv, const double* __restrict__ t,
    const uint32_t* __restrict__ self, const uint32_t* __restrict__ p_ptr,
    const uint32_t* __restrict__ ap_map, double* __restrict__ p, double omega, uint32_t n) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t pb = p_ptr[i], pe = p_ptr[i + 1];
    for (uint32_t q = pb; q < pe; ++q) p[q] = 0.0;
    // positions index triples of the row (checked: the host built them so)
    const uint32_t sq = self[i];
    if (sq != AMG_NONE && sq >= pb && sq + 3 <= pe)
        for (int m = 0; m < 3; ++m) p[sq + m] += t[3 * (size_t)i + m];
    const double od = omega * dinv[i];
    for (uint32_t e = a_ptr[i]; e < a_ptr[i + 1]; ++e) {
        const uint32_t q = ap_map[e];
        if (q == AMG_NONE || q < pb || q + 3 > pe) continue;
        const double coef = od * a[e];
        const uint32_t j = a_col[e];
        for (int m = 0; m < 3; ++m) p[q + m] -= coef * t[3 * (size_t)j + m];
    }
}

// W = A P and A_c = P^T W take one wavefront per output row.  Each lane owns
// up to 64 consecutive entries of the row at a time; the wavefront walks the
// row's terms in order, loads the rows it multiplies 64 entries at a time,
// and every lane picks (by broadcast of each entry) the ones whose column it
// owns.  So each output entry sums its products in the pattern's order, with
// no atomics and no lookups; the host built every output row to contain
// every product's column.
__device__ __forceinline__ uint32_t bcast_u32(uint32_t v, int t) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, t);
}
__device__ __forceinline__ double bcast_f64(double v, int t) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, t);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), t);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// acc += s * (the row [b, e) of (col, val) restricted to column mycol)
__device__ __forceinline__ double pick_row(const uint32_t* __restrict__ col,
                                           const double* __restrict__ val, uint32_t b, uint32_t e,
                                           uint32_t mycol, double s, double acc, int lane) {
    for (uint32_t o = b; o < e; o += WAVE) {
        const uint32_t q = o + (uint32_t)lane;
        const uint32_t qc = q < e ? col[q] : AMG_NONE;
        const double qv = q < e ? val[q] : 0.0;
        const int cnt = (int)((e - o) < (uint32_t)WAVE ? (e - o) : (uint32_t)WAVE);
        for (int t = 0; t < cnt; ++t) {
            const uint32_t ct = bcast_u32(qc, t);
            const double vt = bcast_f64(qv, t);
            if (mycol == ct) acc += s * vt;
        }
    }
    return acc;
}

__global__ __launch_bounds__(BLOCK) void k_amg_w(
    const uint32_t* __restrict__ a_ptr, const uint32_t* __restrict__ a_col,
    const double* __restrict__ a, const uint32_t* __restrict__ p_ptr,
    const uint32_t* __restrict__ p_col, const double* __restrict__ p,
    const uint32_t* __restrict__ w_ptr, const uint32_t* __restrict__ w_col,
    double* __restrict__ w, uint32_t n) {
    const uint32_t i = (uint32_t)(((size_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE);
    const int lane = (int)(threadIdx.x % WAVE);
    if (i >= n) return;   // uniform in the wavefront
    const uint32_t wb = w_ptr[i], we = w_ptr[i + 1];
    for (uint32_t o = wb; o < we; o += WAVE) {
        const uint32_t my = o + (uint32_t)lane;
        const uint32_t mycol = my < we ? w_col[my] : AMG_NONE;
        double acc = 0.0;
        for (uint32_t e = a_ptr[i]; e < a_ptr[i + 1]; ++e) {
            const uint32_t j = a_col[e];
            acc = pick_row(p_col, p, p_ptr[j], p_ptr[j + 1], mycol, a[e], acc, lane);
        }
        if (my < we) w[my] = acc;
    }
}

__global__ __launch_bounds__(BLOCK) void k_amg_galerkin(
    const uint32_t* __restrict__ pt_ptr, const uint32_t* __restrict__ pt_row,
    const uint32_t* __restrict__ pt_pos, const double* __restrict__ p,
    const uint32_t* __restrict__ w_ptr, const uint32_t* __restrict__ w_col,
    const double* __restrict__ w, const uint32_t* __restrict__ c_ptr,
    const uint32_t* __restrict__ c_col, double* __restrict__ ac, uint32_t nc) {
    const uint32_t c = (uint32_t)(((size_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE);
    const int lane = (int)(threadIdx.x % WAVE);
    if (c >= nc) return;   // uniform in the wavefront
    const uint32_t cb = c_ptr[c], ce = c_ptr[c + 1];
    for (uint32_t o = cb; o < ce; o += WAVE) {
        const uint32_t my = o + (uint32_t)lane;
        const uint32_t mycol = my < ce ? c_col[my] : AMG_NONE;
        double acc = 0.0;
        for (uint32_t k = pt_ptr[c]; k < pt_ptr[c + 1]; ++k) {
            const uint32_t i = pt_row[k];
            acc = pick_row(w_col, w, w_ptr[i], w_ptr[i + 1], mycol, p[pt_pos[k]], acc, lane);
        }
        if (my < ce) ac[my] = acc;
    }
}

// ---- V-cycle (FP32) ----------------------------------------------------------
// Rows are taken by groups of LANES consecutive lanes: each lane sums every
// LANES-th entry of its row, the group adds the lanes' sums by a butterfly,
// and the group's first lane does the row's update.  A row's entries are
// contiguous, so a group's loads are too.  The order of the additions is
// fixed by the pattern, so runs are reproducible.

template <int LANES>
__device__ __forceinline__ float group_sum(float s) {
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LANES);
    return s;
}

template <int LANES>
__device__ __forceinline__ float csr_row(const uint32_t* __restrict__ ptr,
                                         const uint32_t* __restrict__ col,
                                         const float* __restrict__ a, const float* __restrict__ x,
                                         uint32_t i, uint32_t lane, bool valid) {
    float s = 0.f;
    if (valid)
        for (uint32_t e = ptr[i] + lane; e < ptr[i + 1]; e += LANES) s += a[e] * x[col[e]];
    return group_sum<LANES>(s);
}

// the group's row and lane; `valid` is false for the groups past the last row
// (they still take part in the butterfly, with zeros)
#define AMG_ROW(LANES, n)                                                           \
    const size_t gid_ = (size_t)blockIdx.x * BLOCK + threadIdx.x;                  \
    const uint32_t i = (uint32_t)(gid_ / LANES), lane = (uint32_t)(gid_ % LANES); \
    const bool valid = i < (n)

// smoothing start: r = D^-1 (b - A x) (ZERO: x = 0, r = D^-1 b), d = r / theta;
// ZERO also sets x = d (otherwise the first step adds d)
template <bool ZERO, int LANES>
__global__ __launch_bounds__(BLOCK) void k_amg_sm_init(
    const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ col,
    const float* __restrict__ a, const float* __restrict__ dinv, const float* __restrict__ b,
    float* __restrict__ x, float* __restrict__ r, float* __restrict__ d, float inv_theta,
    const PcgState* st, uint32_t n) {
    if (st->done) return;
    AMG_ROW(LANES, n);
    const float ax = ZERO ? 0.f : csr_row<LANES>(ptr, col, a, x, i, lane, valid);
    if (!valid || lane != 0) return;
    const float ri = dinv[i] * (b[i] - ax);
    const float di = ri * inv_theta;
    r[i] = ri;
    d[i] = di;
    if (ZERO) x[i] = di;
}

// one step: r -= D^-1 A d_in; d_out = c1 d_in + c2 r; x += d_out (and d_in
// when ADD_IN: the first step after a start that left x alone)
template <bool ADD_IN, int LANES>
__global__ __launch_bounds__(BLOCK) void k_amg_sm_step(
    const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ col,
    const float* __restrict__ a, const float* __restrict__ dinv, const float* __restrict__ din,
    float* __restrict__ dout, float* __restrict__ r, float* __restrict__ x, float c1, float c2,
    const PcgState* st, uint32_t n) {
    if (st->done) return;
    AMG_ROW(LANES, n);
    const float ad = csr_row<LANES>(ptr, col, a, din, i, lane, valid);
    if (!valid || lane != 0) return;
    const float di = din[i];
    const float ri = r[i] - dinv[i] * ad;
    const float dn = c1 * di + c2 * ri;
    r[i] = ri;
    dout[i] = dn;
    x[i] += ADD_IN ? di + dn : dn;
}

template <int LANES>
__global__ __launch_bounds__(BLOCK) void k_amg_residual(
    const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ col,
    const float* __restrict__ a, const float* __restrict__ b, const float* __restrict__ x,
    float* __restrict__ r, const PcgState* st, uint32_t n) {
    if (st->done) return;
    AMG_ROW(LANES, n);
    const float ax = csr_row<LANES>(ptr, col, a, x, i, lane, valid);
    if (valid && lane == 0) r[i] = b[i] - ax;
}

// b_c = P^T r, coarse row c from its (row, position) list
template <int LANES>
__global__ __launch_bounds__(BLOCK) void k_amg_restrict(
    const uint32_t* __restrict__ pt_ptr, const uint32_t* __restrict__ pt_row,
    const uint32_t* __restrict__ pt_pos, const float* __restrict__ p, const float* __restrict__ r,
    float* __restrict__ bc, const PcgState* st, uint32_t nc) {
    if (st->done) return;
    AMG_ROW(LANES, nc);
    float s = 0.f;
    if (valid)
        for (uint32_t k = pt_ptr[i] + lane; k < pt_ptr[i + 1]; k += LANES)
            s += p[pt_pos[k]] * r[pt_row[k]];
    s = group_sum<LANES>(s);
    if (valid && lane == 0) bc[i] = s;
}

// x += P x_c (ADD), or x = P x_c
template <bool ADD, int LANES>
__global__ __launch_bounds__(BLOCK) void k_amg_prolong(
    const uint32_t* __restrict__ p_ptr, const uint32_t* __restrict__ p_col,
    const float* __restrict__ p, const float* __restrict__ xc, float* __restrict__ x,
    const PcgState* st, uint32_t n) {
    if (st->done) return;
    AMG_ROW(LANES, n);
    const float px = csr_row<LANES>(p_ptr, p_col, p, xc, i, lane, valid);
    if (valid && lane == 0) x[i] = ADD ? x[i] + px : px;
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

namespace {

// lanes per row for a row length (entries per row on average)
int pick_lanes(size_t entries, uint32_t rows) {
    const double avg = rows ? (double)entries / rows : 0.0;
    return avg <= 6.0 ? 4 : avg <= 14.0 ? 8 : avg <= 40.0 ? 16 : 32;
}

template <class F>
void with_lanes(int lanes, F&& f) {
    switch (lanes) {
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        default: f(std::integral_constant<int, 32>{}); break;
    }
}

unsigned rows_grid(uint32_t rows, int lanes) { return grid_for((size_t)rows * (size_t)lanes); }

}  // namespace

void amg_free(AmgDev* A) {
    if (!A) return;
    for (AmgDevLevel& L : A->lv) {
        void* ps[] = {L.a_ptr, L.a_col, L.a64, L.a32, L.dinv64, L.dinv32, L.t, L.self, L.p_ptr,
                      L.p_col, L.p64, L.p32, L.ap_map, L.w_ptr, L.w_col, L.w64, L.pt_ptr,
                      L.pt_row, L.pt_pos, L.b, L.x, L.r, L.d0, L.d1};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
    void* ps[] = {A->fine_map, A->cinv, A->ones, A->pow_x, A->pow_y, A->scal, A->partials};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete A;
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
    std::vector<uint32_t> rp(n + 1), col(nnz);
    std::vector<double> val(nnz), xy(2 * (size_t)S.nn);
    NBX_TRY(fem_download_csr(S, rp.data(), col.data(), val.data()));
    NBX_HIP(copy_on(xy.data(), S.nod, xy.size() * sizeof(double), hipMemcpyDeviceToHost, S.s));
    AmgHost H;
    if (amg_build_host(n, rp.data(), col.data(), val.data(), xy.data(), H)) {
        set_last_error("AMG: the matrix is not a 2-dof node pattern");
        return NBX_ERR_ARG;
    }
    if (H.lv.back().n > AMG_COARSEST) {
        set_last_error("AMG: coarsening stopped at %u unknowns (at most %u are solved directly)",
                       H.lv.back().n, AMG_COARSEST);
        return NBX_ERR_ARG;
    }
    // index of every CSR entry's value in the 2x2-block layout: row 2v+k,
    // block idx (the k-th block of node v) -> ent(idx, k, l)
    std::vector<uint32_t> map(nnz);
    {
        const Sell2& K = S.K.s;
        std::vector<uint32_t> sp(K.n_slices + 1);
        std::vector<uint8_t> len(S.nn + 1);
        NBX_HIP(copy_on(sp.data(), K.slice_ptr, (K.n_slices + 1) * 4, hipMemcpyDeviceToHost, S.s));
        if (S.nn) NBX_HIP(copy_on(len.data(), K.rowlen, S.nn, hipMemcpyDeviceToHost, S.s));
        for (uint32_t v = 0; v < S.nn; ++v) {
            size_t idx = (size_t)sp[v >> 6] + (v & 63);
            for (uint32_t k = 0; k < len[v]; ++k, idx += SLICE)
                for (int i = 0; i < 2; ++i) {
                    const uint32_t o = rp[2 * v + i] + 2 * k;
                    map[o] = (uint32_t)ent(K.n_blocks, idx, i, 0);
                    map[o + 1] = (uint32_t)ent(K.n_blocks, idx, i, 1);
                }
        }
    }
    AmgDev* A = new AmgDev();
    S.K.amg = A;
    NBX_TRY(up(&A->fine_map, map, A->bytes));
    A->lv.resize(H.lv.size());
    for (size_t l = 0; l < H.lv.size(); ++l) {
        const AmgHostLevel& h = H.lv[l];
        AmgDevLevel& L = A->lv[l];
        L.n = h.n;
        L.nc = h.nc;
        L.a_nnz = h.a_col.size();
        NBX_TRY(up(&L.a_ptr, h.a_ptr, A->bytes));
        NBX_TRY(up(&L.a_col, h.a_col, A->bytes));
        NBX_TRY(alloc(&L.a64, L.a_nnz, A->bytes));
        NBX_TRY(alloc(&L.a32, L.a_nnz, A->bytes));
        NBX_TRY(alloc(&L.dinv64, L.n, A->bytes));
        NBX_TRY(alloc(&L.dinv32, L.n, A->bytes));
        NBX_TRY(alloc(&L.r, L.n, A->bytes));
        NBX_TRY(alloc(&L.d0, L.n, A->bytes));
        NBX_TRY(alloc(&L.d1, L.n, A->bytes));
        if (l > 0) {
            NBX_TRY(alloc(&L.b, L.n, A->bytes));
            NBX_TRY(alloc(&L.x, L.n, A->bytes));
        }
        L.la = pick_lanes(L.a_nnz, L.n);
        if (l + 1 < H.lv.size()) {
            L.p_nnz = h.p_col.size();
            L.lp = pick_lanes(L.p_nnz, L.n);
            L.lr = pick_lanes(L.p_nnz, L.nc);
            L.w_nnz = h.w_col.size();
            NBX_TRY(up(&L.t, h.t, A->bytes));
            NBX_TRY(up(&L.self, h.self, A->bytes));
            NBX_TRY(up(&L.p_ptr, h.p_ptr, A->bytes));
            NBX_TRY(up(&L.p_col, h.p_col, A->bytes));
            NBX_TRY(alloc(&L.p64, L.p_nnz, A->bytes));
            NBX_TRY(alloc(&L.p32, L.p_nnz, A->bytes));
            NBX_TRY(up(&L.ap_map, h.ap_map, A->bytes));
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

// Numeric phase for the current values of K.
static int amg_numeric(FemSys& S) {
    AmgDev* A = S.K.amg;
    const hipStream_t s = S.s;
    AmgDevLevel& L0 = A->lv[0];
    hipLaunchKernelGGL(k_amg_fine_vals, dim3(grid_for(L0.a_nnz)), dim3(BLOCK), 0, s, S.K.s.val,
                       A->fine_map, L0.a64, L0.a_nnz);
    NBX_HIP(hipGetLastError());
    for (size_t l = 0; l < A->lv.size(); ++l) {
        AmgDevLevel& L = A->lv[l];
        const unsigned g = grid_for(L.n);
        hipLaunchKernelGGL(k_amg_dinv, dim3(g), dim3(BLOCK), 0, s, L.a_ptr, L.a_col, L.a64,
                           L.dinv64, L.dinv32, L.n);
        if (!(l == 0 && A->scaled16))   // (that level's products read the fp16 copy)
            hipLaunchKernelGGL(k_amg_to32, dim3(grid_for(L.a_nnz)), dim3(BLOCK), 0, s, L.a64,
                               L.a32, L.a_nnz);
        // lambda_max(D^-1 A): 15 power steps from a fixed start
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
            for (int k = 1; k < AMG_DEGREE; ++k) {
                const double rn = 1.0 / (2.0 * sigma - rho);
                L.c1[k] = (float)(rn * rho);
                L.c2[k] = (float)(2.0 * rn / delta);
                rho = rn;
            }
        }
        if (l + 1 == A->lv.size()) {
            // coarsest: its inverse, on the host (at most AMG_COARSEST^2 entries)
            std::vector<uint32_t> ptr(L.n + 1), col(L.a_nnz);
            std::vector<double> a(L.a_nnz);
            NBX_HIP(copy_on(ptr.data(), L.a_ptr, ptr.size() * 4, hipMemcpyDeviceToHost, s));
            NBX_HIP(copy_on(col.data(), L.a_col, col.size() * 4, hipMemcpyDeviceToHost, s));
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
        hipLaunchKernelGGL(k_amg_smooth_p, dim3(g), dim3(BLOCK), 0, s, L.a_ptr, L.a_col, L.a64,
                           L.dinv64, L.t, L.self, L.p_ptr, L.ap_map, L.p64, omega, L.n);
        if (l == 0 && A->scaled16)
            hipLaunchKernelGGL(k_amg_p_scaled, dim3(g), dim3(BLOCK), 0, s, L.p_ptr, L.p64, S.K.dis,
                               L.p32, L.n);
        else
            hipLaunchKernelGGL(k_amg_to32, dim3(grid_for(L.p_nnz)), dim3(BLOCK), 0, s, L.p64,
                               L.p32, L.p_nnz);
        hipLaunchKernelGGL(k_amg_w, dim3(grid_for((size_t)L.n * WAVE)), dim3(BLOCK), 0, s, L.a_ptr,
                           L.a_col, L.a64, L.p_ptr, L.p_col, L.p64, L.w_ptr, L.w_col, L.w64, L.n);
        hipLaunchKernelGGL(k_amg_galerkin, dim3(grid_for((size_t)L.nc * WAVE)), dim3(BLOCK), 0, s,
                           L.pt_ptr, L.pt_row, L.pt_pos, L.p64, L.w_ptr, L.w_col, L.w64, C.a_ptr,
                           C.a_col, C.a64, L.nc);
        NBX_HIP(hipGetLastError());
    }
    return NBX_OK;
}

int amg_prepare(FemSys& S) {
    if (!S.K.amg || !S.amg_symbolic_current) {
        NBX_TRY(amg_symbolic(S));
        S.amg_symbolic_current = true;
    }
    if (!S.K.amg_current) {
        // the finest level's smoothing reads the fp16 scaled copy when the
        // matrix allows one (values in fp16 range), its own FP32 CSR copy
        // otherwise
        AmgDev* A = S.K.amg;
        A->scaled16 = false;
        if (A->lv.size() > 1) {
            const int rc = mat_make_off16(S.K, S.s);
            if (rc != NBX_OK && rc != NBX_ERR_ARG) return rc;
            A->scaled16 = rc == NBX_OK && S.K.has_off16 && S.K.has_scale;
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
    f.degree = AMG_DEGREE;
    const double hi = 1.1 * L.lmax, lo = L.lmax / 30.0;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    f.inv_theta = 1.0 / theta;
    for (int k = 1; k < AMG_DEGREE; ++k) {
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


// The V-cycle from level l0 down, x = M^-1 b on that level (b, x: that
// level's arrays).
static void cycle_from(AmgDev* A, size_t l0, const float* b, float* x, const PcgState* st,
                       hipStream_t s) {
    const size_t L = A->lv.size();
    for (size_t l = l0; l + 1 < L; ++l) {
        AmgDevLevel& V = A->lv[l];
        const float* bl = l == l0 ? b : V.b;
        float* xl = l == l0 ? x : V.x;
        hipLaunchKernelGGL((k_amg_sm_init<true, 1>), dim3(rows_grid(V.n, 1)), dim3(BLOCK), 0, s,
                           V.a_ptr, V.a_col, V.a32, V.dinv32, bl, xl, V.r, V.d0, V.inv_theta, st,
                           V.n);
        float *din = V.d0, *dout = V.d1;
        with_lanes(V.la, [&](auto ln) {
            constexpr int LN = decltype(ln)::value;
            for (int k = 1; k < AMG_DEGREE; ++k) {
                hipLaunchKernelGGL((k_amg_sm_step<false, LN>), dim3(rows_grid(V.n, LN)), dim3(BLOCK),
                                   0, s, V.a_ptr, V.a_col, V.a32, V.dinv32, din, dout, V.r, xl,
                                   V.c1[k], V.c2[k], st, V.n);
                std::swap(din, dout);
            }
            hipLaunchKernelGGL((k_amg_residual<LN>), dim3(rows_grid(V.n, LN)), dim3(BLOCK), 0, s,
                               V.a_ptr, V.a_col, V.a32, bl, xl, V.r, st, V.n);
        });
        with_lanes(V.lr, [&](auto ln) {
            constexpr int LN = decltype(ln)::value;
            hipLaunchKernelGGL((k_amg_restrict<LN>), dim3(rows_grid(V.nc, LN)), dim3(BLOCK), 0, s,
                               V.pt_ptr, V.pt_row, V.pt_pos, V.p32, V.r, A->lv[l + 1].b, st, V.nc);
        });
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
        with_lanes(V.lp, [&](auto ln) {
            constexpr int LN = decltype(ln)::value;
            hipLaunchKernelGGL((k_amg_prolong<true, LN>), dim3(rows_grid(V.n, LN)), dim3(BLOCK), 0,
                               s, V.p_ptr, V.p_col, V.p32, A->lv[l + 1].x, xl, st, V.n);
        });
        with_lanes(V.la, [&](auto ln) {
            constexpr int LN = decltype(ln)::value;
            hipLaunchKernelGGL((k_amg_sm_init<false, LN>), dim3(rows_grid(V.n, LN)), dim3(BLOCK), 0,
                               s, V.a_ptr, V.a_col, V.a32, V.dinv32, bl, xl, V.r, V.d0, V.inv_theta,
                               st, V.n);
            float *din = V.d0, *dout = V.d1;
            for (int k = 1; k < AMG_DEGREE; ++k) {
                if (k == 1)
                    hipLaunchKernelGGL((k_amg_sm_step<true, LN>), dim3(rows_grid(V.n, LN)),
                                       dim3(BLOCK), 0, s, V.a_ptr, V.a_col, V.a32, V.dinv32, din,
                                       dout, V.r, xl, V.c1[k], V.c2[k], st, V.n);
                else
                    hipLaunchKernelGGL((k_amg_sm_step<false, LN>), dim3(rows_grid(V.n, LN)),
                                       dim3(BLOCK), 0, s, V.a_ptr, V.a_col, V.a32, V.dinv32, din,
                                       dout, V.r, xl, V.c1[k], V.c2[k], st, V.n);
                std::swap(din, dout);
            }
        });
    }
}

// z = V-cycle(r) on the finest level; b and x are FP32 arrays of the
// finest level's length
int amg_vcycle(AmgDev* A, const float* b, float* x, const PcgState* st, hipStream_t s) {
    cycle_from(A, 0, b, x, st, s);
    NBX_HIP(hipGetLastError());
    return NBX_OK;
}

int amg_coarse_correction(AmgDev* A, const float* r, float* c, const PcgState* st,
                          hipStream_t s) {
    AmgDevLevel& V = A->lv[0];
    AmgDevLevel& W = A->lv[1];
    with_lanes(V.lr, [&](auto ln) {
        constexpr int LN = decltype(ln)::value;
        hipLaunchKernelGGL((k_amg_restrict<LN>), dim3(rows_grid(V.nc, LN)), dim3(BLOCK), 0, s,
                           V.pt_ptr, V.pt_row, V.pt_pos, V.p32, r, W.b, st, V.nc);
    });
    cycle_from(A, 1, W.b, W.x, st, s);
    with_lanes(V.lp, [&](auto ln) {
        constexpr int LN = decltype(ln)::value;
        hipLaunchKernelGGL((k_amg_prolong<false, LN>), dim3(rows_grid(V.n, LN)), dim3(BLOCK), 0,
                           s, V.p_ptr, V.p_col, V.p32, W.x, c, st, V.n);
    });
    NBX_HIP(hipGetLastError());
    return NBX_OK;
}

}  // namespace nbx

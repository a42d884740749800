// Device matrices for gfx950: upload of a host CSR into the 2x2-block
// sliced-ELL format (or scalar CSR when the pattern is not 2-dof blocked),
// diagonal extraction, the Jacobi-scaled FP32 copy, SpMV and its benchmark,
// and solver workspaces.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "rows.hpp"

namespace nbx {

static thread_local char g_err[512];   // per thread: handles are used from several threads
void set_last_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
const char* last_error() { return g_err; }


// y = A x, one node (or scalar row) per thread
template <class Op>
__global__ __launch_bounds__(BLOCK) void k_spmv(Op op, const typename Op::V* __restrict__ x,
                                                typename Op::V* __restrict__ y, uint32_t n) {
    const Chunk c = my_chunk(n);
    for (size_t u = c.beg + threadIdx.x; u < c.end; u += BLOCK) y[u] = op.row((uint32_t)u, x);
}

template <int BITS>
__global__ __launch_bounds__(BLOCK) void k_spmv_scaled(Sell2 A, const float2* __restrict__ x,
                                                       float2* __restrict__ y, uint32_t n) {
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK)
        y[i] = row_scaled<BITS>(A, (uint32_t)i, x);
}

__global__ void k_norm2(const double* __restrict__ v, size_t n, double* partials) {
    double acc[1] = {0.0};
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK)
        acc[0] += v[i] * v[i];
    __shared__ double red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    put_partials<1>(acc, partials);
}

__global__ __launch_bounds__(FIN) void k_fin_out(const double* partials, unsigned np, double* out) {
    double tot[1];
    fin_sum<1>(partials, np, tot);
    if (threadIdx.x == 0) *out = tot[0];
}

__global__ void k_diag_sell2(Sell2 A, double* __restrict__ diag) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    size_t idx = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    double d0 = 0.0, d1 = 0.0;
    for (int k = 0; k < A.rowlen[v]; ++k, idx += SLICE) {
        if (A.bcol[idx] == v) {
            d0 = A.val[ent(A.n_blocks, idx, 0, 0)];
            d1 = A.val[ent(A.n_blocks, idx, 1, 1)];
        }
    }
    diag[2 * v] = d0;
    diag[2 * v + 1] = d1;
}

// ---- symmetric copy ---------------------------------------------------------
// Per row: the number of block columns below the diagonal (lower) and above
// it (upper).  A row whose columns are not strictly ascending, or that lacks
// its diagonal block, sets *bad (the symmetric layout then is not used).
__global__ void k_sym_count(Sell2 A, uint8_t* __restrict__ ulen, uint8_t* __restrict__ llen,
                            unsigned* bad) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    const size_t g = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const int len = A.rowlen[v];
    int lo = 0, dg = 0;
    int64_t prev = -1;
    for (int k = 0; k < len; ++k) {
        const uint32_t w = A.bcol[g + (size_t)k * SLICE];
        if ((int64_t)w <= prev) *bad = 1u;
        prev = w;
        lo += w < v;
        dg += w == v;
    }
    if (dg != 1) *bad = 1u;
    ulen[v] = (uint8_t)(len - lo - dg);
    llen[v] = (uint8_t)lo;
}

// Upper columns, and for every lower block (v, w) the pair (w, position of
// the upper block (w, v)); *bad when row w lacks column v.  Row v of the full
// layout holds lo lower blocks, the diagonal block, then the upper ones.
// Padding: upper (column v), lower (v, 0); never read.
__global__ void k_sym_pattern(Sell2 A, Sym2 Y, unsigned* bad) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    const size_t g = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const int lo = Y.l_rowlen[v], up = Y.u_rowlen[v];
    const uint32_t s = v >> 6;
    const size_t u = (size_t)Y.u_slice_ptr[s] + (v & 63);
    const uint32_t uw = (Y.u_slice_ptr[s + 1] - Y.u_slice_ptr[s]) / SLICE;
    for (uint32_t k = 0; k < uw; ++k)
        Y.u_bcol[u + (size_t)k * SLICE] =
            (int)k < up ? A.bcol[g + (size_t)(lo + 1 + (int)k) * SLICE] : v;
    const size_t l = (size_t)Y.l_slice_ptr[s] + (v & 63);
    const uint32_t lw = (Y.l_slice_ptr[s + 1] - Y.l_slice_ptr[s]) / SLICE;
    for (uint32_t k = 0; k < lw; ++k) {
        uint32_t w = v, pos = 0;
        if ((int)k < lo) {
            w = A.bcol[g + (size_t)k * SLICE];
            const size_t gw = (size_t)A.slice_ptr[w >> 6] + (w & 63);
            const int lenw = A.rowlen[w], low = Y.l_rowlen[w];
            int t = -1;
            for (int kk = 0; kk < lenw; ++kk) {
                if (A.bcol[gw + (size_t)kk * SLICE] == v) {
                    t = kk;
                    break;
                }
            }
            if (t <= low) {
                *bad = 1u;
                t = low + 1;
            }
            pos = Y.u_slice_ptr[w >> 6] + (w & 63) + (uint32_t)(t - low - 1) * SLICE;
        }
        Y.l_idx[2 * (l + (size_t)k * SLICE)] = w;
        Y.l_idx[2 * (l + (size_t)k * SLICE) + 1] = pos;
    }
}

// Copy the diagonal and upper blocks' values from the full layout.
__global__ void k_sym_values(Sell2 A, Sym2 Y) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    const size_t g = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const size_t u = (size_t)Y.u_slice_ptr[v >> 6] + (v & 63);
    const int lo = Y.l_rowlen[v], up = Y.u_rowlen[v];
    const size_t gd = g + (size_t)lo * SLICE;
    Y.d_val[2 * (size_t)v] = A.val[ent(A.n_blocks, gd, 0, 0)];
    Y.d_val[2 * (size_t)v + 1] = A.val[ent(A.n_blocks, gd, 0, 1)];
    Y.d_val[2 * (size_t)Y.n_nodes + v] = A.val[ent(A.n_blocks, gd, 1, 1)];
    for (int k = 0; k < up; ++k) {
        const size_t gi = g + (size_t)(lo + 1 + k) * SLICE, ui = u + (size_t)k * SLICE;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                Y.u_val[ent(Y.n_upper, ui, i, j)] = A.val[ent(A.n_blocks, gi, i, j)];
    }
}

// *bad unless every lower block K(v, w) equals the transpose of the stored
// upper block K(w, v), and the diagonal block its own transpose, bit for bit.
__global__ void k_sym_check(Sell2 A, Sym2 Y, unsigned* bad) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    const size_t g = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const size_t l = (size_t)Y.l_slice_ptr[v >> 6] + (v & 63);
    const int lo = Y.l_rowlen[v];
    const size_t gd = g + (size_t)lo * SLICE;
    bool ok = __double_as_longlong(A.val[ent(A.n_blocks, gd, 0, 1)]) ==
              __double_as_longlong(A.val[ent(A.n_blocks, gd, 1, 0)]);
    for (int k = 0; k < lo; ++k) {
        const size_t gi = g + (size_t)k * SLICE;
        const size_t pos = Y.l_idx[2 * (l + (size_t)k * SLICE) + 1];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                ok = ok && __double_as_longlong(A.val[ent(A.n_blocks, gi, i, j)]) ==
                               __double_as_longlong(Y.u_val[ent(Y.n_upper, pos, j, i)]);
    }
    if (!ok) *bad = 1u;
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

// dis = 1/sqrt(diag); val32 = fl32(a_ij * dis_i * dis_j) and
// val16 = fl16(a_ij * dis_i * dis_j), the products in FP64.
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
        const double* a = A.val + ent(A.n_blocks, idx, 0, 0);
        const double* b = A.val + ent(A.n_blocks, idx, 1, 0);
        float4 o;
        o.x = (float)(a[0] * r0 * c0);
        o.y = (float)(a[1] * r0 * c1);
        o.z = (float)(b[0] * r1 * c0);
        o.w = (float)(b[1] * r1 * c1);
        reinterpret_cast<float4*>(A.val32)[idx] = o;
    }
}

// Stored block idx (row v, column w) of D^-1/2 K D^-1/2 in fp16.
__device__ __forceinline__ hv4 scaled16(const Sell2& A, size_t idx, uint32_t w, double r0,
                                        double r1, const double* __restrict__ dis) {
    const double c0 = dis[2 * w], c1 = dis[2 * w + 1];
    const double* a = A.val + ent(A.n_blocks, idx, 0, 0);
    const double* b = A.val + ent(A.n_blocks, idx, 1, 0);
    hv4 o;
    o.x = (_Float16)(float)(a[0] * r0 * c0);
    o.y = (_Float16)(float)(a[1] * r0 * c1);
    o.z = (_Float16)(float)(b[0] * r1 * c0);
    o.w = (_Float16)(float)(b[1] * r1 * c1);
    return o;
}

__global__ void k_scale16(Sell2 A, const double* __restrict__ dis) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    size_t idx = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const int len = A.rowlen[v];
    const double r0 = dis[2 * v], r1 = dis[2 * v + 1];
    for (int k = 0; k < len; ++k, idx += SLICE)
        reinterpret_cast<hv4*>(A.val16)[idx] = scaled16(A, idx, A.bcol[idx], r0, r1, dis);
}

// ---- fp16 scaled copy without the diagonal blocks ------------------------
// Per row: the number of off-diagonal blocks; *bad unless the row holds
// exactly one diagonal block.
__global__ void k_off_count(Sell2 A, uint8_t* __restrict__ olen, unsigned* bad) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    const size_t g = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const int len = A.rowlen[v];
    int dg = 0;
    for (int k = 0; k < len; ++k) dg += A.bcol[g + (size_t)k * SLICE] == v;
    if (dg != 1) *bad = 1u;
    olen[v] = (uint8_t)(dg == 1 ? len - 1 : 0);
}

// Off-diagonal block columns in row order.  Padding: column v, never read.
__global__ void k_off_pattern(Sell2 A, Off16 O) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    const size_t g = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    const size_t o = (size_t)O.slice_ptr[v >> 6] + (v & 63);
    const uint32_t ow = (O.slice_ptr[(v >> 6) + 1] - O.slice_ptr[v >> 6]) / SLICE;
    const int len = A.rowlen[v];
    uint32_t k2 = 0;
    for (int k = 0; k < len; ++k) {
        const uint32_t w = A.bcol[g + (size_t)k * SLICE];
        if (w != v && k2 < ow) O.bcol[o + (size_t)k2++ * SLICE] = w;
    }
    for (; k2 < ow; ++k2) O.bcol[o + (size_t)k2 * SLICE] = v;
}

// O's column deltas w - v, two blocks per word (padding: 0); *bad when one
// does not fit in int16.
__global__ void k_off_pack(Off16 O, unsigned* bad) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= O.n_nodes) return;
    const uint32_t sl = v >> 6;
    const size_t o = (size_t)O.slice_ptr[sl] + (v & 63);
    const size_t pb = (size_t)O.pslice_ptr[sl] + (v & 63);
    const uint32_t pwid = (O.pslice_ptr[sl + 1] - O.pslice_ptr[sl]) / SLICE;
    const int len = O.rowlen[v];
    for (uint32_t j = 0; j < pwid; ++j) {
        uint32_t word = 0;
        for (int h = 0; h < 2; ++h) {
            const int k = 2 * (int)j + h;
            if (k < len) {
                const int64_t d = (int64_t)O.bcol[o + (size_t)k * SLICE] - (int64_t)v;
                if (d < -32768 || d > 32767) *bad = 1u;
                word |= (uint32_t)(uint16_t)(int16_t)d << (16 * h);
            }
        }
        O.pcol[pb + (size_t)j * SLICE] = word;
    }
}

// The same fp16 values as k_scale16: off-diagonal blocks into O's ELL, the
// diagonal block's (K01, K10) into O.od; *bad unless its K00 and K11 are 1.
__global__ void k_scale16_off(Sell2 A, Off16 O, const double* __restrict__ dis, unsigned* bad) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= A.n_nodes) return;
    size_t idx = (size_t)A.slice_ptr[v >> 6] + (v & 63);
    size_t o = (size_t)O.slice_ptr[v >> 6] + (v & 63);
    const int len = A.rowlen[v], olen = O.rowlen[v];
    const double r0 = dis[2 * v], r1 = dis[2 * v + 1];
    int k2 = 0;
    for (int k = 0; k < len; ++k, idx += SLICE) {
        const uint32_t w = A.bcol[idx];
        const hv4 h = scaled16(A, idx, w, r0, r1, dis);
        if (w == v) {
            if (!(h.x == (_Float16)1.0f && h.w == (_Float16)1.0f)) *bad = 1u;
            O.od[2 * (size_t)v] = h.y;
            O.od[2 * (size_t)v + 1] = h.z;
        } else if (k2 < olen) {
            reinterpret_cast<hv4*>(O.val)[o] = h;
            o += SLICE;
            ++k2;
        }
    }
}

// Gershgorin bound of the scaled matrix: max over rows of sum_j |a_ij dis_i dis_j|.
__global__ void k_gersh(Sell2 A, const double* __restrict__ dis, double* partials) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    double m = 0.0;
    if (v < A.n_nodes) {
        size_t idx = (size_t)A.slice_ptr[v >> 6] + (v & 63);
        const double r0 = dis[2 * v], r1 = dis[2 * v + 1];
        double s0 = 0.0, s1 = 0.0;
        for (int k = 0; k < A.rowlen[v]; ++k, idx += SLICE) {
            const uint32_t w = A.bcol[idx];
            const double c0 = dis[2 * w], c1 = dis[2 * w + 1];
            const double* a = A.val + ent(A.n_blocks, idx, 0, 0);
            const double* b = A.val + ent(A.n_blocks, idx, 1, 0);
            s0 += fabs(a[0] * r0 * c0) + fabs(a[1] * r0 * c1);
            s1 += fabs(b[0] * r1 * c0) + fabs(b[1] * r1 * c1);
        }
        m = fmax(s0, s1);
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, WAVE));
    __shared__ double red[BLOCK / WAVE];
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / WAVE; ++w) m = fmax(m, red[w]);
        partials[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(FIN) void k_fin_max(const double* partials, unsigned np, double* out) {
    double m = 0.0;
    for (unsigned b = threadIdx.x; b < np; b += FIN) m = fmax(m, partials[b]);
    for (int off = WAVE / 2; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, WAVE));
    __shared__ double red[FIN / WAVE];
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < FIN / WAVE; ++w) m = fmax(m, red[w]);
        *out = m;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
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
    w.st = nullptr;
    w.h_st = nullptr;
    w.n = 0;
}

static void sym_free(DevMatrix& m) {
    Sym2& Y = m.y;
    void* ps[] = {Y.d_val,       Y.u_slice_ptr, Y.u_rowlen, Y.u_bcol, Y.u_val,
                  Y.l_slice_ptr, Y.l_rowlen,    Y.l_idx,    m.sym_flag};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    Y = Sym2();
    m.sym_flag = nullptr;
    m.sym_pattern = m.sym_checked = m.has_sym = false;
}

static void off_free(DevMatrix& m) {
    Off16& O = m.o;
    void* ps[] = {O.slice_ptr, O.rowlen, O.bcol, O.val, O.od, O.pslice_ptr, O.pcol, m.off_flag};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    O = Off16();
    m.off_flag = nullptr;
    m.off_pattern = m.has_off16 = m.off_packed = false;
}

void mat_free(DevMatrix& m) {
    sym_free(m);
    off_free(m);
    amg_free(m.amg);
    void* ps[] = {m.s.slice_ptr, m.s.rowlen, m.s.bcol, m.s.val, m.s.val32, m.s.val16,
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
    const size_t nb = sptr.back();
    for (uint32_t v = 0; v < nn; ++v) {
        const uint32_t a = rp[2 * v], b = rp[2 * v + 1], len = (b - a) / 2;
        size_t idx = (size_t)sptr[v >> 6] + (v & 63);
        for (uint32_t k = 0; k < len; ++k, idx += SLICE) {
            hv[ent(nb, idx, 0, 0)] = val[a + 2 * k];
            hv[ent(nb, idx, 0, 1)] = val[a + 2 * k + 1];
            hv[ent(nb, idx, 1, 0)] = val[b + 2 * k];
            hv[ent(nb, idx, 1, 1)] = val[b + 2 * k + 1];
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
            if ((uint64_t)sptr[sl] + (uint64_t)wdt * SLICE > 0xffffffffull) {
                set_last_error("pattern exceeds 32-bit slice offsets");
                return NBX_ERR_PATTERN;
            }
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
        mat_scaled_stale(m);
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
    mat_scaled_stale(m);
    return NBX_OK;
}

void mat_scaled_stale(DevMatrix& m) {
    m.amg_current = false;
    m.has_scale = false;
    m.has_val32 = false;
    m.has_val16 = false;
    m.has_off16 = false;
    m.sym_checked = false;
    m.has_sym = false;
}

static int read_flag(unsigned* dflag, hipStream_t s, unsigned* out) {
    NBX_HIP(hipMemcpyAsync(out, dflag, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    return NBX_OK;
}

// Slice offsets of a layout whose row v holds len[v] blocks; false when
// they exceed 32 bits.
static bool slice_offsets(const std::vector<uint8_t>& len, uint32_t ns,
                          std::vector<uint32_t>& sp) {
    const uint32_t nn = (uint32_t)len.size();
    sp.assign(ns + 1, 0);
    for (uint32_t sl = 0; sl < ns; ++sl) {
        uint32_t wdt = 0;
        for (uint32_t v = sl * SLICE; v < nn && v < (sl + 1) * SLICE; ++v)
            if (len[v] > wdt) wdt = len[v];
        if ((uint64_t)sp[sl] + (uint64_t)wdt * SLICE > 0xffffffffull) return false;
        sp[sl + 1] = sp[sl] + wdt * SLICE;
    }
    return true;
}

// The symmetric layout's pattern from the full one.  Leaves m.sym_pattern
// false (and nothing allocated) when the pattern does not allow it: rows
// not strictly ascending, a lower block without its transposed partner,
// offsets beyond 32 bits, or no device memory for the copy.
static int sym_build_pattern(DevMatrix& m, hipStream_t s) {
    const Sell2& A = m.s;
    Sym2& Y = m.y;
    const uint32_t nn = A.n_nodes, ns = A.n_slices;
    auto give_up = [&]() {
        (void)hipGetLastError();
        sym_free(m);
        return NBX_OK;
    };
    if (hipMalloc(&m.sym_flag, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&Y.u_rowlen, (size_t)nn + 1) != hipSuccess ||
        hipMalloc(&Y.l_rowlen, (size_t)nn + 1) != hipSuccess)
        return give_up();
    NBX_HIP(hipMemsetAsync(m.sym_flag, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_sym_count, dim3(grid_for(nn)), dim3(BLOCK), 0, s, A, Y.u_rowlen,
                       Y.l_rowlen, m.sym_flag);
    NBX_HIP(hipGetLastError());
    std::vector<uint8_t> ul(nn), ll(nn);
    NBX_HIP(hipMemcpyAsync(ul.data(), Y.u_rowlen, nn, hipMemcpyDeviceToHost, s));
    NBX_HIP(hipMemcpyAsync(ll.data(), Y.l_rowlen, nn, hipMemcpyDeviceToHost, s));
    unsigned bad = 0;
    NBX_TRY(read_flag(m.sym_flag, s, &bad));
    std::vector<uint32_t> usp, lsp;
    if (bad || !slice_offsets(ul, ns, usp) || !slice_offsets(ll, ns, lsp)) return give_up();
    Y.n_nodes = nn;
    Y.n_slices = ns;
    Y.n_upper = usp[ns];
    Y.n_lower = lsp[ns];
    if (hipMalloc(&Y.u_slice_ptr, (ns + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&Y.l_slice_ptr, (ns + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&Y.u_bcol, (Y.n_upper + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&Y.u_val, (4 * Y.n_upper + 4) * sizeof(double)) != hipSuccess ||
        hipMalloc(&Y.d_val, (3 * (size_t)nn + 2) * sizeof(double)) != hipSuccess ||
        hipMalloc(&Y.l_idx, (2 * Y.n_lower + 2) * sizeof(uint32_t)) != hipSuccess)
        return give_up();
    NBX_HIP(hipMemcpyAsync(Y.u_slice_ptr, usp.data(), (ns + 1) * sizeof(uint32_t),
                           hipMemcpyHostToDevice, s));
    NBX_HIP(hipMemcpyAsync(Y.l_slice_ptr, lsp.data(), (ns + 1) * sizeof(uint32_t),
                           hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_sym_pattern, dim3(grid_for(nn)), dim3(BLOCK), 0, s, A, Y, m.sym_flag);
    NBX_HIP(hipGetLastError());
    NBX_TRY(read_flag(m.sym_flag, s, &bad));
    if (bad) return give_up();
    m.sym_pattern = true;
    return NBX_OK;
}

int mat_prepare_sym(DevMatrix& m, hipStream_t s) {
    if (!m.blocked || m.n == 0) return NBX_OK;
    if (!m.sym_tried) {
        m.sym_tried = true;
        NBX_TRY(sym_build_pattern(m, s));
    }
    if (!m.sym_pattern || m.sym_checked) return NBX_OK;
    const unsigned g = grid_for(m.s.n_nodes);
    NBX_HIP(hipMemsetAsync(m.sym_flag, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_sym_values, dim3(g), dim3(BLOCK), 0, s, m.s, m.y);
    hipLaunchKernelGGL(k_sym_check, dim3(g), dim3(BLOCK), 0, s, m.s, m.y, m.sym_flag);
    NBX_HIP(hipGetLastError());
    unsigned bad = 0;
    NBX_TRY(read_flag(m.sym_flag, s, &bad));
    m.sym_checked = true;
    m.has_sym = bad == 0;
    return NBX_OK;
}

// dis = 1/sqrt(diag) and the Gershgorin bound of the scaled matrix, once per
// set of values.
static int ensure_scale(DevMatrix& m, hipStream_t s) {
    if (m.has_scale) return NBX_OK;
    if (!m.dis) NBX_HIP(hipMalloc(&m.dis, (size_t)m.n * sizeof(double)));
    const unsigned gg = grid_for(m.s.n_nodes);
    unsigned* bad = nullptr;
    double *part = nullptr, *dmax = nullptr;
    NBX_HIP(hipMalloc(&bad, sizeof(unsigned)));
    NBX_HIP(hipMalloc(&part, (size_t)gg * sizeof(double)));
    NBX_HIP(hipMalloc(&dmax, sizeof(double)));
    NBX_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_dis, dim3(grid_for(m.n)), dim3(BLOCK), 0, s, m.diag, m.dis, m.n, bad);
    // Gershgorin bound of the scaled matrix (Chebyshev preconditioner)
    hipLaunchKernelGGL(k_gersh, dim3(gg), dim3(BLOCK), 0, s, m.s, m.dis, part);
    hipLaunchKernelGGL(k_fin_max, dim3(1), dim3(FIN), 0, s, part, gg, dmax);
    NBX_HIP(hipGetLastError());
    unsigned hbad = 0;
    NBX_HIP(hipMemcpyAsync(&hbad, bad, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    NBX_HIP(hipMemcpyAsync(&m.lambda_max, dmax, sizeof(double), hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    (void)hipFree(bad);
    (void)hipFree(part);
    (void)hipFree(dmax);
    if (hbad) {
        set_last_error("matrix diagonal has %u non-positive entries", hbad);
        return NBX_ERR_ARG;
    }
    m.has_scale = true;
    return NBX_OK;
}

int mat_make_scaled(DevMatrix& m, hipStream_t s, int bits) {
    if (!m.blocked) {
        set_last_error("mixed precision needs a 2-dof node matrix");
        return NBX_ERR_ARG;
    }
    if (bits != 16 && bits != 32 && bits != 64) {
        set_last_error("scaled copy of %d-bit values (16, 32 or 64)", bits);
        return NBX_ERR_ARG;
    }
    // 64: the FP64 preconditioner reads the FP64 matrix itself and needs
    // only the scaling and its bound
    NBX_TRY(ensure_scale(m, s));
    if (bits == 32 && !m.has_val32) {
        if (!m.s.val32) NBX_HIP(hipMalloc(&m.s.val32, (4 * m.s.n_blocks + 4) * sizeof(float)));
        hipLaunchKernelGGL(k_scale32, dim3(grid_for(m.s.n_nodes)), dim3(BLOCK), 0, s, m.s, m.dis);
        NBX_HIP(hipGetLastError());
        m.has_val32 = true;
    }
    if (bits == 16 && !m.has_val16) {
        if (!m.s.val16) NBX_HIP(hipMalloc(&m.s.val16, (4 * m.s.n_blocks + 4) * sizeof(_Float16)));
        hipLaunchKernelGGL(k_scale16, dim3(grid_for(m.s.n_nodes)), dim3(BLOCK), 0, s, m.s, m.dis);
        NBX_HIP(hipGetLastError());
        m.has_val16 = true;
    }
    return NBX_OK;
}

// O's pattern from the full one.  Leaves m.off_pattern false (and nothing
// allocated) when a row lacks its diagonal block, the offsets exceed 32
// bits, or there is no device memory for it.
static int off_build_pattern(DevMatrix& m, hipStream_t s) {
    const Sell2& A = m.s;
    Off16& O = m.o;
    const uint32_t nn = A.n_nodes, ns = A.n_slices;
    auto give_up = [&]() {
        (void)hipGetLastError();
        off_free(m);
        return NBX_OK;
    };
    if (hipMalloc(&m.off_flag, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&O.rowlen, (size_t)nn + 1) != hipSuccess)
        return give_up();
    NBX_HIP(hipMemsetAsync(m.off_flag, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_off_count, dim3(grid_for(nn)), dim3(BLOCK), 0, s, A, O.rowlen, m.off_flag);
    NBX_HIP(hipGetLastError());
    std::vector<uint8_t> ol(nn);
    NBX_HIP(hipMemcpyAsync(ol.data(), O.rowlen, nn, hipMemcpyDeviceToHost, s));
    unsigned bad = 0;
    NBX_TRY(read_flag(m.off_flag, s, &bad));
    std::vector<uint32_t> osp;
    if (bad || !slice_offsets(ol, ns, osp)) return give_up();
    // Every slice as wide as the widest row when that adds at most 1/16 of
    // padding: a row's entries then sit where its index alone says, and a
    // product reads them without first loading the slice offset and the
    // row length.
    uint32_t wmax = 0;
    for (uint8_t l : ol) wmax = wmax > l ? wmax : l;
    const uint64_t uni = (uint64_t)ns * SLICE * wmax;
    O.width = 0;
    if (wmax > 0 && uni <= (uint64_t)osp[ns] + osp[ns] / 16 && uni <= 0xffffffffull) {
        for (uint32_t sl = 0; sl <= ns; ++sl) osp[sl] = sl * SLICE * wmax;
        O.width = wmax;
    }
    O.n_nodes = nn;
    O.n_slices = ns;
    O.n_blocks = osp[ns];
    if (hipMalloc(&O.slice_ptr, (ns + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&O.bcol, (O.n_blocks + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&O.val, (4 * O.n_blocks + 4) * sizeof(_Float16)) != hipSuccess ||
        hipMalloc(&O.od, (2 * (size_t)nn + 2) * sizeof(_Float16)) != hipSuccess)
        return give_up();
    NBX_HIP(hipMemcpyAsync(O.slice_ptr, osp.data(), (ns + 1) * sizeof(uint32_t),
                           hipMemcpyHostToDevice, s));
    // padding values stay 0 (the value kernels write the rows' own blocks)
    NBX_HIP(hipMemsetAsync(O.val, 0, (4 * O.n_blocks + 4) * sizeof(_Float16), s));
    hipLaunchKernelGGL(k_off_pattern, dim3(grid_for(nn)), dim3(BLOCK), 0, s, A, O);
    NBX_HIP(hipGetLastError());
    m.off_pattern = true;
    // the columns as packed 16-bit deltas, when every one fits (otherwise,
    // or without memory for them, the rows read O.bcol)
    std::vector<uint32_t> psp(ns + 1, 0);
    for (uint32_t sl = 0; sl < ns; ++sl)
        psp[sl + 1] = psp[sl] + ((osp[sl + 1] - osp[sl]) / SLICE + 1) / 2 * SLICE;
    O.n_pairs = psp[ns];
    if (hipMalloc(&O.pslice_ptr, (ns + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&O.pcol, (O.n_pairs + 1) * sizeof(uint32_t)) != hipSuccess) {
        (void)hipGetLastError();
        return NBX_OK;
    }
    NBX_HIP(hipMemcpyAsync(O.pslice_ptr, psp.data(), (ns + 1) * sizeof(uint32_t),
                           hipMemcpyHostToDevice, s));
    NBX_HIP(hipMemsetAsync(m.off_flag, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_off_pack, dim3(grid_for(nn)), dim3(BLOCK), 0, s, O, m.off_flag);
    NBX_HIP(hipGetLastError());
    NBX_TRY(read_flag(m.off_flag, s, &bad));
    m.off_packed = bad == 0;
    return NBX_OK;
}

int mat_make_off16(DevMatrix& m, hipStream_t s) {
    if (!m.blocked || m.n == 0) return NBX_OK;
    NBX_TRY(ensure_scale(m, s));
    if (!m.off_tried) {
        m.off_tried = true;
        NBX_TRY(off_build_pattern(m, s));
    }
    if (!m.off_pattern || m.has_off16) return NBX_OK;
    NBX_HIP(hipMemsetAsync(m.off_flag, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_scale16_off, dim3(grid_for(m.s.n_nodes)), dim3(BLOCK), 0, s, m.s, m.o,
                       m.dis, m.off_flag);
    NBX_HIP(hipGetLastError());
    unsigned bad = 0;
    NBX_TRY(read_flag(m.off_flag, s, &bad));
    m.has_off16 = bad == 0;
    return NBX_OK;
}

int mat_prepare_mixed(DevMatrix& m, hipStream_t s, const MixedOpts& o) {
    // the fp16 Chebyshev preconditioner reads the copy without diagonal
    // blocks when the matrix allows it, the full fp16 copy otherwise
    if (m.blocked && o.variant == CHEB_FCG && o.precond_bits == 16) {
        NBX_TRY(mat_make_off16(m, s));
        if (m.has_off16) return NBX_OK;
    }
    return mat_make_scaled(m, s, o.variant == CHEB_FCG ? o.precond_bits : 32);
}

int spmv(const DevMatrix& m, const double* x, double* y, hipStream_t s) {
    if (m.n == 0) return NBX_OK;
    if (m.blocked && m.has_sym) {
        OpSym op{m.y};
        hipLaunchKernelGGL(k_spmv<OpSym>, dim3(map_grid(m.s.n_nodes)), dim3(BLOCK), 0, s, op,
                           (const double2*)x, (double2*)y, m.s.n_nodes);
    } else if (m.blocked) {
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
    if ((prec == 32 && !m.has_val32) || (prec == 16 && !m.has_val16)) {
        set_last_error("spmv_bench: no scaled %d-bit copy", prec);
        return NBX_ERR_ARG;
    }
    if (prec == 65 && !m.has_sym) {
        set_last_error("spmv_bench: the matrix has no symmetric copy");
        return NBX_ERR_ARG;
    }
    if (prec != 64 && prec != 65 && prec != 32 && prec != 16) {
        set_last_error("spmv_bench: precision %d (64, 65, 32 or 16)", prec);
        return NBX_ERR_ARG;
    }
    auto launch = [&]() {
        if (prec == 32)
            hipLaunchKernelGGL(k_spmv_scaled<32>, dim3(G), dim3(BLOCK), 0, s, m.s,
                               (const float2*)w.v32[0], (float2*)w.v32[1], nu);
        else if (prec == 16)
            hipLaunchKernelGGL(k_spmv_scaled<16>, dim3(G), dim3(BLOCK), 0, s, m.s,
                               (const float2*)w.v32[0], (float2*)w.v32[1], nu);
        else if (prec == 65)
            hipLaunchKernelGGL(k_spmv<OpSym>, dim3(G), dim3(BLOCK), 0, s, OpSym{m.y},
                               (const double2*)w.v64[0], (double2*)w.v64[1], nu);
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
    hipLaunchKernelGGL(k_norm2, dim3(g), dim3(BLOCK), 0, s, v, (size_t)n, w.partials);
    hipLaunchKernelGGL(k_fin_out, dim3(1), dim3(FIN), 0, s, w.partials, g, out);
    double h = 0.0;
    if (hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        *err = NBX_ERR_HIP;
        return 0.0;
    }
    return h;
}

}  // namespace nbx

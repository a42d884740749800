// SIMP topology optimisation on the device: density filter, sensitivity
// filter and optimality-criteria update around the device T3 system.
//
// The reference has no topology-optimisation code.  The algorithm is the
// density-filter variant of Andreassen, Clausen, Schevenels, Lazarov and
// Sigmund, "Efficient topology optimization in MATLAB using 88 lines of
// code", Struct. Multidisc. Optim. 43 (2011) 1-16:
//   H_ej = max(0, rmin - |c_e - c_j|) over element centroids c,
//   xPhys = (H x) ./ Hs,  dc <- H (dc ./ Hs),  dv <- H (dv ./ Hs),
//   OC: x_new = clamp(x sqrt(-dc / dv / lmid), x - move, x + move, 0, 1),
//   bisection on lmid from (l1, l2) = (0, 1e9) while (l2-l1)/(l1+l2) > 1e-3,
//   l1 = lmid when the filtered volume exceeds the target, else l2 = lmid.
// Triangles of any size: volume V = sum_e a_e xPhys_e (a = element area), so
// dv_e = a_e.  Every filter sum runs over a row's entries in ascending
// column order with separate multiply and add (-ffp-contract=off), so H x
// and H (dc ./ Hs) are bit-identical with a sequential CPU loop.  The
// bisection tests V through the identity V = x_new . H (a ./ Hs) (H is
// symmetric), which needs no filter pass per step; only the final design
// is filtered.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "simp.hpp"

namespace nbx {

// ---- kernels ----------------------------------------------------------------
// sum_k H_ek v_k, v = x or x ./ Hs
template <bool DIV_IN>
__device__ __forceinline__ double filt_row(const DevFilter& H, uint32_t e,
                                           const double* __restrict__ x) {
    size_t i = (size_t)H.slice_ptr[e >> 6] + (e & 63);
    const int len = H.rowlen[e];
    double s = 0.0;
    for (int k = 0; k < len; ++k, i += SLICE) {
        const uint32_t j = H.idx[i];
        const double v = DIV_IN ? x[j] / H.Hs[j] : x[j];
        s += H.val[i] * v;
    }
    return s;
}

// out_j = (H (d ./ Hs))_j     (sensitivity filter; H is symmetric)
__global__ void k_filter_sens(DevFilter H, const double* __restrict__ d,
                              double* __restrict__ out) {
    const uint32_t j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= H.ne) return;
    out[j] = filt_row<true>(H, j, d);
}

__global__ void k_simp_moduli(const double* __restrict__ xphys, double penal, double E0,
                              double Emin, uint32_t ne, double* __restrict__ E) {
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= ne) return;
    E[e] = Emin + pow(xphys[e], penal) * (E0 - Emin);
}

__global__ void k_oc_init(OcState* st, double target) {
    if (threadIdx.x != 0) return;
    OcState s;
    s.l1 = 0.0;
    s.l2 = 1e9;
    s.target = target;
    s.steps = 0;
    s.pad0 = s.pad1 = 0;
    s.done = (s.l2 - s.l1) / (s.l1 + s.l2) > 1e-3 ? 0u : 1u;
    s.lmid = 0.5 * (s.l2 + s.l1);
    *st = s;
}

// the OC candidate of one element for the multiplier lmid, from
// q = -dc / dv (-dc / dv / lmid evaluates as q / lmid)
__device__ __forceinline__ double oc_candidate_q(double xe, double q, double lmid, double move) {
    double v = xe * sqrt(q / lmid);
    if (v > xe + move) v = xe + move;
    if (v > 1.0) v = 1.0;
    if (v < xe - move) v = xe - move;
    if (v < 0.0) v = 0.0;
    return v;
}
__device__ __forceinline__ double oc_candidate(double xe, double dce, double dve, double lmid,
                                               double move) {
    return oc_candidate_q(xe, -dce / dve, lmid, move);
}

// The bisection's next three levels in one pass.  From the bracket (l1, l2)
// and its midpoint c0 the next two steps can test only 0.5 (c0 + l1) or
// 0.5 (l2 + c0), then one of four midpoints; the pass evaluates all seven
// candidates, each multiplier formed exactly as the step that would test it
// forms it (0.5 (l2 + l1) of that step's bracket), and k_oc_fin3 walks the
// tree one step at a time with the sequential test.  Candidate k's volume is
// the filtered volume of its design, sum_e a_e (H x_new)_e / Hs_e =
// x_new . (H (a ./ Hs)) = x_new . dv because H is symmetric, so a pass reads
// x, dc and dv once and filters nothing (per-workgroup partials, seven
// sums each in the order a one-candidate pass would take).
constexpr int OC_CAND = 7;   // c0; c1 = (L, R); c2 = (LL, LR, RL, RR)
__device__ __forceinline__ void oc_candidates(const OcState& s, double (&c)[OC_CAND]) {
    // L: the volume did not exceed the target (l2 = mid); R: it did (l1 = mid)
    c[0] = s.lmid;
    c[1] = 0.5 * (c[0] + s.l1);    // bracket (l1, c0)
    c[2] = 0.5 * (s.l2 + c[0]);    // bracket (c0, l2)
    c[3] = 0.5 * (c[1] + s.l1);    // (l1, c1L)
    c[4] = 0.5 * (c[0] + c[1]);    // (c1L, c0)
    c[5] = 0.5 * (c[2] + c[0]);    // (c0, c1R)
    c[6] = 0.5 * (s.l2 + c[2]);    // (c1R, l2)
}
__global__ __launch_bounds__(BLOCK) void k_oc_step3(const OcState* __restrict__ st,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ dc,
                                                    const double* __restrict__ dv, double move,
                                                    uint32_t ne, double* partials) {
    const OcState s = *st;
    if (s.done) return;   // uniform over the grid
    double c[OC_CAND];
    oc_candidates(s, c);
    double acc[OC_CAND] = {};
    for (uint32_t e = blockIdx.x * BLOCK + threadIdx.x; e < ne; e += gridDim.x * BLOCK) {
        const double xe = x[e], dve = dv[e], q = -dc[e] / dve;
#pragma unroll
        for (int k = 0; k < OC_CAND; ++k) acc[k] += dve * oc_candidate_q(xe, q, c[k], move);
    }
    __shared__ double red[OC_CAND][BLOCK / WAVE];
    block_sum<OC_CAND>(acc, red);
    put_partials<OC_CAND>(acc, partials);
}

// the design of the last multiplier the bisection tried, and its filtered
// densities
__global__ void k_oc_x(const OcState* __restrict__ st, const double* __restrict__ x,
                       const double* __restrict__ dc, const double* __restrict__ dv,
                       double move, uint32_t ne, double* __restrict__ xnew) {
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= ne) return;
    xnew[e] = oc_candidate(x[e], dc[e], dv[e], st->lmid, move);
}
__global__ void k_oc_phys(DevFilter H, const double* __restrict__ xnew,
                          double* __restrict__ xphys) {
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= H.ne) return;
    xphys[e] = filt_row<false>(H, e, xnew) / H.Hs[e];
}

// up to three bisection steps on the pass's volumes: each the volume test,
// the next bracket and multiplier (or the end of the bisection)
__global__ __launch_bounds__(FIN) void k_oc_fin3(OcState* st, const double* partials,
                                                 unsigned np) {
    if (st->done) return;
    double tot[OC_CAND];
    fin_sum<OC_CAND>(partials, np, tot);
    if (threadIdx.x == 0) {
        OcState s = *st;
        int k = 0;   // the candidate this step tests (s.lmid)
        for (int level = 0; level < 3 && !s.done; ++level) {
            const bool above = tot[k] > s.target;
            if (above)
                s.l1 = s.lmid;
            else
                s.l2 = s.lmid;
            if ((s.l2 - s.l1) / (s.l1 + s.l2) > 1e-3)
                s.lmid = 0.5 * (s.l2 + s.l1);
            else
                s.done = 1;
            s.steps += 1;
            k = 2 * k + (above ? 2 : 1);   // the child candidate
        }
        *st = s;
    }
}

__global__ __launch_bounds__(BLOCK) void k_absdiff_max(const double* __restrict__ a,
                                                       const double* __restrict__ b, uint32_t n,
                                                       double* partials) {
    double m = 0.0;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK)
        m = fmax(m, fabs(a[i] - b[i]));
    for (int off = WAVE / 2; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, WAVE));
    __shared__ double red[BLOCK / WAVE];
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / WAVE; ++w) m = fmax(m, red[w]);
        partials[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(FIN) void k_fin_absmax(const double* partials, unsigned np,
                                                    double* out) {
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

// ---- host -------------------------------------------------------------------
namespace {

// H by uniform bins of size rmin over the centroid bounding box: candidates
// of element e are the elements of the 3x3 bins around its own; a row lists
// its columns in ascending order.
void filter_csr(uint32_t ne, const std::vector<double>& cx, const std::vector<double>& cy,
                double rmin, std::vector<uint32_t>& ptr, std::vector<uint32_t>& idx,
                std::vector<double>& val, std::vector<double>& hs) {
    double x0 = cx[0], x1 = cx[0], y0 = cy[0], y1 = cy[0];
    for (uint32_t e = 0; e < ne; ++e) {
        x0 = std::min(x0, cx[e]);
        x1 = std::max(x1, cx[e]);
        y0 = std::min(y0, cy[e]);
        y1 = std::max(y1, cy[e]);
    }
    const uint32_t nbx = (uint32_t)((x1 - x0) / rmin) + 1;
    const uint32_t nby = (uint32_t)((y1 - y0) / rmin) + 1;
    const size_t nb = (size_t)nbx * nby;
    std::vector<uint32_t> bin(ne), bptr(nb + 1, 0), bidx(ne);
    for (uint32_t e = 0; e < ne; ++e) {
        uint32_t bx = (uint32_t)((cx[e] - x0) / rmin);
        uint32_t by = (uint32_t)((cy[e] - y0) / rmin);
        bx = std::min(bx, nbx - 1);
        by = std::min(by, nby - 1);
        bin[e] = by * nbx + bx;
        bptr[bin[e] + 1]++;
    }
    for (size_t b = 0; b < nb; ++b) bptr[b + 1] += bptr[b];
    {
        std::vector<uint32_t> fill(bptr.begin(), bptr.end() - 1);
        for (uint32_t e = 0; e < ne; ++e) bidx[fill[bin[e]]++] = e;
    }
    ptr.assign(ne + 1, 0);
    idx.clear();
    val.clear();
    hs.assign(ne, 0.0);
    std::vector<uint32_t> cand;
    for (uint32_t e = 0; e < ne; ++e) {
        const int64_t bx = bin[e] % nbx, by = bin[e] / nbx;
        cand.clear();
        for (int64_t jy = std::max<int64_t>(by - 1, 0); jy <= std::min<int64_t>(by + 1, nby - 1); ++jy)
            for (int64_t jx = std::max<int64_t>(bx - 1, 0); jx <= std::min<int64_t>(bx + 1, nbx - 1);
                 ++jx) {
                const size_t b = (size_t)jy * nbx + jx;
                for (uint32_t t = bptr[b]; t < bptr[b + 1]; ++t) {
                    const uint32_t j = bidx[t];
                    const double dx = cx[e] - cx[j], dy = cy[e] - cy[j];
                    if (rmin - std::sqrt(dx * dx + dy * dy) > 0.0) cand.push_back(j);
                }
            }
        std::sort(cand.begin(), cand.end());
        double s = 0.0;
        for (uint32_t j : cand) {
            const double dx = cx[e] - cx[j], dy = cy[e] - cy[j];
            const double h = rmin - std::sqrt(dx * dx + dy * dy);
            idx.push_back(j);
            val.push_back(h);
            s += h;
        }
        hs[e] = s;
        ptr[e + 1] = (uint32_t)idx.size();
    }
}

template <class T>
int upload(T** d, const std::vector<T>& h, hipStream_t s) {
    NBX_HIP(hipMalloc(d, h.size() * sizeof(T) + 8));
    if (!h.empty()) NBX_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return NBX_OK;
}

}  // namespace

int simp_create(Simp& o, FemSys& S, double rmin, double volfrac, double penal, double E0,
                double Emin, double move, const double* x0_host) {
    const uint32_t ne = S.ne;
    const hipStream_t s = S.s;
    if (ne == 0 || !(rmin > 0.0) || !(volfrac > 0.0 && volfrac <= 1.0) || !(move > 0.0)) {
        set_last_error("simp: need elements, rmin > 0, 0 < volfrac <= 1, move > 0");
        return NBX_ERR_ARG;
    }
    if (!S.F_load) {
        set_last_error("simp: set the boundary conditions of the system first");
        return NBX_ERR_ARG;
    }
    o.rmin = rmin;
    o.volfrac = volfrac;
    o.penal = penal;
    o.E0 = E0;
    o.Emin = Emin;
    o.move = move;
    // centroids and areas from the system's mesh
    std::vector<double> nod(2 * (size_t)S.nn);
    std::vector<uint32_t> conn(3 * (size_t)ne);
    NBX_HIP(copy_on(nod.data(), S.nod, nod.size() * 8, hipMemcpyDeviceToHost, s));
    NBX_HIP(copy_on(conn.data(), S.conn, conn.size() * 4, hipMemcpyDeviceToHost, s));
    std::vector<double> cx(ne), cy(ne), area(ne);
    for (uint32_t e = 0; e < ne; ++e) {
        const uint32_t a = conn[3 * (size_t)e], b = conn[3 * (size_t)e + 1], c = conn[3 * (size_t)e + 2];
        const double xa = nod[2 * (size_t)a], ya = nod[2 * (size_t)a + 1];
        const double xb = nod[2 * (size_t)b], yb = nod[2 * (size_t)b + 1];
        const double xc = nod[2 * (size_t)c], yc = nod[2 * (size_t)c + 1];
        cx[e] = (xa + xb + xc) / 3.0;
        cy[e] = (ya + yb + yc) / 3.0;
        area[e] = 0.5 * ((xb - xa) * (yc - ya) - (xc - xa) * (yb - ya));
    }
    double atot = 0.0;
    for (uint32_t e = 0; e < ne; ++e) atot += area[e];
    o.atot = atot;
    filter_csr(ne, cx, cy, rmin, o.h_ptr, o.h_idx, o.h_val, o.h_hs);
    // sliced ELL
    DevFilter& H = o.H;
    H.ne = ne;
    H.n_slices = (ne + SLICE - 1) / SLICE;
    H.n_ent = o.h_idx.size();
    std::vector<uint32_t> sptr(H.n_slices + 1, 0);
    std::vector<uint16_t> rowlen(ne);
    size_t tot = 0;
    for (uint32_t sl = 0; sl < H.n_slices; ++sl) {
        uint32_t w = 0;
        for (uint32_t e = sl * SLICE; e < std::min<uint32_t>(ne, (sl + 1) * SLICE); ++e) {
            const uint32_t len = o.h_ptr[e + 1] - o.h_ptr[e];
            if (len > 0xffffu) {
                set_last_error("simp: filter row of %u entries (rmin too large)", len);
                return NBX_ERR_ARG;
            }
            rowlen[e] = (uint16_t)len;
            w = std::max(w, len);
        }
        sptr[sl] = (uint32_t)tot;
        tot += (size_t)w * SLICE;
        if (tot > 0xffffffffull) {
            set_last_error("simp: filter too large");
            return NBX_ERR_ARG;
        }
    }
    sptr[H.n_slices] = (uint32_t)tot;
    H.n_stored = tot;
    std::vector<uint32_t> eidx(tot);
    std::vector<double> eval(tot, 0.0);
    for (uint32_t e = 0; e < ne; ++e) {
        const size_t base = sptr[e / SLICE] + (e % SLICE);
        const uint32_t len = rowlen[e];
        const uint32_t w = (sptr[e / SLICE + 1] - sptr[e / SLICE]) / SLICE;
        for (uint32_t k = 0; k < w; ++k) {
            eidx[base + (size_t)k * SLICE] = k < len ? o.h_idx[o.h_ptr[e] + k] : e;
            eval[base + (size_t)k * SLICE] = k < len ? o.h_val[o.h_ptr[e] + k] : 0.0;
        }
    }
    NBX_TRY(upload(&H.slice_ptr, sptr, s));
    NBX_TRY(upload(&H.rowlen, rowlen, s));
    NBX_TRY(upload(&H.idx, eidx, s));
    NBX_TRY(upload(&H.val, eval, s));
    NBX_TRY(upload(&H.Hs, o.h_hs, s));
    NBX_TRY(upload(&o.area, area, s));
    // design, filtered densities (xPhys = x at the start), sensitivities
    std::vector<double> x0(ne, volfrac);
    if (x0_host) std::copy(x0_host, x0_host + ne, x0.begin());
    NBX_TRY(upload(&o.x, x0, s));
    NBX_TRY(upload(&o.xphys, x0, s));
    NBX_HIP(hipMalloc(&o.xnew, (size_t)ne * 8 + 8));
    NBX_HIP(hipMalloc(&o.dvf, (size_t)ne * 8 + 8));
    NBX_HIP(hipMalloc(&o.dcf, (size_t)ne * 8 + 8));
    NBX_HIP(hipMalloc(&o.st, sizeof(OcState)));
    NBX_HIP(hipMalloc(&o.scal, 64));
    hipLaunchKernelGGL(k_filter_sens, dim3(grid_for(ne)), dim3(BLOCK), 0, s, H, o.area, o.dvf);
    // system buffers the loop writes
    if (!S.elemE) NBX_HIP(hipMalloc(&S.elemE, (size_t)(ne + 1) * sizeof(double)));
    if (!S.dc) NBX_HIP(hipMalloc(&S.dc, (size_t)ne * 8 + 8));
    if (!S.ce) NBX_HIP(hipMalloc(&S.ce, (size_t)ne * 8 + 8));
    NBX_HIP(hipMemsetAsync(S.u, 0, 2 * (size_t)S.nn * 8, s));
    NBX_HIP(hipGetLastError());
    NBX_HIP(hipStreamSynchronize(s));
    o.iter = 0;
    return NBX_OK;
}

void simp_free(Simp& o) {
    void* ps[] = {o.H.slice_ptr, o.H.rowlen, o.H.idx, o.H.val, o.H.Hs, o.x, o.xnew, o.xphys,
                  o.area, o.dvf, o.dcf, o.st, o.scal};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    o = Simp();
}

int simp_moduli(Simp& o, FemSys& S) {
    hipLaunchKernelGGL(k_simp_moduli, dim3(grid_for(S.ne)), dim3(BLOCK), 0, S.s, o.xphys, o.penal,
                       o.E0, o.Emin, S.ne, S.elemE);
    NBX_HIP(hipGetLastError());
    S.P.elemE = S.elemE;
    return NBX_OK;
}

int simp_update(Simp& o, FemSys& S, double* change) {
    const hipStream_t s = S.s;
    const uint32_t ne = S.ne;
    const DevFilter& H = o.H;
    hipLaunchKernelGGL(k_filter_sens, dim3(grid_for(ne)), dim3(BLOCK), 0, s, H, S.dc, o.dcf);
    hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(WAVE), 0, s, o.st, o.volfrac * o.atot);
    unsigned g = grid_for(ne);
    if (g > (unsigned)MAX_GRID) g = MAX_GRID;
    // Bisection passes (three steps each) are launched in batches, the
    // first as long as the previous update's bisection; kernels of a
    // finished bisection return at once.  From (0, 1e9) the bracket reaches
    // any multiplier above 1e-300 within about 1050 halvings.
    constexpr int MAX_STEPS = 2048;
    int batch = o.last_oc_steps ? (int)(o.last_oc_steps + 2) / 3 : 16;
    OcState h{};
    // (seven partials per workgroup in the workspace's 4 MAX_GRID)
    const unsigned g3 = std::min(g, (unsigned)(4 * MAX_GRID / OC_CAND));
    for (int done_steps = 0;; done_steps += 3 * batch, batch = 4) {
        if (done_steps >= MAX_STEPS) {
            set_last_error("simp: OC bisection did not close in %d steps", MAX_STEPS);
            return NBX_ERR_BREAKDOWN;
        }
        for (int k = 0; k < batch; ++k) {
            hipLaunchKernelGGL(k_oc_step3, dim3(g3), dim3(BLOCK), 0, s, o.st, o.x, o.dcf, o.dvf,
                               o.move, ne, S.ws.partials);
            hipLaunchKernelGGL(k_oc_fin3, dim3(1), dim3(FIN), 0, s, o.st, S.ws.partials, g3);
        }
        NBX_HIP(hipGetLastError());
        NBX_HIP(hipMemcpyAsync(&h, o.st, sizeof h, hipMemcpyDeviceToHost, s));
        NBX_HIP(hipStreamSynchronize(s));
        if (h.done) break;
    }
    o.last_oc_steps = h.steps;
    hipLaunchKernelGGL(k_oc_x, dim3(grid_for(ne)), dim3(BLOCK), 0, s, o.st, o.x, o.dcf, o.dvf,
                       o.move, ne, o.xnew);
    hipLaunchKernelGGL(k_oc_phys, dim3(grid_for(ne)), dim3(BLOCK), 0, s, H, o.xnew, o.xphys);
    hipLaunchKernelGGL(k_absdiff_max, dim3(g), dim3(BLOCK), 0, s, o.xnew, o.x, ne, S.ws.partials);
    hipLaunchKernelGGL(k_fin_absmax, dim3(1), dim3(FIN), 0, s, S.ws.partials, g, o.scal);
    NBX_HIP(hipGetLastError());
    double ch = 0.0;
    NBX_HIP(hipMemcpyAsync(&ch, o.scal, sizeof ch, hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    std::swap(o.x, o.xnew);
    o.iter += 1;
    if (change) *change = ch;
    return NBX_OK;
}

}  // namespace nbx

// Gauss-Seidel on the device: the damage solver's fallback when its
// Jacobi-PCG does not converge (static_damage2D.c:414-427), following
// solver_bot/sparse/solvers/gauss_seidel.c:13-63 and the forward
// substitution of triangular.c:14-35.
//
// A sweep forms r = b - A x (error = r.r) and c = b - U x with the old x in
// one pass, stops when error <= tol^2, else solves (D + L) x = c.  The
// forward substitution is sequential in the row index; here it runs by
// dependency levels: row i's level is one more than the highest level of
// the rows left of its diagonal, so the rows of a level depend only on
// earlier levels and go in parallel, one thread per row, each summing its
// entries in the reference's column order.  x after a sweep equals the
// sequential substitution bit for bit; only the error's sum is in another
// order.  One workgroup runs every sweep, the levels separated by barriers.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gauss_seidel.hpp"

namespace nbx {

namespace {

constexpr int GS_BLOCK = 1024;

__device__ double gs_block_total(double v, double* red) {
    v = wave_sum(v);
    const unsigned lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < GS_BLOCK / WAVE; ++w) t += red[w];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(GS_BLOCK) void k_gauss_seidel(
    const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
    const double* __restrict__ val, const double* __restrict__ b, double* x, double* c,
    const uint32_t* __restrict__ lrows, const uint32_t* __restrict__ lptr, uint32_t nlev,
    uint32_t n, uint32_t max_iter, double tol2, uint32_t* niter, double* error) {
    __shared__ double red[GS_BLOCK / WAVE];
    uint32_t k = 0;
    double err = 1e10;
    while (k < max_iter) {
        double e = 0.0;
        for (uint32_t i = threadIdx.x; i < n; i += GS_BLOCK) {
            double ci = b[i], ei = b[i];
            for (uint32_t s = rp[i]; s < rp[i + 1]; ++s) {
                const uint32_t j = col[s];
                const double p = val[s] * x[j];
                ei -= p;
                if (j > i) ci -= p;
            }
            c[i] = ci;
            e += ei * ei;
        }
        err = gs_block_total(e, red);   // also orders the c writes before the reads
        if (err <= tol2) break;
        for (uint32_t L = 0; L < nlev; ++L) {
            for (uint32_t t = lptr[L] + threadIdx.x; t < lptr[L + 1]; t += GS_BLOCK) {
                const uint32_t i = lrows[t];
                const uint32_t e_ = rp[i + 1];
                double xi = c[i];
                uint32_t s = rp[i];
                while (s < e_ && col[s] < i) {
                    xi -= val[s] * x[col[s]];
                    ++s;
                }
                // the row's last entry when that is the diagonal, else the
                // diagonal found in the row, else 0 (triangular.c:30-33)
                double d = 0.0;
                if (e_ > rp[i] && col[e_ - 1] == i)
                    d = val[e_ - 1];
                else if (s < e_ && col[s] == i)
                    d = val[s];
                x[i] = xi / d;
            }
            __syncthreads();
        }
        ++k;
    }
    if (threadIdx.x == 0) {
        *niter = k;
        *error = err;
    }
}

}  // namespace

void gs_free(GsSystem& g) {
    void* ps[] = {g.rp, g.col, g.val, g.lrows, g.lptr, g.c, g.out};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    g = GsSystem();
}

int gs_prepare(GsSystem& g, uint32_t n, const uint32_t* rp, const uint32_t* col,
               const double* val, hipStream_t s) {
    const size_t nnz = rp[n];
    if (g.n != n || g.nnz != nnz) {
        gs_free(g);
        // dependency levels of the forward substitution
        std::vector<uint32_t> level(n, 0);
        uint32_t nlev = 0;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t l = 0;
            for (uint32_t t = rp[i]; t < rp[i + 1] && col[t] < i; ++t)
                l = std::max(l, level[col[t]] + 1);
            level[i] = l;
            nlev = std::max(nlev, l + 1);
        }
        std::vector<uint32_t> lptr(nlev + 1, 0), lrows(n);
        for (uint32_t i = 0; i < n; ++i) lptr[level[i] + 1]++;
        for (uint32_t l = 0; l < nlev; ++l) lptr[l + 1] += lptr[l];
        std::vector<uint32_t> fill(lptr.begin(), lptr.end() - 1);
        for (uint32_t i = 0; i < n; ++i) lrows[fill[level[i]]++] = i;
        NBX_HIP(hipMalloc(&g.rp, (n + 1) * sizeof(uint32_t)));
        NBX_HIP(hipMalloc(&g.col, (nnz + 1) * sizeof(uint32_t)));
        NBX_HIP(hipMalloc(&g.val, (nnz + 1) * sizeof(double)));
        NBX_HIP(hipMalloc(&g.lrows, (n + 1) * sizeof(uint32_t)));
        NBX_HIP(hipMalloc(&g.lptr, (nlev + 1) * sizeof(uint32_t)));
        NBX_HIP(hipMalloc(&g.c, (n + 1) * sizeof(double)));
        NBX_HIP(hipMalloc(&g.out, 2 * sizeof(double)));
        NBX_HIP(hipMemcpyAsync(g.rp, rp, (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        if (nnz) NBX_HIP(hipMemcpyAsync(g.col, col, nnz * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        if (n) NBX_HIP(hipMemcpyAsync(g.lrows, lrows.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        NBX_HIP(hipMemcpyAsync(g.lptr, lptr.data(), (nlev + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                               s));
        g.n = n;
        g.nnz = nnz;
        g.nlev = nlev;
    }
    if (nnz) NBX_HIP(hipMemcpyAsync(g.val, val, nnz * sizeof(double), hipMemcpyHostToDevice, s));
    NBX_HIP(hipStreamSynchronize(s));
    return NBX_OK;
}

int gauss_seidel(const GsSystem& g, const double* b, double* x, uint32_t max_iter, double tol,
                 uint32_t* niter, double* tol_reached, hipStream_t s) {
    uint32_t* it = (uint32_t*)g.out;
    double* err = g.out + 1;
    hipLaunchKernelGGL(k_gauss_seidel, dim3(1), dim3(GS_BLOCK), 0, s, g.rp, g.col, g.val, b, x,
                       g.c, g.lrows, g.lptr, g.nlev, g.n, max_iter, tol * tol, it, err);
    NBX_HIP(hipGetLastError());
    double h[2];
    NBX_HIP(copy_on(h, g.out, sizeof h, hipMemcpyDeviceToHost, s));
    uint32_t k;
    memcpy(&k, &h[0], sizeof k);
    if (niter) *niter = k;
    if (tol_reached) *tol_reached = sqrt(h[1]);
    return h[1] > tol * tol ? 1 : 0;
}

}  // namespace nbx

// Device Gauss-Seidel (gauss_seidel.c:13-63) on a scalar CSR copy, with the
// forward substitution run by dependency levels (gauss_seidel.hip).
#pragma once

#include "common.hpp"

namespace nbx {

struct GsSystem {
    uint32_t n = 0, nlev = 0;
    size_t nnz = 0;
    uint32_t* rp = nullptr;     // [n + 1]
    uint32_t* col = nullptr;    // [nnz], ascending per row
    double* val = nullptr;      // [nnz]
    uint32_t* lrows = nullptr;  // rows grouped by level, ascending within a level
    uint32_t* lptr = nullptr;   // [nlev + 1]
    double* c = nullptr;        // [n] b - U x of the sweep
    double* out = nullptr;      // sweeps done, final error
};

// Uploads the matrix (pattern and levels only when the pattern changed).
int gs_prepare(GsSystem& g, uint32_t n, const uint32_t* rp, const uint32_t* col,
               const double* val, hipStream_t s);
void gs_free(GsSystem& g);
// x (device, in/out: the initial guess) after at most max_iter sweeps; b on
// the device.  Returns 0 when ||b - A x|| <= tol, 1 after max_iter sweeps,
// like the reference; *niter the sweeps, *tol_reached the last ||b - A x||.
int gauss_seidel(const GsSystem& g, const double* b, double* x, uint32_t max_iter, double tol,
                 uint32_t* niter, double* tol_reached, hipStream_t s);

}  // namespace nbx

// nbx_sparse_solve_CG_precond_Jacobi_noalloc: the Jacobi-PCG of
// nb_sparse_solve_CG_precond_Jacobi (cg_precond_jacobi.c:13-90, the same
// kernels as the drop-in) on a scalar CSR matrix and vectors the caller
// already holds in device memory, for callers that cannot allocate.  Its
// work vectors, reduction partials and solver state live in `scratch`, a
// fixed array of this library's code object: every GPU it runs on holds its
// own copy from module load until the process ends.
//
// Host code never names `scratch` in a device role: the array's name is a
// host-side stand-in.  The first call on a device (that device current)
// takes the array's device address with hipGetSymbolAddress and keeps it
// per device; kernels and copies get that address.  Nothing in the array
// has to start at zero: the solver state is written from the host at the
// start of each call, and every vector and partial is written before it is
// read.  Calls that overlap on one GPU share its array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/nbots_amd.h"
#include "internal.hpp"

namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t ST_BYTES = align256(2 * sizeof(nbx::PcgState));
constexpr size_t PART_BYTES = align256(4 * (size_t)nbx::MAX_GRID * sizeof(double));
constexpr size_t VEC_BYTES = align256(((size_t)NBX_NOALLOC_MAX_N + 2) * sizeof(double));
constexpr int NVEC = 7;   // diag, g, w, the four direction buffers
constexpr size_t SCRATCH_BYTES = ST_BYTES + PART_BYTES + NVEC * VEC_BYTES;

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char* g_scratch[MAX_DEVICES] = {};

}  // namespace

// 16.8 MiB per GPU (NBX_NOALLOC_MAX_N = 262,144 rows)
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

// The device address of `scratch` on device dev (current), taken once.
int scratch_on(int dev, unsigned char** out) {
    if (dev < 0 || dev >= MAX_DEVICES) {
        nbx::set_last_error("device %d outside the no-allocation table", dev);
        return nbx::NBX_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            nbx::set_last_error("hipGetSymbolAddress(scratch) on device %d: %s", dev,
                                hipGetErrorString(e));
            return nbx::NBX_ERR_HIP;
        }
        g_scratch[dev] = (unsigned char*)p;
    }
    *out = g_scratch[dev];
    return nbx::NBX_OK;
}

}  // namespace

extern "C" int nbx_sparse_solve_CG_precond_Jacobi_noalloc(
    uint32_t n, const uint32_t* rowptr_dev, const uint32_t* col_dev, const double* val_dev,
    const double* b_dev, double* x_dev, uint32_t max_iter, double tolerance,
    uint32_t* niter_performed, double* tolerance_reached, void* stream) {
    if (n > NBX_NOALLOC_MAX_N) {
        nbx::set_last_error("%u rows exceed the no-allocation maximum of %u", n,
                            (unsigned)NBX_NOALLOC_MAX_N);
        return NBX_ERR_ARG;
    }
    if (n && (!rowptr_dev || !col_dev || !val_dev || !b_dev || !x_dev)) {
        nbx::set_last_error("nbx_sparse_solve_CG_precond_Jacobi_noalloc: null device pointer");
        return NBX_ERR_ARG;
    }
    const hipStream_t s = (hipStream_t)stream;
    int dev = 0;
    NBX_HIP(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    NBX_TRY(scratch_on(dev, &base));
    unsigned char* vec = base + ST_BYTES + PART_BYTES;
    auto slot = [&](int k) { return (double*)(vec + (size_t)k * VEC_BYTES); };

    nbx::DevMatrix m;   // a view: nothing here is freed
    m.blocked = false;
    m.n = n;
    m.c.n = n;
    m.c.rowptr = const_cast<uint32_t*>(rowptr_dev);
    m.c.col = const_cast<uint32_t*>(col_dev);
    m.c.val = const_cast<double*>(val_dev);
    m.diag = slot(0);
    NBX_TRY(nbx::mat_extract_diag(m, s));

    nbx::PcgState host_state{};
    nbx::Workspace w;   // a view: no events, no pinned mirror, no captured graphs
    w.n = n;
    w.replay = false;
    w.v64[0] = slot(1);   // g
    w.v64[1] = slot(2);   // p_j, j % 4 = 0
    w.v64[2] = slot(3);   // w
    w.v64[6] = slot(4);   // p_j, j % 4 = 1..3
    w.v64[7] = slot(5);
    w.v64[8] = slot(6);
    w.partials = (double*)(base + ST_BYTES);
    w.st = (nbx::PcgState*)base;
    w.h_st = &host_state;
    const int st = nbx::pcg_jacobi_csr_view(m.c, m.diag, b_dev, x_dev, max_iter, tolerance,
                                            niter_performed, tolerance_reached, w, s);
    return st < 0 ? -st : st;
}

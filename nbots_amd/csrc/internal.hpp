// Host-side interfaces between the C-ABI layer and the HIP kernels.
#pragma once

#include "common.hpp"

namespace nbx {

// Scalar state of a device-resident CG run; written only by the last
// workgroup of a reduction kernel, read by the following kernels and polled
// by the host.
struct AmgDev;   // smoothed-aggregation hierarchy (amg.hip)

struct PcgState {
    double rz;            // g.q (Jacobi-PCG) or r.r (scaled CG) of the current iterate
    double gg_cur;        // ||g_k||^2
    double gg_prev;       // ||g_{k-1}||^2, the value the reference's stop test sees
    double pw;            // p.w of the current iteration
    double alpha, beta;
    double tol2;          // stop threshold on gg (absolute) or rz (relative, inner)
    double tol_reached;
    uint32_t k, max_iter;
    uint32_t done, status;
    uint32_t breakdown, hold;      // hold: paused for a reliable update
    double maxrr;                  // largest r.r since the last reliable update
    double conv_rr;                // r.r at which the true residual should meet tol
    uint32_t updates, pad;
    double gg0;                    // squared residual norm at the start (fresh b - A x)
    double alpha_hist[4];          // alpha of iteration j at [j % 4] (grouped x updates)
};

// A device matrix: 2-dof blocked sliced ELL when the pattern allows it,
// otherwise scalar CSR.
struct DevMatrix {
    bool blocked = false;
    uint32_t n = 0;            // scalar rows
    Sell2 s;
    DevCsr c;
    double* diag = nullptr;    // [n] diagonal A(i,i) (0.0 if not stored)
    double* dis = nullptr;     // [n] 1/sqrt(diag) for the scaled copies
    bool has_scale = false;    // dis and lambda_max are current
    bool has_val32 = false;
    bool has_val16 = false;
    double lambda_max = 0.0;   // Gershgorin bound of the scaled matrix
    AmgDev* amg = nullptr;     // the AMG preconditioner's hierarchy (FEM systems build it)
    bool amg_current = false;  // ... and its numeric phase follows the current values
    // symmetric copy of the FP64 values (blocked matrices): its pattern is
    // built once per pattern; its values are refreshed, and the symmetry of
    // the values checked, when the solver first needs them after a change
    Sym2 y;
    bool sym_pattern = false;  // y's pattern is built (false: not yet, or the
                               // pattern is not structurally symmetric)
    bool sym_tried = false;    // y's pattern has been attempted
    bool sym_checked = false;  // y's values reflect the current values
    bool has_sym = false;      // ... and the values are symmetric: use y
    unsigned* sym_flag = nullptr;
    // fp16 scaled copy without the diagonal blocks (the Chebyshev
    // preconditioner's matrix): pattern once per pattern, values with the scale
    Off16 o;
    bool off_tried = false;    // o's pattern has been attempted
    bool off_pattern = false;  // o's pattern is built
    bool has_off16 = false;    // o's values reflect the current scale
    bool off_packed = false;   // o.pcol holds the columns as 16-bit deltas
    unsigned* off_flag = nullptr;
};

// Per-matrix solver scratch (device).  Sized for n scalar rows plus two
// spare entries.  Slots a run reserves (a caller must not pass them as the
// run's b or x):
//   FP64 Jacobi-PCG (pcg_run):  v64[0] g, v64[1] p_j (j % 4 = 0), v64[2] w,
//                               v64[6..8] p_j (j % 4 = 1..3); plain CG also
//                               v64[5] (the unit diagonal)
//   flexible CG (CHEB_FCG):     v64[0] r, v64[1] p, v64[2] w; v32[0..3]
//                               residual, two directions and z of the
//                               Chebyshev steps (FP64 preconditioner:
//                               v64[5..8] instead)
//   FP32 CG (CG_RESTART/RELIABLE): v32[0..3]; a fallback to FP64 PCG takes
//                               the FP64 slots above
//   free for callers: v64[3], v64[4] (the drop-in's b and x, the system's
//   residual) and the spare entries past the end of v64[5] (one-value
//   reductions)
struct Workspace {
    uint32_t n = 0;
    double* v64[9] = {};       // FP64 vectors
    float* v32[4] = {};        // FP32 vectors
    double* partials = nullptr;
    float* partials32 = nullptr;
    PcgState* st = nullptr;    // device state (2 slots: outer, inner)
    PcgState* h_st = nullptr;  // pinned host mirror
    hipEvent_t ev[4] = {};
    static constexpr int NSAMP = 32;    // sampled SpMV launches per solve, one every
    static constexpr int SAMPLE_EVERY = 64;   // SAMPLE_EVERY iterations (PCG)
    hipEvent_t sev[2 * NSAMP] = {};
    // Replay a solve's repeated steps as captured HIP graphs (one host call
    // per step instead of one per kernel; same kernels, same results).
    bool replay = false;
    // Outer iterations of this system's last converged flexible-CG solve
    // (first pass): the next solve launches most of that many before it
    // first polls, then polls every few iterations.  Host cadence only: the
    // iterates and results do not depend on it.
    uint32_t fcg_hint = 0;
};

struct SolveStats {
    uint32_t iters = 0;        // CG iterations (inner FP32 iterations for mixed)
    uint32_t outer = 0;        // refinement sweeps (mixed)
    uint32_t fp64_iters = 0;   // FP64 PCG iterations (incl. fallback)
    double rel_res = 0.0;      // final ||b - A x|| / ||b|| (FP64)
    double ms = 0.0;           // device time of the solve
    double ms_spmv_avg = 0.0;  // average duration of the dominant SpMV kernel
    int fallback = 0;          // 1 if the FP32 sweeps stalled and FP64 PCG finished
    int layout = 0;            // FP64 matrix layout used: 0 full, 1 symmetric
    double res0 = 0.0;         // ||b - A x|| at the start, computed afresh
};

// Mixed-precision settings.
//  CG_RESTART: FP32 CG sweeps to inner_rtol inside FP64 iterative refinement
//  CG_RELIABLE: FP32 CG with FP64 reliable updates (no restarts)
//  CHEB_FCG: FP64 flexible CG preconditioned by a fixed-degree Chebyshev
//            iteration on the FP32 Jacobi-scaled matrix
//  AMG_FCG: FP64 flexible CG preconditioned by one FP32 smoothed-aggregation
//           V-cycle (FEM systems: the hierarchy needs node coordinates)
enum MixedVariant { CG_RESTART = 0, CG_RELIABLE = 1, CHEB_FCG = 2, AMG_FCG = 3 };
struct MixedOpts {
    MixedVariant variant = CHEB_FCG;
    double inner_rtol = 1e-4;
    double delta = 0.1;        // reliable update when r.r < delta^2 max r.r
    uint32_t cheb_degree = 40;  // SpMVs per preconditioner application
    double cheb_ratio = 3e4;    // Chebyshev interval [lambda_max/ratio, lambda_max]
    int precond_bits = 16;     // Chebyshev preconditioner: fp16 or FP32 scaled values (FP64
                               // sums, FP32 vectors), or 64: the FP64 matrix, FP64 vectors
    bool true_residual = false; // stop on the true residual (restarts), FP64 and CHEB_FCG
    bool replay = false;        // replay captured steps as HIP graphs (else kernel by kernel)
};

int ws_alloc(Workspace& w, uint32_t n);

void amg_free(AmgDev* a);
// z = M^-1 r for a one-level hierarchy (at most 200 unknowns: the dense
// coarsest solve alone; FP32 arrays of its length); returns at once when
// st->done
int amg_vcycle(AmgDev* a, const float* r, float* z, const PcgState* st, hipStream_t s);
const double* amg_ones(const AmgDev* a);   // [n] 1.0: the flexible CG's unit scaling
uint32_t amg_levels(const AmgDev* a);
// The finest level smoothed on the matrix's fp16 Jacobi-scaled copy (the
// Chebyshev preconditioner's kernels, in scaled variables): its Chebyshev
// coefficients and the vector for the coarse correction; `scaled16` false
// for a one-level hierarchy (amg_vcycle).
struct AmgFine {
    bool scaled16 = false;
    double inv_theta = 0.0;
    double c1[8] = {}, c2[8] = {};   // steps 1 .. degree - 1
    uint32_t degree = 0;
    float* c = nullptr;              // [finest n]
};
AmgFine amg_fine(const AmgDev* a);
// The coarse part of the scaled-variable cycle: b_1 = P~^T r (r the scaled
// finest residual, P~ = D^1/2 P), the V-cycle on levels 1.., c = P~ x_1.
int amg_coarse_correction(AmgDev* a, const float* r, float* c, const PcgState* st,
                          hipStream_t s);
void ws_free(Workspace& w);
void mat_free(DevMatrix& m);

// Upload a host CSR (row-sorted columns) into a device matrix, blocked when
// possible.  Allocates on first use; `values_only` refreshes values.
int mat_upload_csr(DevMatrix& m, uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                   const double* val, bool values_only, hipStream_t s);
int mat_extract_diag(DevMatrix& m, hipStream_t s);
// The values changed: the scaled copies and their bound are stale.
void mat_scaled_stale(DevMatrix& m);
// Build or refresh the symmetric FP64 copy when the matrix allows it
// (m.has_sym tells); a no-op when it is current.
int mat_prepare_sym(DevMatrix& m, hipStream_t s);
// Jacobi-scaled copy D^-1/2 A D^-1/2 with 16- or 32-bit values (and the
// Gershgorin bound lambda_max); a no-op when current.
int mat_make_scaled(DevMatrix& m, hipStream_t s, int bits);
// fp16 scaled copy without the diagonal blocks (m.has_off16 tells whether
// the matrix allows it); a no-op when current.
int mat_make_off16(DevMatrix& m, hipStream_t s);
// the scaled copy a mixed-precision solve with these options reads
int mat_prepare_mixed(DevMatrix& m, hipStream_t s, const MixedOpts& o);

int spmv(const DevMatrix& m, const double* x, double* y, hipStream_t s);

// Reference-semantics Jacobi PCG (absolute tolerance on ||Ax-b||, lagged stop
// test); x holds the initial guess.  Returns 0 converged / 1 not, <0 error.
// unpreconditioned: the reference's plain CG (conjugate_gradient.c:13-77).
int pcg_jacobi(const DevMatrix& m, const double* b, double* x, uint32_t max_iter,
               double tol, uint32_t* niter, double* tol_reached, Workspace& w,
               hipStream_t s, SolveStats* stats, bool unpreconditioned = false);

// The same on a scalar CSR already in device memory with a caller-made
// workspace (vectors v64[0..2], v64[6..8], partials, st, a host mirror
// h_st; no events): nothing is allocated.  diag holds A(i,i).
int pcg_jacobi_csr_view(const DevCsr& c, const double* diag, const double* b, double* x,
                        uint32_t max_iter, double tol, uint32_t* niter, double* tol_reached,
                        Workspace& w, hipStream_t s);

// FP32 inner CG on the Jacobi-scaled FP32 copy inside FP64 iterative
// refinement; stops when ||b - A x||_2 <= rtol * ||b||_2.
int solve_mixed(const DevMatrix& m, const double* b, double* x, double rtol,
                const MixedOpts& opt, uint32_t max_iter, Workspace& w, hipStream_t s,
                SolveStats* stats);

// FP64 Jacobi PCG with a relative tolerance (x0 = x on input).
int solve_fp64(const DevMatrix& m, const double* b, double* x, double rtol, bool true_residual,
               uint32_t max_iter, Workspace& w, hipStream_t s, SolveStats* stats);

int spmv_bench(const DevMatrix& m, int prec, uint32_t reps, Workspace& w, hipStream_t s,
               double* ms_avg);

double device_norm2(uint32_t n, const double* v, Workspace& w, hipStream_t s, int* err);

}  // namespace nbx

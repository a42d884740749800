// Drop-in nb_sparse_t: the reference's sparse-matrix API
// (headers/nb/solver_bot/sparse/sparse.h:28-78) with SpMV and Jacobi-PCG on
// the GPU.
//
// Host storage is one flat CSR (rowptr, col, val) instead of the reference's
// arrays of row pointers (solver_bot/sparse/sparse_struct.h:6-11); entry
// access keeps the reference's semantics (binary search in a sorted row,
// exit(1) on an entry outside the pattern, sparse.c:134-219).  The device
// copy is refreshed lazily: any host-side write marks the values stale, and
// the next GPU operation uploads them.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/nbots_amd.h"
#include "internal.hpp"
#include "opts.hpp"

struct nb_sparse_s {
    uint32_t N = 0;
    std::vector<uint32_t> rowptr;
    std::vector<uint32_t> col;
    std::vector<double> val;
    // device mirror
    mutable nbx::DevMatrix dev;
    mutable nbx::Workspace ws;
    mutable hipStream_t stream = nullptr;
    mutable bool dev_pattern = false;
    mutable bool dev_values = false;
};

namespace {


int64_t find_slot(const nb_sparse_t* A, uint32_t i, uint32_t j) {
    if (i >= A->N) return -1;
    const uint32_t* b = A->col.data() + A->rowptr[i];
    const uint32_t* e = A->col.data() + A->rowptr[i + 1];
    const uint32_t* it = std::lower_bound(b, e, j);
    if (it == e || *it != j) return -1;
    return it - A->col.data();
}

[[noreturn]] void not_allocated(const char* fn, const nb_sparse_t* A, uint32_t i, uint32_t j,
                                double value) {
    // same message and exit as the reference (sparse.c:140-144)
    printf("ERROR: Entry of sparse matrix is not allocated.\n");
    printf("    -> %s(*A=%p, i=%i, j=%i, value=%lf)\n", fn, (const void*)A, i, j, value);
    exit(1);
}

// Bring the device copy up to date; returns an nbx status.
int sync_device(const nb_sparse_t* A) {
    if (!A->stream) {
        hipError_t e = hipStreamCreate(&A->stream);
        if (e != hipSuccess) {
            nbx::set_last_error("hipStreamCreate: %s", hipGetErrorString(e));
            return nbx::NBX_ERR_HIP;
        }
    }
    if (A->dev_pattern && A->dev_values) return nbx::NBX_OK;
    int rc = nbx::mat_upload_csr(A->dev, A->N, A->rowptr.data(), A->col.data(), A->val.data(),
                                 A->dev_pattern, A->stream);
    if (rc) return rc;
    A->dev_pattern = true;
    A->dev_values = true;
    return nbx::ws_alloc(A->ws, A->N);
}

void stale(nb_sparse_t* A) { A->dev_values = false; }

}  // namespace

static int sparse_solve(const nb_sparse_t* A, const double* b, double* x, int mode, double tol,
                        uint32_t max_iter, const nbx::MixedOpts& mo, nbx_solve_info_t* info);
// solve mode of the drop-in nb_sparse_solve_conjugate_gradient (not public)
constexpr int MODE_PLAIN_CG = 0x100;

extern "C" {

// ---- graph (graph.c:20-40) -------------------------------------------------
uint32_t nb_graph_get_memsize(void) { return sizeof(nb_graph_t); }
void nb_graph_init(nb_graph_t* g) { memset(g, 0, sizeof *g); }
void nb_graph_clear(nb_graph_t* g) {
    if (g->adj) {
        for (uint32_t i = 0; i < g->N; ++i) free(g->adj[i]);
        free(g->adj);
    }
    free(g->N_adj);
    if (g->wij) {
        for (uint32_t i = 0; i < g->N; ++i) free(g->wij[i]);
        free(g->wij);
    }
    free(g->wi);
    memset(g, 0, sizeof *g);
}
void nb_graph_finish(nb_graph_t* g) { nb_graph_clear(g); }

// ---- creation --------------------------------------------------------------
// sparse.c:20-60: row perm[i]*vpn+k1 (or i*vpn+k1) holds the columns
// {adj(i), i} x vpn in ascending order; columns are not permuted.
nb_sparse_t* nb_sparse_create(const nb_graph_t* const graph, const uint32_t* const perm,
                              uint32_t vpn) {
    nb_sparse_t* A = new nb_sparse_t;
    const uint32_t N = graph->N * vpn;
    A->N = N;
    std::vector<uint32_t> size(N, 0);
    for (uint32_t i = 0; i < graph->N; ++i)
        for (uint32_t k1 = 0; k1 < vpn; ++k1) {
            const uint32_t r = perm ? perm[i] * vpn + k1 : i * vpn + k1;
            size[r] = (graph->N_adj[i] + 1) * vpn;
        }
    A->rowptr.assign(N + 1, 0);
    for (uint32_t r = 0; r < N; ++r) A->rowptr[r + 1] = A->rowptr[r] + size[r];
    A->col.assign(A->rowptr[N], 0);
    A->val.assign(A->rowptr[N], 0.0);
    for (uint32_t i = 0; i < graph->N; ++i)
        for (uint32_t k1 = 0; k1 < vpn; ++k1) {
            const uint32_t r = perm ? perm[i] * vpn + k1 : i * vpn + k1;
            uint32_t* row = A->col.data() + A->rowptr[r];
            for (uint32_t j = 0; j < graph->N_adj[i]; ++j)
                for (uint32_t k2 = 0; k2 < vpn; ++k2) row[j * vpn + k2] = graph->adj[i][j] * vpn + k2;
            for (uint32_t k2 = 0; k2 < vpn; ++k2) row[graph->N_adj[i] * vpn + k2] = i * vpn + k2;
            std::sort(row, row + size[r]);
        }
    return A;
}

nb_sparse_t* nb_sparse_clone(nb_sparse_t* A) {
    nb_sparse_t* B = new nb_sparse_t;
    B->N = A->N;
    B->rowptr = A->rowptr;
    B->col = A->col;
    B->val = A->val;
    return B;
}

void nb_sparse_save(const nb_sparse_t* const A, const char* filename) {
    // text format of sparse.c:97-112
    FILE* fp = fopen(filename, "w");
    if (!fp) {
        printf("Error: Opening the file %s to save the matrix.\n", filename);
        return;
    }
    fprintf(fp, "%i %i\n", A->N, A->N);
    for (uint32_t i = 0; i < A->N; ++i)
        for (uint32_t s = A->rowptr[i]; s < A->rowptr[i + 1]; ++s)
            fprintf(fp, "%i %i %e \n", i, A->col[s], A->val[s]);
    fclose(fp);
}

// ---- MATLAB v4 files (solver_bot/matlab_v4.c) --------------------------------
// Record header: five int32 {type, rows, cols, imag, len(label)+1}, the label
// with its '\0', then the doubles column by column.  A sparse matrix is type 2
// with nnz+1 rows and 3 columns: 1-based row indices and N, 1-based column
// indices and N, values and a closing 0, entries ordered by column then row
// (matlab_v4.c:186-245).  A vector is a full N x 1 matrix, type 0
// (matlab_v4.c:493-513).  Writers append and are byte-identical with the
// reference's.  Readers take the record whose label is equal and fill the
// caller's object or buffer; the reference skips only records for which
// strcmp() returns 1 (matlab_v4.c:84) and its vector reader fills a buffer of
// its own (matlab_v4.c:478), so it agrees with these readers on the files its
// writers produce with distinct labels.

static void mat4_fail(const char* msg) {
    printf("ERROR: %s\n", msg);
    exit(1);
}

static void mat4_header(FILE* fp, int32_t type, int32_t rows, int32_t cols, const char* label) {
    const int32_t len = (int32_t)strlen(label);
    const int32_t info[5] = {type, rows, cols, 0, len + 1};
    const char end = '\0';
    fwrite(info, 4, 5, fp);
    fwrite(label, 1, (size_t)len, fp);
    fwrite(&end, 1, 1, fp);
}

// Positions fp after the header and label of the record named label; returns
// its info, or false at the end of the file.
static bool mat4_find(FILE* fp, const char* label, int32_t info[5]) {
    while (fread(info, 4, 5, fp) == 5) {
        if (info[4] <= 0) mat4_fail("The matlab v4 file is corrupted.");
        std::vector<char> name((size_t)info[4] + 1, '\0');
        if (fread(name.data(), 1, (size_t)info[4], fp) != (size_t)info[4])
            mat4_fail("The matlab v4 file is corrupted.");
        if (info[0] != 0 && info[0] != 2)
            mat4_fail("Support only for complete and sparse matrix with double precision.");
        if (strcmp(name.data(), label) == 0) return true;
        const long skip = (long)info[1] * (long)info[2] * (long)sizeof(double);
        if (fseek(fp, skip, SEEK_CUR) != 0) mat4_fail("The matlab v4 file is corrupted.");
    }
    return false;
}

void nb_sparse_save_mat4(const nb_sparse_t* const A, const char* url, char* label) {
    FILE* fp = fopen(url, "ab");
    if (!fp) mat4_fail("Impossible to open matlab v4 file.");
    const uint32_t nnz = (uint32_t)A->col.size();
    mat4_header(fp, 2, (int32_t)nnz + 1, 3, label);
    std::vector<uint32_t> order(nnz), row(nnz);
    for (uint32_t i = 0; i < A->N; ++i)
        for (uint32_t s = A->rowptr[i]; s < A->rowptr[i + 1]; ++s) row[s] = i;
    for (uint32_t s = 0; s < nnz; ++s) order[s] = s;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return A->col[a] != A->col[b] ? A->col[a] < A->col[b] : row[a] < row[b];
    });
    std::vector<double> out(3 * (size_t)nnz + 3);
    const double N = (double)A->N;
    for (uint32_t k = 0; k < nnz; ++k) {
        out[k] = (double)row[order[k]] + 1.0;
        out[nnz + 1 + k] = (double)A->col[order[k]] + 1.0;
        out[2 * (size_t)nnz + 2 + k] = A->val[order[k]];
    }
    out[nnz] = N;
    out[2 * (size_t)nnz + 1] = N;
    out[3 * (size_t)nnz + 2] = 0.0;
    fwrite(out.data(), sizeof(double), out.size(), fp);
    fclose(fp);
}

void nb_sparse_read_mat4(nb_sparse_t* A, const char* url, const char* label) {
    FILE* fp = fopen(url, "rb");
    if (!fp) mat4_fail("Impossible to open matlab v4 file.");
    int32_t info[5];
    if (!mat4_find(fp, label, info) || info[0] != 2) {
        fclose(fp);
        printf("ERROR: Sparse matrix \"%s\" not found in \"%s\".\n", label, url);
        exit(1);
    }
    if (info[2] != 3) mat4_fail("Complex numbers unsupported.");
    const uint32_t nnz = (uint32_t)(info[1] - 1);
    std::vector<double> d(3 * (size_t)nnz + 3);
    if (fread(d.data(), sizeof(double), d.size(), fp) != d.size())
        mat4_fail("The matlab v4 file is corrupted.");
    fclose(fp);
    const uint32_t N = (uint32_t)d[nnz];
    std::vector<uint32_t> rowptr(N + 1, 0), col(nnz);
    std::vector<double> val(nnz);
    for (uint32_t k = 0; k < nnz; ++k) {
        const double r = d[k];
        if (!(r >= 1.0 && r <= (double)N)) mat4_fail("The matlab v4 file is corrupted.");
        rowptr[(uint32_t)r]++;
    }
    for (uint32_t i = 0; i < N; ++i) rowptr[i + 1] += rowptr[i];
    std::vector<uint32_t> fill(rowptr.begin(), rowptr.end() - 1);
    for (uint32_t k = 0; k < nnz; ++k) {
        const uint32_t i = (uint32_t)d[k] - 1;
        const double c = d[nnz + 1 + k];
        if (!(c >= 1.0 && c <= (double)N)) mat4_fail("The matlab v4 file is corrupted.");
        col[fill[i]] = (uint32_t)c - 1;
        val[fill[i]++] = d[2 * (size_t)nnz + 2 + k];
    }
    // columns ascending within a row, values moved with them
    for (uint32_t i = 0; i < N; ++i) {
        std::vector<std::pair<uint32_t, double>> e;
        for (uint32_t s = rowptr[i]; s < rowptr[i + 1]; ++s) e.emplace_back(col[s], val[s]);
        std::stable_sort(e.begin(), e.end(),
                         [](const std::pair<uint32_t, double>& a,
                            const std::pair<uint32_t, double>& b) { return a.first < b.first; });
        for (uint32_t s = rowptr[i], t = 0; s < rowptr[i + 1]; ++s, ++t) {
            col[s] = e[t].first;
            val[s] = e[t].second;
        }
    }
    A->N = N;
    A->rowptr.swap(rowptr);
    A->col.swap(col);
    A->val.swap(val);
    A->dev_pattern = false;
    A->dev_values = false;
}

void nb_mat4_clear(const char* url) {
    FILE* fp = fopen(url, "wb");
    if (fp) fclose(fp);
}

short nb_mat4_exist(const char* url, char* label) {
    FILE* fp = fopen(url, "rb");
    if (!fp) mat4_fail("Impossible to open matlab v4 file.");
    int32_t info[5];
    const bool found = mat4_find(fp, label, info);
    fclose(fp);
    return found ? 1 : 0;
}

void nb_mat4_save_vec(const char* url, char* label, const double* const x, uint32_t N) {
    FILE* fp = fopen(url, "ab");
    if (!fp) mat4_fail("Impossible to open matlab v4 file.");
    mat4_header(fp, 0, (int32_t)N, 1, label);
    fwrite(x, sizeof(double), N, fp);
    fclose(fp);
}

void nb_mat4_read_vec(const char* url, char* label, double* _x) {
    FILE* fp = fopen(url, "rb");
    if (!fp) mat4_fail("Impossible to open matlab v4 file.");
    int32_t info[5];
    if (!mat4_find(fp, label, info) || info[0] != 0 || (info[1] != 1 && info[2] != 1)) {
        fclose(fp);
        printf("ERROR: Vector \"%s\" not found in \"%s\".\n", label, url);
        exit(1);
    }
    if (info[3] == 1) mat4_fail("Complex numbers unsupported.");
    const size_t n = (size_t)(info[1] == 1 ? info[2] : info[1]);
    if (fread(_x, sizeof(double), n, fp) != n) mat4_fail("The matlab v4 file is corrupted.");
    fclose(fp);
}

void nb_sparse_destroy(nb_sparse_t* A) {
    if (!A) return;
    nbx::mat_free(A->dev);
    nbx::ws_free(A->ws);
    if (A->stream) (void)hipStreamDestroy(A->stream);
    delete A;
}

void nb_sparse_reset(nb_sparse_t* A) {
    std::fill(A->val.begin(), A->val.end(), 0.0);
    stale(A);
}

// ---- entry access (sparse.c:134-219) ----------------------------------------
void nb_sparse_set(nb_sparse_t* A, uint32_t i, uint32_t j, double value) {
    const int64_t s = find_slot(A, i, j);
    if (s < 0) not_allocated("nb_sparse_set", A, i, j, value);
    A->val[s] = value;
    stale(A);
}

void nb_sparse_set_identity_row(nb_sparse_t* A, uint32_t row) {
    std::fill(A->val.begin() + A->rowptr[row], A->val.begin() + A->rowptr[row + 1], 0.0);
    nb_sparse_set(A, row, row, 1.0);
}

void nb_sparse_make_diagonal(nb_sparse_t* A, double diag_val) {
    for (uint32_t i = 0; i < A->N; ++i) nb_sparse_set_identity_row(A, i), nb_sparse_set(A, i, i, diag_val);
}

double nb_sparse_get(const nb_sparse_t* const A, uint32_t i, uint32_t j) {
    const int64_t s = find_slot(A, i, j);
    return s < 0 ? 0.0 : A->val[s];
}

double nb_sparse_get_and_set(nb_sparse_t* A, uint32_t i, uint32_t j, double value) {
    const int64_t s = find_slot(A, i, j);
    if (s < 0) not_allocated("nb_sparse_get_and_set", A, i, j, value);
    const double old = A->val[s];
    A->val[s] = value;
    stale(A);
    return old;
}

bool nb_sparse_is_non_zero(const nb_sparse_t* const A, uint32_t i, uint32_t j) {
    return find_slot(A, i, j) >= 0;
}

void nb_sparse_add(nb_sparse_t* A, uint32_t i, uint32_t j, double value) {
    const int64_t s = find_slot(A, i, j);
    if (s < 0) not_allocated("nb_sparse_add", A, i, j, value);
    A->val[s] += value;
    stale(A);
}

void nb_sparse_scale(nb_sparse_t* A, double factor) {
    for (double& v : A->val) v *= factor;
    stale(A);
}

void nb_sparse_multiply_scalar(nb_sparse_t* A, double scalar, uint32_t) {
    nb_sparse_scale(A, scalar);
}

uint32_t nb_sparse_get_size(const nb_sparse_t* const A) { return A->N; }
uint32_t nb_sparse_get_nnz(const nb_sparse_t* const A) { return A->rowptr[A->N]; }

double nb_sparse_get_asym(const nb_sparse_t* const A) {
    // sparse.c:362-379
    double sum = 0;
    for (uint32_t i = 0; i < A->N; ++i)
        for (uint32_t s = A->rowptr[i]; s < A->rowptr[i + 1] && A->col[s] < i; ++s) {
            const double aji = nb_sparse_get(A, A->col[s], i);
            sum += (A->val[s] - aji) * (A->val[s] - aji);
        }
    return sqrt(sum);
}

double nb_sparse_get_frobenius_norm(const nb_sparse_t* const A) {
    double sum = 0;
    for (double v : A->val) sum += v * v;
    return sqrt(sum);
}

// sparse.c:416-430: zero row idx and column idx, diagonal 1, RHS[idx] = value
// and RHS[j] -= A[j][idx] * value for the other rows j of the column.
void nb_sparse_set_Dirichlet_condition(nb_sparse_t* A, double* RHS, uint32_t idx, double value) {
    for (uint32_t s = A->rowptr[idx]; s < A->rowptr[idx + 1]; ++s) {
        const uint32_t jdx = A->col[s];
        if (idx == jdx) {
            A->val[s] = 1.0;
            RHS[idx] = value;
        } else {
            A->val[s] = 0.0;
            const double var = nb_sparse_get_and_set(A, jdx, idx, 0.0);
            RHS[jdx] -= var * value;
        }
    }
    stale(A);
}

// ---- GPU operations -----------------------------------------------------------
// sparse.c:405-414, on the device (bit-identical row sums).
void nb_sparse_multiply_vector(const nb_sparse_t* A, const double* in, double* out, uint32_t) {
    if (A->N == 0) return;
    int rc = sync_device(A);
    double* dx = A->ws.v64[3];
    double* dy = A->ws.v64[4];
    if (rc == 0 && hipMemcpyAsync(dx, in, (size_t)A->N * 8, hipMemcpyHostToDevice, A->stream) != hipSuccess)
        rc = nbx::NBX_ERR_HIP;
    if (rc == 0) rc = nbx::spmv(A->dev, dx, dy, A->stream);
    if (rc == 0 && (hipMemcpyAsync(out, dy, (size_t)A->N * 8, hipMemcpyDeviceToHost, A->stream) != hipSuccess ||
                    hipStreamSynchronize(A->stream) != hipSuccess))
        rc = nbx::NBX_ERR_HIP;
    if (rc) {
        // the reference has no error path here; report and leave NaNs
        fprintf(stderr, "nb_sparse_multiply_vector: GPU failure: %s\n", nbx::last_error());
        for (uint32_t i = 0; i < A->N; ++i) out[i] = NAN;
    }
}

int nb_sparse_solve_CG_precond_Jacobi(const nb_sparse_t* const A, const double* const b,
                                      double* _x, uint32_t max_iter, double tolerance,
                                      uint32_t* niter_performed, double* tolerance_reached,
                                      uint32_t) {
    nbx_solve_info_t info;
    int rc = sparse_solve(A, b, _x, NBX_SOLVE_REFERENCE, tolerance, max_iter, nbx::MixedOpts(), &info);
    if (rc != NBX_OK && rc != NBX_NOT_CONVERGED) {
        fprintf(stderr, "nb_sparse_solve_CG_precond_Jacobi: GPU failure: %s\n", nbx::last_error());
        return -1;
    }
    if (niter_performed) *niter_performed = info.iters;
    if (tolerance_reached) *tolerance_reached = info.tol_reached;
    return rc;
}

// conjugate_gradient.c:13-77 (declared in solvers/conjugate_gradient.h:8-15):
// the Jacobi-PCG kernels with a unit diagonal, same stop test and outputs.
int nb_sparse_solve_conjugate_gradient(const nb_sparse_t* const A, const double* const b,
                                       double* _x, uint32_t max_iter, double tolerance,
                                       uint32_t* niter_performed, double* tolerance_reached,
                                       uint32_t) {
    nbx_solve_info_t info;
    int rc = sparse_solve(A, b, _x, MODE_PLAIN_CG, tolerance, max_iter, nbx::MixedOpts(), &info);
    if (rc != NBX_OK && rc != NBX_NOT_CONVERGED) {
        fprintf(stderr, "nb_sparse_solve_conjugate_gradient: GPU failure: %s\n", nbx::last_error());
        return -1;
    }
    if (niter_performed) *niter_performed = info.iters;
    if (tolerance_reached) *tolerance_reached = info.tol_reached;
    return rc;
}

const char* nbx_last_error(void) { return nbx::last_error(); }

int nbx_device_info(char* name, uint32_t len, uint32_t* n_cu, double* hbm_gb) {
    hipDeviceProp_t p;
    NBX_HIP(hipGetDeviceProperties(&p, 0));
    if (name && len) {
        strncpy(name, p.gcnArchName, len - 1);
        name[len - 1] = 0;
    }
    if (n_cu) *n_cu = (uint32_t)p.multiProcessorCount;
    if (hbm_gb) *hbm_gb = (double)p.totalGlobalMem / 1e9;
    return 0;
}

// Bulk copies in the matrix's CSR order (no reference counterpart: the
// reference exposes entries one at a time, sparse.h:36-43).
int nbx_sparse_export_csr(const nb_sparse_t* A, uint32_t* rowptr, uint32_t* col, double* val) {
    if (rowptr) memcpy(rowptr, A->rowptr.data(), (A->N + 1) * sizeof(uint32_t));
    if (col) memcpy(col, A->col.data(), A->col.size() * sizeof(uint32_t));
    if (val) memcpy(val, A->val.data(), A->val.size() * sizeof(double));
    return NBX_OK;
}

int nbx_sparse_import_values(nb_sparse_t* A, const double* val) {
    memcpy(A->val.data(), val, A->val.size() * sizeof(double));
    stale(A);
    return NBX_OK;
}

int nbx_sparse_is_blocked(const nb_sparse_t* A) {
    if (sync_device(A)) return -1;
    return A->dev.blocked ? 1 : 0;
}

int nbx_sparse_spmv_dev(const nb_sparse_t* A, const double* x_dev, double* y_dev) {
    int rc = sync_device(A);
    if (rc) return rc;
    return nbx::spmv(A->dev, x_dev, y_dev, A->stream);
}

static int sparse_solve(const nb_sparse_t* A, const double* b, double* x, int mode, double tol,
                        uint32_t max_iter, const nbx::MixedOpts& mo, nbx_solve_info_t* info) {
    nbx_solve_info_t tmp;
    if (!info) info = &tmp;
    memset(info, 0, sizeof *info);
    if (A->N == 0) return NBX_OK;
    int rc = sync_device(A);
    if (rc) return rc;
    NBX_TRY(nbx::mat_prepare_sym(A->dev, A->stream));
    info->layout = A->dev.has_sym ? 1 : 0;
    double* db = A->ws.v64[3];
    double* dx = A->ws.v64[4];
    NBX_HIP(hipMemcpyAsync(db, b, (size_t)A->N * 8, hipMemcpyHostToDevice, A->stream));
    NBX_HIP(hipMemcpyAsync(dx, x, (size_t)A->N * 8, hipMemcpyHostToDevice, A->stream));
    nbx::SolveStats st;
    int status = 0;
    if (mode == NBX_SOLVE_REFERENCE || mode == MODE_PLAIN_CG) {
        uint32_t it = 0;
        double tr = 0;
        status = nbx::pcg_jacobi(A->dev, db, dx, max_iter, tol, &it, &tr, A->ws, A->stream, &st,
                                 mode == MODE_PLAIN_CG);
        if (status < 0) return -status;
        info->iters = it;
        info->fp64_iters = it;
        info->tol_reached = tr;
    } else if (mode == NBX_SOLVE_FP64) {
        status = nbx::solve_fp64(A->dev, db, dx, tol, mo.true_residual, max_iter, A->ws, A->stream,
                                 &st);
        if (status > 1) return status;
        info->iters = st.iters;
        info->fp64_iters = st.fp64_iters;
        info->rel_res = st.rel_res;
    } else if (mode == NBX_SOLVE_MIXED) {
        rc = nbx::mat_prepare_mixed(A->dev, A->stream, mo);
        if (rc) return rc;
        status = nbx::solve_mixed(A->dev, db, dx, tol, mo, max_iter,
                                  A->ws, A->stream, &st);
        if (status > 1) return status;
        info->layout |= st.layout & 6;   // the preconditioner's fp16 copy
        info->iters = st.iters;
        info->outer = st.outer;
        info->fp64_iters = st.fp64_iters;
        info->rel_res = st.rel_res;
        info->fallback = st.fallback;
    } else {
        nbx::set_last_error("unknown solve mode %d", mode);
        return NBX_ERR_ARG;
    }
    info->ms = st.ms;
    info->ms_spmv = st.ms_spmv_avg;
    info->status = status;
    NBX_HIP(hipMemcpyAsync(x, dx, (size_t)A->N * 8, hipMemcpyDeviceToHost, A->stream));
    NBX_HIP(hipStreamSynchronize(A->stream));
    return status;
}

int nbx_sparse_solve(const nb_sparse_t* A, const double* b, double* x, int mode, double tol,
                     uint32_t max_iter, double inner_rtol, nbx_solve_info_t* info) {
    return sparse_solve(A, b, x, mode, tol, max_iter, nbx::mixed_opts(inner_rtol), info);
}

int nbx_sparse_solve_ex(const nb_sparse_t* A, const double* b, double* x, int mode, double tol,
                        uint32_t max_iter, const nbx_solve_opts_t* opts, nbx_solve_info_t* info) {
    nbx::MixedOpts mo;
    NBX_TRY(nbx::mixed_opts(opts, &mo));
    return sparse_solve(A, b, x, mode, tol, max_iter, mo, info);
}

}  // extern "C"

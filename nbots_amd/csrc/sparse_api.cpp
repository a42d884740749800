// Drop-in nb_sparse_t: the reference's sparse-matrix API
// (headers/nb/solver_bot/sparse/sparse.h:28-78) with SpMV and Jacobi-PCG on
// the GPU.
//
// Host storage is the reference's own layout (solver_bot/sparse/
// sparse_struct.h:6-11): arrays of row pointers to sorted column indices and
// values, each row allocated with malloc/calloc and released with free, as
// nb_sparse_allocate/nb_sparse_destroy do (sparse_struct.c:23-31,
// sparse.c:114-125).  A matrix built here can therefore go to any libnbots
// function (LU, Cholesky, transposes, permutations), and one built by
// libnbots can come here: the two libraries' nb_sparse_* functions are
// interchangeable on either.
//
// The device copy lives in a registry beside the handle, not in it.  The
// host rows are the truth: every GPU operation compares the device copy's
// pattern and values with the rows (one host pass over them) and uploads
// what changed, so writes made by any library between calls are seen.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/nbots_amd.h"
#include "gauss_seidel.hpp"
#include "internal.hpp"
#include "opts.hpp"

// sparse_struct.h:6-11
struct nb_sparse_s {
    double** rows_values;
    uint32_t** rows_index;
    uint32_t* rows_size;
    uint32_t N;
};

namespace {

// Device copy of one matrix and the host pattern it was built from.
struct Mirror {
    nbx::DevMatrix dev;
    nbx::Workspace ws;
    hipStream_t stream = nullptr;
    bool built = false;
    std::vector<uint32_t> rowptr, col;   // pattern of the device copy
    std::vector<double> val;             // values of the device copy
    nbx::GsSystem gs;                    // scalar CSR copy for Gauss-Seidel
};

std::mutex g_mu;
std::unordered_map<const nb_sparse_t*, Mirror*> g_mirrors;

void drop_mirror(const nb_sparse_t* A) {
    Mirror* m = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_mirrors.find(A);
        if (it == g_mirrors.end()) return;
        m = it->second;
        g_mirrors.erase(it);
    }
    nbx::mat_free(m->dev);
    nbx::ws_free(m->ws);
    nbx::gs_free(m->gs);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

// sparse_struct.c:23-31
nb_sparse_t* allocate(uint32_t N) {
    nb_sparse_t* A = (nb_sparse_t*)malloc(sizeof *A);
    A->rows_values = (double**)malloc((size_t)N * sizeof(double*));
    A->rows_index = (uint32_t**)malloc((size_t)N * sizeof(uint32_t*));
    A->rows_size = (uint32_t*)calloc(N ? N : 1, sizeof(uint32_t));
    A->N = N;
    return A;
}

void free_rows(nb_sparse_t* A) {
    for (uint32_t i = 0; i < A->N; ++i) {
        free(A->rows_values[i]);
        free(A->rows_index[i]);
    }
    free(A->rows_values);
    free(A->rows_index);
    free(A->rows_size);
}

// Position of column j in row i, or -1 (rows hold ascending columns).
int64_t find_slot(const nb_sparse_t* A, uint32_t i, uint32_t j) {
    if (i >= A->N) return -1;
    const uint32_t* b = A->rows_index[i];
    const uint32_t* e = b + A->rows_size[i];
    const uint32_t* it = std::lower_bound(b, e, j);
    if (it == e || *it != j) return -1;
    return it - b;
}

[[noreturn]] void not_allocated(const char* fn, const nb_sparse_t* A, uint32_t i, uint32_t j,
                                double value) {
    // same message and exit as the reference (sparse.c:140-144)
    printf("ERROR: Entry of sparse matrix is not allocated.\n");
    printf("    -> %s(*A=%p, i=%i, j=%i, value=%lf)\n", fn, (const void*)A, i, j, value);
    exit(1);
}

// The mirror of A, brought up to date with A's rows (NBX status in *rc).
Mirror* sync_device(const nb_sparse_t* A, int* rc) {
    Mirror* m;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        Mirror*& slot = g_mirrors[A];
        if (!slot) slot = new Mirror;
        m = slot;
    }
    *rc = nbx::NBX_OK;
    if (!m->stream) {
        const hipError_t e = hipStreamCreate(&m->stream);
        if (e != hipSuccess) {
            nbx::set_last_error("hipStreamCreate: %s", hipGetErrorString(e));
            *rc = nbx::NBX_ERR_HIP;
            return nullptr;
        }
    }
    const uint32_t N = A->N;
    // pattern: the same row sizes and columns as the device copy's?
    bool same = m->built && m->rowptr.size() == (size_t)N + 1;
    for (uint32_t i = 0; same && i < N; ++i) {
        const uint32_t a = m->rowptr[i], n = A->rows_size[i];
        same = m->rowptr[i + 1] - a == n &&
               (n == 0 || memcmp(A->rows_index[i], m->col.data() + a, n * sizeof(uint32_t)) == 0);
    }
    if (!same) {
        uint64_t nnz = 0;
        for (uint32_t i = 0; i < N; ++i) nnz += A->rows_size[i];
        if (nnz >= 0xffffffffull) {
            nbx::set_last_error("matrix has %llu entries (32-bit row offsets)",
                                (unsigned long long)nnz);
            *rc = nbx::NBX_ERR_ARG;
            return nullptr;
        }
        m->rowptr.assign((size_t)N + 1, 0);
        m->col.resize(nnz);
        m->val.resize(nnz);
        for (uint32_t i = 0; i < N; ++i) {
            const uint32_t n = A->rows_size[i];
            m->rowptr[i + 1] = m->rowptr[i] + n;
            if (n) memcpy(m->col.data() + m->rowptr[i], A->rows_index[i], n * sizeof(uint32_t));
        }
    }
    // values: copy the rows into the staging array, noting any change
    bool vals_same = same;
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t n = A->rows_size[i];
        if (!n) continue;
        double* dst = m->val.data() + m->rowptr[i];
        if (vals_same && memcmp(dst, A->rows_values[i], n * sizeof(double)) == 0) continue;
        vals_same = false;
        memcpy(dst, A->rows_values[i], n * sizeof(double));
    }
    if (!same || !vals_same) {
        *rc = nbx::mat_upload_csr(m->dev, N, m->rowptr.data(), m->col.data(), m->val.data(),
                                  same, m->stream);
        if (!*rc) *rc = nbx::ws_alloc(m->ws, N);
        m->built = *rc == nbx::NBX_OK;
        if (*rc) return nullptr;
    }
    return m;
}

}  // namespace

static int sparse_solve(const nb_sparse_t* A, const double* b, double* x, int mode, double tol,
                        uint32_t max_iter, const nbx::MixedOpts& mo, nbx_solve_info_t* info);
// solve mode of the drop-in nb_sparse_solve_conjugate_gradient (not public)
constexpr int MODE_PLAIN_CG = 0x100;

extern "C" {

// ---- creation (sparse.c:20-95) ------------------------------------------------
// Row perm[i]*vpn+k1 (or i*vpn+k1) holds the columns {adj(i), i} x vpn in
// ascending order; columns are not permuted.  Values start at zero.
nb_sparse_t* nb_sparse_create(const nb_graph_t* const graph, const uint32_t* const perm,
                              uint32_t vpn) {
    nb_sparse_t* A = allocate(graph->N * vpn);
    for (uint32_t i = 0; i < graph->N; ++i) {
        const uint32_t size = (graph->N_adj[i] + 1) * vpn;
        for (uint32_t k1 = 0; k1 < vpn; ++k1) {
            const uint32_t r = (perm ? perm[i] : i) * vpn + k1;
            A->rows_size[r] = size;
            A->rows_values[r] = (double*)calloc(size ? size : 1, sizeof(double));
            uint32_t* row = (uint32_t*)malloc((size ? size : 1) * sizeof(uint32_t));
            A->rows_index[r] = row;
            for (uint32_t j = 0; j < graph->N_adj[i]; ++j)
                for (uint32_t k2 = 0; k2 < vpn; ++k2) row[j * vpn + k2] = graph->adj[i][j] * vpn + k2;
            for (uint32_t k2 = 0; k2 < vpn; ++k2) row[graph->N_adj[i] * vpn + k2] = i * vpn + k2;
            std::sort(row, row + size);
        }
    }
    return A;
}

nb_sparse_t* nb_sparse_clone(nb_sparse_t* A) {
    nb_sparse_t* B = allocate(A->N);
    for (uint32_t i = 0; i < A->N; ++i) {
        const uint32_t n = A->rows_size[i];
        B->rows_size[i] = n;
        B->rows_index[i] = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
        B->rows_values[i] = (double*)calloc(n ? n : 1, sizeof(double));
        if (n) {
            memcpy(B->rows_index[i], A->rows_index[i], n * sizeof(uint32_t));
            memcpy(B->rows_values[i], A->rows_values[i], n * sizeof(double));
        }
    }
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
        for (uint32_t s = 0; s < A->rows_size[i]; ++s)
            fprintf(fp, "%i %i %e \n", i, A->rows_index[i][s], A->rows_values[i][s]);
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
    std::vector<uint32_t> row, col;
    std::vector<double> val;
    for (uint32_t i = 0; i < A->N; ++i)
        for (uint32_t s = 0; s < A->rows_size[i]; ++s) {
            row.push_back(i);
            col.push_back(A->rows_index[i][s]);
            val.push_back(A->rows_values[i][s]);
        }
    const uint32_t nnz = (uint32_t)col.size();
    mat4_header(fp, 2, (int32_t)nnz + 1, 3, label);
    std::vector<uint32_t> order(nnz);
    for (uint32_t s = 0; s < nnz; ++s) order[s] = s;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return col[a] != col[b] ? col[a] < col[b] : row[a] < row[b];
    });
    std::vector<double> out(3 * (size_t)nnz + 3);
    const double N = (double)A->N;
    for (uint32_t k = 0; k < nnz; ++k) {
        out[k] = (double)row[order[k]] + 1.0;
        out[nnz + 1 + k] = (double)col[order[k]] + 1.0;
        out[2 * (size_t)nnz + 2 + k] = val[order[k]];
    }
    out[nnz] = N;
    out[2 * (size_t)nnz + 1] = N;
    out[3 * (size_t)nnz + 2] = 0.0;
    fwrite(out.data(), sizeof(double), out.size(), fp);
    fclose(fp);
}

// matlab_v4.c:25-184: A takes the record's N and rows (its previous rows are
// released; the reference overwrites the row arrays without releasing them).
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
    std::vector<uint32_t> size(N, 0);
    for (uint32_t k = 0; k < nnz; ++k) {
        const double r = d[k], c = d[nnz + 1 + k];
        if (!(r >= 1.0 && r <= (double)N) || !(c >= 1.0 && c <= (double)N))
            mat4_fail("The matlab v4 file is corrupted.");
        size[(uint32_t)r - 1]++;
    }
    drop_mirror(A);
    free_rows(A);
    nb_sparse_t* B = allocate(N);
    for (uint32_t i = 0; i < N; ++i) {
        B->rows_size[i] = 0;
        B->rows_index[i] = (uint32_t*)calloc(size[i] ? size[i] : 1, sizeof(uint32_t));
        B->rows_values[i] = (double*)calloc(size[i] ? size[i] : 1, sizeof(double));
    }
    for (uint32_t k = 0; k < nnz; ++k) {
        const uint32_t i = (uint32_t)d[k] - 1;
        const uint32_t s = B->rows_size[i]++;
        B->rows_index[i][s] = (uint32_t)d[nnz + 1 + k] - 1;
        B->rows_values[i][s] = d[2 * (size_t)nnz + 2 + k];
    }
    // columns ascending within a row, values moved with them
    for (uint32_t i = 0; i < N; ++i) {
        std::vector<std::pair<uint32_t, double>> e(B->rows_size[i]);
        for (uint32_t s = 0; s < B->rows_size[i]; ++s)
            e[s] = {B->rows_index[i][s], B->rows_values[i][s]};
        std::stable_sort(e.begin(), e.end(),
                         [](const std::pair<uint32_t, double>& a,
                            const std::pair<uint32_t, double>& b) { return a.first < b.first; });
        for (uint32_t s = 0; s < B->rows_size[i]; ++s) {
            B->rows_index[i][s] = e[s].first;
            B->rows_values[i][s] = e[s].second;
        }
    }
    *A = *B;
    free(B);
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

// sparse.c:114-125, plus the device copy if there is one.  Works on a matrix
// that libnbots allocated (nb_sparse_allocate, the LU factors).
void nb_sparse_destroy(nb_sparse_t* A) {
    if (!A) return;
    drop_mirror(A);
    free_rows(A);
    free(A);
}

void nb_sparse_reset(nb_sparse_t* A) {
    for (uint32_t i = 0; i < A->N; ++i)
        if (A->rows_size[i]) memset(A->rows_values[i], 0, A->rows_size[i] * sizeof(double));
}

// ---- entry access (sparse.c:134-219) ----------------------------------------
void nb_sparse_set(nb_sparse_t* A, uint32_t i, uint32_t j, double value) {
    const int64_t s = find_slot(A, i, j);
    if (s < 0) not_allocated("nb_sparse_set", A, i, j, value);
    A->rows_values[i][s] = value;
}

void nb_sparse_set_identity_row(nb_sparse_t* A, uint32_t row) {
    if (A->rows_size[row]) memset(A->rows_values[row], 0, A->rows_size[row] * sizeof(double));
    nb_sparse_set(A, row, row, 1.0);
}

void nb_sparse_make_diagonal(nb_sparse_t* A, double diag_val) {
    for (uint32_t i = 0; i < A->N; ++i) {
        if (A->rows_size[i]) memset(A->rows_values[i], 0, A->rows_size[i] * sizeof(double));
        nb_sparse_set(A, i, i, diag_val);
    }
}

double nb_sparse_get(const nb_sparse_t* const A, uint32_t i, uint32_t j) {
    const int64_t s = find_slot(A, i, j);
    return s < 0 ? 0.0 : A->rows_values[i][s];
}

double nb_sparse_get_and_set(nb_sparse_t* A, uint32_t i, uint32_t j, double value) {
    const int64_t s = find_slot(A, i, j);
    if (s < 0) not_allocated("nb_sparse_get_and_set", A, i, j, value);
    const double old = A->rows_values[i][s];
    A->rows_values[i][s] = value;
    return old;
}

bool nb_sparse_is_non_zero(const nb_sparse_t* const A, uint32_t i, uint32_t j) {
    return find_slot(A, i, j) >= 0;
}

void nb_sparse_add(nb_sparse_t* A, uint32_t i, uint32_t j, double value) {
    const int64_t s = find_slot(A, i, j);
    if (s < 0) not_allocated("nb_sparse_add", A, i, j, value);
    A->rows_values[i][s] += value;
}

void nb_sparse_scale(nb_sparse_t* A, double factor) {
    for (uint32_t i = 0; i < A->N; ++i)
        for (uint32_t s = 0; s < A->rows_size[i]; ++s) A->rows_values[i][s] *= factor;
}

void nb_sparse_multiply_scalar(nb_sparse_t* A, double scalar, uint32_t) {
    nb_sparse_scale(A, scalar);
}

uint32_t nb_sparse_get_size(const nb_sparse_t* const A) { return A->N; }

uint32_t nb_sparse_get_nnz(const nb_sparse_t* const A) {
    uint32_t nnz = 0;
    for (uint32_t i = 0; i < A->N; ++i) nnz += A->rows_size[i];
    return nnz;
}

double nb_sparse_get_asym(const nb_sparse_t* const A) {
    // sparse.c:362-379: the entries left of the diagonal against their mirrors
    double sum = 0;
    for (uint32_t i = 0; i < A->N; ++i)
        for (uint32_t s = 0; s < A->rows_size[i] && A->rows_index[i][s] < i; ++s) {
            const double aji = nb_sparse_get(A, A->rows_index[i][s], i);
            const double d = A->rows_values[i][s] - aji;
            sum += d * d;
        }
    return sqrt(sum);
}

double nb_sparse_get_frobenius_norm(const nb_sparse_t* const A) {
    double sum = 0;
    for (uint32_t i = 0; i < A->N; ++i)
        for (uint32_t s = 0; s < A->rows_size[i]; ++s) sum += A->rows_values[i][s] * A->rows_values[i][s];
    return sqrt(sum);
}

// sparse.c:416-430: zero row idx and column idx, diagonal 1, RHS[idx] = value
// and RHS[j] -= A[j][idx] * value for the other rows j of the column.
void nb_sparse_set_Dirichlet_condition(nb_sparse_t* A, double* RHS, uint32_t idx, double value) {
    for (uint32_t s = 0; s < A->rows_size[idx]; ++s) {
        const uint32_t jdx = A->rows_index[idx][s];
        if (idx == jdx) {
            A->rows_values[idx][s] = 1.0;
            RHS[idx] = value;
        } else {
            A->rows_values[idx][s] = 0.0;
            const double var = nb_sparse_get_and_set(A, jdx, idx, 0.0);
            RHS[jdx] -= var * value;
        }
    }
}

// ---- GPU operations -----------------------------------------------------------
// sparse.c:405-414, on the device (bit-identical row sums).
void nb_sparse_multiply_vector(const nb_sparse_t* A, const double* in, double* out, uint32_t) {
    if (A->N == 0) return;
    int rc = 0;
    Mirror* m = sync_device(A, &rc);
    if (!rc) {
        double* dx = m->ws.v64[3];
        double* dy = m->ws.v64[4];
        if (hipMemcpyAsync(dx, in, (size_t)A->N * 8, hipMemcpyHostToDevice, m->stream) != hipSuccess)
            rc = nbx::NBX_ERR_HIP;
        if (!rc) rc = nbx::spmv(m->dev, dx, dy, m->stream);
        if (!rc && (hipMemcpyAsync(out, dy, (size_t)A->N * 8, hipMemcpyDeviceToHost, m->stream) !=
                        hipSuccess ||
                    hipStreamSynchronize(m->stream) != hipSuccess))
            rc = nbx::NBX_ERR_HIP;
    }
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

// solvers/gauss_seidel.c:13-63 (declared in solvers/gauss_seidel.h:8-15) on
// the device: the same sweeps, the forward substitution by dependency
// levels with each row summed in the reference's order (gauss_seidel.hip).
int nb_sparse_solve_Gauss_Seidel(const nb_sparse_t* const A, const double* const b, double* _x,
                                 uint32_t max_iter, double tolerance, uint32_t* niter_performed,
                                 double* tolerance_reached, uint32_t) {
    const uint32_t N = A->N;
    int rc = 0;
    Mirror* m = N ? sync_device(A, &rc) : nullptr;
    uint32_t it = 0;
    double err = 0.0;
    int status = 0;
    if (!rc && N) rc = nbx::gs_prepare(m->gs, N, m->rowptr.data(), m->col.data(), m->val.data(),
                                       m->stream);
    if (!rc && N) {
        double* db = m->ws.v64[3];
        double* dx = m->ws.v64[4];
        if (hipMemcpyAsync(db, b, (size_t)N * 8, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
            hipMemcpyAsync(dx, _x, (size_t)N * 8, hipMemcpyHostToDevice, m->stream) != hipSuccess)
            rc = nbx::NBX_ERR_HIP;
        if (!rc) {
            status = nbx::gauss_seidel(m->gs, db, dx, max_iter, tolerance, &it, &err, m->stream);
            if (status > 1) rc = status;
        }
        if (!rc && nbx::copy_on(_x, dx, (size_t)N * 8, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
            rc = nbx::NBX_ERR_HIP;
    } else if (!rc) {
        // no rows: the first sweep's error is the empty sum 0, which meets
        // any tolerance; without a sweep it stays 1e10 (gauss_seidel.c:27-41)
        err = max_iter ? 0.0 : 1e5;
        status = max_iter ? 0 : (1e10 > tolerance * tolerance ? 1 : 0);
    }
    if (rc) {
        fprintf(stderr, "nb_sparse_solve_Gauss_Seidel: GPU failure: %s\n", nbx::last_error());
        return -1;
    }
    if (niter_performed) *niter_performed = it;
    if (tolerance_reached) *tolerance_reached = err;
    return status;
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

// Bulk copies in the matrix's row order (no reference counterpart: the
// reference exposes entries one at a time, sparse.h:36-43).
int nbx_sparse_export_csr(const nb_sparse_t* A, uint32_t* rowptr, uint32_t* col, double* val) {
    uint32_t off = 0;
    if (rowptr) rowptr[0] = 0;
    for (uint32_t i = 0; i < A->N; ++i) {
        const uint32_t n = A->rows_size[i];
        if (col && n) memcpy(col + off, A->rows_index[i], n * sizeof(uint32_t));
        if (val && n) memcpy(val + off, A->rows_values[i], n * sizeof(double));
        off += n;
        if (rowptr) rowptr[i + 1] = off;
    }
    return NBX_OK;
}

int nbx_sparse_import_values(nb_sparse_t* A, const double* val) {
    uint32_t off = 0;
    for (uint32_t i = 0; i < A->N; ++i) {
        const uint32_t n = A->rows_size[i];
        if (n) memcpy(A->rows_values[i], val + off, n * sizeof(double));
        off += n;
    }
    return NBX_OK;
}

int nbx_sparse_is_blocked(const nb_sparse_t* A) {
    int rc = 0;
    Mirror* m = sync_device(A, &rc);
    if (rc) return -1;
    return m->dev.blocked ? 1 : 0;
}

int nbx_sparse_spmv_dev(const nb_sparse_t* A, const double* x_dev, double* y_dev) {
    int rc = 0;
    Mirror* m = sync_device(A, &rc);
    if (rc) return rc;
    return nbx::spmv(m->dev, x_dev, y_dev, m->stream);
}

static int sparse_solve(const nb_sparse_t* A, const double* b, double* x, int mode, double tol,
                        uint32_t max_iter, const nbx::MixedOpts& mo, nbx_solve_info_t* info) {
    nbx_solve_info_t tmp;
    if (!info) info = &tmp;
    memset(info, 0, sizeof *info);
    if (A->N == 0) return NBX_OK;
    int rc = 0;
    Mirror* m = sync_device(A, &rc);
    if (rc) return rc;
    NBX_TRY(nbx::mat_prepare_sym(m->dev, m->stream));
    m->ws.replay = mo.replay;
    info->layout = m->dev.has_sym ? 1 : 0;
    double* db = m->ws.v64[3];
    double* dx = m->ws.v64[4];
    NBX_HIP(hipMemcpyAsync(db, b, (size_t)A->N * 8, hipMemcpyHostToDevice, m->stream));
    NBX_HIP(hipMemcpyAsync(dx, x, (size_t)A->N * 8, hipMemcpyHostToDevice, m->stream));
    nbx::SolveStats st;
    int status = 0;
    if (mode == NBX_SOLVE_REFERENCE || mode == MODE_PLAIN_CG) {
        uint32_t it = 0;
        double tr = 0;
        status = nbx::pcg_jacobi(m->dev, db, dx, max_iter, tol, &it, &tr, m->ws, m->stream, &st,
                                 mode == MODE_PLAIN_CG);
        if (status < 0) return -status;
        info->iters = it;
        info->fp64_iters = it;
        info->tol_reached = tr;
    } else if (mode == NBX_SOLVE_FP64) {
        status = nbx::solve_fp64(m->dev, db, dx, tol, mo.true_residual, max_iter, m->ws, m->stream,
                                 &st);
        if (status > 1) return status;
        info->iters = st.iters;
        info->fp64_iters = st.fp64_iters;
        info->rel_res = st.rel_res;
    } else if (mode == NBX_SOLVE_MIXED) {
        rc = nbx::mat_prepare_mixed(m->dev, m->stream, mo);
        if (rc) return rc;
        status = nbx::solve_mixed(m->dev, db, dx, tol, mo, max_iter, m->ws, m->stream, &st);
        if (status > 1) return status;
        info->layout |= st.layout & 30;   // the preconditioner's fp16 copy
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
    NBX_HIP(hipMemcpyAsync(x, dx, (size_t)A->N * 8, hipMemcpyDeviceToHost, m->stream));
    NBX_HIP(hipStreamSynchronize(m->stream));
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

// Device-resident T3 plane-stress system for gfx950: sparsity pattern,
// stiffness assembly, boundary conditions, strain/stress and SIMP element
// energies.
//
// Pattern.  The reference links nodes that share a mesh edge
// (geometric_bot/mesh/mesh2D/load_graph.c:230-307) and gives node v the
// scalar rows 2v, 2v+1 with columns {2w, 2w+1 : w in adj(v) u {v}} sorted
// (solver_bot/sparse/sparse.c:20-60).  For triangles the edge neighbours of
// v are exactly the other vertices of the triangles around v, so the device
// builds node->element incidence by count / exclusive scan / fill / per-node
// sort, then each node's sorted, de-duplicated neighbour set.  The result
// is bit-identical to the reference's pattern.
//
// Assembly.  The reference adds element matrices into K in element-id order
// (pipeline.c:60-69, :232-264), so every entry of K is a sequential sum over
// the elements that touch it, in ascending id.  Here one thread owns one
// stored 2x2 block (v,w) and walks the (sorted) elements around v, adding
// the (v,w) part of each element matrix containing w.  With the element
// math written in the reference's operation order and no FMA contraction,
// K is bit-identical to the reference's, and no atomics are needed.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "fem.hpp"
#include "rows.hpp"

namespace nbx {

constexpr uint32_t FREE = 0xffffffffu;
constexpr int MAX_ADJ = 96;   // neighbour-set capacity per node

// ---------------------------------------------------------------------------
// element math (reference operation order)
// ---------------------------------------------------------------------------
// T3 tables: pde_bot/finite_element/element.c:54-73
__device__ __constant__ double c_dpsi[3] = {-1.0, 1.0, 0.0};
__device__ __constant__ double c_deta[3] = {-1.0, 0.0, 1.0};
constexpr double T3_W = 0.5;
constexpr double T3_N = 0.33333333333333333333333333333;

// nb_fem_get_jacobian + nb_fem_get_derivatives (pde_bot/finite_element/
// utils.c:9-60).  Returns false for a distorted element (detJ < 0).
__device__ __forceinline__ bool t3_geom(const double2* __restrict__ nod, const uint32_t* n3,
                                        double& detJ, double dx[3], double dy[3]) {
    double dx_dpsi = 0.0, dy_dpsi = 0.0, dx_deta = 0.0, dy_deta = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double2 p = nod[n3[i]];
        dx_dpsi += c_dpsi[i] * p.x;
        dx_deta += c_deta[i] * p.x;
        dy_dpsi += c_dpsi[i] * p.y;
        dy_deta += c_deta[i] * p.y;
    }
    detJ = dx_dpsi * dy_deta - dy_dpsi * dx_deta;
    const double j0 = dy_deta / detJ, j1 = -dy_dpsi / detJ;
    const double j2 = -dx_deta / detJ, j3 = dx_dpsi / detJ;
    if (detJ < 0) return false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        dx[i] = j0 * c_dpsi[i] + j1 * c_deta[i];
        dy[i] = j2 * c_dpsi[i] + j3 * c_deta[i];
    }
    return true;
}

// plane stress D (pde_bot/common_solid_mechanics/formulas.c:48-54)
__device__ __forceinline__ void plane_stress(double E, double nu, double D[4]) {
    D[0] = E / (1.0 - nu * nu);
    D[1] = nu * D[0];
    D[2] = D[0];
    D[3] = E / (2.0 * (1.0 + nu));
}

__device__ __forceinline__ void elem_D(uint32_t e, const ElemParams& P, double D[4]) {
    const bool en = P.enabled ? (P.enabled[e] != 0) : true;
    if (!en) {
        D[0] = D[1] = D[2] = D[3] = 1e-6;   // pipeline.c:93
    } else {
        plane_stress(P.elemE ? P.elemE[e] : P.E, P.nu, D);
    }
    if (P.dmg) {   // static_damage2D.c:534-540
        const double d = P.dmg[e];
        D[0] *= (1.0 - d);
        D[1] *= (1.0 - d);
        D[2] *= (1.0 - d);
        D[3] *= (1.0 - d);
    }
}

// The 2x2 part (node a, node b) of the element matrix, each entry as the
// reference forms it: Ke starts at 0 (pipeline.c:142) and receives one
// Gauss-point term (pipeline.c:197-212).
__device__ __forceinline__ void ke_block(const double D[4], double t, double detJ,
                                         const double dx[3], const double dy[3], int a, int b,
                                         double k[4]) {
    const double wp = T3_W;
    k[0] = 0.0 + (dx[a] * dx[b] * D[0] + dy[a] * dy[b] * D[3]) * detJ * t * wp;
    k[1] = 0.0 + (dx[a] * dy[b] * D[1] + dy[a] * dx[b] * D[3]) * detJ * t * wp;
    k[2] = 0.0 + (dy[a] * dx[b] * D[1] + dx[a] * dy[b] * D[3]) * detJ * t * wp;
    k[3] = 0.0 + (dy[a] * dy[b] * D[2] + dx[a] * dx[b] * D[3]) * detJ * t * wp;
}

// ---------------------------------------------------------------------------
// exclusive scan of uint32 (three-phase, any length)
// ---------------------------------------------------------------------------
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;

__global__ void k_scan_tile(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n,
                            uint32_t* __restrict__ tile_sums) {
    __shared__ uint32_t sh[BLOCK];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        s += v[i];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < BLOCK; off <<= 1) {   // Hillis-Steele inclusive
        uint32_t t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = sh[threadIdx.x] - s;           // exclusive prefix of this thread
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == BLOCK - 1 && tile_sums) tile_sums[blockIdx.x] = sh[BLOCK - 1];
}

__global__ void k_scan_add(uint32_t* __restrict__ out, size_t n, const uint32_t* __restrict__ offs) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] += offs[i / SCAN_TILE];
}

// out[0..n) = exclusive scan of in; out[n] = total.  out has n+1 entries.
static int scan_u32(const uint32_t* in, uint32_t* out, size_t n, hipStream_t s) {
    const size_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles <= 1) {
        uint32_t* sums = nullptr;
        NBX_HIP(hipMalloc(&sums, sizeof(uint32_t)));
        hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(BLOCK), 0, s, in, out, n, sums);
        NBX_HIP(hipMemcpyAsync(out + n, sums, sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        NBX_HIP(hipStreamSynchronize(s));
        (void)hipFree(sums);
        return NBX_OK;
    }
    uint32_t *sums = nullptr, *soff = nullptr;
    NBX_HIP(hipMalloc(&sums, tiles * sizeof(uint32_t)));
    NBX_HIP(hipMalloc(&soff, (tiles + 1) * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)tiles), dim3(BLOCK), 0, s, in, out, n, sums);
    int rc = scan_u32(sums, soff, tiles, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scan_add, dim3(grid_for(n)), dim3(BLOCK), 0, s, out, n, soff);
    NBX_HIP(hipMemcpyAsync(out + n, soff + tiles, sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    NBX_HIP(hipStreamSynchronize(s));
    (void)hipFree(sums);
    (void)hipFree(soff);
    return NBX_OK;
}

// ---------------------------------------------------------------------------
// pattern
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double* p, size_t n, double v) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_count_inc(const uint32_t* __restrict__ conn, uint32_t ne, uint32_t nn,
                            uint32_t* cnt, unsigned* err) {
    const size_t t = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= 3 * (size_t)ne) return;
    const uint32_t v = conn[t];
    if (v >= nn) {
        atomicOr(err, 1u);
        return;
    }
    atomicAdd(&cnt[v], 1u);
}

__global__ void k_fill_inc(const uint32_t* __restrict__ conn, uint32_t ne,
                           const uint32_t* __restrict__ ptr, uint32_t* pos, uint32_t* n2e) {
    const size_t t = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= 3 * (size_t)ne) return;
    const uint32_t v = conn[t];
    const uint32_t k = atomicAdd(&pos[v], 1u);
    n2e[ptr[v] + k] = (uint32_t)(t / 3);
}

// ascending element ids per node (insertion sort; lists are short)
__global__ void k_sort_inc(const uint32_t* __restrict__ ptr, uint32_t* n2e, uint32_t nn) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= nn) return;
    const uint32_t b = ptr[v], e = ptr[v + 1];
    for (uint32_t i = b + 1; i < e; ++i) {
        const uint32_t key = n2e[i];
        uint32_t j = i;
        while (j > b && n2e[j - 1] > key) {
            n2e[j] = n2e[j - 1];
            --j;
        }
        n2e[j] = key;
    }
}

// sorted, de-duplicated {v} u {vertices of the elements around v}
__device__ int node_adj(uint32_t v, const uint32_t* __restrict__ ptr,
                        const uint32_t* __restrict__ n2e, const uint32_t* __restrict__ conn,
                        uint32_t* list) {
    int n = 0;
    list[n++] = v;
    for (uint32_t i = ptr[v]; i < ptr[v + 1]; ++i) {
        const uint32_t e = n2e[i];
        for (int a = 0; a < 3; ++a) {
            const uint32_t w = conn[3 * e + a];
            int j = n;
            bool dup = false;
            for (int q = 0; q < n; ++q)
                if (list[q] == w) {
                    dup = true;
                    break;
                }
            if (dup) continue;
            if (n >= MAX_ADJ) return -1;
            while (j > 0 && list[j - 1] > w) {
                list[j] = list[j - 1];
                --j;
            }
            list[j] = w;
            ++n;
        }
    }
    return n;
}

__global__ void k_row_len(const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ n2e,
                          const uint32_t* __restrict__ conn, uint32_t nn, uint8_t* rowlen,
                          unsigned* err) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= nn) return;
    uint32_t list[MAX_ADJ];
    const int n = node_adj(v, ptr, n2e, conn, list);
    if (n < 0 || n > 255) {
        atomicOr(err, 2u);
        rowlen[v] = 0;
        return;
    }
    rowlen[v] = (uint8_t)n;
}

// slice width (max row length over 64 rows) * 64 stored blocks
__global__ void k_slice_blocks(const uint8_t* __restrict__ rowlen, uint32_t nn, uint32_t ns,
                               uint32_t* sblk) {
    const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= ns) return;
    uint32_t w = 0;
    for (uint32_t v = s * SLICE; v < nn && v < (s + 1) * SLICE; ++v) w = max(w, (uint32_t)rowlen[v]);
    sblk[s] = w * SLICE;
}

__global__ void k_fill_cols(const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ n2e,
                            const uint32_t* __restrict__ conn, uint32_t nn,
                            const uint32_t* __restrict__ sptr, uint32_t* bcol) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= nn) return;
    uint32_t list[MAX_ADJ];
    const int n = node_adj(v, ptr, n2e, conn, list);
    const uint32_t s = v >> 6;
    const uint32_t width = (sptr[s + 1] - sptr[s]) / SLICE;
    size_t idx = (size_t)sptr[s] + (v & 63);
    for (uint32_t k = 0; k < width; ++k, idx += SLICE) bcol[idx] = (int)k < n ? list[k] : v;
}

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_assemble(Sell2 K, const double2* __restrict__ nod,
                                                    const uint32_t* __restrict__ conn,
                                                    const uint32_t* __restrict__ ptr,
                                                    const uint32_t* __restrict__ n2e, ElemParams P,
                                                    unsigned* err) {
    const size_t idx = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= K.n_blocks) return;
    // locate (slice, k, lane) of this stored block
    uint32_t lo = 0, hi = K.n_slices;       // slice: largest s with sptr[s] <= idx
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (K.slice_ptr[mid] <= idx) lo = mid;
        else hi = mid;
    }
    const uint32_t s = lo;
    const size_t off = idx - K.slice_ptr[s];
    const uint32_t k = (uint32_t)(off / SLICE), lane = (uint32_t)(off % SLICE);
    const uint32_t v = s * SLICE + lane;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (v < K.n_nodes && k < K.rowlen[v]) {
        const uint32_t w = K.bcol[idx];
        for (uint32_t i = ptr[v]; i < ptr[v + 1]; ++i) {
            const uint32_t e = n2e[i];
            const uint32_t n3[3] = {conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]};
            int a = -1, b = -1;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (n3[q] == v && a < 0) a = q;
                if (n3[q] == w && b < 0) b = q;
            }
            if (b < 0) continue;
            double detJ, dx[3], dy[3];
            if (!t3_geom(nod, n3, detJ, dx, dy)) {
                atomicOr(err, 4u);
                continue;
            }
            double D[4], kb[4];
            elem_D(e, P, D);
            ke_block(D, P.thickness, detJ, dx, dy, a, b, kb);
            acc[0] += kb[0];
            acc[1] += kb[1];
            acc[2] += kb[2];
            acc[3] += kb[3];
        }
    }
    reinterpret_cast<double2*>(K.val + ent(K.n_blocks, idx, 0, 0))[0] = make_double2(acc[0], acc[1]);
    reinterpret_cast<double2*>(K.val + ent(K.n_blocks, idx, 1, 0))[0] = make_double2(acc[2], acc[3]);
}

// F from self weight: Fe (pipeline.c:224-228) gathered per node in element
// order (pipeline.c:260-262), starting from F = 0 (pipeline.c:58).
__global__ void k_body_force(const double2* __restrict__ nod, const uint32_t* __restrict__ conn,
                             const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ n2e,
                             ElemParams P, double gx, double gy, double density,
                             double* __restrict__ F, uint32_t nn) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= nn) return;
    double f0 = 0.0, f1 = 0.0;
    for (uint32_t i = ptr[v]; i < ptr[v + 1]; ++i) {
        const uint32_t e = n2e[i];
        const uint32_t n3[3] = {conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]};
        double detJ, dx[3], dy[3];
        if (!t3_geom(nod, n3, detJ, dx, dy)) continue;
        const bool en = P.enabled ? (P.enabled[e] != 0) : true;
        const double rho = en ? density : 1e-6;
        const double fx = gx * rho, fy = gy * rho;
        const double integral = T3_N * detJ * P.thickness * T3_W;
        for (int a = 0; a < 3; ++a)
            if (n3[a] == v) {
                f0 += 0.0 + integral * fx;
                f1 += 0.0 + integral * fy;
            }
    }
    F[2 * v] = f0;
    F[2 * v + 1] = f1;
}

// ---------------------------------------------------------------------------
// boundary conditions
// ---------------------------------------------------------------------------
// Neumann: per touched dof, its additions in the reference's order.
__global__ void k_neumann(double* __restrict__ F, const uint32_t* __restrict__ dof,
                          const uint32_t* __restrict__ aptr, const double* __restrict__ add,
                          uint32_t m) {
    const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= m) return;
    double f = F[dof[t]];
    for (uint32_t i = aptr[t]; i < aptr[t + 1]; ++i) f += add[i];
    F[dof[t]] = f;
}

__global__ void k_bc_scatter(const uint32_t* __restrict__ dof, const uint32_t* __restrict__ rank,
                             const double* __restrict__ vf, const double* __restrict__ vl,
                             uint32_t m, uint32_t* brank, double* bvf, double* bvl) {
    const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= m) return;
    brank[dof[t]] = rank[t];
    bvf[dof[t]] = vf[t];
    bvl[dof[t]] = vl[t];
}

// RHS of the symmetric elimination (sparse.c:416-430, repeated in the
// order of set_bconditions.c:52-61): a constrained row takes its last
// value; a free row r subtracts A(r,c) * value(c) for each constrained
// column c, in the order the columns were first eliminated.
__global__ void k_bc_rhs(Sell2 K, const uint32_t* __restrict__ rank, const double* __restrict__ vf,
                         const double* __restrict__ vl, double* __restrict__ F) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= K.n_nodes) return;
    const size_t base = (size_t)K.slice_ptr[v >> 6] + (v & 63);
    const int len = K.rowlen[v];
    for (int i = 0; i < 2; ++i) {
        const uint32_t r = 2 * v + i;
        if (rank[r] != FREE) {
            F[r] = vl[r];
            continue;
        }
        double f = F[r];
        int64_t last = -1;
        for (;;) {
            uint32_t best = FREE;
            double a = 0.0, val = 0.0;
            size_t idx = base;
            for (int k = 0; k < len; ++k, idx += SLICE) {
                const uint32_t w = K.bcol[idx];
                for (int j = 0; j < 2; ++j) {
                    const uint32_t c = 2 * w + j, rk = rank[c];
                    if (rk != FREE && (int64_t)rk > last && (best == FREE || rk < best)) {
                        best = rk;
                        a = K.val[ent(K.n_blocks, idx, i, j)];
                        val = vf[c];
                    }
                }
            }
            if (best == FREE) break;
            f -= a * val;
            last = best;
        }
        F[r] = f;
    }
}

__global__ void k_bc_matrix(Sell2 K, const uint32_t* __restrict__ rank) {
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= K.n_nodes) return;
    size_t idx = (size_t)K.slice_ptr[v >> 6] + (v & 63);
    const int len = K.rowlen[v];
    for (int k = 0; k < len; ++k, idx += SLICE) {
        const uint32_t w = K.bcol[idx];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                const uint32_t r = 2 * v + i, c = 2 * w + j;
                double* a = K.val + ent(K.n_blocks, idx, i, j);
                if (rank[r] != FREE) *a = (r == c) ? 1.0 : 0.0;
                else if (rank[c] != FREE) *a = 0.0;
            }
    }
}

__global__ void k_zero_constrained(double* __restrict__ x, const uint32_t* __restrict__ rank,
                                   uint32_t n) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && rank[i] != FREE) x[i] = 0.0;
}

// ---------------------------------------------------------------------------
// post-processing and SIMP
// ---------------------------------------------------------------------------
// strain: pipeline.c:266-319; stress: static_elasticity2D.c:99-127
__global__ void k_strain_stress(const double2* __restrict__ nod, const uint32_t* __restrict__ conn,
                                const double* __restrict__ u, ElemParams P, uint32_t ne,
                                double* __restrict__ strain, double* __restrict__ stress) {
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= ne) return;
    const uint32_t n3[3] = {conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]};
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    double detJ, dx[3], dy[3];
    if (t3_geom(nod, n3, detJ, dx, dy)) {
        for (int i = 0; i < 3; ++i) {
            const double ux = u[2 * n3[i]], uy = u[2 * n3[i] + 1];
            s0 += dx[i] * ux;
            s1 += dy[i] * uy;
            s2 += (dy[i] * ux + dx[i] * uy);
        }
    }
    strain[3 * e] = s0;
    strain[3 * e + 1] = s1;
    strain[3 * e + 2] = s2;
    if (stress) {
        double D[4];
        elem_D(e, P, D);
        stress[3 * e] = s0 * D[0] + s1 * D[1];
        stress[3 * e + 1] = s0 * D[1] + s1 * D[2];
        stress[3 * e + 2] = s2 * D[3];
    }
}

// ce_e = w_e' K0_e u_e with K0_e the unit-modulus element matrix (w = u for
// compliance); dc_e = -p rho^(p-1) (E0 - Emin) ce_e; C = sum_e E_e ce_e.
__global__ __launch_bounds__(BLOCK) void k_compliance(const double2* __restrict__ nod,
                                                      const uint32_t* __restrict__ conn,
                                                      const double* __restrict__ u,
                                                      const double* __restrict__ w, double nu,
                                                      double thickness, const double* __restrict__ rho,
                                                      double penal, double E0, double Emin,
                                                      const double* __restrict__ elemE, uint32_t ne,
                                                      double* __restrict__ ce, double* __restrict__ dc,
                                                      double* partials) {
    double acc[1] = {0.0};
    double D[4];
    plane_stress(1.0, nu, D);
    for (uint32_t e = blockIdx.x * BLOCK + threadIdx.x; e < ne; e += gridDim.x * BLOCK) {
        const uint32_t n3[3] = {conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]};
        double detJ, dx[3], dy[3];
        double c = 0.0;
        if (t3_geom(nod, n3, detJ, dx, dy)) {
            double ue[6], we[6];
            for (int i = 0; i < 3; ++i) {
                ue[2 * i] = u[2 * n3[i]];
                ue[2 * i + 1] = u[2 * n3[i] + 1];
                we[2 * i] = w[2 * n3[i]];
                we[2 * i + 1] = w[2 * n3[i] + 1];
            }
            for (int r = 0; r < 6; ++r) {
                double t = 0.0;
                for (int q = 0; q < 3; ++q) {
                    double kb[4];
                    ke_block(D, thickness, detJ, dx, dy, r / 2, q, kb);
                    t += kb[2 * (r & 1)] * ue[2 * q];
                    t += kb[2 * (r & 1) + 1] * ue[2 * q + 1];
                }
                c += we[r] * t;
            }
        }
        if (ce) ce[e] = c;
        double Ee;
        if (rho) {
            const double r = rho[e];
            Ee = Emin + pow(r, penal) * (E0 - Emin);
            if (dc) dc[e] = -penal * pow(r, penal - 1.0) * (E0 - Emin) * c;
        } else {
            Ee = elemE ? elemE[e] : E0;
        }
        acc[0] += Ee * c;
    }
    __shared__ double red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    put_partials<1>(acc, partials);
}

__global__ __launch_bounds__(FIN) void k_fin_sum1(const double* partials, unsigned np, double* out) {
    double tot[1];
    fin_sum<1>(partials, np, tot);
    if (threadIdx.x == 0) *out = tot[0];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int read_err(unsigned* d_err, hipStream_t s, unsigned* h) {
    NBX_HIP(hipMemcpyAsync(h, d_err, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    return NBX_OK;
}

int fem_create(FemSys& F, uint32_t nn, const double* nod, uint32_t ne, const uint32_t* conn) {
    F.nn = nn;
    F.ne = ne;
    NBX_HIP(hipStreamCreate(&F.s));
    const hipStream_t s = F.s;
    NBX_HIP(hipMalloc(&F.err, sizeof(unsigned)));
    NBX_HIP(hipMemsetAsync(F.err, 0, sizeof(unsigned), s));
    NBX_HIP(hipMalloc(&F.nod, (size_t)(nn ? nn : 1) * sizeof(double2)));
    NBX_HIP(hipMalloc(&F.conn, (3 * (size_t)ne + 1) * sizeof(uint32_t)));
    if (nn) NBX_HIP(hipMemcpyAsync(F.nod, nod, (size_t)nn * 16, hipMemcpyHostToDevice, s));
    if (ne) NBX_HIP(hipMemcpyAsync(F.conn, conn, 3 * (size_t)ne * 4, hipMemcpyHostToDevice, s));
    // node -> element incidence
    uint32_t* cnt = nullptr;
    NBX_HIP(hipMalloc(&cnt, (size_t)(nn + 1) * sizeof(uint32_t)));
    NBX_HIP(hipMalloc(&F.n2e_ptr, (size_t)(nn + 1) * sizeof(uint32_t)));
    NBX_HIP(hipMalloc(&F.n2e, (3 * (size_t)ne + 1) * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nn + 1)), dim3(BLOCK), 0, s, cnt, (size_t)nn + 1, 0u);
    if (ne)
        hipLaunchKernelGGL(k_count_inc, dim3(grid_for(3 * (size_t)ne)), dim3(BLOCK), 0, s, F.conn, ne,
                           nn, cnt, F.err);
    unsigned herr = 0;
    NBX_TRY(read_err(F.err, s, &herr));
    if (herr) {
        (void)hipFree(cnt);
        set_last_error("connectivity references a node >= N_nod");
        return NBX_ERR_ARG;
    }
    NBX_TRY(scan_u32(cnt, F.n2e_ptr, nn, s));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nn + 1)), dim3(BLOCK), 0, s, cnt, (size_t)nn + 1, 0u);
    if (ne)
        hipLaunchKernelGGL(k_fill_inc, dim3(grid_for(3 * (size_t)ne)), dim3(BLOCK), 0, s, F.conn, ne,
                           F.n2e_ptr, cnt, F.n2e);
    if (nn) hipLaunchKernelGGL(k_sort_inc, dim3(grid_for(nn)), dim3(BLOCK), 0, s, F.n2e_ptr, F.n2e, nn);
    NBX_HIP(hipGetLastError());
    (void)hipFree(cnt);
    // 2x2-block sliced-ELL pattern
    DevMatrix& K = F.K;
    Sell2& S = K.s;
    K.blocked = true;
    K.n = 2 * nn;
    S.n_nodes = nn;
    S.n_slices = (nn + SLICE - 1) / SLICE;
    NBX_HIP(hipMalloc(&S.rowlen, (size_t)nn + 1));
    if (nn)
        hipLaunchKernelGGL(k_row_len, dim3(grid_for(nn)), dim3(BLOCK), 0, s, F.n2e_ptr, F.n2e, F.conn,
                           nn, S.rowlen, F.err);
    NBX_TRY(read_err(F.err, s, &herr));
    if (herr) {
        set_last_error("a node has more than %d neighbours (the device pattern's capacity)",
                       MAX_ADJ - 1);
        return NBX_ERR_PATTERN;
    }
    // row lengths on the host: 64-bit total of the stored blocks, which the
    // 32-bit slice offsets must hold
    std::vector<uint8_t> hl(nn);
    if (nn) NBX_HIP(hipMemcpyAsync(hl.data(), S.rowlen, nn, hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    {
        uint64_t stored = 0;
        for (uint32_t sl = 0; sl < S.n_slices; ++sl) {
            uint32_t wdt = 0;
            for (uint32_t v = sl * SLICE; v < nn && v < (sl + 1) * SLICE; ++v)
                wdt = hl[v] > wdt ? hl[v] : wdt;
            stored += (uint64_t)wdt * SLICE;
        }
        if (stored > 0xffffffffull) {
            set_last_error("pattern of %llu stored 2x2 blocks exceeds 32-bit slice offsets",
                           (unsigned long long)stored);
            return NBX_ERR_PATTERN;
        }
    }
    uint32_t* sblk = nullptr;
    NBX_HIP(hipMalloc(&sblk, (size_t)(S.n_slices + 1) * sizeof(uint32_t)));
    NBX_HIP(hipMalloc(&S.slice_ptr, (size_t)(S.n_slices + 1) * sizeof(uint32_t)));
    if (S.n_slices)
        hipLaunchKernelGGL(k_slice_blocks, dim3(grid_for(S.n_slices)), dim3(BLOCK), 0, s, S.rowlen, nn,
                           S.n_slices, sblk);
    NBX_TRY(scan_u32(sblk, S.slice_ptr, S.n_slices, s));
    (void)hipFree(sblk);
    uint32_t nb = 0;
    NBX_HIP(hipMemcpyAsync(&nb, S.slice_ptr + S.n_slices, 4, hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    S.n_blocks = nb;
    NBX_HIP(hipMalloc(&S.bcol, ((size_t)nb + 1) * sizeof(uint32_t)));
    NBX_HIP(hipMalloc(&S.val, (4 * (size_t)nb + 4) * sizeof(double)));
    if (nn)
        hipLaunchKernelGGL(k_fill_cols, dim3(grid_for(nn)), dim3(BLOCK), 0, s, F.n2e_ptr, F.n2e, F.conn,
                           nn, S.slice_ptr, S.bcol);
    NBX_HIP(hipGetLastError());
    // pattern statistics
    size_t valid = 0;
    uint32_t mx = 0;
    for (uint8_t l : hl) {
        valid += l;
        mx = l > mx ? l : mx;
    }
    S.n_valid = valid;
    S.max_rowlen = mx;
    F.nnz = 4 * valid;
    NBX_HIP(hipMalloc(&K.diag, (size_t)(2 * nn + 1) * sizeof(double)));
    NBX_HIP(hipMalloc(&F.F, (size_t)(2 * nn + 1) * sizeof(double)));
    NBX_HIP(hipMalloc(&F.u, (size_t)(2 * nn + 1) * sizeof(double)));
    NBX_HIP(hipMalloc(&F.bc_rank, (size_t)(2 * nn + 1) * sizeof(uint32_t)));
    NBX_HIP(hipMalloc(&F.bc_vf, (size_t)(2 * nn + 1) * sizeof(double)));
    NBX_HIP(hipMalloc(&F.bc_vl, (size_t)(2 * nn + 1) * sizeof(double)));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(2 * (size_t)nn)), dim3(BLOCK), 0, s, F.bc_rank,
                       2 * (size_t)nn, FREE);
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(2 * (size_t)nn)), dim3(BLOCK), 0, s, F.u,
                       2 * (size_t)nn, 0.0);
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(2 * (size_t)nn)), dim3(BLOCK), 0, s, F.F,
                       2 * (size_t)nn, 0.0);
    NBX_HIP(hipGetLastError());
    NBX_HIP(hipStreamSynchronize(s));
    return ws_alloc(F.ws, 2 * nn);
}

void fem_free(FemSys& F) {
    mat_free(F.K);
    ws_free(F.ws);
    void* ps[] = {F.nod, F.conn, F.n2e_ptr, F.n2e, F.F, F.F_load, F.u, F.aux, F.aux2, F.bc_rank,
                  F.bc_vf, F.bc_vl, F.elemE, F.enabled, F.err, F.rho, F.dc, F.ce};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (F.s) (void)hipStreamDestroy(F.s);
    F = FemSys();
}

int fem_assemble(FemSys& F, double E, double nu, double thickness, const double* elemE,
                 const uint8_t* enabled, int self_weight, double density, const double* gravity) {
    const hipStream_t s = F.s;
    if (elemE) {
        if (!F.elemE) NBX_HIP(hipMalloc(&F.elemE, (size_t)(F.ne + 1) * sizeof(double)));
        NBX_HIP(hipMemcpyAsync(F.elemE, elemE, (size_t)F.ne * 8, hipMemcpyHostToDevice, s));
    }
    if (enabled) {
        if (!F.enabled) NBX_HIP(hipMalloc(&F.enabled, (size_t)F.ne + 1));
        NBX_HIP(hipMemcpyAsync(F.enabled, enabled, F.ne, hipMemcpyHostToDevice, s));
    }
    F.P.E = E;
    F.P.nu = nu;
    F.P.thickness = thickness;
    F.P.elemE = elemE ? F.elemE : nullptr;
    F.P.enabled = enabled ? F.enabled : nullptr;
    return fem_assemble_device(F);
}

int fem_assemble_device(FemSys& F) {
    const hipStream_t s = F.s;
    NBX_HIP(hipMemsetAsync(F.err, 0, sizeof(unsigned), s));
    const Sell2& S = F.K.s;
    if (S.n_blocks)
        hipLaunchKernelGGL(k_assemble, dim3(grid_for(S.n_blocks)), dim3(BLOCK), 0, s, S, F.nod, F.conn,
                           F.n2e_ptr, F.n2e, F.P, F.err);
    NBX_HIP(hipGetLastError());
    unsigned herr = 0;
    NBX_TRY(read_err(F.err, s, &herr));
    mat_scaled_stale(F.K);
    F.bc_applied = false;
    if (herr & 4u) {
        set_last_error("distorted element (detJ < 0)");
        return NBX_ERR_DISTORTED;
    }
    return mat_extract_diag(F.K, s);
}

int fem_body_force(FemSys& F, int self_weight, double density, const double* gravity) {
    const hipStream_t s = F.s;
    const uint32_t nn = F.nn;
    if (self_weight && nn) {
        hipLaunchKernelGGL(k_body_force, dim3(grid_for(nn)), dim3(BLOCK), 0, s, F.nod, F.conn,
                           F.n2e_ptr, F.n2e, F.P, gravity[0], gravity[1], density, F.F, nn);
    } else {
        hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(2 * (size_t)nn)), dim3(BLOCK), 0, s, F.F,
                           2 * (size_t)nn, 0.0);
    }
    NBX_HIP(hipGetLastError());
    return NBX_OK;
}

template <class T>
static int upload(T** d, const std::vector<T>& h, hipStream_t s) {
    NBX_HIP(hipMalloc(d, (h.size() + 1) * sizeof(T)));
    if (!h.empty()) NBX_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return NBX_OK;
}

int fem_apply_bc(FemSys& F, const BcPlan& plan) {
    const hipStream_t s = F.s;
    // Neumann additions per dof, in order
    if (!plan.n_dof.empty()) {
        uint32_t *dof = nullptr, *aptr = nullptr;
        double* add = nullptr;
        NBX_TRY(upload(&dof, plan.n_dof, s));
        NBX_TRY(upload(&aptr, plan.n_ptr, s));
        NBX_TRY(upload(&add, plan.n_add, s));
        const uint32_t m = (uint32_t)plan.n_dof.size();
        hipLaunchKernelGGL(k_neumann, dim3(grid_for(m)), dim3(BLOCK), 0, s, F.F, dof, aptr, add, m);
        NBX_HIP(hipGetLastError());
        NBX_HIP(hipStreamSynchronize(s));
        (void)hipFree(dof);
        (void)hipFree(aptr);
        (void)hipFree(add);
    }
    // the load before elimination, kept so that a reassembled K (new moduli)
    // can take the same boundary conditions without the host
    if (!F.F_load) NBX_HIP(hipMalloc(&F.F_load, 2 * (size_t)F.nn * 8 + 8));
    NBX_HIP(hipMemcpyAsync(F.F_load, F.F, 2 * (size_t)F.nn * 8, hipMemcpyDeviceToDevice, s));
    // Dirichlet: per-dof rank of first elimination, first and last values
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(2 * (size_t)F.nn)), dim3(BLOCK), 0, s, F.bc_rank,
                       2 * (size_t)F.nn, FREE);
    if (!plan.d_dof.empty()) {
        uint32_t *dof = nullptr, *rank = nullptr;
        double *vf = nullptr, *vl = nullptr;
        NBX_TRY(upload(&dof, plan.d_dof, s));
        NBX_TRY(upload(&rank, plan.d_rank, s));
        NBX_TRY(upload(&vf, plan.d_vfirst, s));
        NBX_TRY(upload(&vl, plan.d_vlast, s));
        const uint32_t m = (uint32_t)plan.d_dof.size();
        hipLaunchKernelGGL(k_bc_scatter, dim3(grid_for(m)), dim3(BLOCK), 0, s, dof, rank, vf, vl, m,
                           F.bc_rank, F.bc_vf, F.bc_vl);
        if (F.nn) {
            hipLaunchKernelGGL(k_bc_rhs, dim3(grid_for(F.nn)), dim3(BLOCK), 0, s, F.K.s, F.bc_rank,
                               F.bc_vf, F.bc_vl, F.F);
            hipLaunchKernelGGL(k_bc_matrix, dim3(grid_for(F.nn)), dim3(BLOCK), 0, s, F.K.s, F.bc_rank);
        }
        NBX_HIP(hipGetLastError());
        NBX_HIP(hipStreamSynchronize(s));
        (void)hipFree(dof);
        (void)hipFree(rank);
        (void)hipFree(vf);
        (void)hipFree(vl);
    }
    F.bc_applied = true;
    F.n_constrained = (uint32_t)plan.d_dof.size();
    F.amg_symbolic_current = false;   // the eliminated unknowns may differ
    F.K.amg_current = false;
    return mat_extract_diag(F.K, s);
}

// The boundary conditions of the last fem_apply_bc on a freshly assembled K:
// the same kernels, from the saved pre-elimination load.
int fem_reapply_bc(FemSys& F) {
    const hipStream_t s = F.s;
    if (!F.F_load) {
        set_last_error("no boundary conditions to re-apply");
        return NBX_ERR_ARG;
    }
    NBX_HIP(hipMemcpyAsync(F.F, F.F_load, 2 * (size_t)F.nn * 8, hipMemcpyDeviceToDevice, s));
    if (F.n_constrained && F.nn) {
        hipLaunchKernelGGL(k_bc_rhs, dim3(grid_for(F.nn)), dim3(BLOCK), 0, s, F.K.s, F.bc_rank,
                           F.bc_vf, F.bc_vl, F.F);
        hipLaunchKernelGGL(k_bc_matrix, dim3(grid_for(F.nn)), dim3(BLOCK), 0, s, F.K.s, F.bc_rank);
        NBX_HIP(hipGetLastError());
    }
    F.bc_applied = true;
    F.K.amg_current = false;
    return mat_extract_diag(F.K, s);
}

int fem_zero_constrained(FemSys& F, double* x) {
    if (F.nn)
        hipLaunchKernelGGL(k_zero_constrained, dim3(grid_for(2 * (size_t)F.nn)), dim3(BLOCK), 0, F.s, x,
                           F.bc_rank, 2 * F.nn);
    NBX_HIP(hipGetLastError());
    return NBX_OK;
}

int fem_strain_stress(FemSys& F, double* strain, double* stress) {
    if (F.ne)
        hipLaunchKernelGGL(k_strain_stress, dim3(grid_for(F.ne)), dim3(BLOCK), 0, F.s, F.nod, F.conn, F.u,
                           F.P, F.ne, strain, stress);
    NBX_HIP(hipGetLastError());
    return NBX_OK;
}

int fem_compliance(FemSys& F, const double* rho, double penal, double E0, double Emin,
                   double* ce, double* dc, double* C_host) {
    double* out = F.ws.v64[5] + F.ws.n;
    unsigned g = grid_for(F.ne);
    if (g > (unsigned)MAX_GRID) g = MAX_GRID;
    hipLaunchKernelGGL(k_compliance, dim3(g), dim3(BLOCK), 0, F.s, F.nod, F.conn, F.u, F.u, F.P.nu,
                       F.P.thickness, rho, penal, E0, Emin, F.P.elemE, F.ne, ce, dc, F.ws.partials);
    hipLaunchKernelGGL(k_fin_sum1, dim3(1), dim3(FIN), 0, F.s, F.ws.partials, g, out);
    NBX_HIP(hipGetLastError());
    NBX_HIP(hipMemcpyAsync(C_host, out, sizeof(double), hipMemcpyDeviceToHost, F.s));
    NBX_HIP(hipStreamSynchronize(F.s));
    return NBX_OK;
}

// out_e = a_e' K0_e b_e (device vectors of 2 nn values)
int fem_elem_cross(FemSys& F, const double* a, const double* b, double* out) {
    unsigned g = grid_for(F.ne);
    if (g > (unsigned)MAX_GRID) g = MAX_GRID;
    hipLaunchKernelGGL(k_compliance, dim3(g), dim3(BLOCK), 0, F.s, F.nod, F.conn, b, a, F.P.nu,
                       F.P.thickness, nullptr, 3.0, 1.0, 0.0, nullptr, F.ne, out, nullptr,
                       F.ws.partials);
    NBX_HIP(hipGetLastError());
    return NBX_OK;
}

// ---- quasi-static damage ------------------------------------------------------
// Strain of each T3 from u (DMG_pipeline_compute_strain,
// static_damage2D.c:571-631) and tension_damage (:134-185): effective stress
// D eps, positive principal stresses by Mohr's circle, tau = sqrt(s+' D^-1
// s+), r = max(r_prev, tau), d = 1 - (r0/r) exp(A (1 - r/r0)) with r0 =
// ft/sqrt(E), A = 1/((Gf/l)(E/ft^2) - 1/2), l = sqrt(2 area) (get_clfd, :633-647).
__global__ void k_damage_update(const double2* __restrict__ nod, const uint32_t* __restrict__ conn,
                                const double* __restrict__ u, uint32_t ne, double E, double nu,
                                double Gf, double ft, const double* __restrict__ r_prev,
                                double* __restrict__ r, double* __restrict__ dmg,
                                double* __restrict__ strain, unsigned* err) {
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= ne) return;
    const uint32_t n3[3] = {conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]};
    double detJ, dx[3], dy[3];
    if (!t3_geom(nod, n3, detJ, dx, dy)) {
        atomicOr(err, 4u);
        return;
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double ux = u[2 * n3[i]], uy = u[2 * n3[i] + 1];
        s0 += dx[i] * ux;
        s1 += dy[i] * uy;
        s2 += (dy[i] * ux + dx[i] * uy);
    }
    strain[3 * e] = s0;
    strain[3 * e + 1] = s1;
    strain[3 * e + 2] = s2;
    const double2 a = nod[n3[0]], b = nod[n3[1]], c = nod[n3[2]];
    const double clfd = sqrt((b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x));
    double D[4];
    plane_stress(E, nu, D);
    const double r0 = ft / sqrt(E);
    const double e0 = D[0] * s0 + D[1] * s1;
    const double e1 = D[1] * s0 + D[2] * s1;
    const double e2 = D[3] * s2;
    const double avg = 0.5 * (e0 + e1);
    const double R = sqrt(0.25 * ((e0 - e1) * (e0 - e1)) + e2 * e2);
    const double p1 = 0.0 > avg + R ? 0.0 : avg + R;
    const double p2 = 0.0 > avg - R ? 0.0 : avg - R;
    const double det = D[0] * D[2] - D[1] * D[1];
    const double i11 = D[2] / det, i12 = -D[1] / det, i22 = D[0] / det;
    const double sTs = i11 * (p1 * p1) + 2 * i12 * p1 * p2 + i22 * (p2 * p2);
    const double tau = sqrt(sTs);
    const double rp = r_prev[e];
    const double re = rp > tau ? rp : tau;
    r[e] = re;
    const double div = (Gf / clfd) * (E / (ft * ft));
    const double A = 1.0 / (div - 0.5);
    dmg[e] = 1.0 - (r0 / re) * exp(A * (1.0 - (re / r0)));
}

// res = F - K u per row (the row product as nb_sparse_multiply_vector forms
// it), partial sums of res^2
__global__ __launch_bounds__(BLOCK) void k_residual64(Sell2 K, const double2* __restrict__ F,
                                                      const double2* __restrict__ u,
                                                      double2* __restrict__ res, double* partials,
                                                      uint32_t n) {
    double acc[1] = {0.0};
    const Chunk c = my_chunk(n);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const double2 ku = row64(K, (uint32_t)i, u);
        const double2 f = F[i];
        const double2 rv = make_double2(f.x - ku.x, f.y - ku.y);
        res[i] = rv;
        acc[0] += rv.x * rv.x + rv.y * rv.y;
    }
    __shared__ double red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    put_partials<1>(acc, partials);
}

// r = F - K u per row with compensated arithmetic (Dot2 of Ogita, Rump and
// Oishi 2005: each product's rounding error from an FMA, each sum's from
// TwoSum, the errors summed alongside), so r is accurate to about one
// rounding of r itself instead of eps * sum |K_ij u_j|; partials of r.r.
__device__ __forceinline__ void two_sum64(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void dot2_add(double a, double x, double& s, double& c) {
    const double p = a * x;
    const double pe = fma(a, x, -p);
    double t, e;
    two_sum64(s, p, t, e);
    s = t;
    c = c + (e + pe);
}
__global__ __launch_bounds__(BLOCK) void k_residual_dot2(Sell2 K, const double2* __restrict__ F,
                                                         const double2* __restrict__ u,
                                                         double* partials, uint32_t n) {
    double acc[1] = {0.0};
    const Chunk c = my_chunk(n);
    const dv2* __restrict__ pa = reinterpret_cast<const dv2*>(K.val);
    const dv2* __restrict__ pb = reinterpret_cast<const dv2*>(K.val + 2 * K.n_blocks);
    for (size_t i = c.beg + threadIdx.x; i < c.end; i += BLOCK) {
        const double2 f = F[i];
        double s0 = -f.x, c0 = 0.0, s1 = -f.y, c1 = 0.0;
        size_t idx = (size_t)K.slice_ptr[i >> 6] + (i & 63);
        const int len = K.rowlen[i];
        for (int k = 0; k < len; ++k, idx += SLICE) {
            const double2 x = u[K.bcol[idx]];
            const dv2 a = pa[idx], b = pb[idx];
            dot2_add(a.x, x.x, s0, c0);
            dot2_add(a.y, x.y, s0, c0);
            dot2_add(b.x, x.x, s1, c1);
            dot2_add(b.y, x.y, s1, c1);
        }
        const double r0 = s0 + c0, r1 = s1 + c1;   // K u - F
        acc[0] += r0 * r0 + r1 * r1;
    }
    __shared__ double red[1][BLOCK / WAVE];
    block_sum<1>(acc, red);
    put_partials<1>(acc, partials);
}

__global__ void k_add_f64(double* __restrict__ x, const double* __restrict__ dx, size_t n) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) x[i] += dx[i];
}

// stress from given strain (static_elasticity2D.c:99-127)
__global__ void k_stress_only(const double* __restrict__ strain, ElemParams P, uint32_t ne,
                              double* __restrict__ stress) {
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= ne) return;
    double D[4];
    elem_D(e, P, D);
    const double s0 = strain[3 * e], s1 = strain[3 * e + 1], s2 = strain[3 * e + 2];
    stress[3 * e] = s0 * D[0] + s1 * D[1];
    stress[3 * e + 1] = s0 * D[1] + s1 * D[2];
    stress[3 * e + 2] = s2 * D[3];
}

int stress_from_strain(uint32_t ne, double E, double nu, const uint8_t* enabled_host,
                       const double* strain_host, double* stress_host) {
    if (ne == 0) return NBX_OK;
    double *ds = nullptr, *dt = nullptr;
    uint8_t* den = nullptr;
    const size_t bytes = 3 * (size_t)ne * sizeof(double);
    NBX_HIP(hipMalloc(&ds, bytes));
    NBX_HIP(hipMalloc(&dt, bytes));
    NBX_HIP(hipMemcpy(ds, strain_host, bytes, hipMemcpyHostToDevice));
    ElemParams P;
    P.E = E;
    P.nu = nu;
    if (enabled_host) {
        NBX_HIP(hipMalloc(&den, ne));
        NBX_HIP(hipMemcpy(den, enabled_host, ne, hipMemcpyHostToDevice));
        P.enabled = den;
    }
    hipLaunchKernelGGL(k_stress_only, dim3(grid_for(ne)), dim3(BLOCK), 0, 0, ds, P, ne, dt);
    NBX_HIP(hipGetLastError());
    NBX_HIP(hipMemcpy(stress_host, dt, bytes, hipMemcpyDeviceToHost));
    (void)hipFree(ds);
    (void)hipFree(dt);
    if (den) (void)hipFree(den);
    return NBX_OK;
}

// K in the reference's scalar CSR layout: row 2v+i lists, for each block
// (v,w) in ascending w, the columns 2w and 2w+1.
int fem_download_csr(const FemSys& F, uint32_t* rowptr, uint32_t* col, double* val) {
    const Sell2& S = F.K.s;
    const uint32_t nn = S.n_nodes;
    std::vector<uint32_t> sp(S.n_slices + 1), bc(S.n_blocks + 1);
    std::vector<uint8_t> len(nn + 1);
    std::vector<double> hv(4 * S.n_blocks + 4);
    NBX_HIP(copy_on(sp.data(), S.slice_ptr, (S.n_slices + 1) * 4, hipMemcpyDeviceToHost, F.s));
    if (nn) NBX_HIP(copy_on(len.data(), S.rowlen, nn, hipMemcpyDeviceToHost, F.s));
    NBX_HIP(copy_on(bc.data(), S.bcol, (S.n_blocks + 1) * 4, hipMemcpyDeviceToHost, F.s));
    NBX_HIP(copy_on(hv.data(), S.val, (4 * S.n_blocks + 4) * 8, hipMemcpyDeviceToHost, F.s));
    rowptr[0] = 0;
    for (uint32_t v = 0; v < nn; ++v) {
        rowptr[2 * v + 1] = rowptr[2 * v] + 2 * len[v];
        rowptr[2 * v + 2] = rowptr[2 * v + 1] + 2 * len[v];
    }
    for (uint32_t v = 0; v < nn; ++v) {
        size_t idx = (size_t)sp[v >> 6] + (v & 63);
        for (uint32_t k = 0; k < len[v]; ++k, idx += SLICE) {
            for (int i = 0; i < 2; ++i) {
                const uint32_t o = rowptr[2 * v + i] + 2 * k;
                col[o] = 2 * bc[idx];
                col[o + 1] = 2 * bc[idx] + 1;
                val[o] = hv[ent(S.n_blocks, idx, i, 0)];
                val[o + 1] = hv[ent(S.n_blocks, idx, i, 1)];
            }
        }
    }
    return NBX_OK;
}

int fem_residual(FemSys& F, double* res, double* norm) {
    const uint32_t nu = F.K.s.n_nodes;
    const unsigned G = red_grid(nu);
    double* out = F.ws.v64[5] + F.ws.n;
    hipLaunchKernelGGL(k_residual64, dim3(G), dim3(BLOCK), 0, F.s, F.K.s, (const double2*)F.F,
                       (const double2*)F.u, (double2*)res, F.ws.partials, nu);
    hipLaunchKernelGGL(k_fin_sum1, dim3(1), dim3(FIN), 0, F.s, F.ws.partials, G, out);
    NBX_HIP(hipGetLastError());
    double h = 0.0;
    NBX_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, F.s));
    NBX_HIP(hipStreamSynchronize(F.s));
    *norm = sqrt(h);
    return NBX_OK;
}

int fem_residual_dot2(FemSys& F, double* norm) {
    const uint32_t nu = F.K.s.n_nodes;
    const unsigned G = red_grid(nu);
    double* out = F.ws.v64[5] + F.ws.n;
    hipLaunchKernelGGL(k_residual_dot2, dim3(G), dim3(BLOCK), 0, F.s, F.K.s, (const double2*)F.F,
                       (const double2*)F.u, F.ws.partials, nu);
    hipLaunchKernelGGL(k_fin_sum1, dim3(1), dim3(FIN), 0, F.s, F.ws.partials, G, out);
    NBX_HIP(hipGetLastError());
    double h = 0.0;
    NBX_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, F.s));
    NBX_HIP(hipStreamSynchronize(F.s));
    *norm = sqrt(h);
    return NBX_OK;
}

int fem_add(FemSys& F, double* x, const double* dx) {
    const size_t n = 2 * (size_t)F.nn;
    hipLaunchKernelGGL(k_add_f64, dim3(grid_for(n)), dim3(BLOCK), 0, F.s, x, dx, n);
    NBX_HIP(hipGetLastError());
    return NBX_OK;
}

int fem_damage_update(FemSys& F, double E, double nu, double Gf, double ft, const double* r_prev,
                      double* r, double* dmg, double* strain) {
    NBX_HIP(hipMemsetAsync(F.err, 0, sizeof(unsigned), F.s));
    hipLaunchKernelGGL(k_damage_update, dim3(grid_for(F.ne)), dim3(BLOCK), 0, F.s, F.nod, F.conn, F.u,
                       F.ne, E, nu, Gf, ft, r_prev, r, dmg, strain, F.err);
    NBX_HIP(hipGetLastError());
    unsigned herr = 0;
    NBX_TRY(read_err(F.err, F.s, &herr));
    if (herr & 4u) {
        set_last_error("distorted element (detJ < 0)");
        return NBX_ERR_DISTORTED;
    }
    return NBX_OK;
}

}  // namespace nbx

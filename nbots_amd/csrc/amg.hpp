// Smoothed-aggregation algebraic multigrid for the 2-dof T3 systems: the
// preconditioner of the NBX_MIXED_AMG_FCG solve (an FP64 flexible CG whose
// preconditioner is one FP32 V-cycle).
//
// Method (Vanek, Mandel & Brezina 1996, smoothed aggregation; the near
// null space of plane elasticity is the three rigid-body modes):
//   * nodes are aggregated greedily on the matrix graph (a root with its
//     whole free neighbourhood, then the rest joins a neighbouring
//     aggregate); a node whose every unknown is eliminated by a Dirichlet
//     condition has no coarse part;
//   * the tentative prolongator T maps an aggregate's three coarse unknowns
//     to the orthonormalised rigid-body modes on its nodes (QR per aggregate;
//     the R factors are the next level's modes);
//   * P = (I - omega D^-1 A) T with omega = 4 / (3 lambda_max(D^-1 A));
//   * A_c = P^T A P, recursively, until at most AMG_COARSEST unknowns remain,
//     whose matrix is inverted (dense, FP64) and applied as a dense product;
//   * one V-cycle: Chebyshev smoothing in D^-1 A on [lambda/30, 1.1 lambda]
//     before and after the coarse correction, of degree AMG_FINE_DEGREE on
//     the finest level (degree 1: one Jacobi step weighted 1/theta, theta
//     the interval's centre), AMG_LEVEL1_DEGREE on the next, then
//     AMG_LEVEL2_DEGREE, and AMG_DEGREE below.
// The host builds every pattern as scalar CSR, grouped by node (a node's
// rows share one pattern of complete column groups).  The aggregation, T
// and every pattern depend on the matrix's graph, its Dirichlet rows and the
// node coordinates only, so they are built once on the host
// (amg_setup.cpp); P, A_c and the coarsest inverse follow the values on the
// device whenever K changes (amg.hip, which also keeps the V-cycle's block
// copies).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <memory>
#include <new>
#include <type_traits>
#include <utility>
#include <vector>

namespace nbx {

// An allocator whose value-initialisation leaves trivial values
// uninitialised: the host setup's large arrays are written once by its
// threads, so they are neither zeroed first nor first touched (page-faulted)
// by one thread.
template <class T, class A = std::allocator<T>>
struct NoInitAlloc : A {
    using A::A;
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U, typename std::allocator_traits<A>::template rebind_alloc<U>>;
    };
    template <class U>
    void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... Args>
    void construct(U* p, Args&&... args) {
        std::allocator_traits<A>::construct(static_cast<A&>(*this), p, std::forward<Args>(args)...);
    }
};
template <class T>
using HostVec = std::vector<T, NoInitAlloc<T>>;

constexpr uint32_t AMG_NONE = 0xffffffffu;
// Chebyshev smoothing steps before and after the coarse correction on the
// levels below level 2, a few thousand unknowns and fewer, where a step
// costs the latency of its launch and dependent gathers (3.5-5 us) rather
// than its bytes: degree 5 there and 4 on level 2 take 11-20 % fewer outer
// iterations than 3 (4.2M-DOF cantilever: 57 against 64; 1.05M-DOF SIMP
// beam: 74 against 92 per solve) for 2-3 steps more per level and cycle
// (C3 solve 42.4 -> 38.7 ms, SIMP solve 23.2 -> 21.3 ms per iteration)
constexpr int AMG_DEGREE = 5;
constexpr int AMG_LEVEL2_DEGREE = 4;
// ... on the finest level, whose products are the cycle's largest: with
// degree 1 the solves take as many outer iterations as with 3 (4.2M-DOF
// cantilever: 64 against 66; 1.05M-DOF SIMP beam: 94-95 against 90 per
// solve) for a third of the finest products (C3 solve 0.055 -> 0.046 s)
constexpr int AMG_FINE_DEGREE = 1;
// ... and on level 1, the largest coarse level (a third of the finest one's
// nonzeros with six products per cycle at degree 3): degree 2 takes as many
// outer iterations (4.2M DOF: 64; 1.05M-DOF SIMP beam: 92 against 95) for
// two products fewer (C3 solve 0.045 -> 0.042 s, SIMP 39.4 -> 36.8 ms)
constexpr int AMG_LEVEL1_DEGREE = 2;
constexpr int AMG_MAX_DEGREE = 8;      // the coefficient arrays' length
// the smoothing degree of level l
constexpr int amg_level_degree(size_t l) {
    return l == 0 ? AMG_FINE_DEGREE : l == 1 ? AMG_LEVEL1_DEGREE
                                    : l == 2 ? AMG_LEVEL2_DEGREE : AMG_DEGREE;
}
constexpr uint32_t AMG_COARSEST = 200; // unknowns of the directly solved level
constexpr int AMG_MAX_LEVELS = 12;
static_assert(AMG_FINE_DEGREE >= 1 && AMG_FINE_DEGREE <= AMG_MAX_DEGREE,
              "AmgFine holds the coefficients of at most AMG_MAX_DEGREE steps");
static_assert(AMG_LEVEL1_DEGREE >= 2 && AMG_LEVEL2_DEGREE >= 2 && AMG_DEGREE >= 2 &&
              AMG_LEVEL1_DEGREE <= AMG_MAX_DEGREE && AMG_LEVEL2_DEGREE <= AMG_MAX_DEGREE &&
              AMG_DEGREE <= AMG_MAX_DEGREE,
              "a coarse level keeps its coefficients in arrays of AMG_MAX_DEGREE and adds its "
              "last direction to x itself (degree >= 2)");

// One level of the host-built symbolic hierarchy.
struct AmgHostLevel {
    uint32_t n = 0;                   // scalar unknowns
    uint32_t bs = 0;                  // unknowns per node (2 on the finest level, 3 below)
    HostVec<uint32_t> a_ptr, a_col;   // this level's operator pattern
    // to the next coarser level (empty on the coarsest)
    uint32_t nc = 0;                  // coarse unknowns (3 per aggregate)
    HostVec<double> t;                // [3 n] T's values of row i: columns agg(i)*3 + m
    HostVec<uint32_t> agg;            // [n / bs] each node's aggregate, or NONE
    HostVec<uint32_t> p_ptr, p_col;   // P pattern (n x nc)
    HostVec<uint32_t> w_ptr, w_col;   // W = A P pattern (n x nc)
    HostVec<uint32_t> pt_ptr, pt_row, pt_pos;   // P^T: coarse column c -> (row, position)
};

struct AmgHost {
    std::vector<AmgHostLevel> lv;     // lv[0] finest; lv.back() is the coarsest
};

// Build the symbolic hierarchy from the finest operator: CSR pattern
// (rowptr[n+1], col ascending per row), which of its entries hold a nonzero
// value (nz[e] = 1: the Dirichlet-eliminated unknowns are the rows whose
// off-diagonal entries are all 0, and the elimination's zeros are no graph
// edges) and the coordinates of the n/2 nodes (x, y interleaved).  Returns
// 0, -1 when the pattern is not a 2-dof node pattern, or -2 when host memory
// ran out (the setup runs on up to 16 threads).
int amg_build_host(uint32_t n, const uint32_t* rowptr, const uint32_t* col, const uint8_t* nz,
                   const double* xy, AmgHost& out);

// The scalar CSR pattern of a 2x2-block sliced-ELL matrix of `nodes` block
// rows (slices of 64 rows: block k of row v sits in slot slice_ptr[v / 64] +
// v % 64 + 64 k, rowlen[v] blocks, column node bcol[slot]): row 2v + i holds
// columns 2 bcol + 0, 2 bcol + 1 of each block in slot order, and nz[e] is bit
// (2 i + j) of nzbits[slot], 1 where entry (i, j) of the block is nonzero.
// Threaded like the setup.
void amg_expand_blocks(uint32_t nodes, const uint32_t* slice_ptr, const uint8_t* rowlen,
                       const uint32_t* bcol, const uint8_t* nzbits, HostVec<uint32_t>& rowptr,
                       HostVec<uint32_t>& col, HostVec<uint8_t>& nz);

// The power iteration's start vector: +-1 from a hash of the index, so that
// every frequency is present (a smooth start underestimates lambda_max, and
// a Chebyshev smoother whose interval ends below the spectrum diverges).
inline double amg_sign(uint32_t i) {
    uint32_t h = i * 2654435761u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    return (h & 1u) ? 1.0 : -1.0;
}

// In-place inverse of a dense symmetric positive semi-definite n x n matrix
// (row-major, FP64) by Cholesky with a pivot guard: an unknown whose pivot
// falls below 1e-13 times the largest diagonal entry is dropped (its row and
// column of the inverse are 0).  Returns the number of dropped unknowns.
int amg_dense_inverse(uint32_t n, double* a);

}  // namespace nbx

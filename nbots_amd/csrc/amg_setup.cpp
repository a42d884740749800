// Host side of the smoothed-aggregation hierarchy: aggregation, tentative
// prolongators and every pattern and index map the device numeric phase
// uses (amg.hpp).  Plain C++: no HIP, so it also builds and runs under the
// host sanitizers.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <exception>
#include <new>
#include <thread>

#include "amg.hpp"

namespace nbx {

namespace {

struct Graph {
    HostVec<uint32_t> ptr, adj;
};

// Host threads for the setup (at most 16: a one-GPU job's share of a host).
unsigned setup_threads(uint32_t items) {
    const unsigned h = std::thread::hardware_concurrency();
    const unsigned t = std::min(h == 0 ? 1u : h, 16u);
    return std::max(1u, std::min(t, items / 4096u));
}

// Runs body(t) for t in [0, nt) on nt threads (the caller's among them);
// an exception in any of them (out of host memory) is rethrown here after
// all have joined.
template <class B>
void run_threads(unsigned nt, B body) {
    if (nt <= 1) {
        body(0u);
        return;
    }
    std::vector<std::exception_ptr> err(nt);
    auto guarded = [&](unsigned t) {
        try {
            body(t);
        } catch (...) {
            err[t] = std::current_exception();
        }
    };
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (unsigned t = 1; t < nt; ++t) th.emplace_back(guarded, t);
    guarded(0);
    for (std::thread& x : th) x.join();
    for (const std::exception_ptr& e : err)
        if (e) std::rethrow_exception(e);
}

// f(lo, hi) over contiguous ranges of [0, n), one per thread.
template <class F>
void par_for(uint32_t n, F f) {
    const unsigned nt = setup_threads(n);
    run_threads(nt, [&](unsigned t) {
        f((uint32_t)((uint64_t)n * t / nt), (uint32_t)((uint64_t)n * (t + 1) / nt));
    });
}

// Per-item lists built in parallel: f(i, out) appends item i's list to
// out.  The result (ptr[n + 1], data) does not depend on the threads.
template <class F>
void build_lists(uint32_t n, F f, HostVec<uint32_t>& ptr, HostVec<uint32_t>& data) {
    const unsigned nt = setup_threads(n);
    std::vector<std::vector<uint32_t>> part(nt);
    std::vector<uint32_t> len(n);
    std::vector<uint32_t> lo(nt + 1);
    for (unsigned t = 0; t <= nt; ++t) lo[t] = (uint32_t)((uint64_t)n * t / nt);
    auto fill = [&](unsigned t) {
        std::vector<uint32_t>& out = part[t];
        for (uint32_t i = lo[t]; i < lo[t + 1]; ++i) {
            const size_t before = out.size();
            f(i, out);
            len[i] = (uint32_t)(out.size() - before);
        }
    };
    run_threads(nt, fill);
    ptr.resize(n + 1);
    ptr[0] = 0;
    for (uint32_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + len[i];
    data.resize(ptr[n]);
    run_threads(nt, [&](unsigned t) {
        std::copy(part[t].begin(), part[t].end(), data.begin() + ptr[lo[t]]);
    });
}


// Node graph of a level: node I's neighbours are the nodes J != I that hold
// a column of one of I's rows.  An unknown marked dead (Dirichlet-eliminated
// on the finest level, a rank-deficient coarse mode below) has no edges, and
// on the finest level an edge needs a nonzero value (the elimination zeroes
// the entries it decouples).
Graph node_graph(const AmgHostLevel& L, const std::vector<uint8_t>& dead, const uint8_t* nz) {
    const uint32_t bs = L.bs, nodes = L.n / bs;
    Graph g;
    build_lists(
        nodes,
        [&](uint32_t I, std::vector<uint32_t>& out) {
            const size_t b = out.size();
            for (uint32_t i = I * bs; i < I * bs + bs; ++i) {
                if (dead[i]) continue;
                for (uint32_t e = L.a_ptr[i]; e < L.a_ptr[i + 1]; ++e) {
                    const uint32_t j = L.a_col[e], J = j / bs;
                    if (J == I || dead[j]) continue;
                    if (nz && !nz[e]) continue;
                    out.push_back(J);
                }
            }
            std::sort(out.begin() + (ptrdiff_t)b, out.end());
            out.erase(std::unique(out.begin() + (ptrdiff_t)b, out.end()), out.end());
        },
        g.ptr, g.adj);
    return g;
}

// Greedy aggregation: (1) a node whose neighbourhood is nonempty and free
// becomes a root and takes it; (2) a remaining node joins the aggregate of
// its first neighbour aggregated in (1); (3) what is left forms aggregates
// with its free neighbours.  Nodes without neighbours stay unaggregated.
uint32_t aggregate(const Graph& g, std::vector<uint32_t>& agg) {
    const uint32_t nodes = (uint32_t)g.ptr.size() - 1;
    agg.assign(nodes, AMG_NONE);
    uint32_t na = 0;
    for (uint32_t I = 0; I < nodes; ++I) {
        const uint32_t b = g.ptr[I], e = g.ptr[I + 1];
        if (agg[I] != AMG_NONE || b == e) continue;
        bool free_ = true;
        for (uint32_t k = b; k < e && free_; ++k) free_ = agg[g.adj[k]] == AMG_NONE;
        if (!free_) continue;
        agg[I] = na;
        for (uint32_t k = b; k < e; ++k) agg[g.adj[k]] = na;
        ++na;
    }
    const std::vector<uint32_t> first = agg;
    for (uint32_t I = 0; I < nodes; ++I) {
        if (agg[I] != AMG_NONE) continue;
        for (uint32_t k = g.ptr[I]; k < g.ptr[I + 1]; ++k)
            if (first[g.adj[k]] != AMG_NONE) {
                agg[I] = first[g.adj[k]];
                break;
            }
    }
    for (uint32_t I = 0; I < nodes; ++I) {
        if (agg[I] != AMG_NONE || g.ptr[I] == g.ptr[I + 1]) continue;
        agg[I] = na;
        for (uint32_t k = g.ptr[I]; k < g.ptr[I + 1]; ++k)
            if (agg[g.adj[k]] == AMG_NONE) agg[g.adj[k]] = na;
        ++na;
    }
    return na;
}

uint32_t find(const uint32_t* b, const uint32_t* e, uint32_t c) {
    const uint32_t* p = std::lower_bound(b, e, c);
    return (p != e && *p == c) ? (uint32_t)(p - b) : AMG_NONE;
}

// One coarsening step: fills L's prolongator fields and returns the coarse
// level's pattern, modes B (3 per coarse unknown) and dead unknowns.
void coarsen(AmgHostLevel& L, const std::vector<double>& B, const std::vector<uint8_t>& dead,
             const uint8_t* nz, AmgHostLevel& C, std::vector<double>& Bc,
             std::vector<uint8_t>& dead_c) {
    const uint32_t bs = L.bs, n = L.n, nodes = n / bs;
    const Graph g = node_graph(L, dead, nz);
    std::vector<uint32_t> agg;
    const uint32_t na = aggregate(g, agg);
    L.nc = 3 * na;
    // members of each aggregate, ascending
    std::vector<uint32_t> mptr(na + 1, 0), mem(nodes);
    for (uint32_t I = 0; I < nodes; ++I)
        if (agg[I] != AMG_NONE) ++mptr[agg[I] + 1];
    for (uint32_t a = 0; a < na; ++a) mptr[a + 1] += mptr[a];
    {
        std::vector<uint32_t> fill(mptr.begin(), mptr.end() - 1);
        for (uint32_t I = 0; I < nodes; ++I)
            if (agg[I] != AMG_NONE) mem[fill[agg[I]]++] = I;
    }
    // tentative prolongator: modified Gram-Schmidt (twice) on each
    // aggregate's rows of B; a column that vanishes is a dead coarse unknown
    L.t.resize(3 * (size_t)n);
    par_for(n, [&](uint32_t i0, uint32_t i1) {   // rows of no aggregate keep 0
        std::fill(L.t.begin() + 3 * (size_t)i0, L.t.begin() + 3 * (size_t)i1, 0.0);
    });
    Bc.assign(3 * (size_t)L.nc, 0.0);
    dead_c.assign(L.nc, 0);
    par_for(na, [&](uint32_t a0, uint32_t a1) {
    std::vector<double> q;
    for (uint32_t a = a0; a < a1; ++a) {
        const uint32_t m = (mptr[a + 1] - mptr[a]) * bs;
        q.assign(3 * (size_t)m, 0.0);   // q[c * m + r]
        for (uint32_t k = 0; k < mptr[a + 1] - mptr[a]; ++k)
            for (uint32_t d = 0; d < bs; ++d) {
                const uint32_t i = mem[mptr[a] + k] * bs + d;
                for (int c = 0; c < 3; ++c) q[c * (size_t)m + k * bs + d] = dead[i] ? 0.0 : B[3 * (size_t)i + c];
            }
        double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int c = 0; c < 3; ++c) {
            double* v = &q[c * (size_t)m];
            double n0 = 0.0;
            for (uint32_t r = 0; r < m; ++r) n0 += v[r] * v[r];
            n0 = sqrt(n0);
            for (int pass = 0; pass < 2; ++pass)
                for (int k = 0; k < c; ++k) {
                    const double* u = &q[k * (size_t)m];
                    double s = 0.0;
                    for (uint32_t r = 0; r < m; ++r) s += u[r] * v[r];
                    for (uint32_t r = 0; r < m; ++r) v[r] -= s * u[r];
                    R[k][c] += s;
                }
            double nv = 0.0;
            for (uint32_t r = 0; r < m; ++r) nv += v[r] * v[r];
            nv = sqrt(nv);
            if (n0 > 0.0 && nv > 1e-10 * n0) {
                for (uint32_t r = 0; r < m; ++r) v[r] /= nv;
                R[c][c] = nv;
            } else {
                for (uint32_t r = 0; r < m; ++r) v[r] = 0.0;
                dead_c[3 * a + c] = 1;
            }
        }
        for (uint32_t k = 0; k < mptr[a + 1] - mptr[a]; ++k)
            for (uint32_t d = 0; d < bs; ++d) {
                const uint32_t i = mem[mptr[a] + k] * bs + d;
                for (int c = 0; c < 3; ++c) L.t[3 * (size_t)i + c] = q[c * (size_t)m + k * bs + d];
            }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) Bc[3 * (size_t)(3 * a + r) + c] = dead_c[3 * a + r] ? 0.0 : R[r][c];
    }
    });
    // P pattern per node (shared by its rows): the aggregates of the nodes
    // its rows reach, three columns each
    HostVec<uint32_t> anode_ptr, anode;   // the aggregates, ascending
    build_lists(
        nodes,
        [&](uint32_t I, std::vector<uint32_t>& out) {
            const size_t b = out.size();
            for (uint32_t i = I * bs; i < I * bs + bs; ++i)
                for (uint32_t e = L.a_ptr[i]; e < L.a_ptr[i + 1]; ++e) {
                    const uint32_t A = agg[L.a_col[e] / bs];
                    if (A != AMG_NONE) out.push_back(A);
                }
            std::sort(out.begin() + (ptrdiff_t)b, out.end());
            out.erase(std::unique(out.begin() + (ptrdiff_t)b, out.end()), out.end());
        },
        anode_ptr, anode);
    L.p_ptr.resize(n + 1);
    L.p_ptr[0] = 0;
    for (uint32_t i = 0; i < n; ++i)
        L.p_ptr[i + 1] = L.p_ptr[i] + 3 * (anode_ptr[i / bs + 1] - anode_ptr[i / bs]);
    L.p_col.resize(L.p_ptr[n]);
    par_for(nodes, [&](uint32_t I0, uint32_t I1) {
        for (uint32_t I = I0; I < I1; ++I) {
            const uint32_t* b = &anode[anode_ptr[I]];
            const uint32_t* e = &anode[0] + anode_ptr[I + 1];
            for (uint32_t i = I * bs; i < I * bs + bs; ++i) {
                uint32_t q = L.p_ptr[i];
                for (const uint32_t* A = b; A != e; ++A)
                    for (uint32_t c = 0; c < 3; ++c) L.p_col[q++] = 3 * *A + c;
            }
        }
    });
    L.agg.assign(agg.begin(), agg.end());
    // W = A P pattern per node: the union of the P rows its rows reach
    // (aggregates; three columns each)
    HostVec<uint32_t> wnode_ptr, wnode;
    build_lists(
        nodes,
        [&](uint32_t I, std::vector<uint32_t>& out) {
            const size_t b = out.size();
            uint32_t prev = AMG_NONE;
            for (uint32_t i = I * bs; i < I * bs + bs; ++i)
                for (uint32_t e = L.a_ptr[i]; e < L.a_ptr[i + 1]; ++e) {
                    const uint32_t J = L.a_col[e] / bs;
                    if (J == prev) continue;   // a node's columns are adjacent
                    prev = J;
                    out.insert(out.end(), anode.begin() + anode_ptr[J],
                               anode.begin() + anode_ptr[J + 1]);
                }
            std::sort(out.begin() + (ptrdiff_t)b, out.end());
            out.erase(std::unique(out.begin() + (ptrdiff_t)b, out.end()), out.end());
        },
        wnode_ptr, wnode);
    L.w_ptr.resize(n + 1);
    L.w_ptr[0] = 0;
    for (uint32_t i = 0; i < n; ++i)
        L.w_ptr[i + 1] = L.w_ptr[i] + 3 * (wnode_ptr[i / bs + 1] - wnode_ptr[i / bs]);
    L.w_col.resize(L.w_ptr[n]);
    par_for(n, [&](uint32_t i0, uint32_t i1) {
        for (uint32_t i = i0; i < i1; ++i) {
            uint32_t q = L.w_ptr[i];
            for (uint32_t k = wnode_ptr[i / bs]; k < wnode_ptr[i / bs + 1]; ++k)
                for (uint32_t c = 0; c < 3; ++c) L.w_col[q++] = 3 * wnode[k] + c;
        }
    });
    // the nodes whose P rows hold each aggregate, ascending
    std::vector<uint32_t> hptr(na + 1, 0), hold(anode.size());
    for (uint32_t A : anode) ++hptr[A + 1];
    for (uint32_t a = 0; a < na; ++a) hptr[a + 1] += hptr[a];
    {
        std::vector<uint32_t> fill(hptr.begin(), hptr.end() - 1);
        for (uint32_t I = 0; I < nodes; ++I)
            for (uint32_t k = anode_ptr[I]; k < anode_ptr[I + 1]; ++k) hold[fill[anode[k]]++] = I;
    }
    // P^T lists: coarse column c -> the rows i holding it (ascending) and
    // its position in P
    L.pt_ptr.resize(L.nc + 1);
    L.pt_ptr[0] = 0;
    for (uint32_t c = 0; c < L.nc; ++c)
        L.pt_ptr[c + 1] = L.pt_ptr[c] + bs * (hptr[c / 3 + 1] - hptr[c / 3]);
    L.pt_row.resize(L.p_ptr[n]);
    L.pt_pos.resize(L.p_ptr[n]);
    par_for(na, [&](uint32_t a0, uint32_t a1) {
        for (uint32_t a = a0; a < a1; ++a)
            for (uint32_t m = 0; m < 3; ++m) {
                uint32_t q = L.pt_ptr[3 * a + m];
                for (uint32_t k = hptr[a]; k < hptr[a + 1]; ++k) {
                    const uint32_t I = hold[k];
                    const uint32_t pos = 3 * find(&anode[anode_ptr[I]], &anode[0] + anode_ptr[I + 1], a) + m;
                    for (uint32_t i = I * bs; i < I * bs + bs; ++i) {
                        L.pt_row[q] = i;
                        L.pt_pos[q++] = L.p_ptr[i] + pos;
                    }
                }
            }
    });
    // coarse pattern per aggregate: the union of the W rows of the nodes
    // whose P rows hold it
    C.n = L.nc;
    C.bs = 3;
    HostVec<uint32_t> cnode_ptr, cnode;
    build_lists(
        na,
        [&](uint32_t a, std::vector<uint32_t>& out) {
            const size_t b = out.size();
            for (uint32_t k = hptr[a]; k < hptr[a + 1]; ++k) {
                const uint32_t I = hold[k];
                out.insert(out.end(), wnode.begin() + wnode_ptr[I], wnode.begin() + wnode_ptr[I + 1]);
            }
            std::sort(out.begin() + (ptrdiff_t)b, out.end());
            out.erase(std::unique(out.begin() + (ptrdiff_t)b, out.end()), out.end());
        },
        cnode_ptr, cnode);
    C.a_ptr.resize(C.n + 1);
    C.a_ptr[0] = 0;
    for (uint32_t c = 0; c < C.n; ++c)
        C.a_ptr[c + 1] = C.a_ptr[c] + 3 * (cnode_ptr[c / 3 + 1] - cnode_ptr[c / 3]);
    C.a_col.resize(C.a_ptr[C.n]);
    par_for(C.n, [&](uint32_t c0, uint32_t c1) {
        for (uint32_t c = c0; c < c1; ++c) {
            uint32_t q = C.a_ptr[c];
            for (uint32_t k = cnode_ptr[c / 3]; k < cnode_ptr[c / 3 + 1]; ++k)
                for (uint32_t m = 0; m < 3; ++m) C.a_col[q++] = 3 * cnode[k] + m;
        }
    });
}

}  // namespace

static int build_host(uint32_t n, const uint32_t* rowptr, const uint32_t* col, const uint8_t* nz,
                      const double* xy, AmgHost& out) {
    out.lv.clear();
    if (n == 0 || n % 2) return -1;
    AmgHostLevel L0;
    L0.n = n;
    L0.bs = 2;
    L0.a_ptr.resize(n + 1);
    L0.a_col.resize(rowptr[n]);
    par_for(n + 1, [&](uint32_t i0, uint32_t i1) {
        std::copy(rowptr + i0, rowptr + i1, L0.a_ptr.begin() + i0);
    });
    par_for(rowptr[n], [&](uint32_t e0, uint32_t e1) {
        std::copy(col + e0, col + e1, L0.a_col.begin() + e0);
    });
    // Dirichlet-eliminated unknowns: every off-diagonal value of the row is 0
    std::vector<uint8_t> dead(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        bool any = false;
        for (uint32_t e = rowptr[i]; e < rowptr[i + 1] && !any; ++e)
            any = col[e] != i && nz[e];
        dead[i] = !any;
    }
    // rigid-body modes, centred on the mesh's centroid
    const uint32_t nodes = n / 2;
    double cx = 0.0, cy = 0.0;
    for (uint32_t I = 0; I < nodes; ++I) {
        cx += xy[2 * I];
        cy += xy[2 * I + 1];
    }
    cx /= nodes;
    cy /= nodes;
    std::vector<double> B(3 * (size_t)n, 0.0);
    for (uint32_t I = 0; I < nodes; ++I) {
        const double x = xy[2 * I] - cx, y = xy[2 * I + 1] - cy;
        double* bx = &B[3 * (size_t)(2 * I)];
        double* by = &B[3 * (size_t)(2 * I + 1)];
        bx[0] = 1.0;
        bx[1] = 0.0;
        bx[2] = -y;
        by[0] = 0.0;
        by[1] = 1.0;
        by[2] = x;
    }
    out.lv.push_back(std::move(L0));
    const uint8_t* v = nz;
    while (out.lv.back().n > AMG_COARSEST && (int)out.lv.size() < AMG_MAX_LEVELS) {
        AmgHostLevel C;
        std::vector<double> Bc;
        std::vector<uint8_t> dead_c;
        coarsen(out.lv.back(), B, dead, v, C, Bc, dead_c);
        if (C.n == 0 || C.n >= out.lv.back().n) {
            // nothing (or no reduction) to coarsen: this level is the coarsest
            AmgHostLevel& L = out.lv.back();
            L.nc = 0;
            L.t.clear(); L.agg.clear(); L.p_ptr.clear(); L.p_col.clear();
            L.w_ptr.clear(); L.w_col.clear(); L.pt_ptr.clear(); L.pt_row.clear(); L.pt_pos.clear();
            break;
        }
        out.lv.push_back(std::move(C));
        B.swap(Bc);
        dead.swap(dead_c);
        v = nullptr;   // below the finest level every structural entry counts
    }
    return 0;
}

int amg_build_host(uint32_t n, const uint32_t* rowptr, const uint32_t* col, const uint8_t* nz,
                   const double* xy, AmgHost& out) {
    try {
        return build_host(n, rowptr, col, nz, xy, out);
    } catch (const std::bad_alloc&) {
        out.lv.clear();
        return -2;
    }
}

// x . y over len contiguous entries, four partial sums in a fixed order
static double dense_dot(const double* x, const double* y, uint32_t len) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    uint32_t p = 0;
    for (; p + 4 <= len; p += 4) {
        s0 += x[p] * y[p];
        s1 += x[p + 1] * y[p + 1];
        s2 += x[p + 2] * y[p + 2];
        s3 += x[p + 3] * y[p + 3];
    }
    for (; p < len; ++p) s0 += x[p] * y[p];
    return (s0 + s1) + (s2 + s3);
}

void amg_expand_blocks(uint32_t nodes, const uint32_t* slice_ptr, const uint8_t* rowlen,
                       const uint32_t* bcol, const uint8_t* nzbits, HostVec<uint32_t>& rowptr,
                       HostVec<uint32_t>& col, HostVec<uint8_t>& nz) {
    rowptr.resize(2 * (size_t)nodes + 1);
    rowptr[0] = 0;
    for (uint32_t v = 0; v < nodes; ++v) {
        rowptr[2 * (size_t)v + 1] = rowptr[2 * (size_t)v] + 2u * rowlen[v];
        rowptr[2 * (size_t)v + 2] = rowptr[2 * (size_t)v + 1] + 2u * rowlen[v];
    }
    col.resize(rowptr[2 * (size_t)nodes]);
    nz.resize(col.size());
    par_for(nodes, [&](uint32_t v0, uint32_t v1) {
        for (uint32_t v = v0; v < v1; ++v) {
            size_t slot = (size_t)slice_ptr[v / 64] + v % 64;
            for (uint32_t k = 0; k < rowlen[v]; ++k, slot += 64) {
                const uint32_t c = bcol[slot];
                const uint8_t bits = nzbits[slot];
                for (uint32_t i = 0; i < 2; ++i) {
                    const size_t o = rowptr[2 * (size_t)v + i] + 2 * (size_t)k;
                    col[o] = 2 * c;
                    col[o + 1] = 2 * c + 1;
                    nz[o] = (bits >> (2 * i)) & 1u;
                    nz[o + 1] = (bits >> (2 * i + 1)) & 1u;
                }
            }
        }
    });
}

int amg_dense_inverse(uint32_t n, double* a) {
    // Cholesky, lower triangle in place (every inner product over
    // contiguous rows)
    double dmax = 0.0;
    for (uint32_t i = 0; i < n; ++i) dmax = fmax(dmax, a[(size_t)i * n + i]);
    std::vector<uint8_t> drop(n, 0);
    int dropped = 0;
    for (uint32_t k = 0; k < n; ++k) {
        double* rk = a + (size_t)k * n;
        const double s = rk[k] - dense_dot(rk, rk, k);
        if (!(s > 1e-13 * dmax)) {
            drop[k] = 1;
            ++dropped;
            for (uint32_t i = k; i < n; ++i) a[(size_t)i * n + k] = 0.0;
            continue;
        }
        const double lkk = sqrt(s);
        rk[k] = lkk;
        for (uint32_t i = k + 1; i < n; ++i) {
            double* ri = a + (size_t)i * n;
            ri[k] = (ri[k] - dense_dot(ri, rk, k)) / lkk;
        }
    }
    // MT = (L^-1)^T: row j of MT is column j of L^-1 (entries j .. n-1);
    // the rows of L^-1 of dropped unknowns are 0
    std::vector<double> MT((size_t)n * n, 0.0);
    for (uint32_t j = 0; j < n; ++j) {
        double* mj = MT.data() + (size_t)j * n;
        for (uint32_t i = j; i < n; ++i) {
            if (drop[i]) continue;
            const double* li = a + (size_t)i * n;
            const double t = (i == j ? 1.0 : 0.0) - dense_dot(li + j, mj + j, i - j);
            mj[i] = t / li[i];
        }
    }
    // A^-1 = (L^-1)^T L^-1: entry (i, j), i >= j, sums over p >= i
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j <= i; ++j) {
            const double t = dense_dot(MT.data() + (size_t)i * n + i, MT.data() + (size_t)j * n + i,
                                       n - i);
            a[(size_t)i * n + j] = t;
            a[(size_t)j * n + i] = t;
        }
    return dropped;
}

}  // namespace nbx

// Host check of the smoothed-aggregation hierarchy (nbots_amd/csrc/
// amg_setup.cpp), built under AddressSanitizer + UBSan by tests/amg/Makefile.
//
// Reads a system (CSR + node coordinates + right-hand side) from a binary
// file, builds the symbolic hierarchy, then runs the numeric phase and the
// V-cycle in FP64 with the loops of the device kernels (nbots_amd/csrc/
// amg.hip), as the preconditioner of a flexible CG, and prints one line:
//   levels=<L> sizes=<n0,n1,...> iters=<k> rel_res=<r> dropped=<d>
// Exit 0 when the solve reached the tolerance.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "amg.hpp"

using namespace nbx;

namespace {

struct Num {
    std::vector<double> a, dinv, p, w;
    double lmax = 0.0;
    std::vector<double> inv;   // coarsest
};

template <class T>
bool rd(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return fread(v.data(), sizeof(T), n, f) == n;
}

size_t pos_of(const HostVec<uint32_t>& col, size_t b, size_t e, uint32_t c) {
    size_t lo = b, hi = e;
    while (lo < hi) {
        const size_t m = lo + (hi - lo) / 2;
        if (col[m] < c) lo = m + 1; else hi = m;
    }
    if (lo == e || col[lo] != c) {
        fprintf(stderr, "pattern miss\n");
        exit(3);
    }
    return lo;
}

void spmv(const AmgHostLevel& L, const std::vector<double>& a, const std::vector<double>& x,
          std::vector<double>& y) {
    for (uint32_t i = 0; i < L.n; ++i) {
        double s = 0.0;
        for (uint32_t e = L.a_ptr[i]; e < L.a_ptr[i + 1]; ++e) s += a[e] * x[L.a_col[e]];
        y[i] = s;
    }
}

struct Cheb {
    int degree;
    double theta, c1[8], c2[8];
};
Cheb cheb(double lmax, int degree) {
    Cheb c;
    c.degree = degree;
    const double hi = 1.1 * lmax, lo = lmax / 30.0;
    c.theta = 0.5 * (hi + lo);
    const double delta = 0.5 * (hi - lo), sigma = c.theta / delta;
    double rho = 1.0 / sigma;
    for (int k = 1; k < degree; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        c.c1[k] = rn * rho;
        c.c2[k] = 2.0 * rn / delta;
        rho = rn;
    }
    return c;
}

// degree amg_level_degree(l): AMG_FINE_DEGREE on the finest level, then
// AMG_LEVEL1_DEGREE, AMG_LEVEL2_DEGREE and AMG_DEGREE
void smooth(const AmgHostLevel& L, const Num& N, size_t l, const std::vector<double>& b,
            std::vector<double>& x, bool zero) {
    const Cheb c = cheb(N.lmax, amg_level_degree(l));
    std::vector<double> r(L.n), d(L.n), y(L.n);
    if (!zero) spmv(L, N.a, x, y);
    for (uint32_t i = 0; i < L.n; ++i) {
        r[i] = N.dinv[i] * (b[i] - (zero ? 0.0 : y[i]));
        d[i] = r[i] / c.theta;
        x[i] = (zero ? 0.0 : x[i]) + d[i];
    }
    for (int k = 1; k < c.degree; ++k) {
        spmv(L, N.a, d, y);
        for (uint32_t i = 0; i < L.n; ++i) {
            r[i] -= N.dinv[i] * y[i];
            d[i] = c.c1[k] * d[i] + c.c2[k] * r[i];
            x[i] += d[i];
        }
    }
}

void vcycle(const AmgHost& H, const std::vector<Num>& N, size_t l, const std::vector<double>& b,
            std::vector<double>& x) {
    const AmgHostLevel& L = H.lv[l];
    x.assign(L.n, 0.0);
    if (l + 1 == H.lv.size()) {
        for (uint32_t i = 0; i < L.n; ++i) {
            double s = 0.0;
            for (uint32_t j = 0; j < L.n; ++j) s += N[l].inv[(size_t)i * L.n + j] * b[j];
            x[i] = s;
        }
        return;
    }
    smooth(L, N[l], l, b, x, true);
    std::vector<double> res(L.n), y(L.n);
    spmv(L, N[l].a, x, y);
    for (uint32_t i = 0; i < L.n; ++i) res[i] = b[i] - y[i];
    std::vector<double> bc(L.nc, 0.0), xc;
    for (uint32_t c = 0; c < L.nc; ++c) {
        double s = 0.0;
        for (uint32_t k = L.pt_ptr[c]; k < L.pt_ptr[c + 1]; ++k) s += N[l].p[L.pt_pos[k]] * res[L.pt_row[k]];
        bc[c] = s;
    }
    vcycle(H, N, l + 1, bc, xc);
    for (uint32_t i = 0; i < L.n; ++i) {
        double s = 0.0;
        for (uint32_t q = L.p_ptr[i]; q < L.p_ptr[i + 1]; ++q) s += N[l].p[q] * xc[L.p_col[q]];
        x[i] += s;
    }
    smooth(L, N[l], l, b, x, false);
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

}  // namespace

// --dense: amg_dense_inverse on two matrices with known answers, one line
//   spd_err=<max |A X - I|> lap_dropped=<d> lap_err=<max |A_KK X_KK - I|>
// (an SPD matrix of order 200, and the path-graph Laplacian of order 150,
// singular with the constants as its null space: the Cholesky pivot guard
// drops its last unknown, and the inverse of the rest is exact)
int dense_check() {
    auto max_err = [](uint32_t n, const std::vector<double>& A, const std::vector<double>& X,
                      const std::vector<uint8_t>& keep) {
        double e = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            if (!keep[i]) continue;
            for (uint32_t j = 0; j < n; ++j) {
                if (!keep[j]) continue;
                double s = 0.0;
                for (uint32_t k = 0; k < n; ++k)
                    if (keep[k]) s += A[(size_t)i * n + k] * X[(size_t)k * n + j];
                e = fmax(e, fabs(s - (i == j ? 1.0 : 0.0)));
            }
        }
        return e;
    };
    uint32_t n = 200;
    std::vector<double> A((size_t)n * n);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < n; ++j)
            A[(size_t)i * n + j] = 1.0 / (1.0 + fabs((double)i - (double)j)) + (i == j ? 3.0 : 0.0);
    std::vector<double> X = A;
    const int d0 = amg_dense_inverse(n, X.data());
    const double e0 = max_err(n, A, X, std::vector<uint8_t>(n, 1));
    n = 150;
    std::vector<double> Lp((size_t)n * n, 0.0);
    for (uint32_t i = 0; i + 1 < n; ++i) {
        Lp[(size_t)i * n + i] += 1.0;
        Lp[(size_t)(i + 1) * n + i + 1] += 1.0;
        Lp[(size_t)i * n + i + 1] -= 1.0;
        Lp[(size_t)(i + 1) * n + i] -= 1.0;
    }
    std::vector<double> Y = Lp;
    const int d1 = amg_dense_inverse(n, Y.data());
    std::vector<uint8_t> keep(n, 1);
    keep[n - 1] = 0;
    double zero = 0.0;   // the dropped row and column of the inverse
    for (uint32_t k = 0; k < n; ++k)
        zero = fmax(zero, fmax(fabs(Y[(size_t)(n - 1) * n + k]), fabs(Y[(size_t)k * n + n - 1])));
    const double e1 = max_err(n, Lp, Y, keep);
    printf("spd_dropped=%d spd_err=%.3e lap_dropped=%d lap_err=%.3e lap_dropped_entries=%.3e\n", d0,
           e0, d1, e1, zero);
    return d0 == 0 && e0 < 1e-12 && d1 == 1 && e1 < 1e-9 && zero == 0.0 ? 0 : 1;
}

int main(int argc, char** argv) {
    if (argc == 2 && std::string(argv[1]) == "--dense") return dense_check();
    if (argc < 2) {
        fprintf(stderr, "usage: amg_host_check <system.bin> [max_iter] [omega_scale]\n");
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    uint64_t hdr[2];
    if (fread(hdr, 8, 2, f) != 2) return 2;
    const uint32_t n = (uint32_t)hdr[0];
    const size_t nnz = hdr[1];
    std::vector<uint32_t> rp, col;
    std::vector<double> val, xy, b;
    if (!rd(f, rp, n + 1) || !rd(f, col, nnz) || !rd(f, val, nnz) || !rd(f, xy, n) || !rd(f, b, n))
        return 2;
    fclose(f);
    // argv[3]: a factor on the prolongator's smoothing weight (1 = the
    // method; 0 = plain aggregation, for the check that smoothing matters)
    const double omega_scale = argc > 3 ? atof(argv[3]) : 1.0;
    AmgHost H;
    std::vector<uint8_t> nz(nnz);
    for (size_t e = 0; e < nnz; ++e) nz[e] = val[e] != 0.0;
    if (amg_build_host(n, rp.data(), col.data(), nz.data(), xy.data(), H)) return 4;
    // numeric phase, with the device kernels' loops
    std::vector<Num> N(H.lv.size());
    N[0].a = val;
    int dropped = 0;
    for (size_t l = 0; l < H.lv.size(); ++l) {
        const AmgHostLevel& L = H.lv[l];
        Num& M = N[l];
        M.dinv.assign(L.n, 0.0);
        for (uint32_t i = 0; i < L.n; ++i)
            for (uint32_t e = L.a_ptr[i]; e < L.a_ptr[i + 1]; ++e)
                if (L.a_col[e] == i && M.a[e] > 0.0) M.dinv[i] = 1.0 / M.a[e];
        // power iteration on D^-1 A from a fixed +-1 start (every frequency)
        std::vector<double> x(L.n), y(L.n);
        for (uint32_t i = 0; i < L.n; ++i) x[i] = amg_sign(i);
        double s = sqrt(dot(x, x));
        for (uint32_t i = 0; i < L.n; ++i) x[i] /= s;
        for (int it = 0; it < 15; ++it) {
            spmv(L, M.a, x, y);
            for (uint32_t i = 0; i < L.n; ++i) y[i] *= M.dinv[i];
            s = sqrt(dot(y, y));
            if (!(s > 0.0)) break;
            for (uint32_t i = 0; i < L.n; ++i) x[i] = y[i] / s;
        }
        M.lmax = s;
        if (l + 1 == H.lv.size()) {
            M.inv.assign((size_t)L.n * L.n, 0.0);
            for (uint32_t i = 0; i < L.n; ++i)
                for (uint32_t e = L.a_ptr[i]; e < L.a_ptr[i + 1]; ++e)
                    M.inv[(size_t)i * L.n + L.a_col[e]] = M.a[e];
            dropped = amg_dense_inverse(L.n, M.inv.data());
            break;
        }
        const double omega = omega_scale * 4.0 / (3.0 * M.lmax);
        // P row i's position of aggregate a's columns (its groups ascend)
        auto pos_in_row = [&](uint32_t i, uint32_t a) {
            for (uint32_t q = L.p_ptr[i]; q < L.p_ptr[i + 1]; q += 3)
                if (L.p_col[q] == 3 * a) return q;
            fprintf(stderr, "aggregate %u not in P row %u\n", a, i);
            exit(3);
        };
        M.p.assign(L.p_ptr[L.n], 0.0);
        for (uint32_t i = 0; i < L.n; ++i) {
            const uint32_t own = L.agg[i / L.bs];
            if (own != AMG_NONE) {
                const uint32_t q = pos_in_row(i, own);
                for (int m = 0; m < 3; ++m) M.p[q + m] += L.t[3 * (size_t)i + m];
            }
            for (uint32_t e = L.a_ptr[i]; e < L.a_ptr[i + 1]; ++e) {
                const uint32_t j = L.a_col[e], a = L.agg[j / L.bs];
                if (a == AMG_NONE) continue;
                const double coef = omega * M.dinv[i] * M.a[e];
                const uint32_t q = pos_in_row(i, a);
                for (int m = 0; m < 3; ++m) M.p[q + m] -= coef * L.t[3 * (size_t)j + m];
            }
        }
        M.w.assign(L.w_ptr[L.n], 0.0);
        for (uint32_t i = 0; i < L.n; ++i)
            for (uint32_t e = L.a_ptr[i]; e < L.a_ptr[i + 1]; ++e) {
                const uint32_t j = L.a_col[e];
                for (uint32_t q = L.p_ptr[j]; q < L.p_ptr[j + 1]; ++q)
                    M.w[pos_of(L.w_col, L.w_ptr[i], L.w_ptr[i + 1], L.p_col[q])] += M.a[e] * M.p[q];
            }
        const AmgHostLevel& C = H.lv[l + 1];
        N[l + 1].a.assign(C.a_ptr[C.n], 0.0);
        for (uint32_t c = 0; c < L.nc; ++c)
            for (uint32_t k = L.pt_ptr[c]; k < L.pt_ptr[c + 1]; ++k) {
                const uint32_t i = L.pt_row[k];
                const double coef = M.p[L.pt_pos[k]];
                for (uint32_t q = L.w_ptr[i]; q < L.w_ptr[i + 1]; ++q)
                    N[l + 1].a[pos_of(C.a_col, C.a_ptr[c], C.a_ptr[c + 1], L.w_col[q])] += coef * M.w[q];
            }
    }
    // flexible CG with the V-cycle
    const uint32_t max_iter = argc > 2 ? (uint32_t)atoi(argv[2]) : 500;
    std::vector<double> x(n, 0.0), r = b, z, p, w(n), Ap(n);
    const double bn = sqrt(dot(b, b));
    double pw = 0.0;
    uint32_t k = 0;
    while (sqrt(dot(r, r)) > 1e-8 * bn && k < max_iter) {
        vcycle(H, N, 0, r, z);
        if (k == 0) {
            p = z;
        } else {
            const double beta = -dot(z, w) / pw;
            for (uint32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
        spmv(H.lv[0], N[0].a, p, w);
        pw = dot(p, w);
        const double alpha = dot(p, r) / pw;
        for (uint32_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * w[i];
        }
        ++k;
    }
    spmv(H.lv[0], N[0].a, x, Ap);
    double rr = 0.0;
    for (uint32_t i = 0; i < n; ++i) rr += (b[i] - Ap[i]) * (b[i] - Ap[i]);
    printf("levels=%zu sizes=", H.lv.size());
    for (size_t l = 0; l < H.lv.size(); ++l) printf("%s%u", l ? "," : "", H.lv[l].n);
    printf(" iters=%u rel_res=%.3e dropped=%d\n", k, sqrt(rr) / bn, dropped);
    return sqrt(rr) <= 1.0000001e-8 * bn ? 0 : 1;
}

// Host checks of the multigrid setup (puncpp_amd/csrc/amg.cpp) on 7-point
// Laplacians: the structure of the hierarchy (R = P^T, Galerkin operators
// A_c = R A P, symmetric, Dirichlet rows outside every aggregate, P 1 = 1 for
// the singular operator, the coarsest inverse) and its quality as a CG
// preconditioner, with a host V-cycle of the same form as the device one
// (damped Jacobi before and after each coarse correction).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "amg.h"

using namespace punc_gpu;

static int failures = 0;
#define CHECK(cond, ...)                 \
    do {                                 \
        if (!(cond)) {                   \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");           \
            ++failures;                  \
        }                                \
    } while (0)

// 7-point Laplacian on an n^3 grid.  periodic: wrap-around, singular;
// otherwise the boundary nodes are identity rows and their columns are
// eliminated from the interior rows (the field solve's Dirichlet form).
static Csr laplacian(int n, bool periodic, std::vector<uint8_t>& skip) {
    const int64_t N = (int64_t)n * n * n;
    Csr A;
    A.nrow = A.ncol = N;
    A.rp.assign((size_t)N + 1, 0);
    skip.assign((size_t)N, 0);
    auto id = [n](int i, int j, int k) { return (int64_t)((k * n + j) * n + i); };
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const int64_t r = id(i, j, k);
                const bool bnd = !periodic && (i == 0 || j == 0 || k == 0 || i == n - 1 || j == n - 1 || k == n - 1);
                std::vector<std::pair<int64_t, double>> e;
                if (bnd) {
                    skip[r] = 1;
                    e.push_back({r, 1.0});
                } else {
                    e.push_back({r, 6.0});
                    const int d[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
                    for (auto& o : d) {
                        int a = i + o[0], b = j + o[1], c = k + o[2];
                        if (periodic) {
                            a = (a + n) % n;
                            b = (b + n) % n;
                            c = (c + n) % n;
                        } else if (a == 0 || b == 0 || c == 0 || a == n - 1 || b == n - 1 || c == n - 1) {
                            continue;  // eliminated Dirichlet column
                        }
                        e.push_back({id(a, b, c), -1.0});
                    }
                }
                std::sort(e.begin(), e.end());
                for (auto& p : e) {
                    A.ci.push_back((int32_t)p.first);
                    A.v.push_back(p.second);
                }
                A.rp[r + 1] = (int64_t)A.ci.size();
            }
    return A;
}

static double entry(const Csr& A, int64_t i, int64_t j) {
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
        if (A.ci[k] == j) return A.v[k];
    return 0.0;
}

static void vcycle(const AmgHierarchy& H, size_t l, const std::vector<double>& b, std::vector<double>& y) {
    if (l == H.levels.size()) {
        const int64_t n = H.ncoarse;
        y.assign((size_t)n, 0.0);
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j < n; ++j) y[i] += H.coarse_inv[i * n + j] * b[j];
        return;
    }
    const AmgLevel& L = H.levels[l];
    const int64_t n = L.A.nrow;
    std::vector<double> x((size_t)n), t((size_t)n), r((size_t)n), bc((size_t)L.R.nrow), yc;
    for (int64_t i = 0; i < n; ++i) x[i] = L.omega * L.dinv[i] * b[i];
    csr_apply(L.A, x.data(), t.data());
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - t[i];
    csr_apply(L.R, r.data(), bc.data());
    vcycle(H, l + 1, bc, yc);
    csr_apply(L.P, yc.data(), t.data());
    for (int64_t i = 0; i < n; ++i) x[i] += t[i];
    csr_apply(L.A, x.data(), t.data());
    y.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) y[i] = x[i] + L.omega * L.dinv[i] * (b[i] - t[i]);
}

// PCG iterations to |r| <= 1e-12 |b| (M: 0 Jacobi, 1 the V-cycle).
static int pcg(const Csr& A, const AmgHierarchy& H, const std::vector<double>& b, int M, std::vector<double>& x) {
    const int64_t n = A.nrow;
    std::vector<double> r = b, z, p, q((size_t)n);
    x.assign((size_t)n, 0.0);
    auto prec = [&](const std::vector<double>& in, std::vector<double>& out) {
        if (M == 1) {
            vcycle(H, 0, in, out);
        } else {
            out.resize((size_t)n);
            for (int64_t i = 0; i < n; ++i) out[i] = in[i] / entry(A, i, i);
        }
    };
    prec(r, z);
    p = z;
    double rz = 0.0, bb = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        rz += r[i] * z[i];
        bb += b[i] * b[i];
    }
    for (int it = 1; it <= 5000; ++it) {
        csr_apply(A, p.data(), q.data());
        double pq = 0.0;
        for (int64_t i = 0; i < n; ++i) pq += p[i] * q[i];
        const double a = rz / pq;
        double rr = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += a * p[i];
            r[i] -= a * q[i];
            rr += r[i] * r[i];
        }
        if (std::sqrt(rr) <= 1e-12 * std::sqrt(bb)) return it;
        prec(r, z);
        double rz2 = 0.0;
        for (int64_t i = 0; i < n; ++i) rz2 += r[i] * z[i];
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + (rz2 / rz) * p[i];
        rz = rz2;
    }
    return -1;
}

static void check_structure(const AmgHierarchy& H, const std::vector<uint8_t>& skip, bool singular) {
    uint64_t s = 12345;
    auto rnd = [&s]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    };
    for (size_t l = 0; l < H.levels.size(); ++l) {
        const AmgLevel& L = H.levels[l];
        // R = P^T, entry by entry
        for (int64_t i = 0; i < L.P.nrow; ++i)
            for (int64_t k = L.P.rp[i]; k < L.P.rp[i + 1]; ++k)
                CHECK(entry(L.R, L.P.ci[k], i) == L.P.v[k], "level %zu: R != P^T", l);
        CHECK(L.R.ci.size() == L.P.ci.size(), "level %zu: nnz(R) != nnz(P)", l);
        // Galerkin: A_next x = R (A (P x)) for a random x
        const Csr& An = l + 1 < H.levels.size() ? H.levels[l + 1].A : Csr();
        const int64_t nc = L.P.ncol;
        std::vector<double> x((size_t)nc), px((size_t)L.A.nrow), apx((size_t)L.A.nrow), rapx((size_t)nc);
        for (auto& v : x) v = rnd();
        csr_apply(L.P, x.data(), px.data());
        csr_apply(L.A, px.data(), apx.data());
        csr_apply(L.R, apx.data(), rapx.data());
        double err = 0.0, scale = 0.0;
        if (l + 1 < H.levels.size()) {
            std::vector<double> ax((size_t)nc);
            csr_apply(An, x.data(), ax.data());
            for (int64_t i = 0; i < nc; ++i) {
                err = std::max(err, std::fabs(ax[i] - rapx[i]));
                scale = std::max(scale, std::fabs(rapx[i]));
            }
            CHECK(err <= 1e-12 * scale, "level %zu: A_c != R A P (%g)", l + 1, err / scale);
            for (int64_t i = 0; i < An.nrow; ++i)  // symmetric
                for (int64_t k = An.rp[i]; k < An.rp[i + 1]; ++k)
                    CHECK(std::fabs(entry(An, An.ci[k], i) - An.v[k]) <= 1e-12 * std::fabs(An.v[k]),
                          "level %zu: A_c not symmetric", l + 1);
        }
        if (l == 0) {
            for (int64_t i = 0; i < L.P.nrow; ++i)
                if (!skip.empty() && skip[i]) CHECK(L.P.rp[i + 1] == L.P.rp[i], "Dirichlet row %ld aggregated", (long)i);
            if (singular)  // the constant is preserved: P 1 = 1
                for (int64_t i = 0; i < L.P.nrow; ++i) {
                    double rs = 0.0;
                    for (int64_t k = L.P.rp[i]; k < L.P.rp[i + 1]; ++k) rs += L.P.v[k];
                    CHECK(std::fabs(rs - 1.0) <= 1e-12, "row %ld of P sums to %g", (long)i, rs);
                }
        }
        CHECK(L.omega > 0.0 && L.omega <= 2.0 / 3.0, "level %zu: smoother weight %g", l, L.omega);
    }
    CHECK(H.ncoarse > 0 && H.ncoarse <= 256, "coarsest size %ld", (long)H.ncoarse);
    // the coarsest inverse: C (R A P) e_j = e_j (singular: C inverts R A P + g 1 1^T,
    // so the check is made on vectors orthogonal to the constant)
    if (!H.levels.empty()) {
        const AmgLevel& L = H.levels.back();
        const int64_t nc = H.ncoarse;
        double worst = 0.0;
        for (int64_t j = 0; j < nc; ++j) {
            std::vector<double> e((size_t)nc, 0.0), px((size_t)L.A.nrow), apx((size_t)L.A.nrow), ac((size_t)nc);
            e[j] = 1.0;
            if (singular)
                for (int64_t i = 0; i < nc; ++i) e[i] -= 1.0 / (double)nc;
            csr_apply(L.P, e.data(), px.data());
            csr_apply(L.A, px.data(), apx.data());
            csr_apply(L.R, apx.data(), ac.data());
            for (int64_t i = 0; i < nc; ++i) {
                double y = 0.0;
                for (int64_t k = 0; k < nc; ++k) y += H.coarse_inv[i * nc + k] * ac[k];
                worst = std::max(worst, std::fabs(y - e[i]));
            }
        }
        CHECK(worst <= 1e-9, "coarsest inverse: |C A_c e - e| = %g", worst);
    }
}

int main() {
    // Dirichlet
    {
        std::vector<uint8_t> skip;
        const Csr A = laplacian(26, false, skip);
        AmgHierarchy H;
        std::string err;
        CHECK(amg_build(A, skip, false, H, err), "build: %s", err.c_str());
        CHECK(H.levels.size() >= 2, "levels %zu", H.levels.size());
        check_structure(H, skip, false);
        std::vector<double> b((size_t)A.nrow), x;
        for (int64_t i = 0; i < A.nrow; ++i) b[i] = skip[i] ? 0.0 : std::sin(0.37 * (double)i) + 0.1;
        const int it_m = pcg(A, H, b, 1, x);
        const int it_j = pcg(A, H, b, 0, x);
        std::printf("dirichlet: %ld unknowns, %zu levels + dense %ld, multigrid %d iterations, jacobi %d\n",
                    (long)A.nrow, H.levels.size(), (long)H.ncoarse, it_m, it_j);
        CHECK(it_m > 0 && it_m <= 30, "multigrid iterations %d", it_m);
        CHECK(it_j > 3 * it_m, "jacobi iterations %d", it_j);
        std::vector<double> ax((size_t)A.nrow);
        csr_apply(A, x.data(), ax.data());
        double rn = 0.0, bn = 0.0;
        for (int64_t i = 0; i < A.nrow; ++i) {
            rn += (ax[i] - b[i]) * (ax[i] - b[i]);
            bn += b[i] * b[i];
        }
        CHECK(std::sqrt(rn) <= 1e-11 * std::sqrt(bn), "residual %g", std::sqrt(rn / bn));
    }
    // Periodic (singular)
    {
        std::vector<uint8_t> skip;
        const Csr A = laplacian(20, true, skip);
        AmgHierarchy H;
        std::string err;
        CHECK(amg_build(A, {}, true, H, err), "build: %s", err.c_str());
        check_structure(H, {}, true);
        std::vector<double> b((size_t)A.nrow), x;
        double mean = 0.0;
        for (int64_t i = 0; i < A.nrow; ++i) mean += (b[i] = std::cos(0.11 * (double)i));
        mean /= (double)A.nrow;
        for (auto& v : b) v -= mean;  // consistent right-hand side
        const int it_m = pcg(A, H, b, 1, x);
        std::printf("periodic: %ld unknowns, %zu levels + dense %ld, multigrid %d iterations\n", (long)A.nrow,
                    H.levels.size(), (long)H.ncoarse, it_m);
        CHECK(it_m > 0 && it_m <= 30, "multigrid iterations %d", it_m);
    }
    // Too small for a hierarchy
    {
        std::vector<uint8_t> skip;
        const Csr A = laplacian(5, false, skip);
        AmgHierarchy H;
        std::string err;
        CHECK(amg_build(A, skip, false, H, err) && H.levels.empty() && H.ncoarse == A.nrow, "small operator");
    }
    std::printf(failures ? "amg_check: %d failures\n" : "amg_check: ok\n", failures);
    return failures ? 1 : 0;
}

// Smoothed-aggregation AMG setup (host).  Per level:
//  * strength: j is a strong neighbour of i when |a_ij| >= theta sqrt(|a_ii a_jj|)
//    (theta 0.08 on the finest level; 0.02 below it, where the Galerkin
//    operators are denser and their couplings individually weaker);
//  * aggregation in node order, three passes: a node whose strong
//    neighbours are all free roots an aggregate with them; a remaining node
//    joins the aggregate of its first aggregated strong neighbour; what is
//    left roots aggregates with its free neighbours;
//  * tentative prolongation T (one 1 per aggregated row: the constant near
//    null space), smoothed by one damped-Jacobi step:
//    P = (I - 4/(3 rho) D^-1 A) T, rho the spectral radius of D^-1 A from a
//    power iteration;
//  * Galerkin coarse operator A_c = P^T A P.
// The V-cycle's smoother is damped Jacobi with weight min(2/3, 4/(3 rho)),
// one sweep before and one after the coarse correction, which keeps the
// preconditioner symmetric positive definite.
#include "amg.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace punc_gpu {

void csr_apply(const Csr& A, const double* x, double* y) {
    for (int64_t i = 0; i < A.nrow; ++i) {
        double s = 0.0;
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) s += A.v[k] * x[A.ci[k]];
        y[i] = s;
    }
}

namespace {

constexpr int64_t COARSE_MAX = 256;   // coarsest size solved densely
constexpr int64_t DENSE_LIMIT = 2048;  // largest dense coarse solve accepted when coarsening stalls
constexpr int MAX_LEVELS = 12;

std::vector<double> diagonal(const Csr& A) {
    std::vector<double> d((size_t)A.nrow, 0.0);
    for (int64_t i = 0; i < A.nrow; ++i)
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] == i) d[i] = A.v[k];
    return d;
}

// Aggregates (agg[i] = aggregate id, -1 for skipped rows); returns their number.
int64_t aggregate(const Csr& A, const std::vector<double>& dg, const std::vector<uint8_t>& skip, double theta,
                  std::vector<int64_t>& agg) {
    const int64_t n = A.nrow;
    auto strong = [&](int64_t i, int64_t k) {
        const int32_t j = A.ci[k];
        return j != i && std::fabs(A.v[k]) >= theta * std::sqrt(std::fabs(dg[i] * dg[j]));
    };
    constexpr int64_t FREE = -1, SKIP = -2;
    agg.assign((size_t)n, FREE);
    for (int64_t i = 0; i < n; ++i)
        if (!skip.empty() && skip[i]) agg[i] = SKIP;
    int64_t na = 0;
    // pass 1: roots whose strong neighbourhood is entirely free
    for (int64_t i = 0; i < n; ++i) {
        if (agg[i] != FREE) continue;
        bool any = false, all_free = true;
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (strong(i, k)) {
                any = true;
                if (agg[A.ci[k]] != FREE) {
                    all_free = false;
                    break;
                }
            }
        if (!any || !all_free) continue;
        agg[i] = na;
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (strong(i, k)) agg[A.ci[k]] = na;
        ++na;
    }
    // pass 2: join the aggregate of the first aggregated strong neighbour
    // (aggregates of pass 1 only, so that the joins do not chain)
    std::vector<int64_t> join((size_t)n, FREE);
    for (int64_t i = 0; i < n; ++i) {
        if (agg[i] != FREE) continue;
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (strong(i, k) && agg[A.ci[k]] >= 0) {
                join[i] = agg[A.ci[k]];
                break;
            }
    }
    for (int64_t i = 0; i < n; ++i)
        if (join[i] >= 0) agg[i] = join[i];
    // pass 3: the rest root aggregates with their free strong neighbours
    for (int64_t i = 0; i < n; ++i) {
        if (agg[i] != FREE) continue;
        agg[i] = na;
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (strong(i, k) && agg[A.ci[k]] == FREE) agg[A.ci[k]] = na;
        ++na;
    }
    for (int64_t i = 0; i < n; ++i)
        if (agg[i] == SKIP) agg[i] = -1;
    return na;
}

// Spectral radius of D^-1 A by power iteration from a fixed start vector.
double spectral_radius(const Csr& A, const std::vector<double>& dinv) {
    const int64_t n = A.nrow;
    std::vector<double> x((size_t)n), y((size_t)n);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (int64_t i = 0; i < n; ++i) {  // splitmix64 values in [0.5, 1.5)
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        x[i] = 0.5 + (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
    double rho = 0.0;
    for (int it = 0; it < 20; ++it) {
        double nx = 0.0;
        for (double v : x) nx += v * v;
        nx = std::sqrt(nx);
        if (nx == 0.0) return 1.0;
        for (double& v : x) v /= nx;
        csr_apply(A, x.data(), y.data());
        double ny = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            y[i] *= dinv[i];
            ny += y[i] * y[i];
        }
        rho = std::sqrt(ny);
        x.swap(y);
    }
    return rho;
}

Csr transpose(const Csr& A) {
    Csr T;
    T.nrow = A.ncol;
    T.ncol = A.nrow;
    T.rp.assign((size_t)T.nrow + 1, 0);
    for (int32_t j : A.ci) T.rp[j + 1]++;
    for (int64_t i = 0; i < T.nrow; ++i) T.rp[i + 1] += T.rp[i];
    T.ci.resize(A.ci.size());
    T.v.resize(A.v.size());
    std::vector<int64_t> pos(T.rp.begin(), T.rp.end() - 1);
    for (int64_t i = 0; i < A.nrow; ++i)  // rows ascending: columns of T come out sorted
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int64_t d = pos[A.ci[k]]++;
            T.ci[d] = (int32_t)i;
            T.v[d] = A.v[k];
        }
    return T;
}

// C = A B (Gustavson, dense accumulator with a marker per column).
Csr multiply(const Csr& A, const Csr& B) {
    Csr C;
    C.nrow = A.nrow;
    C.ncol = B.ncol;
    C.rp.assign((size_t)C.nrow + 1, 0);
    std::vector<int64_t> mark((size_t)B.ncol, -1);
    std::vector<double> acc((size_t)B.ncol, 0.0);
    std::vector<int32_t> cols;
    for (int64_t i = 0; i < A.nrow; ++i) {
        cols.clear();
        for (int64_t ka = A.rp[i]; ka < A.rp[i + 1]; ++ka) {
            const int32_t j = A.ci[ka];
            const double a = A.v[ka];
            for (int64_t kb = B.rp[j]; kb < B.rp[j + 1]; ++kb) {
                const int32_t c = B.ci[kb];
                if (mark[c] != i) {
                    mark[c] = i;
                    acc[c] = 0.0;
                    cols.push_back(c);
                }
                acc[c] += a * B.v[kb];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int32_t c : cols) {
            C.ci.push_back(c);
            C.v.push_back(acc[c]);
        }
        C.rp[i + 1] = (int64_t)C.ci.size();
    }
    return C;
}

// P = T - w D^-1 A T for the aggregates agg (T: 1 at (i, agg[i])).
Csr smoothed_prolongation(const Csr& A, const std::vector<double>& dinv, const std::vector<int64_t>& agg, int64_t na,
                          double w) {
    Csr P;
    P.nrow = A.nrow;
    P.ncol = na;
    P.rp.assign((size_t)A.nrow + 1, 0);
    std::vector<int64_t> mark((size_t)na, -1);
    std::vector<double> acc((size_t)na, 0.0);
    std::vector<int32_t> cols;
    for (int64_t i = 0; i < A.nrow; ++i) {
        cols.clear();
        auto add = [&](int64_t c, double v) {
            if (mark[c] != i) {
                mark[c] = i;
                acc[c] = 0.0;
                cols.push_back((int32_t)c);
            }
            acc[c] += v;
        };
        if (agg[i] >= 0) add(agg[i], 1.0);
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int64_t a = agg[A.ci[k]];
            if (a >= 0) add(a, -w * dinv[i] * A.v[k]);
        }
        std::sort(cols.begin(), cols.end());
        for (int32_t c : cols)
            if (acc[c] != 0.0) {
                P.ci.push_back(c);
                P.v.push_back(acc[c]);
            }
        P.rp[i + 1] = (int64_t)P.ci.size();
    }
    return P;
}

// Dense inverse of an SPD matrix (row-major, n x n) through its Cholesky factor.
bool spd_inverse(std::vector<double>& a, int64_t n) {
    std::vector<double> L((size_t)(n * n), 0.0);
    for (int64_t j = 0; j < n; ++j) {
        double d = a[j * n + j];
        for (int64_t k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
        if (!(d > 0.0)) return false;
        const double ljj = std::sqrt(d);
        L[j * n + j] = ljj;
        for (int64_t i = j + 1; i < n; ++i) {
            double s = a[i * n + j];
            for (int64_t k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
            L[i * n + j] = s / ljj;
        }
    }
    std::vector<double> col((size_t)n);
    for (int64_t c = 0; c < n; ++c) {  // solve L L^T x = e_c
        for (int64_t i = 0; i < n; ++i) {
            double s = i == c ? 1.0 : 0.0;
            for (int64_t k = 0; k < i; ++k) s -= L[i * n + k] * col[k];
            col[i] = s / L[i * n + i];
        }
        for (int64_t i = n - 1; i >= 0; --i) {
            double s = col[i];
            for (int64_t k = i + 1; k < n; ++k) s -= L[k * n + i] * col[k];
            col[i] = s / L[i * n + i];
        }
        for (int64_t i = 0; i < n; ++i) a[i * n + c] = col[i];
    }
    return true;
}

}  // namespace

bool amg_build(const Csr& A0, const std::vector<uint8_t>& skip0, bool singular, AmgHierarchy& H, std::string& err) {
    H = AmgHierarchy();
    Csr A = A0;
    std::vector<uint8_t> skip = skip0;
    while (A.nrow > COARSE_MAX && (int)H.levels.size() < MAX_LEVELS) {
        AmgLevel L;
        const std::vector<double> dg = diagonal(A);
        L.dinv.resize((size_t)A.nrow);
        for (int64_t i = 0; i < A.nrow; ++i) {
            if (!(dg[i] > 0.0)) {
                err = "operator with a non-positive diagonal entry";
                return false;
            }
            L.dinv[i] = 1.0 / dg[i];
        }
        std::vector<int64_t> agg;
        const int64_t na = aggregate(A, dg, skip, H.levels.empty() ? 0.08 : 0.02, agg);
        if (na == 0 || na > (A.nrow * 7) / 10) break;  // coarsening stalled
        const double rho = spectral_radius(A, L.dinv);
        L.omega = std::min(2.0 / 3.0, 4.0 / (3.0 * rho));
        L.P = smoothed_prolongation(A, L.dinv, agg, na, 4.0 / (3.0 * rho));
        L.R = transpose(L.P);
        Csr Ac = multiply(L.R, multiply(A, L.P));
        L.A = std::move(A);
        H.levels.push_back(std::move(L));
        A = std::move(Ac);
        skip.clear();
    }
    const int64_t n = A.nrow;
    if (n > DENSE_LIMIT) {
        err = "multigrid coarsening stalled at " + std::to_string(n) + " unknowns";
        return false;
    }
    H.ncoarse = n;
    H.coarse_inv.assign((size_t)(n * n), 0.0);
    double tr = 0.0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            H.coarse_inv[i * n + A.ci[k]] = A.v[k];
            if (A.ci[k] == i) tr += A.v[k];
        }
    if (singular && n > 0) {
        const double g = tr / ((double)n * (double)n);
        for (double& v : H.coarse_inv) v += g;
    }
    if (!spd_inverse(H.coarse_inv, n)) {
        err = "coarsest multigrid operator is not positive definite";
        return false;
    }
    return true;
}

}  // namespace punc_gpu

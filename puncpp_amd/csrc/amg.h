// Smoothed-aggregation algebraic multigrid: the hierarchy is built on the host
// once per operator and applied on the device as the CG preconditioner of the
// field solve.  It stands where the reference's PETSc Krylov solver takes
// hypre's AMG preconditioner (poisson.cpp:299-304).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace punc_gpu {

struct Csr {
    int64_t nrow = 0, ncol = 0;
    std::vector<int64_t> rp;  // nrow + 1
    std::vector<int32_t> ci;  // sorted within a row
    std::vector<double> v;
};

struct AmgLevel {
    Csr A;                     // the operator on this level
    Csr P, R;                  // prolongation from the next level, R = P^T
    std::vector<double> dinv;  // 1 / diag(A)
    double omega = 0.0;        // damped-Jacobi weight of the smoother
};

struct AmgHierarchy {
    std::vector<AmgLevel> levels;     // levels with a coarser level below them
    int64_t ncoarse = 0;              // size of the coarsest operator
    std::vector<double> coarse_inv;   // its dense inverse (row-major), ncoarse^2
};

// Builds the hierarchy of the symmetric positive (semi)definite operator A.
// skip: rows left out of every aggregate (rows of the identity, such as
// eliminated Dirichlet dofs, which the smoother solves exactly); may be empty.
// singular: the constant vector is A's null space (no Dirichlet dofs); the
// coarsest solve is then the inverse of A_c + g 1 1^T, which acts as the
// pseudo-inverse on right-hand sides orthogonal to the constant.
bool amg_build(const Csr& A, const std::vector<uint8_t>& skip, bool singular, AmgHierarchy& H, std::string& err);

// y = A x (host; for tests and the setup's own use).
void csr_apply(const Csr& A, const double* x, double* y);

}  // namespace punc_gpu

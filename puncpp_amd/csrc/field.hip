// Field solve on the device (reference PoissonSolver, poisson.cpp:260-376 with
// Potential3D.ufl; ESolver, efield.cpp:41-75 with EField3D.ufl).
//
// Assembly (host, once per mesh): CG1 element matrices on each tetrahedron
// from the gradients g_k of its barycentric coordinates and its volume V:
//   stiffness  K_ij = eps0 V g_i.g_j                 (a = eps0 grad(phi).grad(v) dx)
//   mass       M_ij = V (1 + delta_ij) / 20          (L = rho v dx, and the E projection)
//   gradient   B_c,ij = -(V / 4) (g_j)_c             (L = -grad(phi).v dx, component c)
// summed into matrices with one sparsity (the dof graph of the cells), kept
// on the device in ELL layout for the one-off products (right-hand sides)
// and in CSR for the operators the conjugate-gradient iterations apply.
// Dirichlet dofs are eliminated symmetrically: the operator keeps the free
// rows and columns and has unit rows at the Dirichlet dofs; the right-hand
// side is b - K phi_D on free dofs and phi_D on Dirichlet dofs.  This has
// the solution of the reference's row-replaced system (DirichletBC::apply)
// and keeps the operator symmetric positive definite, so the solves are
// preconditioned conjugate gradients: smoothed-aggregation multigrid for the
// Poisson operator (amg.cpp), Jacobi for the mass matrix.  Every reduction is a set of
// per-block partials that every block of the next kernel adds in block order
// (deterministic; no atomics or fences - the kernel boundary orders the
// memory), and the CG scalars stay on the device: iterations run in batches
// without host synchronisation.
#include "field.h"

#include "amg.h"

#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>

namespace punc_gpu {

namespace {

constexpr int FB = 256;         // threads per block
constexpr int MAX_BLOCKS = 1024;  // also the partials every block of the next kernel adds
constexpr int BATCH = 32;       // CG iterations between convergence checks (Jacobi; even)
constexpr int BATCH_AMG = 8;    // the same with the multigrid preconditioner (even)

struct CgState {
    double rz[2];  // r.z entering iteration k at [k & 1]
    double tol2;   // stop when |r|^2 <= tol2
    double mean;   // last k_finish_mean result
    int64_t it;    // iterations when converged
    int64_t base;  // iterations before the current batch
    int32_t done;
    int32_t pad;
};

__device__ inline double block_reduce(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r += sh[k];
    return r;  // valid in thread 0
}

// Sum of cnt partials part[k * stride + off] in a fixed order; every block
// computes the same value (the partials were written by an earlier kernel,
// so no atomics or fences are needed).  Returned to every thread.
__device__ inline double sum_partials(const double* __restrict__ part, int cnt, int stride, int off, double* sh) {
    __shared__ double bc;
    double s = 0.0;
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) s += part[(int64_t)k * stride + off];
    const double r = block_reduce(s, sh);
    if (threadIdx.x == 0) bc = r;
    __syncthreads();
    return bc;
}

// Sparse matrices are stored as ELL: every row padded to the widest row's w
// entries (padding: the row's own column with value 0), entry k of row i at
// k * n + i.  One thread per row: a wavefront's k-th entries are contiguous,
// and the w products of a row are independent loads.
__device__ inline double row_dot(const int32_t* __restrict__ col, const double* __restrict__ av, int w, int64_t n,
                                 const double* __restrict__ x, int64_t i) {
    double s = 0.0;
    for (int k = 0; k < w; ++k) s += av[k * n + i] * x[col[k * n + i]];
    return s;
}

// CSR matrix on the device (32-bit row pointers).
struct CsrDev {
    const int32_t* rp;  // rows + 1
    const int32_t* ci;
    const double* v;
    int64_t rows;
};

// Row i of M applied to NC interleaved vectors (x[NC * j + c]) by a group of
// G lanes (sub: this lane's place in the group): lane g takes entries
// g, g + G, ... of the row; the group's sums are added with lane shuffles,
// so every lane of the group returns the row's values.
template <int NC, int G>
__device__ inline void group_row_nc(const CsrDev& M, int64_t i, int sub, const double* __restrict__ x, double out[NC]) {
    for (int c = 0; c < NC; ++c) out[c] = 0.0;
    for (int32_t k = M.rp[i] + sub; k < M.rp[i + 1]; k += G) {
        const double a = M.v[k];
        const int64_t j = M.ci[k];
        for (int c = 0; c < NC; ++c) out[c] += a * x[NC * j + c];
    }
    for (int o = G / 2; o > 0; o >>= 1)
        for (int c = 0; c < NC; ++c) out[c] += __shfl_xor(out[c], o, G);
}

// ELL matrix on the device (layout of row_dot).
struct EllDev {
    const int32_t* col;
    const double* v;
    int w;
    int64_t rows;
};

// The same for an ELL matrix, one lane per row (G = 1).
template <int NC, int G>
__device__ inline void group_row_nc(const EllDev& M, int64_t i, int sub, const double* __restrict__ x, double out[NC]) {
    static_assert(G == 1, "ELL rows take one lane");
    for (int c = 0; c < NC; ++c) out[c] = 0.0;
    for (int k = 0; k < M.w; ++k) {
        const double a = M.v[k * M.rows + i];
        const int64_t j = M.col[k * M.rows + i];
        for (int c = 0; c < NC; ++c) out[c] += a * x[NC * j + c];
    }
}

__global__ void __launch_bounds__(FB) k_spmv(const int32_t* __restrict__ col, int w, const double* __restrict__ av,
                                             const double* __restrict__ x, double* __restrict__ y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = row_dot(col, av, w, n, x, i);
}

// phi_D: the exterior dofs take ext (0 without it), object j's dofs its
// potential objpot[j - 1] (0 without them), the free dofs 0.
__global__ void k_dirichlet_values(const int32_t* __restrict__ kind, const double* __restrict__ ext,
                                   const double* __restrict__ objpot, double* __restrict__ phid, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = kind[i];
        phid[i] = k == 0 ? (ext ? ext[i] : 0.0) : (k >= 1 && objpot ? objpot[k - 1] : 0.0);
    }
}

// Poisson right-hand side of the eliminated system: b_i - (K phi_D)_i on free
// dofs, phi_D on Dirichlet dofs; the initial iterate x0 = phi_D there.
__global__ void k_rhs_dirichlet(const double* __restrict__ b, const double* __restrict__ kphid,
                                const double* __restrict__ phid, const uint8_t* __restrict__ isd,
                                const double* __restrict__ prev, double* __restrict__ rhs, double* __restrict__ x,
                                int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        rhs[i] = isd[i] ? phid[i] : b[i] - kphid[i];
        x[i] = isd[i] ? phid[i] : (prev ? prev[i] : 0.0);  // prev: warm start from the last solution
    }
}

// rhs[3i + c] = (B_c phi)_i: the three gradient matrices share the sparsity.
__global__ void __launch_bounds__(FB) k_grad3(const int32_t* __restrict__ col, int w, const double* __restrict__ g0,
                                              const double* __restrict__ g1, const double* __restrict__ g2,
                                              const double* __restrict__ phi, double* __restrict__ rhs, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int k = 0; k < w; ++k) {
            const double f = phi[col[k * n + i]];
            s0 += g0[k * n + i] * f;
            s1 += g1[k * n + i] * f;
            s2 += g2[k * n + i] * f;
        }
        rhs[3 * i] = s0;
        rhs[3 * i + 1] = s1;
        rhs[3 * i + 2] = s2;
    }
}

// out_j = sum_i sum_c (B_c)_ij y[3i + c]: the transpose of k_grad3's product
// (out zeroed first; FP64 atomics, the sum order varies).
__global__ void __launch_bounds__(FB) k_grad3_t(const int32_t* __restrict__ col, int w, const double* __restrict__ g0,
                                                const double* __restrict__ g1, const double* __restrict__ g2,
                                                const double* __restrict__ y, double* __restrict__ out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double y0 = y[3 * i], y1 = y[3 * i + 1], y2 = y[3 * i + 2];
        for (int k = 0; k < w; ++k) {
            const double v = (g0[k * n + i] * y0 + g1[k * n + i] * y1) + g2[k * n + i] * y2;
            if (v != 0.0) unsafeAtomicAdd(&out[col[k * n + i]], v);
        }
    }
}

// Per-block sums of v, then the mean (one block).
__global__ void __launch_bounds__(FB) k_partial_sum(const double* __restrict__ v, int64_t n, double* part) {
    __shared__ double sh[8];
    double a = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a += v[i];
    const double t = block_reduce(a, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(FB) k_finish_mean(const double* __restrict__ part, int cnt, int64_t n, CgState* S) {
    __shared__ double sh[8];
    const double t = sum_partials(part, cnt, 1, 0, sh);
    if (threadIdx.x == 0) S->mean = t / (double)n;
}

// v -= mean (orthogonalisation against the normalised constant vector).
__global__ void k_remove_mean(double* __restrict__ v, int64_t n, const CgState* __restrict__ S) {
    const double m = S->mean;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] -= m;
}

// CG start: r = b - A x (x given), z = D^-1 r, p = z; per-block r.z, r.r, b.b
// (dinv null: r, r.r and b.b; the V-cycle writes z and r.z).  NC: the
// vectors hold NC interleaved components (one CG on the block-diagonal
// system); A and D^-1 are those of one component.
template <int NC, int G, typename Op>
__global__ void __launch_bounds__(FB) k_cg_init(const Op A, const double* __restrict__ dinv,
                                                const double* __restrict__ b, const double* __restrict__ x,
                                                double* __restrict__ r, double* __restrict__ z,
                                                double* __restrict__ p, double* part) {
    __shared__ double sh[8];
    const int sub = threadIdx.x & (G - 1);
    const int64_t step = ((int64_t)gridDim.x * blockDim.x) / G;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; i < A.rows; i += step) {
        double ax[NC];
        group_row_nc<NC, G>(A, i, sub, x, ax);
        if (sub != 0) continue;
        for (int c = 0; c < NC; ++c) {
            const int64_t m = NC * i + c;
            const double ri = b[m] - ax[c];
            r[m] = ri;
            if (dinv) {  // Jacobi; without it the V-cycle writes z and r.z
                const double zi = dinv[i] * ri;
                z[m] = zi;
                p[m] = zi;
                a0 += ri * zi;
            }
            a1 += ri * ri;
            a2 += b[m] * b[m];
        }
    }
    const double t0 = block_reduce(a0, sh), t1 = block_reduce(a1, sh), t2 = block_reduce(a2, sh);
    if (threadIdx.x == 0) {
        if (dinv) part[3 * blockIdx.x] = t0;
        part[3 * blockIdx.x + 1] = t1;
        part[3 * blockIdx.x + 2] = t2;
    }
}

__global__ void __launch_bounds__(FB) k_cg_init_finish(const double* __restrict__ part, int cnt, double rtol,
                                                       double atol, CgState* S) {
    __shared__ double sh[8];
    const double rz = sum_partials(part, cnt, 3, 0, sh);
    const double rr = sum_partials(part, cnt, 3, 1, sh);
    const double bb = sum_partials(part, cnt, 3, 2, sh);
    if (threadIdx.x == 0) {
        S->rz[0] = rz;
        const double tol = fmax(rtol * sqrt(bb), atol);
        S->tol2 = tol * tol;
        S->it = 0;
        S->base = 0;
        S->done = rr <= S->tol2 ? 1 : 0;
    }
}

// Iteration t = base + k (batches hold an even number of iterations, so
// t & 1 = k & 1).  rz[t & 1] holds r.z entering iteration t.
//
// Step 1 (direction and product in one pass): beta = rz_t / rz_{t-1} from the
// previous update's partials (every block adds them), convergence on |r|^2;
// then p = z + beta p and q = A p computed as A z + beta q (the product of
// the updated p needs neighbours' p; A z only needs z, complete since the
// last kernel).  Per-block p.q.  Block 0 records rz_t and, at convergence,
// the flag and the iteration count.
template <int NC, int G, typename Op>
__global__ void __launch_bounds__(FB) k_cg_dir_spmv(int64_t k, const Op A, const double* __restrict__ z,
                                                    double* __restrict__ p, double* __restrict__ q,
                                                    const double* __restrict__ part2, int cnt2, double* part_pq,
                                                    CgState* S) {
    __shared__ double sh[8];
    if (S->done) return;  // uniform: written by an earlier kernel
    const bool first = S->base == 0 && k == 0;
    double beta = 0.0;
    if (!first) {
        const double rz = sum_partials(part2, cnt2, 2, 0, sh);
        const double rr = sum_partials(part2, cnt2, 2, 1, sh);
        if (rr <= S->tol2) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                S->done = 1;
                S->it = S->base + k;
            }
            return;
        }
        beta = rz / S->rz[(k + 1) & 1];
        if (blockIdx.x == 0 && threadIdx.x == 0) S->rz[k & 1] = rz;
    }
    double a = 0.0;
    const int sub = threadIdx.x & (G - 1);
    const int64_t step = ((int64_t)gridDim.x * blockDim.x) / G;
    for (int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; i < A.rows; i += step) {
        double az[NC];
        group_row_nc<NC, G>(A, i, sub, z, az);
        if (sub != 0) continue;
        for (int c = 0; c < NC; ++c) {
            const int64_t m = NC * i + c;
            const double pi = first ? z[m] : z[m] + beta * p[m];
            const double qi = first ? az[c] : az[c] + beta * q[m];
            p[m] = pi;
            q[m] = qi;
            a += pi * qi;
        }
    }
    const double t = block_reduce(a, sh);
    if (threadIdx.x == 0) part_pq[blockIdx.x] = t;
}

// Step 2: alpha = rz_t / p.q (every block adds the partials); x += alpha p,
// r -= alpha q, z = D^-1 r; per-block r.z and r.r (dinv null: r.r only, the
// multigrid V-cycle that follows writes z and r.z).
template <int NC>
__global__ void __launch_bounds__(FB) k_cg_update(int64_t k, double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ z, const double* __restrict__ p,
                                                  const double* __restrict__ q, const double* __restrict__ dinv,
                                                  int64_t n, const double* __restrict__ part_pq, int cnt_pq,
                                                  double* part2, const CgState* __restrict__ S) {
    __shared__ double sh[8];
    if (S->done) return;
    const double alpha = S->rz[k & 1] / sum_partials(part_pq, cnt_pq, 1, 0, sh);
    double a0 = 0.0, a1 = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < NC; ++c) {
            const int64_t m = NC * i + c;
            x[m] += alpha * p[m];
            const double ri = r[m] - alpha * q[m];
            r[m] = ri;
            if (dinv) {  // Jacobi; without it the V-cycle writes z and r.z
                const double zi = dinv[i] * ri;
                z[m] = zi;
                a0 += ri * zi;
            }
            a1 += ri * ri;
        }
    }
    const double t0 = block_reduce(a0, sh), t1 = block_reduce(a1, sh);
    if (threadIdx.x == 0) {
        if (dinv) part2[2 * blockIdx.x] = t0;
        part2[2 * blockIdx.x + 1] = t1;
    }
}

// ---------------------------------------------------------------------------
// Multigrid V-cycle (the preconditioner z = M r of the stiffness solves).
// Level l holds A_l, the prolongation P_l from level l+1 and R_l = P_l^T in
// CSR, and 1/diag(A_l).  Going down, b_l -> x_l = w D^-1 b_l, r_l = b_l -
// A_l x_l, b_{l+1} = R_l r_l; the coarsest level applies its dense inverse;
// going up, x_l += P_l y_{l+1}, y_l = x_l + w D^-1 (b_l - A_l x_l).
// A row is handled by a group of G lanes (G a power of two near the
// matrix's mean row length): lane g of the group takes entries g, g + G, ...
// of the row, contiguous across the group, and the group adds its partial
// sums with lane shuffles.  With S given, every kernel returns at once once
// the solve has converged.
// ---------------------------------------------------------------------------
// The group's sum of v[k] * x[ci[k]] * (SCALE ? s[ci[k]] : 1) over row i.
template <int G, bool SCALE>
__device__ inline double group_row(const CsrDev& M, int64_t i, int sub, const double* __restrict__ x,
                                   const double* __restrict__ sc) {
    double s = 0.0;
    for (int32_t k = M.rp[i] + sub; k < M.rp[i + 1]; k += G) {
        const int32_t j = M.ci[k];
        s += SCALE ? M.v[k] * (sc[j] * x[j]) : M.v[k] * x[j];
    }
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
    return s;
}

// MODE 0 (pre-smooth): x = w D^-1 b, r = b - A x (out = r, aux = x)
// MODE 1 (apply): out = M in
// MODE 2 (prolong): out += M in
template <int G, int MODE>
__global__ void __launch_bounds__(FB) k_amg_rows(const CsrDev M, const double* __restrict__ dinv, double omega,
                                                 const double* __restrict__ in, double* __restrict__ out,
                                                 double* __restrict__ aux, const CgState* __restrict__ S) {
    if (S && S->done) return;
    const int sub = threadIdx.x & (G - 1);
    const int64_t step = ((int64_t)gridDim.x * blockDim.x) / G;
    for (int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; i < M.rows; i += step) {
        if (MODE == 0) {
            const double ax = omega * group_row<G, true>(M, i, sub, in, dinv);
            if (sub == 0) {
                aux[i] = omega * dinv[i] * in[i];
                out[i] = in[i] - ax;
            }
        } else {
            const double t = group_row<G, false>(M, i, sub, in, nullptr);
            if (sub == 0) out[i] = MODE == 2 ? out[i] + t : t;
        }
    }
}

// Post-smooth: y = x + w D^-1 (b - A x); with part, per-block b.y at
// part[stride * block + off].
template <int G>
__global__ void __launch_bounds__(FB) k_amg_post(const CsrDev M, const double* __restrict__ dinv, double omega,
                                                 const double* __restrict__ b, const double* __restrict__ x,
                                                 double* __restrict__ y, double* part, int stride, int off,
                                                 const CgState* __restrict__ S) {
    __shared__ double sh[8];
    if (S && S->done) return;
    const int sub = threadIdx.x & (G - 1);
    const int64_t step = ((int64_t)gridDim.x * blockDim.x) / G;
    double a = 0.0;
    for (int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; i < M.rows; i += step) {
        const double ax = group_row<G, false>(M, i, sub, x, nullptr);
        if (sub == 0) {
            const double yi = x[i] + omega * dinv[i] * (b[i] - ax);
            y[i] = yi;
            a += b[i] * yi;
        }
    }
    if (part) {
        const double t = block_reduce(a, sh);
        if (threadIdx.x == 0) part[(int64_t)stride * blockIdx.x + off] = t;
    }
}

// y = C b for the dense coarsest inverse C (n x n, row-major): one wavefront per row.
__global__ void __launch_bounds__(FB) k_amg_dense(const double* __restrict__ C, int64_t n, const double* __restrict__ b,
                                                  double* __restrict__ y, const CgState* __restrict__ S) {
    if (S && S->done) return;
    const int lane = threadIdx.x & 63;
    const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    if (row >= n) return;  // wavefront-uniform
    double s = 0.0;
    for (int64_t j = lane; j < n; j += 64) s += C[row * n + j] * b[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) y[row] = s;
}

// Object charges from a potential (the Charge form eps0 grad(phi).n ds over
// each object's facets, Charge.ufl): a fixed linear functional per object,
// sum over its entries of w * phi[dof].  One block per object.
__global__ void __launch_bounds__(FB) k_image_charge(const int64_t* __restrict__ off, const int32_t* __restrict__ dof,
                                                     const double* __restrict__ w, const double* __restrict__ phi,
                                                     double* __restrict__ out) {
    __shared__ double sh[8];
    double s = 0.0;
    for (int64_t k = off[blockIdx.x] + threadIdx.x; k < off[blockIdx.x + 1]; k += blockDim.x) s += w[k] * phi[dof[k]];
    const double t = block_reduce(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// End of a batch of nb iterations.
__global__ void k_cg_batch_end(CgState* S, int nb) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !S->done) S->base += nb;
}

template <typename T>
hipError_t dmalloc(T** p, size_t count) {
    return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

}  // namespace

struct FieldSolver {
    int device = 0;
    hipStream_t st = nullptr;
    int64_t n = 0, nnz = 0;
    int nobj = 0;
    bool ext = false, nullspace = false;
    int maxit = 20000;
    bool warm = false;  // start from the previous solution (the reference starts from zero)
    int grid = 1;  // one row / element per thread, grid-stride
    int w = 0;     // ELL width
    std::vector<int32_t> dir_kind;  // per dof: -1 free, 0 exterior, 1+j object j
    int32_t* d_kind = nullptr;      // the same on the device
    double* d_objpot = nullptr;     // object potentials of the current solve
    double* h_objpot = nullptr;     // page-locked staging for them
    CgState* h_state = nullptr;     // page-locked read-back of the CG state
    // device
    int32_t* col = nullptr;  // ELL columns (w * n)
    double *kfull = nullptr, *mass = nullptr, *grad[3] = {nullptr, nullptr, nullptr};
    double *dinv_k = nullptr, *dinv_m = nullptr;
    uint8_t* isd = nullptr;
    double *b = nullptr, *t = nullptr, *phid = nullptr, *rhs = nullptr;
    double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
    double *phi = nullptr, *E = nullptr, *part = nullptr;
    CgState* S = nullptr;
    bool own_stream = false;
    // object charge functionals (CSR over objects) and the circuit
    int64_t* q_off = nullptr;
    int32_t* q_dof = nullptr;
    double* q_w = nullptr;
    double* q_out = nullptr;
    double* zero = nullptr;  // a zero charge density (Laplace solves)
    bool has_circuit = false;
    double dt = 0.0;
    std::vector<CircuitSource> vsrc, isrc;
    std::vector<std::vector<int>> groups, fixed_voltage;
    std::vector<double> invC;        // n x n, row-major
    std::vector<double> inv_circuit; // 2n x 2n
    std::vector<double> circuit_vector;
    // the CG operators: the eliminated stiffness in CSR without its explicit
    // zeros (gk lanes per row), the mass matrix in ELL
    CsrDev kcsr{};
    EllDev mell{};  // the mass matrix's ELL arrays (one lane per row measured faster there)
    int gk = 1;
    int32_t* csr_i[4] = {nullptr, nullptr, nullptr, nullptr};
    double* csr_d[2] = {nullptr, nullptr};
    // multigrid preconditioner of the stiffness solves (levels: see k_amg_rows)
    struct AmgDev {
        int64_t n = 0, nc = 0;  // rows of this level and of the next
        CsrDev A{}, P{}, R{};
        int ga = 1, gp = 1, gr = 1;  // lanes per row of each
        int32_t* mem_i[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        double* mem_d[3] = {nullptr, nullptr, nullptr};
        double* dinv = nullptr;
        double omega = 0.0;
        double *b = nullptr, *x = nullptr, *r = nullptr, *y = nullptr;  // level 0: b, y are the CG r, z
    };
    std::vector<AmgDev> amg;
    double *coarse_inv = nullptr, *cb = nullptr, *cy = nullptr;
    int64_t ncoarse = 0;
    bool amg_ok = false;   // a hierarchy was built
    bool use_amg = false;  // the stiffness solves use it
    std::string amg_note;  // why not, when it was not built
    // one captured batch of iterations per operator (stiffness, mass)
    struct Batch {
        const double* av;
        bool amg;
        int nc;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<Batch> batches;
};

static void free_all(FieldSolver* F) {
    for (auto& b : F->batches) {
        hipGraphExecDestroy(b.exec);
        hipGraphDestroy(b.graph);
    }
    F->batches.clear();
    for (size_t l = 0; l < F->amg.size(); ++l) {
        auto& L = F->amg[l];
        for (auto* m : L.mem_i) hipFree(m);
        for (auto* m : L.mem_d) hipFree(m);
        if (l > 0) {  // level 0's b and y are the CG vectors
            hipFree(L.b);
            hipFree(L.y);
        }
        hipFree(L.dinv);
        hipFree(L.x);
        hipFree(L.r);
    }
    F->amg.clear();
    for (auto* m : F->csr_i) hipFree(m);
    for (auto* m : F->csr_d) hipFree(m);
    hipFree(F->coarse_inv);
    hipFree(F->cb);
    hipFree(F->cy);
    hipFree(F->col);
    hipFree(F->kfull);
    hipFree(F->mass);
    for (auto& g : F->grad) hipFree(g);
    hipFree(F->dinv_k);
    hipFree(F->dinv_m);
    hipFree(F->isd);
    hipFree(F->b);
    hipFree(F->t);
    hipFree(F->phid);
    hipFree(F->d_kind);
    hipFree(F->d_objpot);
    if (F->h_objpot) hipHostFree(F->h_objpot);
    if (F->h_state) hipHostFree(F->h_state);
    hipFree(F->rhs);
    hipFree(F->x);
    hipFree(F->r);
    hipFree(F->z);
    hipFree(F->p);
    hipFree(F->q);
    hipFree(F->phi);
    hipFree(F->E);
    hipFree(F->part);
    hipFree(F->S);
    hipFree(F->q_off);
    hipFree(F->q_dof);
    hipFree(F->q_w);
    hipFree(F->q_out);
    hipFree(F->zero);
    if (F->own_stream && F->st) hipStreamDestroy(F->st);
}

// Barycentric gradients and volume of a tetrahedron: with J = [P1-P0, P2-P0,
// P3-P0] (columns), lambda_k (k = 1..3) has gradient row k-1 of J^-1 and
// lambda_0 minus their sum.
static double tet_gradients(const double* vc, double g[4][3]) {
    double J[3][3];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) J[r][k] = vc[3 * (k + 1) + r] - vc[r];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    // rows of J^-1 = cofactor columns / det
    double inv[3][3];
    inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) g[k + 1][d] = inv[k][d];
    for (int d = 0; d < 3; ++d) g[0][d] = -(g[1][d] + g[2][d] + g[3][d]);
    return fabs(det) / 6.0;
}

// Lanes per row for a CSR matrix: the largest power of two not above its
// mean row length, at most 64.
static int lanes_per_row(const Csr& A) {
    const double mean = A.nrow ? (double)A.ci.size() / (double)A.nrow : 1.0;
    int g = 1;
    while (g < 64 && 2.0 * g <= mean) g *= 2;
    return g;
}

static hipError_t upload_amg(FieldSolver* F, const AmgHierarchy& H) {
    hipError_t e = hipSuccess;
    auto up = [&](auto** d, const auto& h) {
        if (e != hipSuccess) return;
        e = dmalloc(d, h.size());
        if (e == hipSuccess && !h.empty())
            e = hipMemcpy(*d, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice);
    };
    auto vec = [&](double** d, int64_t len) {
        if (e == hipSuccess) e = dmalloc(d, (size_t)len);
    };
    auto csr = [&](const Csr& h, CsrDev& d, int32_t** ip, int32_t** ic, double** dv) {
        if (h.ci.size() >= ((size_t)1 << 31) && e == hipSuccess) e = hipErrorInvalidValue;  // 32-bit row pointers
        std::vector<int32_t> rp(h.rp.begin(), h.rp.end());
        up(ip, rp);
        up(ic, h.ci);
        up(dv, h.v);
        d.rp = *ip;
        d.ci = *ic;
        d.v = *dv;
        d.rows = h.nrow;
    };
    const size_t L = H.levels.size();
    F->amg.resize(L);
    for (size_t l = 0; l < L; ++l) {
        const AmgLevel& h = H.levels[l];
        FieldSolver::AmgDev& d = F->amg[l];
        d.n = h.A.nrow;
        d.nc = h.P.ncol;
        d.omega = h.omega;
        csr(h.A, d.A, &d.mem_i[0], &d.mem_i[1], &d.mem_d[0]);
        csr(h.P, d.P, &d.mem_i[2], &d.mem_i[3], &d.mem_d[1]);
        csr(h.R, d.R, &d.mem_i[4], &d.mem_i[5], &d.mem_d[2]);
        d.ga = lanes_per_row(h.A);
        d.gp = lanes_per_row(h.P);
        d.gr = lanes_per_row(h.R);
        up(&d.dinv, h.dinv);
        vec(&d.x, d.n);
        vec(&d.r, d.n);
        if (l == 0) {
            d.b = F->r;
            d.y = F->z;
        } else {
            vec(&d.b, d.n);
            vec(&d.y, d.n);
        }
    }
    F->ncoarse = H.ncoarse;
    up(&F->coarse_inv, H.coarse_inv);
    vec(&F->cb, H.ncoarse);
    vec(&F->cy, H.ncoarse);
    return e;
}

FieldSolver* field_create(const HostMesh& m, const MeshTopo& t, double eps0, bool dirichlet_exterior, int n_objects,
                          bool remove_null_space, int device, hipStream_t stream, std::string& err) {
    if (n_objects < 0 || n_objects > 61) {
        err = "number of objects out of range";
        return nullptr;
    }
    if (remove_null_space && (dirichlet_exterior || n_objects > 0)) {
        err = "remove_null_space needs a problem without Dirichlet conditions";
        return nullptr;
    }
    const int64_t n = m.ndof, nc = m.nc;
    // sparsity: the dofs sharing a cell
    std::vector<std::vector<int32_t>> adj((size_t)n);
    std::vector<int32_t> cd((size_t)nc * 4);
    for (int64_t c = 0; c < nc; ++c)
        for (int k = 0; k < 4; ++k) cd[c * 4 + k] = m.vdof[m.cells[c * 4 + k]];
    for (int64_t c = 0; c < nc; ++c)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) adj[cd[c * 4 + a]].push_back(cd[c * 4 + b]);
    std::vector<int32_t> rp((size_t)n + 1, 0), ci;
    for (int64_t i = 0; i < n; ++i) {
        auto& v = adj[i];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        rp[i + 1] = rp[i] + (int32_t)v.size();
    }
    const int64_t nnz = rp[n];
    if (nnz >= (int64_t(1) << 31)) {
        err = "matrix too large";
        return nullptr;
    }
    ci.resize((size_t)nnz);
    for (int64_t i = 0; i < n; ++i) std::copy(adj[i].begin(), adj[i].end(), ci.begin() + rp[i]);
    adj.clear();
    auto slot = [&](int32_t i, int32_t j) {
        return (int64_t)(std::lower_bound(ci.begin() + rp[i], ci.begin() + rp[i + 1], j) - ci.begin());
    };
    std::vector<double> K((size_t)nnz, 0.0), Mv((size_t)nnz, 0.0), B[3];
    for (auto& v : B) v.assign((size_t)nnz, 0.0);
    for (int64_t c = 0; c < nc; ++c) {
        double g[4][3];
        const double V = tet_gradients(&t.vcoords[c * 12], g);
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                const int64_t k = slot(cd[c * 4 + a], cd[c * 4 + b]);
                K[k] += eps0 * V * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
                Mv[k] += V * (a == b ? 2.0 : 1.0) / 20.0;
                for (int d = 0; d < 3; ++d) B[d][k] += -(V / 4.0) * g[b][d];
            }
    }
    // Dirichlet dofs: the vertices of boundary facets (adj = -1 exterior,
    // -(2 + j) object j); objects are applied after the exterior.
    std::vector<int32_t> kind((size_t)n, -1);
    static const int fv[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t c = 0; c < nc; ++c)
            for (int f = 0; f < 4; ++f) {
                const int64_t a = t.adj[c * 4 + f];
                int32_t kd = -1;
                if (pass == 0 && a == -1 && dirichlet_exterior) kd = 0;
                if (pass == 1 && a <= -2 && -a - 2 < n_objects) kd = (int32_t)(-a - 1);
                if (kd < 0) continue;
                for (int v = 0; v < 3; ++v) kind[cd[c * 4 + fv[f][v]]] = kd;
            }
    // Charge functionals: on each facet of object j (cell c, local facet f,
    // outward unit normal n_f away from vertex f, area A_f) the form gives
    // sum_k eps0 A_f (g_k . n_f) phi_k over the cell's four dofs.
    std::vector<std::vector<std::pair<int32_t, double>>> qf((size_t)n_objects);
    for (int64_t c = 0; c < nc; ++c)
        for (int f = 0; f < 4; ++f) {
            const int64_t a = t.adj[c * 4 + f];
            if (a > -2 || -a - 2 >= n_objects) continue;
            const double* vc = &t.vcoords[c * 12];
            const double* P0 = vc + 3 * fv[f][0];
            const double* P1 = vc + 3 * fv[f][1];
            const double* P2 = vc + 3 * fv[f][2];
            double e1[3], e2[3], nn[3];
            for (int d = 0; d < 3; ++d) {
                e1[d] = P1[d] - P0[d];
                e2[d] = P2[d] - P0[d];
            }
            nn[0] = e1[1] * e2[2] - e1[2] * e2[1];
            nn[1] = e1[2] * e2[0] - e1[0] * e2[2];
            nn[2] = e1[0] * e2[1] - e1[1] * e2[0];
            const double len = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            const double area = 0.5 * len;
            double side = 0.0;  // orientation: away from the opposite vertex
            for (int d = 0; d < 3; ++d) side += nn[d] * (vc[3 * f + d] - P0[d]);
            const double sg = side > 0 ? -1.0 : 1.0;
            for (int d = 0; d < 3; ++d) nn[d] *= sg / len;
            double g[4][3];
            tet_gradients(vc, g);
            for (int k = 0; k < 4; ++k)
                qf[-a - 2].push_back({cd[c * 4 + k], eps0 * area * (g[k][0] * nn[0] + g[k][1] * nn[1] + g[k][2] * nn[2])});
        }
    std::vector<int64_t> qoff(1, 0);
    std::vector<int32_t> qdof;
    std::vector<double> qw;
    for (auto& l : qf) {
        std::sort(l.begin(), l.end(), [](const std::pair<int32_t, double>& x, const std::pair<int32_t, double>& y) {
            return x.first < y.first;
        });
        for (size_t k = 0; k < l.size();) {  // merge equal dofs in facet order
            double sum = 0.0;
            size_t e = k;
            for (; e < l.size() && l[e].first == l[k].first; ++e) sum += l[e].second;
            qdof.push_back(l[k].first);
            qw.push_back(sum);
            k = e;
        }
        qoff.push_back((int64_t)qdof.size());
    }
    std::vector<double> Kr(K), dk((size_t)n), dm((size_t)n);
    std::vector<uint8_t> isd((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        isd[i] = kind[i] >= 0;
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j) {
            if (isd[i]) Kr[j] = ci[j] == i ? 1.0 : 0.0;
            else if (kind[ci[j]] >= 0) Kr[j] = 0.0;
            if (ci[j] == i) {
                dk[i] = 1.0 / Kr[j];
                dm[i] = 1.0 / Mv[j];
            }
        }
    }
    // ELL layout of the CSR rows
    int w = 0;
    for (int64_t i = 0; i < n; ++i) w = std::max(w, rp[i + 1] - rp[i]);
    std::vector<int32_t> ecol((size_t)w * n);
    auto ell = [&](const std::vector<double>& v) {
        std::vector<double> o((size_t)w * n, 0.0);
        for (int64_t i = 0; i < n; ++i)
            for (int32_t j = rp[i]; j < rp[i + 1]; ++j) o[(size_t)(j - rp[i]) * n + i] = v[j];
        return o;
    };
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < w; ++k) {
            const int32_t j = rp[i] + k;
            ecol[(size_t)k * n + i] = j < rp[i + 1] ? ci[j] : (int32_t)i;
        }
    FieldSolver* F = new FieldSolver();
    F->w = w;
    F->device = device;
    F->st = stream;
    F->n = n;
    F->nnz = nnz;
    F->nobj = n_objects;
    F->ext = dirichlet_exterior;
    F->nullspace = remove_null_space;
    F->dir_kind = kind;
    F->grid = (int)std::min<int64_t>(MAX_BLOCKS, std::max<int64_t>(1, (n + FB - 1) / FB));
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess && !F->st) {
        // a stream of its own (default priority): the iteration batches are
        // captured into graphs, which the legacy default stream cannot do
        e = hipStreamCreateWithFlags(&F->st, hipStreamNonBlocking);
        F->own_stream = e == hipSuccess;
    }
    auto up = [&](auto** d, const auto& h) {
        if (e != hipSuccess) return;
        e = dmalloc(d, h.size());
        if (e == hipSuccess && !h.empty())
            e = hipMemcpy(*d, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice);
    };
    up(&F->col, ecol);
    up(&F->kfull, ell(K));
    up(&F->mass, ell(Mv));
    for (int d = 0; d < 3; ++d) up(&F->grad[d], ell(B[d]));
    up(&F->q_off, qoff);
    up(&F->q_dof, qdof);
    up(&F->q_w, qw);
    up(&F->dinv_k, dk);
    up(&F->dinv_m, dm);
    up(&F->isd, isd);
    up(&F->d_kind, kind);
    if (e == hipSuccess) e = dmalloc(&F->d_objpot, (size_t)std::max(n_objects, 1));
    if (e == hipSuccess) e = hipHostMalloc((void**)&F->h_objpot, (size_t)std::max(n_objects, 1) * 8);
    if (e == hipSuccess) e = hipHostMalloc((void**)&F->h_state, sizeof(CgState));
    double** vecs[] = {&F->b, &F->t, &F->phid, &F->phi};
    for (double** v : vecs)
        if (e == hipSuccess) e = dmalloc(v, (size_t)n);
    double** vecs3[] = {&F->rhs, &F->x, &F->r, &F->z, &F->p, &F->q};  // also the E solve's three components
    for (double** v : vecs3)
        if (e == hipSuccess) e = dmalloc(v, (size_t)n * 3);
    if (e == hipSuccess) e = dmalloc(&F->E, (size_t)n * 3);
    if (e == hipSuccess) e = dmalloc(&F->part, (size_t)F->grid * 5);
    if (e == hipSuccess) e = dmalloc(&F->S, 1);
    if (e == hipSuccess) e = dmalloc(&F->q_out, (size_t)std::max(n_objects, 1));
    if (e == hipSuccess) e = dmalloc(&F->zero, (size_t)n);
    if (e == hipSuccess) e = hipMemset(F->zero, 0, n * 8);
    if (e == hipSuccess) e = hipMemset(F->S, 0, sizeof(CgState));
    if (e == hipSuccess) e = hipMemset(F->phi, 0, n * 8);
    if (e == hipSuccess) e = hipMemset(F->E, 0, n * 24);
    // Multigrid hierarchy of the eliminated stiffness operator (its explicit
    // zeros dropped); the Dirichlet rows are identity rows, left out of the
    // aggregates.  Without one the stiffness solves stay Jacobi-preconditioned.
    Csr A;  // the eliminated stiffness without its explicit zeros
    A.nrow = A.ncol = n;
    A.rp.assign((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j)
            if (Kr[j] != 0.0 || ci[j] == i) {
                A.ci.push_back(ci[j]);
                A.v.push_back(Kr[j]);
            }
        A.rp[i + 1] = (int64_t)A.ci.size();
    }
    auto upcsr = [&](const Csr& h, CsrDev& d, int32_t** ip, int32_t** ic, double** dv) {
        up(ip, std::vector<int32_t>(h.rp.begin(), h.rp.end()));
        up(ic, h.ci);
        up(dv, h.v);
        d.rp = *ip;
        d.ci = *ic;
        d.v = *dv;
        d.rows = h.nrow;
    };
    upcsr(A, F->kcsr, &F->csr_i[0], &F->csr_i[1], &F->csr_d[0]);
    F->gk = std::min(lanes_per_row(A), 16);
    F->mell = EllDev{F->col, F->mass, F->w, n};
    if (e == hipSuccess) {
        AmgHierarchy H;
        std::string note;
        if (amg_build(A, isd, remove_null_space, H, note) && !H.levels.empty()) {
            e = upload_amg(F, H);
            F->amg_ok = F->use_amg = e == hipSuccess;
        } else {
            F->amg_note = H.levels.empty() && note.empty() ? "operator small enough for Jacobi" : note;
        }
    }
    if (e != hipSuccess) {
        err = std::string("device allocation: ") + hipGetErrorString(e);
        free_all(F);
        delete F;
        return nullptr;
    }
    return F;
}

void field_destroy(FieldSolver* F) {
    if (!F) return;
    hipSetDevice(F->device);
    hipStreamSynchronize(F->st);
    free_all(F);
    delete F;
}

int64_t field_ndof(const FieldSolver* F) { return F->n; }
int field_num_objects(const FieldSolver* F) { return F->nobj; }
double* field_phi(FieldSolver* F) { return F->phi; }
double* field_E(FieldSolver* F) { return F->E; }
void field_set_max_iterations(FieldSolver* F, int maxit) { F->maxit = std::max(1, maxit); }
void field_set_warm_start(FieldSolver* F, bool warm) { F->warm = warm; }
bool field_set_preconditioner(FieldSolver* F, int kind, std::string& err) {
    if (kind != 0 && kind != 1) {
        err = "preconditioner: 0 (Jacobi) or 1 (multigrid)";
        return false;
    }
    if (kind == 1 && !F->amg_ok) {
        err = "no multigrid hierarchy for this operator: " + F->amg_note;
        return false;
    }
    F->use_amg = kind == 1;
    return true;
}
int field_preconditioner(const FieldSolver* F, int* levels, int64_t* coarse) {
    if (levels) *levels = (int)F->amg.size();
    if (coarse) *coarse = F->ncoarse;
    return F->use_amg ? 1 : 0;
}

#define FCHK(x)                                                      \
    do {                                                             \
        hipError_t _e = (x);                                         \
        if (_e != hipSuccess) {                                      \
            err = std::string(#x ": ") + hipGetErrorString(_e);      \
            return false;                                            \
        }                                                            \
    } while (0)

template <int MODE>
static void launch_rows(int g, dim3 grid, hipStream_t st, const CsrDev& M, const double* dinv, double omega,
                        const double* in, double* out, double* aux, const CgState* S) {
    const dim3 blk(FB);
    switch (g) {
#define PUNC_AMG_G(G) \
    case G: hipLaunchKernelGGL((k_amg_rows<G, MODE>), grid, blk, 0, st, M, dinv, omega, in, out, aux, S); break;
        PUNC_AMG_G(1) PUNC_AMG_G(2) PUNC_AMG_G(4) PUNC_AMG_G(8) PUNC_AMG_G(16) PUNC_AMG_G(32) PUNC_AMG_G(64)
#undef PUNC_AMG_G
    }
}

static void launch_post(int g, dim3 grid, hipStream_t st, const CsrDev& M, const double* dinv, double omega,
                        const double* b, const double* x, double* y, double* part, int stride, int off,
                        const CgState* S) {
    const dim3 blk(FB);
    switch (g) {
#define PUNC_AMG_G(G) \
    case G: hipLaunchKernelGGL((k_amg_post<G>), grid, blk, 0, st, M, dinv, omega, b, x, y, part, stride, off, S); break;
        PUNC_AMG_G(1) PUNC_AMG_G(2) PUNC_AMG_G(4) PUNC_AMG_G(8) PUNC_AMG_G(16) PUNC_AMG_G(32) PUNC_AMG_G(64)
#undef PUNC_AMG_G
    }
}

// Blocks for rows x lanes threads (grid-stride beyond 4 * MAX_BLOCKS).
static dim3 grid_rows(int64_t rows, int lanes) {
    return dim3((unsigned)std::min<int64_t>(4 * MAX_BLOCKS, std::max<int64_t>(1, (rows * lanes + FB - 1) / FB)));
}

// z = M r by one V-cycle; per-block r.z into part[stride * block + off]
// (level 0's post-smooth runs on F->grid blocks, the grid of the CG kernels,
// whose partials the next CG kernel adds).
static void enqueue_vcycle(FieldSolver* F, const CgState* S, double* part, int stride, int off) {
    const size_t L = F->amg.size();
    for (size_t l = 0; l < L; ++l) {
        const auto& d = F->amg[l];
        double* bnext = l + 1 < L ? F->amg[l + 1].b : F->cb;
        launch_rows<0>(d.ga, grid_rows(d.n, d.ga), F->st, d.A, d.dinv, d.omega, d.b, d.r, d.x, S);
        launch_rows<1>(d.gr, grid_rows(d.nc, d.gr), F->st, d.R, nullptr, 0.0, d.r, bnext, nullptr, S);
    }
    hipLaunchKernelGGL(k_amg_dense, dim3((unsigned)((F->ncoarse + 3) / 4)), dim3(FB), 0, F->st, F->coarse_inv,
                       F->ncoarse, F->cb, F->cy, S);
    for (size_t l = L; l-- > 0;) {
        const auto& d = F->amg[l];
        const double* ynext = l + 1 < L ? F->amg[l + 1].y : F->cy;
        launch_rows<2>(d.gp, grid_rows(d.n, d.gp), F->st, d.P, nullptr, 0.0, ynext, d.x, nullptr, S);
        launch_post(d.ga, l == 0 ? dim3(F->grid) : grid_rows(d.n, d.ga), F->st, d.A, d.dinv, d.omega, d.b, d.x, d.y,
                    l == 0 ? part : nullptr, stride, off, S);
    }
}

// A CG operator: the CSR stiffness (G lanes per row) or the ELL mass matrix.
struct CgOp {
    const CsrDev* csr;
    const EllDev* ell;
    int G;
    const double* key() const { return csr ? csr->v : ell->v; }
};

template <int NC, int G, typename Op>
static void launch_dir_update(FieldSolver* F, int64_t k, const Op& A, const double* dinv, double* part_pq,
                              double* part2) {
    const dim3 g(F->grid), blk(FB);
    hipLaunchKernelGGL((k_cg_dir_spmv<NC, G, Op>), g, blk, 0, F->st, k, A, F->z, F->p, F->q, part2, F->grid, part_pq,
                       F->S);
    hipLaunchKernelGGL(k_cg_update<NC>, g, blk, 0, F->st, k, F->x, F->r, F->z, F->p, F->q, dinv, F->n, part_pq,
                       F->grid, part2, F->S);
}

template <int NC, int G, typename Op>
static void launch_init(FieldSolver* F, const Op& A, const double* dinv, const double* b) {
    hipLaunchKernelGGL((k_cg_init<NC, G, Op>), dim3(F->grid), dim3(FB), 0, F->st, A, dinv, b, F->x, F->r, F->z, F->p,
                       F->part);
}

// Calls fn(components, lanes, operator) with the first two as integral
// constants: ELL with one lane per row, CSR with op.G lanes.
template <int NC, typename Fn>
static void dispatch_nc(const CgOp& op, Fn&& fn) {
    using std::integral_constant;
    if (op.ell) {
        fn(integral_constant<int, NC>(), integral_constant<int, 1>(), *op.ell);
        return;
    }
    switch (op.G) {
        case 1: fn(integral_constant<int, NC>(), integral_constant<int, 1>(), *op.csr); break;
        case 2: fn(integral_constant<int, NC>(), integral_constant<int, 2>(), *op.csr); break;
        case 4: fn(integral_constant<int, NC>(), integral_constant<int, 4>(), *op.csr); break;
        case 8: fn(integral_constant<int, NC>(), integral_constant<int, 8>(), *op.csr); break;
        default: fn(integral_constant<int, NC>(), integral_constant<int, 16>(), *op.csr); break;
    }
}

template <typename Fn>
static void dispatch_op(const CgOp& op, int nc, Fn&& fn) {
    if (nc == 3)
        dispatch_nc<3>(op, fn);
    else
        dispatch_nc<1>(op, fn);
}

// dinv null: the multigrid preconditioner.  nc: interleaved components (1, 3).
static void enqueue_batch(FieldSolver* F, const CgOp& op, const double* dinv, int nc) {
    double* part_pq = F->part;
    double* part2 = F->part + 3 * (size_t)F->grid;
    const int nb = dinv ? BATCH : BATCH_AMG;
    for (int k = 0; k < nb; ++k) {
        dispatch_op(op, nc, [&](auto nc_c, auto g_c, const auto& A) {
            launch_dir_update<decltype(nc_c)::value, decltype(g_c)::value>(F, (int64_t)k, A, dinv, part_pq, part2);
        });
        if (!dinv) enqueue_vcycle(F, F->S, part2, 2, 0);
    }
    hipLaunchKernelGGL(k_cg_batch_end, dim3(1), dim3(64), 0, F->st, F->S, nb);
}

// The captured batch for operator op (captured on first use).
static bool batch_exec(FieldSolver* F, const CgOp& op, const double* dinv, int nc, hipGraphExec_t* out,
                       std::string& err) {
    for (auto& b : F->batches)
        if (b.av == op.key() && b.amg == (dinv == nullptr) && b.nc == nc) {
            *out = b.exec;
            return true;
        }
    FieldSolver::Batch b;
    b.av = op.key();
    b.amg = dinv == nullptr;
    b.nc = nc;
    FCHK(hipStreamBeginCapture(F->st, hipStreamCaptureModeThreadLocal));
    enqueue_batch(F, op, dinv, nc);
    hipError_t e = hipStreamEndCapture(F->st, &b.graph);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("capturing the CG batch: ") + hipGetErrorString(e);
        return false;
    }
    e = hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(b.graph);
        err = std::string("instantiating the CG batch: ") + hipGetErrorString(e);
        return false;
    }
    F->batches.push_back(b);
    *out = b.exec;
    return true;
}

// Preconditioned CG on operator op for rhs b from F->x (in place): the
// multigrid V-cycle for the stiffness operator when its hierarchy is in use,
// Jacobi (dinv) otherwise.  nc = 3: the vectors hold three interleaved
// components, solved as one CG on the block-diagonal system (Jacobi only).
static bool run_cg(FieldSolver* F, const CgOp& op, const double* dinv, const double* b, int* iterations,
                   std::string& err, int nc = 1) {
    if (op.csr == &F->kcsr && F->use_amg && nc == 1) dinv = nullptr;
    hipGraphExec_t exec;
    if (!batch_exec(F, op, dinv, nc, &exec, err)) return false;
    dispatch_op(op, nc, [&](auto nc_c, auto g_c, const auto& A) {
        launch_init<decltype(nc_c)::value, decltype(g_c)::value>(F, A, dinv, b);
    });
    if (!dinv) enqueue_vcycle(F, nullptr, F->part, 3, 0);
    hipLaunchKernelGGL(k_cg_init_finish, dim3(1), dim3(FB), 0, F->st, F->part, F->grid, 1e-12, 1e-14, F->S);
    FCHK(hipGetLastError());
    CgState h;
    for (;;) {
        FCHK(hipGraphLaunch(exec, F->st));
        FCHK(hipMemcpyAsync(F->h_state, F->S, sizeof(CgState), hipMemcpyDeviceToHost, F->st));
        FCHK(hipStreamSynchronize(F->st));
        h = *F->h_state;
        if (h.done || h.base >= F->maxit) break;
    }
    if (iterations) *iterations = (int)(h.done ? h.it : h.base);
    if (!h.done) {
        err = "conjugate gradients did not converge in " + std::to_string(h.base) + " iterations";
        return false;
    }
    return true;
}

bool field_poisson(FieldSolver* F, const double* d_rho, const double* ext_values, const double* object_potentials,
                   int* iterations, std::string& err) {
    FCHK(hipSetDevice(F->device));
    const int64_t n = F->n;
    const dim3 g(F->grid), blk(FB);
    // phi_D on the Dirichlet dofs (exterior values, object potentials), filled
    // on the device: the caller's exterior values are staged in rhs (written
    // only later, by k_rhs_dirichlet), the object potentials through
    // page-locked memory (the stream is synchronized before this returns, so
    // the staging buffer is free again for the next call)
    const double* d_ext = nullptr;
    if (ext_values) {
        FCHK(hipMemcpyAsync(F->rhs, ext_values, n * 8, hipMemcpyHostToDevice, F->st));
        d_ext = F->rhs;
    }
    const double* d_obj = nullptr;
    if (object_potentials && F->nobj > 0) {
        memcpy(F->h_objpot, object_potentials, (size_t)F->nobj * 8);
        FCHK(hipMemcpyAsync(F->d_objpot, F->h_objpot, (size_t)F->nobj * 8, hipMemcpyHostToDevice, F->st));
        d_obj = F->d_objpot;
    }
    hipLaunchKernelGGL(k_dirichlet_values, g, blk, 0, F->st, F->d_kind, d_ext, d_obj, F->phid, n);
    hipLaunchKernelGGL(k_spmv, g, blk, 0, F->st, F->col, F->w, F->mass, d_rho, F->b, n);     // b = M rho
    hipLaunchKernelGGL(k_spmv, g, blk, 0, F->st, F->col, F->w, F->kfull, F->phid, F->t, n);  // K phi_D
    hipLaunchKernelGGL(k_rhs_dirichlet, g, blk, 0, F->st, F->b, F->t, F->phid, F->isd,
                       F->warm ? (const double*)F->phi : nullptr, F->rhs, F->x, n);
    FCHK(hipGetLastError());
    if (F->nullspace) {
        hipLaunchKernelGGL(k_partial_sum, g, blk, 0, F->st, F->rhs, n, F->part);
        hipLaunchKernelGGL(k_finish_mean, dim3(1), blk, 0, F->st, F->part, F->grid, n, F->S);
        hipLaunchKernelGGL(k_remove_mean, g, blk, 0, F->st, F->rhs, n, F->S);
        FCHK(hipGetLastError());
    }
    if (!run_cg(F, CgOp{&F->kcsr, nullptr, F->gk}, F->dinv_k, F->rhs, iterations, err)) return false;
    if (F->nullspace) {
        hipLaunchKernelGGL(k_partial_sum, g, blk, 0, F->st, F->x, n, F->part);
        hipLaunchKernelGGL(k_finish_mean, dim3(1), blk, 0, F->st, F->part, F->grid, n, F->S);
        hipLaunchKernelGGL(k_remove_mean, g, blk, 0, F->st, F->x, n, F->S);
        FCHK(hipGetLastError());
    }
    FCHK(hipMemcpyAsync(F->phi, F->x, n * 8, hipMemcpyDeviceToDevice, F->st));
    FCHK(hipStreamSynchronize(F->st));
    return true;
}

bool field_efield(FieldSolver* F, const double* d_phi, int* iterations, std::string& err) {
    FCHK(hipSetDevice(F->device));
    const int64_t n = F->n;
    const dim3 g(F->grid), blk(FB);
    // the vector right-hand side (-grad(phi).v dx, dof-major like E) and one
    // CG on the vector space's block-diagonal mass matrix, as the reference
    // solves the vector system at once (efield.cpp:70-75)
    hipLaunchKernelGGL(k_grad3, g, blk, 0, F->st, F->col, F->w, F->grad[0], F->grad[1], F->grad[2], d_phi, F->rhs, n);
    if (F->warm)
        FCHK(hipMemcpyAsync(F->x, F->E, n * 24, hipMemcpyDeviceToDevice, F->st));
    else
        FCHK(hipMemsetAsync(F->x, 0, n * 24, F->st));
    FCHK(hipGetLastError());
    int it = 0;
    if (!run_cg(F, CgOp{nullptr, &F->mell, 1}, F->dinv_m, F->rhs, &it, err, 3)) return false;
    FCHK(hipMemcpyAsync(F->E, F->x, n * 24, hipMemcpyDeviceToDevice, F->st));
    FCHK(hipStreamSynchronize(F->st));
    if (iterations) *iterations = it;
    return true;
}

// Reverse of the E projection: E = M^-1 B phi, M symmetric, so
// dJ/dphi = B^T M^-1 dJ/dE: one CG on the vector mass matrix (right-hand side
// d_gE, dof-major like E), then the transposed gradient product into d_gphi.
bool field_efield_adjoint(FieldSolver* F, const double* d_gE, double* d_gphi, int* iterations, std::string& err) {
    FCHK(hipSetDevice(F->device));
    const int64_t n = F->n;
    const dim3 g(F->grid), blk(FB);
    FCHK(hipMemcpyAsync(F->rhs, d_gE, n * 24, hipMemcpyDeviceToDevice, F->st));
    FCHK(hipMemsetAsync(F->x, 0, n * 24, F->st));
    int it = 0;
    if (!run_cg(F, CgOp{nullptr, &F->mell, 1}, F->dinv_m, F->rhs, &it, err, 3)) return false;
    FCHK(hipMemsetAsync(d_gphi, 0, n * 8, F->st));
    hipLaunchKernelGGL(k_grad3_t, g, blk, 0, F->st, F->col, F->w, F->grad[0], F->grad[1], F->grad[2], F->x, d_gphi,
                       n);
    FCHK(hipGetLastError());
    FCHK(hipStreamSynchronize(F->st));
    if (iterations) *iterations = it;
    return true;
}

// ---------------------------------------------------------------------------
// Capacitance-matrix circuit (CircuitCM, object_CM.cpp:53-357; charge-sharing
// sets, object.cpp:68-135).  Small dense algebra on the host; the solves and
// the charge functionals on the device.
// ---------------------------------------------------------------------------
bool field_object_charges(FieldSolver* F, const double* d_phi, double* out, std::string& err) {
    if (F->nobj == 0) return true;
    FCHK(hipSetDevice(F->device));
    hipLaunchKernelGGL(k_image_charge, dim3(F->nobj), dim3(FB), 0, F->st, F->q_off, F->q_dof, F->q_w, d_phi, F->q_out);
    FCHK(hipGetLastError());
    FCHK(hipMemcpyAsync(out, F->q_out, F->nobj * 8, hipMemcpyDeviceToHost, F->st));
    FCHK(hipStreamSynchronize(F->st));
    return true;
}

// Inverse by LU with partial pivoting (the reference: ublas lu_factorize +
// lu_substitute on the identity); false if singular.
static bool invert(std::vector<double> A, int n, std::vector<double>& inv) {
    std::vector<int> piv((size_t)n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i * n + k]) > fabs(A[p * n + k])) p = i;
        if (A[p * n + k] == 0.0) return false;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
        for (int i = k + 1; i < n; ++i) {
            A[i * n + k] /= A[k * n + k];
            for (int j = k + 1; j < n; ++j) A[i * n + j] -= A[i * n + k] * A[k * n + j];
        }
    }
    inv.assign((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c) {
        std::vector<double> b((size_t)n, 0.0);
        b[c] = 1.0;
        for (int k = 0; k < n; ++k) std::swap(b[k], b[piv[k]]);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) b[i] -= A[i * n + j] * b[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) b[i] -= A[i * n + j] * b[j];
            b[i] /= A[i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[i * n + c] = b[i];
    }
    return true;
}

// get_charge_sharing_set / get_charge_sharing_sets (object.cpp:68-135).
static void sharing_set(std::vector<CircuitSource>& vs, int node, std::vector<int>& group) {
    group.push_back(node);
    size_t i = 0;
    while (i < vs.size()) {
        const CircuitSource v = vs[i];
        if (v.node_a == node) {
            vs.erase(vs.begin() + i);
            sharing_set(vs, v.node_b, group);
        } else if (v.node_b == node) {
            vs.erase(vs.begin() + i);
            sharing_set(vs, v.node_a, group);
        } else {
            i += 1;
        }
    }
}

static std::vector<std::vector<int>> sharing_sets(std::vector<CircuitSource> vs, int nobj) {
    std::vector<int> nodes((size_t)nobj);
    for (int i = 0; i < nobj; ++i) nodes[i] = i;
    std::vector<std::vector<int>> groups;
    while (!vs.empty()) {
        std::vector<int> group;
        sharing_set(vs, vs[0].node_a, group);
        groups.push_back(group);
    }
    for (auto& g : groups)
        for (int v : g)
            if (v != -1) nodes.erase(std::remove(nodes.begin(), nodes.end(), v), nodes.end());
    groups.erase(std::remove_if(groups.begin(), groups.end(),
                                [](const std::vector<int>& a) { return std::find(a.begin(), a.end(), -1) != a.end(); }),
                 groups.end());
    for (int v : nodes) groups.push_back({v});
    return groups;
}

bool field_set_circuit(FieldSolver* F, double dt, const std::vector<CircuitSource>& vsources,
                       const std::vector<CircuitSource>& isources, double* inv_capacitance, std::string& err) {
    const int n = F->nobj;
    if (n == 0) {
        err = "a circuit needs objects";
        return false;
    }
    for (const auto& v : vsources)
        if (v.node_b < 0 || v.node_b >= n || v.node_a < -1 || v.node_a >= n) {
            err = "voltage source node out of range";
            return false;
        }
    for (const auto& v : isources)
        if (v.node_b < -1 || v.node_b >= n || v.node_a < -1 || v.node_a >= n) {
            err = "current source node out of range";
            return false;
        }
    // inv_capacitance (object_CM.cpp:53-93): object j at 1, the others at 0,
    // no charge; capacitance(i, j) = charge on object i of that potential.
    std::vector<double> C((size_t)n * n), q((size_t)n), pot((size_t)n);
    const bool warm = F->warm;
    F->warm = false;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) pot[i] = i == j ? 1.0 : 0.0;
        if (!field_poisson(F, F->zero, nullptr, pot.data(), nullptr, err)) {
            F->warm = warm;
            return false;
        }
        if (!field_object_charges(F, F->phi, q.data(), err)) {
            F->warm = warm;
            return false;
        }
        for (int i = 0; i < n; ++i) C[i * n + j] = q[i];
    }
    F->warm = warm;
    if (!invert(C, n, F->invC)) {
        err = "singular capacitance matrix";
        return false;
    }
    if (inv_capacitance) std::copy(F->invC.begin(), F->invC.end(), inv_capacitance);
    F->dt = dt;
    F->vsrc = vsources;
    F->isrc = isources;
    F->groups = sharing_sets(vsources, n);
    // objects held at a fixed voltage relative to ground (object_CM.cpp:129-145)
    F->fixed_voltage.clear();
    for (size_t i = 0; i < vsources.size(); ++i) {
        if (vsources[i].node_a != -1) continue;
        std::vector<int> tmp{vsources[i].node_b};
        int node = vsources[i].node_b;
        for (size_t j = 0; j < vsources.size(); ++j)
            if (vsources[j].node_a == node) {
                tmp.push_back(vsources[j].node_b);
                node = vsources[j].node_b;
            }
        F->fixed_voltage.push_back(tmp);
    }
    F->circuit_vector.assign((size_t)2 * n, 0.0);
    for (const auto& v : vsources) F->circuit_vector[v.node_b] = v.value;
    // assemble_matrix (object_CM.cpp:204-248)
    const auto& iC = F->invC;
    std::vector<double> M((size_t)4 * n * n, 0.0);
    auto at = [&](int r, int c) -> double& { return M[(size_t)r * 2 * n + c]; };
    for (const auto& v : vsources) {
        at(v.node_b, v.node_b) = 1.0;
        if (v.node_a != -1) at(v.node_b, v.node_a) = -1.0;
    }
    for (const auto& g : F->groups) {
        for (int j : g) at(n + g[0], n + j) = 1.0;
        at(g[0], g[0]) = 1.0;
        for (int j : g) at(g[0], n + j) = -1.0 * iC[g[0] * n + j];
        for (size_t j = 1; j < g.size(); ++j) {
            at(n + g[j], g[j]) = 1.0;
            for (int k : g) at(n + g[j], n + k) = -1.0 * iC[g[j] * n + k];
        }
    }
    for (const auto& fvl : F->fixed_voltage)
        for (int j : fvl) {
            at(n + j, j) = 1.0;
            for (int k : fvl) at(n + j, n + k) = -1.0 * iC[j * n + k];
        }
    if (!invert(M, 2 * n, F->inv_circuit)) {
        err = "singular circuit matrix";
        return false;
    }
    F->has_circuit = true;
    return true;
}

bool field_poisson_circuit(FieldSolver* F, const double* d_rho, double* charges, double* potentials, int* iterations,
                           std::string& err) {
    if (!F->has_circuit) {
        err = "no circuit set";
        return false;
    }
    const int n = F->nobj;
    // pre_solve (object_CM.cpp:40-47, 343-357): current sources, objects grounded
    for (const auto& s : F->isrc) {
        const double dQ = s.value * F->dt;
        if (s.node_a != -1) charges[s.node_a] -= dQ;
        if (s.node_b != -1) charges[s.node_b] += dQ;
    }
    std::vector<double> pot((size_t)n, 0.0), image((size_t)n);
    int it0 = 0, it1 = 0;
    if (!field_poisson(F, d_rho, nullptr, pot.data(), &it0, err)) return false;
    // post_solve (object_CM.cpp:315-341): image charges, assemble_vector, potentials and charges
    if (!field_object_charges(F, F->phi, image.data(), err)) return false;
    const auto& iC = F->invC;
    std::vector<double>& cv = F->circuit_vector;
    for (const auto& g : F->groups) {
        double collected = 0.0;
        for (int j : g) collected += charges[j];
        cv[n + g[0]] = collected;
    }
    for (const auto& g : F->groups) {
        double c = 0.0;
        for (int j : g) c += iC[g[0] * n + j] * image[j];
        cv[g[0]] = c;
        for (size_t j = 1; j < g.size(); ++j) {
            c = 0.0;
            for (int k : g) c += iC[g[j] * n + k] * image[k];
            cv[n + g[j]] = c;
        }
    }
    for (const auto& fvl : F->fixed_voltage)
        for (int j : fvl) {
            double c = 0.0;
            for (int k : fvl) c += iC[j * n + k] * image[k];
            cv[n + j] = c;
        }
    for (int i = 0; i < n; ++i) {
        double potential = 0.0, charge = 0.0;
        for (int j = 0; j < 2 * n; ++j) potential += F->inv_circuit[(size_t)i * 2 * n + j] * cv[j];
        for (int j = 0; j < 2 * n; ++j) charge += F->inv_circuit[(size_t)(n + i) * 2 * n + j] * cv[j];
        pot[i] = potential;
        potentials[i] = potential;
        charges[i] = charge;
    }
    // correction_required: the solve with the circuit's potentials
    if (!field_poisson(F, d_rho, nullptr, pot.data(), &it1, err)) return false;
    if (iterations) *iterations = it0 + it1;
    return true;
}

}  // namespace punc_gpu

// Field solve on the device (reference PoissonSolver, poisson.cpp:260-376 with
// Potential3D.ufl; ESolver, efield.cpp:41-75 with EField3D.ufl).
//
// Assembly (host, once per mesh): CG1 element matrices on each tetrahedron
// from the gradients g_k of its barycentric coordinates and its volume V:
//   stiffness  K_ij = eps0 V g_i.g_j                 (a = eps0 grad(phi).grad(v) dx)
//   mass       M_ij = V (1 + delta_ij) / 20          (L = rho v dx, and the E projection)
//   gradient   B_c,ij = -(V / 4) (g_j)_c             (L = -grad(phi).v dx, component c)
// summed into CSR matrices with one sparsity (the dof graph of the cells).
// Dirichlet dofs are eliminated symmetrically: the operator keeps the free
// rows and columns and has unit rows at the Dirichlet dofs; the right-hand
// side is b - K phi_D on free dofs and phi_D on Dirichlet dofs.  This has
// the solution of the reference's row-replaced system (DirichletBC::apply)
// and keeps the operator symmetric positive definite, so the solves are
// Jacobi-preconditioned conjugate gradients.  Every reduction is a per-block
// partial plus a last-block sum in block order (deterministic), and the CG
// scalars stay on the device: iterations run in batches without host
// synchronisation.
#include "field.h"

#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

namespace punc_gpu {

namespace {

constexpr int FB = 256;         // threads per block
constexpr int MAX_BLOCKS = 2048;
constexpr int BATCH = 32;       // CG iterations between convergence checks

struct CgState {
    double rz, pq, alpha, beta, rr, tol2;
    double red[4];  // results of the last generic reduction
    int64_t it;
    int32_t done;
    uint32_t counter;
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

// Every block publishes NV partial sums; the last block to finish adds them
// in block order and returns true (in all its threads) with the totals in tot.
template <int NV>
__device__ bool last_block_sum(const double (&v)[NV], double* part, uint32_t* counter, double (&tot)[NV]) {
    __shared__ double sh[16];
    __shared__ bool last;
    for (int k = 0; k < NV; ++k) {
        const double b = block_reduce(v[k], sh);
        if (threadIdx.x == 0)
            __hip_atomic_store(&part[(int64_t)NV * blockIdx.x + k], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t t = atomicAdd(counter, 1u);
        last = t == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (uint32_t b = threadIdx.x; b < gridDim.x; b += blockDim.x)
            s += __hip_atomic_load(&part[(int64_t)NV * b + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // threads hold strided partial sums; combine them in thread order
        const double r = block_reduce(s, sh);
        if (threadIdx.x == 0) tot[k] = r;
    }
    if (threadIdx.x == 0) *counter = 0;
    return true;
}

__device__ inline double row_dot(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                 const double* __restrict__ av, const double* __restrict__ x, int64_t i) {
    double s = 0.0;
    for (int32_t j = rp[i]; j < rp[i + 1]; ++j) s += av[j] * x[ci[j]];
    return s;
}

__global__ void __launch_bounds__(FB) k_spmv(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                             const double* __restrict__ av, const double* __restrict__ x,
                                             double* __restrict__ y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = row_dot(rp, ci, av, x, i);
}

// Poisson right-hand side of the eliminated system: b_i - (K phi_D)_i on free
// dofs, phi_D on Dirichlet dofs; the initial iterate x0 = phi_D.
__global__ void k_rhs_dirichlet(const double* __restrict__ b, const double* __restrict__ kphid,
                                const double* __restrict__ phid, const uint8_t* __restrict__ isd,
                                double* __restrict__ rhs, double* __restrict__ x, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        rhs[i] = isd[i] ? phid[i] : b[i] - kphid[i];
        x[i] = isd[i] ? phid[i] : 0.0;
    }
}

// Sum of v (red[0]) by the last block.
__global__ void __launch_bounds__(FB) k_sum(const double* __restrict__ v, int64_t n, double* part, CgState* S) {
    double a[1] = {0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[0] += v[i];
    double t[1];
    if (last_block_sum<1>(a, part, &S->counter, t) && threadIdx.x == 0) S->red[0] = t[0];
}

// v -= red[0] / n (orthogonalisation against the normalised constant vector).
__global__ void k_remove_mean(double* __restrict__ v, int64_t n, const CgState* __restrict__ S) {
    const double m = S->red[0] / (double)n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] -= m;
}

// CG start: r = b - A x (x given), z = D^-1 r, p = z; rz, rr, |b|^2 -> tol2.
__global__ void __launch_bounds__(FB) k_cg_init(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                const double* __restrict__ av, const double* __restrict__ dinv,
                                                const double* __restrict__ b, const double* __restrict__ x,
                                                double* __restrict__ r, double* __restrict__ z, double* __restrict__ p,
                                                int64_t n, double rtol, double atol, double* part, CgState* S) {
    double a[3] = {0.0, 0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double ri = b[i] - row_dot(rp, ci, av, x, i);
        const double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        a[0] += ri * zi;
        a[1] += ri * ri;
        a[2] += b[i] * b[i];
    }
    double t[3];
    if (last_block_sum<3>(a, part, &S->counter, t) && threadIdx.x == 0) {
        S->rz = t[0];
        S->rr = t[1];
        const double tol = fmax(rtol * sqrt(t[2]), atol);
        S->tol2 = tol * tol;
        S->it = 0;
        S->done = t[1] <= S->tol2 ? 1 : 0;
    }
}

// q = A p, alpha = rz / (p.q).
__global__ void __launch_bounds__(FB) k_cg_spmv(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                const double* __restrict__ av, const double* __restrict__ p,
                                                double* __restrict__ q, int64_t n, double* part, CgState* S) {
    if (S->done) return;  // uniform: written by an earlier kernel
    double a[1] = {0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double qi = row_dot(rp, ci, av, p, i);
        q[i] = qi;
        a[0] += p[i] * qi;
    }
    double t[1];
    if (last_block_sum<1>(a, part, &S->counter, t) && threadIdx.x == 0) {
        S->pq = t[0];
        S->alpha = S->rz / t[0];
    }
}

// x += alpha p, r -= alpha q, z = D^-1 r; beta = rz' / rz; convergence.
__global__ void __launch_bounds__(FB) k_cg_update(double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ z, const double* __restrict__ p,
                                                  const double* __restrict__ q, const double* __restrict__ dinv,
                                                  int64_t n, double* part, CgState* S) {
    if (S->done) return;
    const double alpha = S->alpha;
    double a[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        const double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        a[0] += ri * zi;
        a[1] += ri * ri;
    }
    double t[2];
    if (last_block_sum<2>(a, part, &S->counter, t) && threadIdx.x == 0) {
        S->beta = t[0] / S->rz;
        S->rz = t[0];
        S->rr = t[1];
        S->it += 1;
        if (t[1] <= S->tol2) S->done = 1;
    }
}

// p = z + beta p.
__global__ void k_cg_direction(double* __restrict__ p, const double* __restrict__ z, int64_t n,
                               const CgState* __restrict__ S) {
    if (S->done) return;
    const double beta = S->beta;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

// E3[3i + c] = Ec[i].
__global__ void k_pack3(const double* __restrict__ e, int c, double* __restrict__ E3, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        E3[3 * i + c] = e[i];
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
    int grid = 1;
    std::vector<int32_t> dir_kind;  // per dof: -1 free, 0 exterior, 1+j object j
    // device
    int32_t *rp = nullptr, *ci = nullptr;
    double *kred = nullptr, *kfull = nullptr, *mass = nullptr, *grad[3] = {nullptr, nullptr, nullptr};
    double *dinv_k = nullptr, *dinv_m = nullptr;
    uint8_t* isd = nullptr;
    double *b = nullptr, *t = nullptr, *phid = nullptr, *rhs = nullptr;
    double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
    double *phi = nullptr, *E = nullptr, *part = nullptr;
    CgState* S = nullptr;
};

static void free_all(FieldSolver* F) {
    hipFree(F->rp);
    hipFree(F->ci);
    hipFree(F->kred);
    hipFree(F->kfull);
    hipFree(F->mass);
    for (auto& g : F->grad) hipFree(g);
    hipFree(F->dinv_k);
    hipFree(F->dinv_m);
    hipFree(F->isd);
    hipFree(F->b);
    hipFree(F->t);
    hipFree(F->phid);
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
    FieldSolver* F = new FieldSolver();
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
    auto up = [&](auto** d, const auto& h) {
        if (e != hipSuccess) return;
        e = dmalloc(d, h.size());
        if (e == hipSuccess && !h.empty())
            e = hipMemcpy(*d, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice);
    };
    up(&F->rp, rp);
    up(&F->ci, ci);
    up(&F->kred, Kr);
    up(&F->kfull, K);
    up(&F->mass, Mv);
    for (int d = 0; d < 3; ++d) up(&F->grad[d], B[d]);
    up(&F->dinv_k, dk);
    up(&F->dinv_m, dm);
    up(&F->isd, isd);
    double** vecs[] = {&F->b, &F->t, &F->phid, &F->rhs, &F->x, &F->r, &F->z, &F->p, &F->q, &F->phi};
    for (double** v : vecs)
        if (e == hipSuccess) e = dmalloc(v, (size_t)n);
    if (e == hipSuccess) e = dmalloc(&F->E, (size_t)n * 3);
    if (e == hipSuccess) e = dmalloc(&F->part, (size_t)F->grid * 4);
    if (e == hipSuccess) e = dmalloc(&F->S, 1);
    if (e == hipSuccess) e = hipMemset(F->S, 0, sizeof(CgState));
    if (e == hipSuccess) e = hipMemset(F->phi, 0, n * 8);
    if (e == hipSuccess) e = hipMemset(F->E, 0, n * 24);
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

#define FCHK(x)                                                      \
    do {                                                             \
        hipError_t _e = (x);                                         \
        if (_e != hipSuccess) {                                      \
            err = std::string(#x ": ") + hipGetErrorString(_e);      \
            return false;                                            \
        }                                                            \
    } while (0)

// Jacobi-preconditioned CG on (rp, ci, av) for rhs b from F->x (in place).
static bool run_cg(FieldSolver* F, const double* av, const double* dinv, const double* b, int* iterations,
                   std::string& err) {
    const int64_t n = F->n;
    const dim3 g(F->grid), blk(FB);
    hipLaunchKernelGGL(k_cg_init, g, blk, 0, F->st, F->rp, F->ci, av, dinv, b, F->x, F->r, F->z, F->p, n, 1e-12,
                       1e-14, F->part, F->S);
    FCHK(hipGetLastError());
    CgState h;
    for (;;) {
        for (int k = 0; k < BATCH; ++k) {
            hipLaunchKernelGGL(k_cg_spmv, g, blk, 0, F->st, F->rp, F->ci, av, F->p, F->q, n, F->part, F->S);
            hipLaunchKernelGGL(k_cg_update, g, blk, 0, F->st, F->x, F->r, F->z, F->p, F->q, dinv, n, F->part, F->S);
            hipLaunchKernelGGL(k_cg_direction, g, blk, 0, F->st, F->p, F->z, n, F->S);
        }
        FCHK(hipGetLastError());
        FCHK(hipMemcpyAsync(&h, F->S, sizeof(h), hipMemcpyDeviceToHost, F->st));
        FCHK(hipStreamSynchronize(F->st));
        if (h.done || h.it >= F->maxit) break;
    }
    if (iterations) *iterations = (int)h.it;
    if (!h.done) {
        err = "conjugate gradients did not converge in " + std::to_string(h.it) + " iterations (|r|^2 = " +
              std::to_string(h.rr) + ")";
        return false;
    }
    if (!std::isfinite(h.rr)) {
        err = "conjugate gradients broke down";
        return false;
    }
    return true;
}

bool field_poisson(FieldSolver* F, const double* d_rho, const double* ext_values, const double* object_potentials,
                   int* iterations, std::string& err) {
    FCHK(hipSetDevice(F->device));
    const int64_t n = F->n;
    const dim3 g(F->grid), blk(FB);
    // phi_D on the Dirichlet dofs (exterior values, object potentials)
    std::vector<double> pd((size_t)n, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t k = F->dir_kind[i];
        if (k == 0) pd[i] = ext_values ? ext_values[i] : 0.0;
        if (k >= 1) pd[i] = object_potentials ? object_potentials[k - 1] : 0.0;
    }
    FCHK(hipMemcpyAsync(F->phid, pd.data(), n * 8, hipMemcpyHostToDevice, F->st));
    hipLaunchKernelGGL(k_spmv, g, blk, 0, F->st, F->rp, F->ci, F->mass, d_rho, F->b, n);       // b = M rho
    hipLaunchKernelGGL(k_spmv, g, blk, 0, F->st, F->rp, F->ci, F->kfull, F->phid, F->t, n);    // K phi_D
    hipLaunchKernelGGL(k_rhs_dirichlet, g, blk, 0, F->st, F->b, F->t, F->phid, F->isd, F->rhs, F->x, n);
    FCHK(hipGetLastError());
    if (F->nullspace) {
        hipLaunchKernelGGL(k_sum, g, blk, 0, F->st, F->rhs, n, F->part, F->S);
        hipLaunchKernelGGL(k_remove_mean, g, blk, 0, F->st, F->rhs, n, F->S);
        FCHK(hipGetLastError());
    }
    if (!run_cg(F, F->kred, F->dinv_k, F->rhs, iterations, err)) return false;
    if (F->nullspace) {
        hipLaunchKernelGGL(k_sum, g, blk, 0, F->st, F->x, n, F->part, F->S);
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
    int total = 0;
    for (int c = 0; c < 3; ++c) {
        hipLaunchKernelGGL(k_spmv, g, blk, 0, F->st, F->rp, F->ci, F->grad[c], d_phi, F->rhs, n);
        FCHK(hipMemsetAsync(F->x, 0, n * 8, F->st));
        int it = 0;
        if (!run_cg(F, F->mass, F->dinv_m, F->rhs, &it, err)) return false;
        total += it;
        hipLaunchKernelGGL(k_pack3, g, blk, 0, F->st, F->x, c, F->E, n);
        FCHK(hipGetLastError());
    }
    FCHK(hipStreamSynchronize(F->st));
    if (iterations) *iterations = total;
    return true;
}

}  // namespace punc_gpu

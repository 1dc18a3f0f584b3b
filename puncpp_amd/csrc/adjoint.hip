// Reverse mode through K steps of Boris push + move + relocation and the CG1
// deposit of the final positions, along the forward run's cell path.
//
// Forward step n (reference boris_cg1, pusher.h:343-368, and move /
// move_periodic, pusher.h:648-681):
//   E_p = sum_i lambda_i(x_n) E_i            (cell c_n)
//   v-  = v_n + a E_p,  a = (0.5 dt) q/m
//   v'  = v- + v- x t,  v+ = v- + v' x s,  v_{n+1} = v+ + a E_p
//   x_{n+1} = x_n + dt v_{n+1}  (wrapped when periodic), c_{n+1} = walk
// Objective input: w_j = sum_p q_p lambda_j(x_K) over particles alive at K
// (distribute_cg1 before the dv_inv scaling, distributor.h:185-204).
// Given g = dJ/dw, the reverse sweep seeds xbar_K = q sum_i g_{dof_i} grad
// lambda_i, vbar_K = 0 and applies per step
//   vbar_{n+1} += dt xbar_{n+1}
//   Ebar = a vbar_{n+1};  v'bar = s x vbar_{n+1};  v-bar = vbar_{n+1} + v'bar + t x v'bar
//   Ebar += a v-bar;  vbar_n = v-bar
//   xbar_n = xbar_{n+1} + sum_i (Ebar . E_i) grad lambda_i
//   dJ/dE_i += lambda_i Ebar
// The position -> cell map is piecewise constant and both the CG1 field and
// the CG1 deposit are continuous across faces, so the path-wise derivative is
// the derivative of J wherever no particle crosses the mesh boundary.
//
// Collected charge.  A particle that leaves through boundary b during step n
// (walk exit through facet f of cell c, plane pi(x) = c0 + n.x, negative
// inside) crosses pi = 0 at the fraction s = pi(x_n) / (pi(x_n) - pi(x_{n+1}))
// of the step.  The objective input Q_b = sum_p q_p (K - n_p - s_p) / K over
// the particles that left through b is the time average over the K steps of
// the charge b has collected: unlike the collected charge itself (piecewise
// constant along a fixed crossing path) it is continuous, also when a
// crossing moves from the end of one step to the start of the next, and its
// derivative flows through s.  Given gq_b = dJ/dQ_b, the reverse sweep seeds
// such a particle at its exit step with
//   xbar_{n+1} = gq_b q (-1/K) ds/dpi_{n+1} n,   vbar_{n+1} = 0,
//   xbar_n    += gq_b q (-1/K) ds/dpi_n n,
//   ds/dpi_n = -pi_{n+1} / D^2,  ds/dpi_{n+1} = pi_n / D^2,  D = pi_n - pi_{n+1},
// and propagates it through the steps before as any other particle.
// A particle absorbed without such a seed carries no sensitivity.
//
// States of one checkpoint segment are recomputed forward from the
// segment's checkpoint and held on the device, so memory is
// (K / every + every + 1) particle states.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/punc_gpu.h"
#include "device.h"
#include "kernels.h"
#include "mesh_host.h"

namespace punc_gpu {

struct AdjState {
    double* x;
    double* y;
    double* z;
    double* vx;
    double* vy;
    double* vz;
    int32_t* c;  // cell, -1 once absorbed
};

struct AdjCoef {
    double a;     // (0.5*dt)*(q/m)
    double t[3];  // tan(((dt*q)/(2.0*m))*B_i)
    double s[3];  // 2 t_i / (1 + |t|^2)
    double q;
};

__device__ inline void adj_bary(const double* A, double x, double y, double z, double l[4]) {
    l[0] = ((A[0] + A[1] * x) + A[2] * y) + A[3] * z;
    l[1] = ((A[4] + A[5] * x) + A[6] * y) + A[7] * z;
    l[2] = ((A[8] + A[9] * x) + A[10] * y) + A[11] * z;
    l[3] = ((1 - l[0]) - l[1]) - l[2];
}

// relocate_fast's walk (population.h:547-565).  Returns the cell, -bnd when
// the walk leaves the mesh (exit: cell * 4 + facet of the exit), or
// WALK_FAILED after the hop cap.
__device__ inline int32_t adj_walk(const DevMesh& M, int32_t c, double x, double y, double z, int32_t& exit) {
    for (int h = 0; h < WALK_MAX_HOPS; ++h) {
        const double* p = M.planes[c].p;
        const int f = plane_proj(p, x, y, z) > 0 ? 0
                      : plane_proj(p + 4, x, y, z) > 0 ? 1
                      : plane_proj(p + 8, x, y, z) > 0 ? 2
                      : plane_proj(p + 12, x, y, z) > 0 ? 3 : -1;
        if (f < 0) return c;
        const int32_t nb = M.rec[c].nbr[f];
        if (nb < 0) {
            exit = c * 4 + f;
            return nb;
        }
        c = nb;
    }
    return WALK_FAILED;
}

// Absorption record of the main forward sweep: abs_info = bnd | step << 8
// (0: not absorbed), abs_exit = exit cell * 4 + facet, abs_w = the particle's
// weight (K - n - s) / K in Q_bnd.
struct AdjAbsorb {
    int32_t* info;
    int32_t* exit;
    double* w;
};

__global__ void __launch_bounds__(256)
k_adj_step(const DevMesh M, AdjState in, AdjState out, const int32_t* __restrict__ sp,
           const AdjCoef* __restrict__ coef, const double4* __restrict__ E, double dt, int periodic,
           double Lx, double Ly, double Lz, int64_t n, unsigned long long* __restrict__ lost, int record,
           int step, int Ksteps, AdjAbsorb ab, unsigned long long* __restrict__ irregular) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = in.c[i];
        double x = in.x[i], y = in.y[i], z = in.z[i];
        double vx = in.vx[i], vy = in.vy[i], vz = in.vz[i];
        int32_t c2 = -1;
        if (c >= 0) {
            const CellRec& R = M.rec[c];
            double l[4];
            adj_bary(R.bary, x, y, z, l);
            const double4 e0 = E[R.dof[0]], e1 = E[R.dof[1]], e2 = E[R.dof[2]], e3 = E[R.dof[3]];
            double Ex = 0.0, Ey = 0.0, Ez = 0.0;
            Ex += l[0] * e0.x; Ex += l[1] * e1.x; Ex += l[2] * e2.x; Ex += l[3] * e3.x;
            Ey += l[0] * e0.y; Ey += l[1] * e1.y; Ey += l[2] * e2.y; Ey += l[3] * e3.y;
            Ez += l[0] * e0.z; Ez += l[1] * e1.z; Ez += l[2] * e2.z; Ez += l[3] * e3.z;
            const AdjCoef K = coef[sp[i]];
            const double mx = vx + K.a * Ex, my = vy + K.a * Ey, mz = vz + K.a * Ez;
            const double cx = my * K.t[2] - mz * K.t[1];
            const double cy = -mx * K.t[2] + mz * K.t[0];
            const double cz = mx * K.t[1] - my * K.t[0];
            const double px = mx + cx, py = my + cy, pz = mz + cz;
            const double dx = py * K.s[2] - pz * K.s[1];
            const double dy = -px * K.s[2] + pz * K.s[0];
            const double dz = px * K.s[1] - py * K.s[0];
            vx = (mx + dx) + K.a * Ex;
            vy = (my + dy) + K.a * Ey;
            vz = (mz + dz) + K.a * Ez;
            x += dt * vx;
            y += dt * vy;
            z += dt * vz;
            if (periodic) {
                x -= Lx * floor(x / Lx);
                y -= Ly * floor(y / Ly);
                z -= Lz * floor(z / Lz);
            }
            int32_t exit = -1;
            c2 = adj_walk(M, c, x, y, z, exit);
            if (c2 == WALK_FAILED) {
                atomicAdd(lost, 1ull);
                c2 = -1;
            } else if (c2 < 0) {
                if (record) {
                    const double* p = M.planes[exit >> 2].p + 4 * (exit & 3);
                    const double pa = plane_proj(p, in.x[i], in.y[i], in.z[i]);
                    const double pb = plane_proj(p, x, y, z);
                    double sc = 0.0;
                    if (pa <= 0.0 && pb > 0.0) {
                        sc = pa / (pa - pb);
                    } else {  // the visibility walk left through a plane the segment does not cross
                        atomicAdd(irregular, 1ull);
                        sc = pa > 0.0 ? 0.0 : 1.0;
                    }
                    ab.info[i] = (-c2) | (step << 8);
                    ab.exit[i] = exit;
                    ab.w[i] = ((double)(Ksteps - step) - sc) / (double)Ksteps;
                }
                c2 = -1;
            }
        }
        out.x[i] = x;
        out.y[i] = y;
        out.z[i] = z;
        out.vx[i] = vx;
        out.vy[i] = vy;
        out.vz[i] = vz;
        out.c[i] = c2;
    }
}

__global__ void __launch_bounds__(256)
k_adj_deposit(const DevMesh M, AdjState st, const int32_t* __restrict__ sp, const AdjCoef* __restrict__ coef,
              int64_t n, double* __restrict__ w) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = st.c[i];
        if (c < 0) continue;
        const CellRec& R = M.rec[c];
        double l[4];
        adj_bary(R.bary, st.x[i], st.y[i], st.z[i], l);
        const double q = coef[sp[i]].q;
        for (int k = 0; k < 4; ++k) unsafeAtomicAdd(&w[R.dof[k]], q * l[k]);
    }
}

// xbar_K = q sum_i g_{dof_i} grad lambda_i; vbar_K = 0.
__global__ void __launch_bounds__(256)
k_adj_seed(const DevMesh M, AdjState st, const int32_t* __restrict__ sp, const AdjCoef* __restrict__ coef,
           const double* __restrict__ g, int64_t n, double* __restrict__ xb, double* __restrict__ vb) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = st.c[i];
        double gx = 0.0, gy = 0.0, gz = 0.0;
        if (c >= 0) {
            const CellRec& R = M.rec[c];
            const double* A = R.bary;
            const double g0 = g[R.dof[0]], g1 = g[R.dof[1]], g2 = g[R.dof[2]], g3 = g[R.dof[3]];
            // grad lambda_3 = -(grad lambda_0 + grad lambda_1 + grad lambda_2)
            const double q = coef[sp[i]].q;
            gx = q * ((g0 - g3) * A[1] + (g1 - g3) * A[5] + (g2 - g3) * A[9]);
            gy = q * ((g0 - g3) * A[2] + (g1 - g3) * A[6] + (g2 - g3) * A[10]);
            gz = q * ((g0 - g3) * A[3] + (g1 - g3) * A[7] + (g2 - g3) * A[11]);
        }
        xb[i] = gx;
        xb[n + i] = gy;
        xb[2 * n + i] = gz;
        vb[i] = 0.0;
        vb[n + i] = 0.0;
        vb[2 * n + i] = 0.0;
    }
}

__global__ void __launch_bounds__(256)
k_adj_reverse(const DevMesh M, AdjState st, AdjState nx, const int32_t* __restrict__ sp,
              const AdjCoef* __restrict__ coef, const double4* __restrict__ E, double dt, int64_t n,
              double* __restrict__ xb, double* __restrict__ vb, double* __restrict__ gE, AdjAbsorb ab,
              const double* __restrict__ gq, int Ksteps) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = st.c[i];
        if (c < 0) continue;  // absorbed earlier
        const AdjCoef K = coef[sp[i]];
        double xbx, xby, xbz;
        double sx0 = 0.0, sy0 = 0.0, sz0 = 0.0;  // seed of xbar_n at an exit step
        if (nx.c[i] < 0) {
            // left the mesh during this step: only a collected-charge seed
            // gives it sensitivity
            const double G = gq ? gq[ab.info[i] & 0xff] * K.q : 0.0;
            if (G == 0.0) continue;
            const int32_t ex = ab.exit[i];
            const double* p = M.planes[ex >> 2].p + 4 * (ex & 3);
            const double pa = plane_proj(p, st.x[i], st.y[i], st.z[i]);
            const double pb = plane_proj(p, nx.x[i], nx.y[i], nx.z[i]);
            if (!(pa <= 0.0 && pb > 0.0)) continue;  // irregular exit (counted by the forward): no seed
            const double D = pa - pb;
            const double dJds = -G / (double)Ksteps;
            const double ga = dJds * (-pb / (D * D)), gb = dJds * (pa / (D * D));
            xbx = gb * p[1];
            xby = gb * p[2];
            xbz = gb * p[3];
            sx0 = ga * p[1];
            sy0 = ga * p[2];
            sz0 = ga * p[3];
            vb[i] = 0.0;
            vb[n + i] = 0.0;
            vb[2 * n + i] = 0.0;
        } else {
            xbx = xb[i];
            xby = xb[n + i];
            xbz = xb[2 * n + i];
        }
        const CellRec& R = M.rec[c];
        double l[4];
        adj_bary(R.bary, st.x[i], st.y[i], st.z[i], l);
        const double4 e[4] = {E[R.dof[0]], E[R.dof[1]], E[R.dof[2]], E[R.dof[3]]};
        // vbar_{n+1} += dt xbar_{n+1}
        const double wx = vb[i] + dt * xbx, wy = vb[n + i] + dt * xby, wz = vb[2 * n + i] + dt * xbz;
        // v'bar = s x vbar_{n+1}
        const double px = K.s[1] * wz - K.s[2] * wy;
        const double py = K.s[2] * wx - K.s[0] * wz;
        const double pz = K.s[0] * wy - K.s[1] * wx;
        // v-bar = vbar_{n+1} + v'bar + t x v'bar
        const double mx = wx + px + (K.t[1] * pz - K.t[2] * py);
        const double my = wy + py + (K.t[2] * px - K.t[0] * pz);
        const double mz = wz + pz + (K.t[0] * py - K.t[1] * px);
        const double Ebx = K.a * (wx + mx), Eby = K.a * (wy + my), Ebz = K.a * (wz + mz);
        vb[i] = mx;
        vb[n + i] = my;
        vb[2 * n + i] = mz;
        // xbar_n = xbar_{n+1} + sum_i (Ebar . E_i) grad lambda_i
        double d[4];
        for (int k = 0; k < 4; ++k) d[k] = Ebx * e[k].x + Eby * e[k].y + Ebz * e[k].z;
        const double* A = R.bary;
        xbx += (d[0] - d[3]) * A[1] + (d[1] - d[3]) * A[5] + (d[2] - d[3]) * A[9] + sx0;
        xby += (d[0] - d[3]) * A[2] + (d[1] - d[3]) * A[6] + (d[2] - d[3]) * A[10] + sy0;
        xbz += (d[0] - d[3]) * A[3] + (d[1] - d[3]) * A[7] + (d[2] - d[3]) * A[11] + sz0;
        xb[i] = xbx;
        xb[n + i] = xby;
        xb[2 * n + i] = xbz;
        for (int k = 0; k < 4; ++k) {
            double* g = gE + 3 * (int64_t)R.dof[k];
            unsafeAtomicAdd(g, l[k] * Ebx);
            unsafeAtomicAdd(g + 1, l[k] * Eby);
            unsafeAtomicAdd(g + 2, l[k] * Ebz);
        }
    }
}

}  // namespace punc_gpu

// ===========================================================================
// C ABI
// ===========================================================================
using namespace punc_gpu;

int punc_internal_mesh_view(punc_mesh_t* m, DevMesh* out, int64_t* ndof, std::string& err);

static thread_local std::string g_adj_err;

struct punc_adj_s {
    DevMesh M;
    int64_t n = 0, ndof = 0;
    int nsp = 0;
    std::vector<double> q, m;
    int32_t* d_sp = nullptr;
    AdjCoef* d_coef = nullptr;
    double4* d_E = nullptr;
    double* d_w = nullptr;
    double* d_gE = nullptr;
    double* d_g = nullptr;
    double* d_xb = nullptr;
    double* d_vb = nullptr;
    unsigned long long* d_lost = nullptr;  // [0] walks over the hop cap, [1] irregular exits
    double* d_gq = nullptr;                // 64: dJ/dQ_b
    AdjAbsorb ab{};                        // absorption record of the forward sweep
    // checkpoints (K/every + 1 states) and one segment (every + 1 states)
    std::vector<AdjState> ckpt, seg;
    std::vector<void*> allocs;
    int K = 0, every = 0;
    double dt = 0.0;
    int periodic = 0;
    double Ld[3] = {0, 0, 0};
    hipStream_t stream = nullptr;
    int grid = 0;
    bool have_forward = false;
};

static int adj_fail(const std::string& s) {
    g_adj_err = s;
    return -1;
}
#define AHIP(call)                                                                   \
    do {                                                                             \
        hipError_t e__ = (call);                                                     \
        if (e__ != hipSuccess) return adj_fail(std::string(#call) + ": " + hipGetErrorString(e__)); \
    } while (0)

extern "C" const char* punc_gpu_adjoint_last_error(void) { return g_adj_err.c_str(); }

static int adj_alloc_state(punc_adj_s* A, AdjState& s) {
    double** f[6] = {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz};
    for (auto p : f) {
        AHIP(hipMalloc((void**)p, std::max<int64_t>(A->n, 1) * 8));
        A->allocs.push_back(*p);
    }
    AHIP(hipMalloc((void**)&s.c, std::max<int64_t>(A->n, 1) * 4));
    A->allocs.push_back(s.c);
    return 0;
}

static void adj_free_states(punc_adj_s* A) {
    for (void* p : A->allocs) hipFree(p);
    A->allocs.clear();
    A->ckpt.clear();
    A->seg.clear();
}

extern "C" void punc_gpu_adjoint_destroy(punc_adj_t* A) {
    if (!A) return;
    if (A->stream) hipStreamSynchronize(A->stream);
    adj_free_states(A);
    hipFree(A->d_sp);
    hipFree(A->d_coef);
    hipFree(A->d_E);
    hipFree(A->d_w);
    hipFree(A->d_gE);
    hipFree(A->d_g);
    hipFree(A->d_xb);
    hipFree(A->d_vb);
    hipFree(A->d_lost);
    hipFree(A->d_gq);
    hipFree(A->ab.info);
    hipFree(A->ab.exit);
    hipFree(A->ab.w);
    if (A->stream) hipStreamDestroy(A->stream);
    delete A;
}

extern "C" int punc_gpu_adjoint_create(punc_mesh_t* m, int nspecies, const double* q, const double* mass,
                                       int64_t n, const int32_t* species, punc_adj_t** out) {
    if (!m || !q || !mass || !out || (n > 0 && !species)) return adj_fail("adjoint_create: null argument");
    if (nspecies < 1 || n < 1) return adj_fail("adjoint_create: need species and particles");
    for (int64_t i = 0; i < n; ++i)
        if (species[i] < 0 || species[i] >= nspecies) return adj_fail("adjoint_create: species out of range");
    auto* A = new punc_adj_s();
    std::string err;
    if (punc_internal_mesh_view(m, &A->M, &A->ndof, err)) {
        delete A;
        return adj_fail("adjoint_create: " + err);
    }
    A->n = n;
    A->nsp = nspecies;
    A->q.assign(q, q + nspecies);
    A->m.assign(mass, mass + nspecies);
    int dev = 0;
    hipDeviceProp_t prop;
#define ACHK(call)                                                                                  \
    do {                                                                                            \
        hipError_t e2 = (call);                                                                     \
        if (e2 != hipSuccess) {                                                                     \
            std::string msg = std::string(#call) + ": " + hipGetErrorString(e2);                   \
            punc_gpu_adjoint_destroy(A);                                                            \
            return adj_fail(msg);                                                                   \
        }                                                                                           \
    } while (0)
    ACHK(hipGetDevice(&dev));
    ACHK(hipGetDeviceProperties(&prop, dev));
    A->grid = prop.multiProcessorCount * 8;
    ACHK(hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking));
    ACHK(hipMalloc((void**)&A->d_sp, n * 4));
    ACHK(hipMemcpy(A->d_sp, species, n * 4, hipMemcpyHostToDevice));
    ACHK(hipMalloc((void**)&A->d_coef, nspecies * sizeof(AdjCoef)));
    ACHK(hipMalloc((void**)&A->d_E, A->ndof * sizeof(double4)));
    ACHK(hipMalloc((void**)&A->d_w, A->ndof * 8));
    ACHK(hipMalloc((void**)&A->d_g, A->ndof * 8));
    ACHK(hipMalloc((void**)&A->d_gE, A->ndof * 3 * 8));
    ACHK(hipMalloc((void**)&A->d_xb, n * 3 * 8));
    ACHK(hipMalloc((void**)&A->d_vb, n * 3 * 8));
    ACHK(hipMalloc((void**)&A->d_lost, 16));
    ACHK(hipMalloc((void**)&A->d_gq, 64 * 8));
    ACHK(hipMalloc((void**)&A->ab.info, n * 4));
    ACHK(hipMalloc((void**)&A->ab.exit, n * 4));
    ACHK(hipMalloc((void**)&A->ab.w, n * 8));
#undef ACHK
    *out = A;
    return 0;
}

static int fill_state(punc_adj_s* A, AdjState& s, const double* x, const double* v, const int64_t* cells) {
    std::vector<double> col(A->n);
    double* dst[6] = {s.x, s.y, s.z, s.vx, s.vy, s.vz};
    for (int k = 0; k < 6; ++k) {
        const double* src = k < 3 ? x : v;
        for (int64_t i = 0; i < A->n; ++i) col[i] = src[3 * i + (k % 3)];
        AHIP(hipMemcpy(dst[k], col.data(), A->n * 8, hipMemcpyHostToDevice));
    }
    std::vector<int32_t> c(A->n);
    for (int64_t i = 0; i < A->n; ++i) {
        if (cells[i] < 0 || cells[i] >= A->M.nc) return adj_fail("adjoint_forward: cell out of range");
        c[i] = (int32_t)cells[i];
    }
    AHIP(hipMemcpy(s.c, c.data(), A->n * 4, hipMemcpyHostToDevice));
    return 0;
}

static int copy_state(punc_adj_s* A, const AdjState& a, AdjState& b) {
    const double* src[6] = {a.x, a.y, a.z, a.vx, a.vy, a.vz};
    double* dst[6] = {b.x, b.y, b.z, b.vx, b.vy, b.vz};
    for (int k = 0; k < 6; ++k) AHIP(hipMemcpyAsync(dst[k], src[k], A->n * 8, hipMemcpyDeviceToDevice, A->stream));
    AHIP(hipMemcpyAsync(b.c, a.c, A->n * 4, hipMemcpyDeviceToDevice, A->stream));
    return 0;
}

// record: the main forward sweep (writes the absorption record); step: the
// step's index in 0..K-1.
static int adj_step(punc_adj_s* A, const AdjState& in, AdjState& out, bool record, int step) {
    hipLaunchKernelGGL(k_adj_step, dim3(A->grid), dim3(256), 0, A->stream, A->M, in, out, A->d_sp, A->d_coef, A->d_E,
                       A->dt, A->periodic, A->Ld[0], A->Ld[1], A->Ld[2], A->n, A->d_lost, record ? 1 : 0, step,
                       A->K, A->ab, A->d_lost + 1);
    AHIP(hipGetLastError());
    return 0;
}

static int download_state(punc_adj_s* A, const AdjState& s, double* x, double* v, int64_t* cells) {
    std::vector<double> col(A->n);
    const double* src[6] = {s.x, s.y, s.z, s.vx, s.vy, s.vz};
    for (int k = 0; k < 6; ++k) {
        double* dst = k < 3 ? x : v;
        if (!dst) continue;
        AHIP(hipMemcpy(col.data(), src[k], A->n * 8, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < A->n; ++i) dst[3 * i + (k % 3)] = col[i];
    }
    if (cells) {
        std::vector<int32_t> c(A->n);
        AHIP(hipMemcpy(c.data(), s.c, A->n * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < A->n; ++i) cells[i] = c[i];
    }
    return 0;
}

extern "C" int punc_gpu_adjoint_forward(punc_adj_t* A, const double* x0, const int64_t* cell0, const double* v0,
                                        const double* E, const double B[3], double dt, const double* Ld, int K,
                                        int every, double* w, double* xK, double* vK, int64_t* cellK) {
    if (!A || !x0 || !cell0 || !v0 || !E || !B) return adj_fail("adjoint_forward: null argument");
    if (K < 1 || every < 1) return adj_fail("adjoint_forward: K and checkpoint interval must be >= 1");
    AHIP(hipStreamSynchronize(A->stream));
    adj_free_states(A);
    A->K = K;
    A->every = every;
    A->dt = dt;
    A->periodic = Ld ? 1 : 0;
    for (int d = 0; d < 3; ++d) A->Ld[d] = Ld ? Ld[d] : 0.0;
    std::vector<AdjCoef> h(A->nsp);
    for (int s = 0; s < A->nsp; ++s) {
        const double q = A->q[s], m = A->m[s];
        AdjCoef& c = h[s];
        memset(&c, 0, sizeof(c));
        c.q = q;
        c.a = 0.5 * dt * (q / m);  // pusher.h:345
        double t_mag2 = 0.0;
        for (int i = 0; i < 3; ++i) {
            c.t[i] = tan((dt * q / (2.0 * m)) * B[i]);  // pusher.h:334
            t_mag2 += c.t[i] * c.t[i];
        }
        for (int i = 0; i < 3; ++i) c.s[i] = 2 * c.t[i] / (1 + t_mag2);
    }
    AHIP(hipMemcpy(A->d_coef, h.data(), A->nsp * sizeof(AdjCoef), hipMemcpyHostToDevice));
    // Contiguous copy into the dJ/dE buffer (3 per dof, free until the reverse
    // sweep), then expanded to one double4 per dof on the device.
    AHIP(hipMemcpy(A->d_gE, E, A->ndof * 3 * sizeof(double), hipMemcpyHostToDevice));
    AHIP(launch_expand_field(A->d_gE, A->d_E, A->ndof, A->stream));
    AHIP(hipStreamSynchronize(A->stream));
    AHIP(hipMemset(A->d_lost, 0, 16));
    AHIP(hipMemset(A->ab.info, 0, A->n * 4));
    AHIP(hipMemset(A->ab.exit, 0xff, A->n * 4));
    AHIP(hipMemset(A->ab.w, 0, A->n * 8));
    const int nck = K / every + 1;
    A->ckpt.resize(nck);
    for (auto& s : A->ckpt)
        if (adj_alloc_state(A, s)) return -1;
    A->seg.resize(every + 1);
    for (auto& s : A->seg)
        if (adj_alloc_state(A, s)) return -1;
    if (fill_state(A, A->ckpt[0], x0, v0, cell0)) return -1;
    // Forward sweep: ping-pong through the segment buffers, keeping a
    // checkpoint every `every` steps.
    if (copy_state(A, A->ckpt[0], A->seg[0])) return -1;
    int cur = 0;
    for (int nstep = 0; nstep < K; ++nstep) {
        const int nxt = 1 - cur;
        if (adj_step(A, A->seg[cur], A->seg[nxt], true, nstep)) return -1;
        cur = nxt;
        if ((nstep + 1) % every == 0 && (nstep + 1) / every < nck)
            if (copy_state(A, A->seg[cur], A->ckpt[(nstep + 1) / every])) return -1;
    }
    AHIP(hipMemsetAsync(A->d_w, 0, A->ndof * 8, A->stream));
    hipLaunchKernelGGL(k_adj_deposit, dim3(A->grid), dim3(256), 0, A->stream, A->M, A->seg[cur], A->d_sp, A->d_coef,
                       A->n, A->d_w);
    AHIP(hipGetLastError());
    AHIP(hipStreamSynchronize(A->stream));
    unsigned long long lost = 0;
    AHIP(hipMemcpy(&lost, A->d_lost, 8, hipMemcpyDeviceToHost));
    if (lost) return adj_fail("adjoint_forward: a relocation walk exceeded the hop cap");
    if (w) AHIP(hipMemcpy(w, A->d_w, A->ndof * 8, hipMemcpyDeviceToHost));
    if (xK || vK || cellK)
        if (download_state(A, A->seg[cur], xK, vK, cellK)) return -1;
    A->have_forward = true;
    return 0;
}

extern "C" int punc_gpu_adjoint_collected(punc_adj_t* A, double* weight, int32_t* bnd, int32_t* step,
                                          int64_t* irregular) {
    if (!A) return adj_fail("adjoint_collected: null argument");
    if (!A->have_forward) return adj_fail("adjoint_collected: run adjoint_forward first");
    if (weight) AHIP(hipMemcpy(weight, A->ab.w, A->n * 8, hipMemcpyDeviceToHost));
    if (bnd || step) {
        std::vector<int32_t> info(A->n);
        AHIP(hipMemcpy(info.data(), A->ab.info, A->n * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < A->n; ++i) {
            if (bnd) bnd[i] = info[i] & 0xff;
            if (step) step[i] = info[i] ? (info[i] >> 8) : -1;
        }
    }
    if (irregular) {
        unsigned long long c[2];
        AHIP(hipMemcpy(c, A->d_lost, 16, hipMemcpyDeviceToHost));
        *irregular = (int64_t)c[1];
    }
    return 0;
}

extern "C" int punc_gpu_adjoint_backward(punc_adj_t* A, const double* g, double* dJdE, double* dJdv0, double* dJdx0) {
    return punc_gpu_adjoint_backward_q(A, g, nullptr, dJdE, dJdv0, dJdx0);
}

extern "C" int punc_gpu_adjoint_backward_q(punc_adj_t* A, const double* g, const double* gq, double* dJdE,
                                           double* dJdv0, double* dJdx0) {
    if (!A || !g) return adj_fail("adjoint_backward: null argument");
    if (!A->have_forward) return adj_fail("adjoint_backward: run adjoint_forward first");
    const int K = A->K, every = A->every;
    AHIP(hipMemcpy(A->d_g, g, A->ndof * 8, hipMemcpyHostToDevice));
    if (gq) AHIP(hipMemcpy(A->d_gq, gq, 64 * 8, hipMemcpyHostToDevice));
    AHIP(hipMemsetAsync(A->d_gE, 0, A->ndof * 3 * 8, A->stream));
    // Segments [s0, s1) from the last to the first.
    bool seeded = false;
    for (int s0 = ((K - 1) / every) * every; s0 >= 0; s0 -= every) {
        const int s1 = std::min(K, s0 + every);
        if (copy_state(A, A->ckpt[s0 / every], A->seg[0])) return -1;
        for (int k = s0; k < s1; ++k)
            if (adj_step(A, A->seg[k - s0], A->seg[k - s0 + 1], false, k)) return -1;
        if (!seeded) {
            hipLaunchKernelGGL(k_adj_seed, dim3(A->grid), dim3(256), 0, A->stream, A->M, A->seg[s1 - s0], A->d_sp,
                               A->d_coef, A->d_g, A->n, A->d_xb, A->d_vb);
            AHIP(hipGetLastError());
            seeded = true;
        }
        for (int k = s1 - 1; k >= s0; --k) {
            hipLaunchKernelGGL(k_adj_reverse, dim3(A->grid), dim3(256), 0, A->stream, A->M, A->seg[k - s0],
                               A->seg[k - s0 + 1], A->d_sp, A->d_coef, A->d_E, A->dt, A->n, A->d_xb, A->d_vb,
                               A->d_gE, A->ab, gq ? (const double*)A->d_gq : nullptr, K);
            AHIP(hipGetLastError());
        }
    }
    AHIP(hipStreamSynchronize(A->stream));
    if (dJdE) AHIP(hipMemcpy(dJdE, A->d_gE, A->ndof * 3 * 8, hipMemcpyDeviceToHost));
    std::vector<double> col(A->n * 3);
    if (dJdv0) {
        AHIP(hipMemcpy(col.data(), A->d_vb, A->n * 3 * 8, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < A->n; ++i)
            for (int d = 0; d < 3; ++d) dJdv0[3 * i + d] = col[d * A->n + i];
    }
    if (dJdx0) {
        AHIP(hipMemcpy(col.data(), A->d_xb, A->n * 3 * 8, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < A->n; ++i)
            for (int d = 0; d < 3; ++d) dJdx0[3 * i + d] = col[d * A->n + i];
    }
    return 0;
}

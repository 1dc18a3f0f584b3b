// HIP kernels for the particle half of the PUNC PIC step on MI355X (gfx950).
//
// All particle data is FP64 structure-of-arrays in HBM; one lane handles one
// particle.  Blocks of the persistent kernels take contiguous particle chunks,
// and the chunk order is arranged so that each XCD (blocks are dispatched
// round-robin over the 8 XCDs) works on one contiguous range of the
// cell-sorted population, keeping a region's cell records and field dofs in
// that XCD's L2.  The library is built with -ffp-contract=off and every
// expression follows the reference's evaluation order, so positions,
// velocities and cell decisions are bit-identical to the CPU path.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.h"
#include "kernels.h"

namespace punc_gpu {

// ---------------------------------------------------------------------------
// Helpers
// ---------------------------------------------------------------------------
// Particle array pointers are read from DevState, so the compiler cannot tell
// their address space and would emit FLAT loads and stores.  FLAT operations
// also count against the LDS/scalar wait counter: every wait for an LDS
// result (deposit table, cross-lane shuffles) would then also wait for all
// outstanding particle loads and stores.  Viewing them through global-address-
// space pointers gives plain global_load/global_store.
#define GLOBAL __attribute__((address_space(1)))
typedef double v2d __attribute__((ext_vector_type(2)));
struct GArrays {
    GLOBAL v2d* xy;   // (x, y)
    GLOBAL v2d* zvx;  // (z, vx)
    GLOBAL v2d* vyz;  // (vy, vz)
    GLOBAL int32_t* cs;
};

__device__ inline GArrays gview(const PopArrays& a) {
    GArrays g;
    g.xy = (GLOBAL v2d*)a.xy;
    g.zvx = (GLOBAL v2d*)a.zvx;
    g.vyz = (GLOBAL v2d*)a.vyz;
    g.cs = (GLOBAL int32_t*)a.cs;
    return g;
}

__device__ inline v2d mk2(double a, double b) {
    v2d r;
    r.x = a;
    r.y = b;
    return r;
}

// Particle i's position (two 16-byte loads; the second also holds vx).
__device__ inline void ld_pos(const GArrays& A, int64_t i, double& x, double& y, double& z) {
    const v2d a = A.xy[i], b = A.zvx[i];
    x = a.x;
    y = a.y;
    z = b.x;
}
// Particle i's velocity (two 16-byte loads; the first also holds z).
__device__ inline void ld_vel(const GArrays& A, int64_t i, double& vx, double& vy, double& vz) {
    const v2d b = A.zvx[i], c = A.vyz[i];
    vx = b.y;
    vy = c.x;
    vz = c.y;
}
__device__ inline void ld_all(const GArrays& A, int64_t i, double& x, double& y, double& z, double& vx, double& vy,
                              double& vz) {
    const v2d a = A.xy[i], b = A.zvx[i], c = A.vyz[i];
    x = a.x;
    y = a.y;
    z = b.x;
    vx = b.y;
    vy = c.x;
    vz = c.y;
}
__device__ inline void st_all(const GArrays& A, int64_t i, double x, double y, double z, double vx, double vy,
                              double vz) {
    A.xy[i] = mk2(x, y);
    A.zvx[i] = mk2(z, vx);
    A.vyz[i] = mk2(vy, vz);
}
// Particle s of S to slot d of D, all fields.
__device__ inline void copy_particle(const GArrays& D, int64_t d, const GArrays& S, int64_t s) {
    D.xy[d] = S.xy[s];
    D.zvx[d] = S.zvx[s];
    D.vyz[d] = S.vyz[s];
    D.cs[d] = S.cs[s];
}

// Chunk of the population for this block, in one of two orders; both take
// the population from the top down, so that its tail - a step's freshly
// injected particles, one boundary cell per lane and the slowest chunks of
// the pass - is dispatched first instead of last.
// - XCD ranges (xcd = true): XCD x = b % 8 (blocks are dispatched
//   round-robin over the 8 XCDs) owns the contiguous range [x G/8, (x+1) G/8),
//   so its cells and field dofs stay in that XCD's L2.
// - Plain (xcd = false): logical block G-1-b, so the 8 XCDs work side by
//   side on one narrow window of the population and the tail is shared out
//   among them.  The fused step uses it except when it reads a fresh cell
//   order (measured: 2-3% faster per step on an ordered population, 9% slower
//   on the step that gathers through a new permutation).
__device__ inline int64_t logical_block(bool xcd, int64_t b, int64_t G) {
    if (xcd && G % 8 == 0) return (b % 8) * (G / 8) + (G / 8 - 1 - b / 8);
    return G - 1 - b;
}

// Block b of G particle blocks (the fused step's grid also holds the
// injection generator's blocks).
__device__ inline void chunk_range(int64_t n, int64_t& s, int64_t& e, bool xcd, int64_t b, int64_t G) {
    int64_t per = (n + G - 1) / G;
    per = (per + blockDim.x - 1) / blockDim.x * blockDim.x;
    s = logical_block(xcd, b, G) * per;
    e = s + per < n ? s + per : n;
}

__device__ inline void chunk_range(int64_t n, int64_t& s, int64_t& e, bool xcd = true) {
    chunk_range(n, s, e, xcd, blockIdx.x, gridDim.x);
}

// Barycentric coordinates (reference population.h:885-893).
__device__ inline void bary4(const double* A, double x, double y, double z, double l[4]) {
    l[0] = ((A[0] + A[1] * x) + A[2] * y) + A[3] * z;
    l[1] = ((A[4] + A[5] * x) + A[6] * y) + A[7] * z;
    l[2] = ((A[8] + A[9] * x) + A[10] * y) + A[11] * z;
    l[3] = ((1 - l[0]) - l[1]) - l[2];
}

// Visibility walk (reference relocate_fast, population.h:539-569): from cell c,
// move through the FIRST facet (in local order) whose plane gives proj > 0;
// return the cell where no facet fires, or -bnd_id when the walk leaves the
// mesh.  Iterative with a hop cap instead of the reference's recursion.
__device__ inline int32_t walk(const DevMesh& M, int32_t c, double x, double y, double z) {
    for (int h = 0; h < WALK_MAX_HOPS; ++h) {
        const double* p = M.planes[c].p;
        const double p0 = plane_proj(p, x, y, z);
        const double p1 = plane_proj(p + 4, x, y, z);
        const double p2 = plane_proj(p + 8, x, y, z);
        const double p3 = plane_proj(p + 12, x, y, z);
        const int f = p0 > 0 ? 0 : (p1 > 0 ? 1 : (p2 > 0 ? 2 : (p3 > 0 ? 3 : -1)));
        if (f < 0) return c;
        const int32_t nb = M.rec[c].nbr[f];
        if (nb < 0) return nb;
        c = nb;
    }
    return WALK_FAILED;
}

// The same walk with the facet tests decided by the barycentric coordinates
// where the mesh margin makes them certain.  With m = inside_tol *
// max(1, sum |lambda_j|): lambda_i > m implies the computed proj_i < 0 and
// lambda_i < -m implies the computed proj_i > 0 (the margin's rounding bounds
// grow at most linearly with sum |lambda_j|, which bounds the exact
// sum |mu_j|); only for |lambda_i| <= m is the stored plane evaluated, as in
// walk().  So every decision, and the result, equals walk()'s.  The record
// of each visited cell (barycentric rows, dofs, neighbours: one 128-byte
// line) serves both the decisions and, at the final cell, the deposit: l and
// dof return lambda and the dofs there.  Points outside the coordinate range
// the margin was derived for take walk() itself.
__device__ inline int32_t walk_bary(const DevMesh& M, int32_t c, double x, double y, double z, double l[4],
                                    int32_t dof[4]) {
    const double X = M.coord_bound;
    if (!(M.inside_tol > 0.0 && fabs(x) <= X && fabs(y) <= X && fabs(z) <= X)) {
        const int32_t c2 = walk(M, c, x, y, z);
        if (c2 >= 0) {
            const CellRec& R = M.rec[c2];
            bary4(R.bary, x, y, z, l);
            for (int k = 0; k < 4; ++k) dof[k] = R.dof[k];
        }
        return c2;
    }
    for (int h = 0; h < WALK_MAX_HOPS; ++h) {
        const CellRec& R = M.rec[c];
        bary4(R.bary, x, y, z, l);
        const double S = (fabs(l[0]) + fabs(l[1])) + (fabs(l[2]) + fabs(l[3]));
        const double m = M.inside_tol * (S > 1.0 ? S : 1.0);
        int f = -1;
        for (int i = 0; i < 4; ++i) {
            bool pos;
            if (l[i] > m)
                pos = false;
            else if (l[i] < -m)
                pos = true;
            else
                pos = plane_proj(M.planes[c].p + 4 * i, x, y, z) > 0.0;
            if (pos) {
                f = i;
                break;
            }
        }
        if (f < 0) {
            for (int k = 0; k < 4; ++k) dof[k] = R.dof[k];
            return c;
        }
        const int32_t nb = R.nbr[f];
        if (nb < 0) return nb;
        c = nb;
    }
    return WALK_FAILED;
}

// Near-face test (counted and reported, never acted on): some facet plane of
// cell c gives |proj| < NEAR_FACE_DIST at the point, proj evaluated as the
// walk evaluates it.  l: the computed barycentric coordinates there; inside
// the margin's coordinate range only points with a coordinate below
// near_lam can be that close, so the planes are read for those alone.
__device__ inline bool near_face_test(const DevMesh& M, int32_t c, double x, double y, double z,
                                      const double l[4], bool in_range) {
    if (in_range) {
        const double lm = fmin(fmin(l[0], l[1]), fmin(l[2], l[3]));
        if (!(lm < M.near_lam)) return false;
    }
    const double* p = M.planes[c].p;
    bool nf = false;
    for (int i = 0; i < 4; ++i) nf = nf || fabs(plane_proj(p + 4 * i, x, y, z)) < NEAR_FACE_DIST;
    return nf;
}

// Adds the wavefront's near-face particles to S->n_near (one atomic per
// wavefront).  Every lane of the wavefront must call this together.
__device__ inline void count_near(DevState* __restrict__ S, bool nf) {
    const unsigned long long m = __ballot(nf);
    if (m && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1)
        atomicAdd((unsigned long long*)&S->n_near, (unsigned long long)__popcll(m));
}

// CG1 deposit of one particle per lane (distribute_cg1, distributor.h:189-203:
// accum_i += q*lambda_i, then rho[dof_i] += accum_i).  key = cell, or -1 for
// lanes with nothing to deposit.  mode 0: four FP64 atomics per particle.
// mode 1: lanes of a wavefront that hold the same cell (runs of equal keys,
// which cell-sorted particles form) first sum their contributions with a
// segmented reduction across lanes, then the run's first lane issues the four
// atomics.  Every lane of the wavefront must call this together.
// Removal bookkeeping for the particles of one wavefront that left the mesh
// (key >= 0): one atomic per wavefront reserves their slots in the hole list,
// and one atomic per distinct (species, boundary) adds to the absorbed
// counts.  Device-scope atomics on a single address serialize, so per-lane
// atomics here would cost a memory round trip per leaving particle.
// Every lane of the wavefront must call this together.
constexpr int32_t ABS_LOST = 0x7fffffff;

__device__ inline void record_leavers(DevState* __restrict__ S, int32_t* __restrict__ holes,
                                      unsigned long long* __restrict__ absorbed, int32_t key, int64_t i) {
    const unsigned long long lmask = __ballot(key >= 0);
    if (lmask == 0ull) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)lmask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd((unsigned long long*)&S->n_holes, (unsigned long long)__popcll(lmask));
    base = __shfl(base, leader, 64);
    if (key >= 0) holes[base + __popcll(lmask & ((1ull << lane) - 1ull))] = (int32_t)i;
    unsigned long long left = lmask;
    while (left) {
        const int l0 = __ffsll((long long)left) - 1;
        const int32_t k0 = __shfl(key, l0, 64);
        const unsigned long long same = __ballot(key == k0) & left;
        if (lane == l0) {
            if (k0 == ABS_LOST)
                atomicAdd((unsigned long long*)&S->n_lost, (unsigned long long)__popcll(same));
            else
                atomicAdd(&absorbed[k0], (unsigned long long)__popcll(same));
        }
        left &= ~same;
    }
}

// Block-local accumulation table in LDS (deposit mode 2): open addressing on
// the dof id, values summed with LDS FP64 atomics, flushed to HBM with one
// global atomic per occupied slot.
constexpr int LDS_SLOTS = 2048;
constexpr int32_t LDS_EMPTY = -1;
constexpr size_t LDS_TABLE_BYTES = LDS_SLOTS * (sizeof(double) + sizeof(int32_t)) + 16;

struct LdsTable {
    double* val;
    int32_t* key;
    int32_t* count;
};

__device__ inline LdsTable lds_table() {
    extern __shared__ double dyn_smem[];
    LdsTable t;
    t.val = dyn_smem;
    t.key = (int32_t*)(dyn_smem + LDS_SLOTS);
    t.count = t.key + LDS_SLOTS;
    return t;
}

__device__ inline void lds_clear(LdsTable t) {
    for (int k = threadIdx.x; k < LDS_SLOTS; k += blockDim.x) {
        t.key[k] = LDS_EMPTY;
        t.val[k] = 0.0;
    }
    if (threadIdx.x == 0) *t.count = 0;
}

__device__ inline void lds_add(LdsTable t, double* __restrict__ rho_raw, int32_t dof, double v) {
    // Linear probing from the hashed slot.  One exit for the common case (the
    // slot found or claimed within the probe limit), so the add below runs
    // once for the whole wavefront.
    uint32_t h = ((uint32_t)dof * 0x9E3779B1u) >> (32 - 11);  // 11 = log2(LDS_SLOTS)
    int probe = 0;
#pragma unroll 1
    for (;;) {
        int32_t k = __hip_atomic_load(&t.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (k == LDS_EMPTY) {
            k = atomicCAS(&t.key[h], LDS_EMPTY, dof);
            if (k == LDS_EMPTY) {
                atomicAdd(t.count, 1);
                k = dof;
            }
        }
        if (k == dof) break;
        h = (h + 1) & (LDS_SLOTS - 1);
        if (++probe == 16) {
            unsafeAtomicAdd(&rho_raw[dof], v);  // table region full: straight to HBM
            return;
        }
    }
    unsafeAtomicAdd(&t.val[h], v);
}

// Every thread of the block must call this (it synchronizes).
__device__ inline void lds_flush(LdsTable t, double* __restrict__ rho_raw) {
    __syncthreads();
    for (int k = threadIdx.x; k < LDS_SLOTS; k += blockDim.x) {
        const int32_t d = t.key[k];
        if (d != LDS_EMPTY) {
            unsafeAtomicAdd(&rho_raw[d], t.val[k]);
            t.key[k] = LDS_EMPTY;
            t.val[k] = 0.0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *t.count = 0;
    __syncthreads();
}

// ready: lpre/dpre already hold lambda at (x, y, z) in cell `key` and its dofs.
// Move of a double from another lane with DPP (both 32-bit halves with the
// same control; lanes without a source read 0).
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Segmented inclusive sums over runs of lanes: afterwards lane L holds the
// sum of v over lanes head..L of its run (head: the run's first lane).  The
// steps of a wavefront scan on GFX9 DPP: shifts by 1, 2, 4, 8 inside each
// 16-lane row, then the last lane of row 0 (2) broadcast into row 1 (3),
// then lane 31 into rows 2 and 3; a step adds only a source lane of the
// same run.  All lanes of the wavefront must be active.
__device__ inline void run_scan4(double v[4], int lane, int head) {
    const int r0 = lane & ~15;  // first lane of this lane's row
#define PUNC_RUN_STEP(CTRL, ROWM, SRC_OK)                               \
    {                                                                   \
        const bool ok = (SRC_OK);                                       \
        for (int k = 0; k < 4; ++k) {                                   \
            const double t = dpp_f64<CTRL, ROWM>(v[k]);                 \
            if (ok) v[k] += t;                                          \
        }                                                               \
    }
    PUNC_RUN_STEP(0x111, 0xf, lane - 1 >= head && lane - 1 >= r0)
    PUNC_RUN_STEP(0x112, 0xf, lane - 2 >= head && lane - 2 >= r0)
    PUNC_RUN_STEP(0x114, 0xf, lane - 4 >= head && lane - 4 >= r0)
    PUNC_RUN_STEP(0x118, 0xf, lane - 8 >= head && lane - 8 >= r0)
    PUNC_RUN_STEP(0x142, 0xa, ((lane >> 4) & 1) && r0 - 1 >= head)
    PUNC_RUN_STEP(0x143, 0xc, lane >= 32 && 31 >= head)
#undef PUNC_RUN_STEP
}

template <int mode>
__device__ inline void deposit_cg1(const DevMesh& M, double* __restrict__ rho_raw, int32_t key, double x,
                                   double y, double z, double q, bool ready, const double* lpre,
                                   const int32_t* dpre) {
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    int32_t dof[4] = {0, 0, 0, 0};
    if (key >= 0) {
        if (ready) {
            for (int k = 0; k < 4; ++k) {
                d[k] = q * lpre[k];
                dof[k] = dpre[k];
            }
        } else {
            const CellRec& R = M.rec[key];
            double l[4];
            bary4(R.bary, x, y, z, l);
            for (int k = 0; k < 4; ++k) {
                d[k] = q * l[k];
                dof[k] = R.dof[k];
            }
        }
    }
    if (mode == 3) {
        if (key < 0) return;
        const LdsTable t = lds_table();
        for (int k = 0; k < 4; ++k) lds_add(t, rho_raw, dof[k], d[k]);
        return;
    }
    if (mode == 4) {
        // Lanes of a quad with the same key are first summed into the first
        // lane of their stretch (two DPP steps inside the quad: from lane+1,
        // then from lane+2), so that fewer lanes of a wavefront add to the
        // same LDS slot at once; those lanes then add as in mode 3.
        const int ql = threadIdx.x & 3;
        const int32_t k1 = __builtin_amdgcn_update_dpp(0, key, 0xF9, 0xf, 0xf, true);  // quad_perm(1,2,3,3)
        const int32_t k2 = __builtin_amdgcn_update_dpp(0, key, 0xEE, 0xf, 0xf, true);  // quad_perm(2,3,2,3)
        const int32_t km = __builtin_amdgcn_update_dpp(0, key, 0x90, 0xf, 0xf, true);  // quad_perm(0,0,1,2)
        const bool s1 = ql < 3 && k1 == key;
        const bool s2 = ql < 2 && s1 && k2 == key;
        for (int k = 0; k < 4; ++k) {
            const double t1 = dpp_f64<0xF9, 0xf>(d[k]);
            d[k] += s1 ? t1 : 0.0;
        }
        for (int k = 0; k < 4; ++k) {
            const double t2 = dpp_f64<0xEE, 0xf>(d[k]);
            d[k] += s2 ? t2 : 0.0;
        }
        const bool head = ql == 0 || km != key;
        if (key < 0 || !head) return;
        const LdsTable t = lds_table();
        for (int k = 0; k < 4; ++k) lds_add(t, rho_raw, dof[k], d[k]);
        return;
    }
    if (mode >= 1) {
        // A run is a maximal stretch of consecutive lanes with one key; the
        // same key may reappear in a later run, which is summed separately.
        // The run's last lane ends up with the run's sums and deposits them.
        const int lane = threadIdx.x & 63;
        const int32_t kprev = __shfl_up(key, 1, 64);
        const bool head = lane == 0 || kprev != key;
        const unsigned long long heads = __ballot(head);
        const unsigned long long upto = lane == 63 ? heads : (heads & ((2ull << lane) - 1ull));
        const int hl = 63 - __clzll((long long)upto);  // first lane of this lane's run
        const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
        run_scan4(d, lane, hl);
        if (key < 0 || !tail) return;
        if (mode == 2) {
            const LdsTable t = lds_table();
            for (int k = 0; k < 4; ++k) lds_add(t, rho_raw, dof[k], d[k]);
            return;
        }
    } else if (key < 0) {
        return;
    }
    for (int k = 0; k < 4; ++k) unsafeAtomicAdd(&rho_raw[dof[k]], d[k]);
}

__device__ inline double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int k = 0; k < nw; ++k) r += sh[k];
    }
    __syncthreads();
    return r;
}

// Per-cell copy of the CG1 field: the 3 components at the cell's 4 dofs in
// local vertex order (12 doubles, 128-byte stride).  Reading it instead of
// E[dof] removes one dependent gather from the push.
constexpr int CELL_FIELD_STRIDE = 16;

__device__ inline void load_cell_field(const double* __restrict__ cellE, int32_t c, double ef[12]) {
    typedef double v2d __attribute__((ext_vector_type(2)));
    const GLOBAL v2d* p = (const GLOBAL v2d*)(cellE + (int64_t)c * CELL_FIELD_STRIDE);
    for (int m = 0; m < 6; ++m) {
        const v2d t = p[m];
        ef[2 * m] = t.x;
        ef[2 * m + 1] = t.y;
    }
}

// One velocity update at barycentric coordinates l in a cell whose field
// values are ef.  PUSH 1: accel_cg1, 2: boris_cg1 (coefficients from SpCoef).
template <int PUSH>
__device__ inline void push_velocity(const SpCoef& K, const double l[4], const double ef[12], double& vx,
                                     double& vy, double& vz, double& ke) {
    // Ei = sum_i coeffs[i]*values[j*4+i] starting from 0.0 (pusher.h:188-192, 324-328)
    double Ex = 0.0, Ey = 0.0, Ez = 0.0;
    Ex += l[0] * ef[0]; Ex += l[1] * ef[3]; Ex += l[2] * ef[6]; Ex += l[3] * ef[9];
    Ey += l[0] * ef[1]; Ey += l[1] * ef[4]; Ey += l[2] * ef[7]; Ey += l[3] * ef[10];
    Ez += l[0] * ef[2]; Ez += l[1] * ef[5]; Ez += l[2] * ef[8]; Ez += l[3] * ef[11];
    if (PUSH == 1) {
        // accel_cg1 (pusher.h:194-196): Ei *= dt*(q/m); KE += 0.5*m*v*(v+Ei); v += Ei
        Ex = Ex * K.a; Ey = Ey * K.a; Ez = Ez * K.a;
        ke += (K.half_m * vx) * (vx + Ex);
        ke += (K.half_m * vy) * (vy + Ey);
        ke += (K.half_m * vz) * (vz + Ez);
        vx += Ex; vy += Ey; vz += Ez;
    } else {
        // boris_cg1 (pusher.h:343-368)
        const double mx = vx + K.a * Ex, my = vy + K.a * Ey, mz = vz + K.a * Ez;
        ke += (K.half_m * mx) * mx;
        ke += (K.half_m * my) * my;
        ke += (K.half_m * mz) * mz;
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
    }
}

// ---------------------------------------------------------------------------
// The particle kernel.  PUSH: 0 none, 1 accel_cg1, 2 boris_cg1.  MOVE: 0 none,
// 1 move, 2 move_periodic.  LOCATE: relocation walk with removal bookkeeping.
// DEPOSIT: CG1 charge deposit q*lambda_i at the (new) cell.
// ---------------------------------------------------------------------------
template <int PUSH, int MOVE, bool LOCATE, bool DEPOSIT, int DMODE>
__global__ void __launch_bounds__(256)
k_particles(const DevMesh M, DevState* __restrict__ S, const SpCoef* __restrict__ coef,
            const double* __restrict__ cellE, const double dt, const double Lx, const double Ly,
            const double Lz, const int cell_bits, double* __restrict__ rho_raw,
            double* __restrict__ ke_part, int32_t* __restrict__ holes,
            unsigned long long* __restrict__ absorbed) {
    __shared__ double sh[8];
    const GArrays A = gview(S->cur);
    const int64_t n = S->n;
    int64_t s, e;
    chunk_range(n, s, e);
    if (s >= e) {  // block-uniform: nothing in this block's chunk
        if (PUSH != 0 && threadIdx.x == 0) ke_part[blockIdx.x] = 0.0;
        return;
    }
    const int32_t cmask = (1 << cell_bits) - 1;
    double ke = 0.0;
    constexpr bool use_lds = DEPOSIT && DMODE >= 2;
    if (use_lds) {
        lds_clear(lds_table());
        __syncthreads();
    }
    // Block-uniform trip count so that whole wavefronts reach the deposit's
    // cross-lane reduction together; lanes past the chunk end are inactive.
    for (int64_t base = s; base < e; base += blockDim.x) {
        const int64_t i = base + threadIdx.x;
        const bool active = i < e;
        int32_t dkey = -1;  // deposit cell of this lane (-1: nothing to deposit)
        double dx_ = 0.0, dy_ = 0.0, dz_ = 0.0, dq = 0.0;
        int32_t leave_key = -1;  // particle left: species*MAX_BND + bnd, or ABS_LOST
        bool dready = false;     // dl/ddof already hold the deposit weights and dofs
        bool near = false;       // relocated near a facet of its cell
        double dl[4] = {0.0, 0.0, 0.0, 0.0};
        int32_t ddof[4] = {0, 0, 0, 0};
        if (active) {
        const int32_t cs = A.cs[i];
        const int sp = cs >> cell_bits;
        int32_t c = cs & cmask;
        double x, y, z;
        double vx = 0, vy = 0, vz = 0;
        if (PUSH != 0 || MOVE != 0)
            ld_all(A, i, x, y, z, vx, vy, vz);
        else
            ld_pos(A, i, x, y, z);
        if (PUSH != 0) {
            const CellRec& R = M.rec[c];
            double l[4], ef[12];
            bary4(R.bary, x, y, z, l);
            load_cell_field(cellE, c, ef);
            push_velocity<PUSH>(coef[sp], l, ef, vx, vy, vz, ke);
        }
        if (MOVE != 0) {
            // move / move_periodic (pusher.h:658, 676-677)
            x += dt * vx;
            y += dt * vy;
            z += dt * vz;
            if (MOVE == 2) {
                x -= Lx * floor(x / Lx);
                y -= Ly * floor(y / Ly);
                z -= Lz * floor(z / Lz);
            }
            A.xy[i] = mk2(x, y);
        }
        // the pairs holding what changed: (z, vx) after a push or a move,
        // (vy, vz) after a push
        if (PUSH != 0 || MOVE != 0) A.zvx[i] = mk2(z, vx);
        if (PUSH != 0) A.vyz[i] = mk2(vy, vz);
        bool alive = true;
        if (LOCATE) {
            // Certainly-inside shortcut: if all barycentric coordinates of the
            // new position in the old cell exceed the mesh's rounding margin,
            // every facet projection is negative and the walk would return
            // the old cell; the coordinates are then also the deposit weights.
            bool stay = false;
            if (M.inside_tol > 0.0) {
                const CellRec& R = M.rec[c];
                bary4(R.bary, x, y, z, dl);
                const double tol = M.inside_tol;
                stay = dl[0] > tol && dl[1] > tol && dl[2] > tol && dl[3] > tol;
                if (stay)
                    for (int k = 0; k < 4; ++k) ddof[k] = R.dof[k];
            }
            const int32_t c2 = stay ? c : walk(M, c, x, y, z);
            dready = stay;
            if (c2 >= 0) {
                const bool in_range = M.inside_tol > 0.0 && fabs(x) <= M.coord_bound &&
                                      fabs(y) <= M.coord_bound && fabs(z) <= M.coord_bound;
                double ln[4];
                if (stay) {
                    for (int k = 0; k < 4; ++k) ln[k] = dl[k];
                } else {
                    bary4(M.rec[c2].bary, x, y, z, ln);
                }
                near = near_face_test(M, c2, x, y, z, ln, in_range);
            }
            if (c2 < 0) {
                alive = false;
                leave_key = (c2 == WALK_FAILED) ? ABS_LOST : sp * MAX_BND + (-c2);
                A.cs[i] = c2;
            } else if (c2 != c) {
                c = c2;
                A.cs[i] = (sp << cell_bits) | c2;
            }
        }
        if (DEPOSIT && alive) {
            dkey = c;
            dx_ = x;
            dy_ = y;
            dz_ = z;
            dq = coef[sp].q;
        }
        }  // active
        if (LOCATE) {
            record_leavers(S, holes, absorbed, leave_key, i);
            count_near(S, near);
        }
        if (DEPOSIT) deposit_cg1<DMODE>(M, rho_raw, dkey, dx_, dy_, dz_, dq, dready, dl, ddof);
        if (use_lds) {
            __syncthreads();
            if (*lds_table().count > LDS_SLOTS * 5 / 8) lds_flush(lds_table(), rho_raw);
        }
    }
    if (use_lds) lds_flush(lds_table(), rho_raw);
    if (PUSH != 0) {
        const double t = block_sum(ke, sh);
        if (threadIdx.x == 0) ke_part[blockIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// Fused step: push, move, relocation with removal bookkeeping, and deposit in
// one pass over the population.
//
// Phase 1, one lane per particle, software-pipelined: while a group of
// particles is processed, the next group's positions, velocities and cell
// fields are already in flight.  Vector-memory results arrive in issue order,
// so the next group's loads are issued after the current group's cell-record
// and field gathers (waiting for those does not wait for the prefetch), and
// nothing the current group needs is loaded after them (species
// coefficients are read from LDS).  A particle whose barycentric coordinates
// at its new position in its old cell all exceed the mesh margin stays there
// (the reference's walk, relocate_fast population.h:539-569, would return
// that cell) and deposits with those coordinates; every other particle goes,
// with its new position, to this block's segment of the walk list (starting
// at the block's chunk start, so it cannot overflow; the counter lives in
// LDS).  There is no block barrier in this loop: the deposit table in LDS is
// only flushed at the end (a full table sends further contributions straight
// to HBM).
// Phase 2: the same block runs the walks of its list, the removal
// bookkeeping of the particles that leave, and their deposit.
// ---------------------------------------------------------------------------
struct WalkRec {
    double x, y, z;
    int32_t i, cs;
};
static_assert(sizeof(WalkRec) == 32, "WalkRec layout");

constexpr int STEP_WALK_LDS = 512;  // most walk records a block keeps in LDS

struct PartRegs {
    double x, y, z, vx, vy, vz;
    int32_t cs;
};

__device__ inline void load_particle(const GArrays& A, int64_t i, PartRegs& p) {
    // Ordinary loads: measured 4% faster for the whole kernel than
    // non-temporal ones on the 64M-particle workload; the stores of the
    // results stay non-temporal (ordinary stores were 7% slower).
    p.cs = A.cs[i];
    ld_all(A, i, p.x, p.y, p.z, p.vx, p.vy, p.vz);
}

// All four barycentric coordinates above the margin: the point is certainly
// inside (every computed facet projection is negative).
__device__ inline bool certainly_inside(const double l[4], double tol) {
    return l[0] > tol && l[1] > tol && l[2] > tol && l[3] > tol;
}

// The fused step's dynamic LDS: the deposit table (mode 2), then the species
// coefficients.
template <int DMODE>
__device__ inline SpCoef* step_coef_lds() {
    extern __shared__ double dyn_smem[];
    return (SpCoef*)(dyn_smem + (DMODE >= 2 ? LDS_TABLE_BYTES / 8 : 0));
}

// The injection generator (defined with the injection kernels below).
template <bool MAXW>
__device__ __forceinline__ void inject_gen_body(const DevMesh& M, const DevInject& I, const DevState* __restrict__ S,
                                                const SpInject* __restrict__ spi, const int64_t* __restrict__ soff,
                                                const int32_t* __restrict__ spair, uint64_t seed, double dt,
                                                int cell_bits, int64_t cand_cap, double* __restrict__ cx,
                                                int32_t* __restrict__ ccs, int32_t* __restrict__ kc,
                                                unsigned long long* __restrict__ errs, int64_t blk, int64_t nblk);

// The Maxwellian generator's arguments when its blocks share the fused
// step's grid: blocks [0, blocks) generate, the others are the step's
// particle blocks in order.  One kernel instead of a kernel on a second
// stream spares the step the join of the two streams (about 6 us per step,
// which pays on small populations).
struct GenArgs {
    DevInject I;
    const SpInject* spi;
    const int64_t* soff;
    const int32_t* spair;
    uint64_t seed;
    double dt;
    int64_t cand_cap;
    double* cx;
    int32_t* ccs;
    int32_t* kc;
    unsigned long long* errs;
    int blocks;
};

// PE: also the particle potential energy at the step's starting positions
// (particle_potential_energy_cg1, diagnostics.h:379-412), phi read from
// slots 12..15 of the cell's field row: per-block sums into pe_part.
template <int PUSH, int MOVE, int DMODE, bool PE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_step_main(const DevMesh M, DevState* __restrict__ S, const SpCoef* __restrict__ gcoef, const int nsp,
            const double* __restrict__ cellE, const double dt, const double Lx, const double Ly,
            const double Lz, const int cell_bits, double* __restrict__ rho_raw,
            double* __restrict__ ke_part, int32_t* __restrict__ holes, unsigned long long* __restrict__ absorbed,
            WalkRec* __restrict__ wrec, double* __restrict__ pe_part, const int wl_cap, const GenArgs G) {
    typedef double v2d __attribute__((ext_vector_type(2)));
    typedef int32_t v4i __attribute__((ext_vector_type(4)));
    __shared__ double sh[8];
    __shared__ int32_t wcnt;
    if ((int64_t)blockIdx.x < G.blocks) {  // block-uniform
        inject_gen_body<true>(M, G.I, S, G.spi, G.soff, G.spair, G.seed, G.dt, cell_bits, G.cand_cap, G.cx, G.ccs,
                              G.kc, G.errs, blockIdx.x, G.blocks);
        return;
    }
    const int64_t blk = blockIdx.x - G.blocks, nblk = gridDim.x - G.blocks;
    // Species coefficients in LDS: read per lane (species differ within a
    // wavefront) without adding to the vector-memory queue that the
    // prefetched particle loads sit in.
    SpCoef* coef = step_coef_lds<DMODE>();
    for (int k = threadIdx.x; k < nsp * (int)(sizeof(SpCoef) / 8); k += blockDim.x)
        ((double*)coef)[k] = ((const double*)gcoef)[k];
    // The block's walk list: its first wl_cap records in LDS (after the
    // coefficients), the rest in the block's segment of wrec.  Phase 1 writes
    // it and phase 2 of the same block reads it, so LDS spares the list two
    // vector-memory stores per wavefront iteration and their reads.
    WalkRec* wl = (WalkRec*)(coef + nsp);
    const GArrays A = gview(S->cur);
    const int64_t n = S->n;
    // A pending cell order: particle j of this step is cur[perm[j]], and the
    // results go to the alternate buffers at j (swapped in after the kernel).
    // That step takes the XCD-range chunk order, every other one the plain.
    const bool permuted = S->perm_pending != 0;
    int64_t s, e;
    chunk_range(n, s, e, permuted, blk, nblk);
    if (s >= e) {  // block-uniform: nothing in this block's chunk
        if (threadIdx.x == 0) {
            ke_part[blk] = 0.0;
            if (PE) pe_part[blk] = 0.0;
        }
        return;
    }
    const GLOBAL uint32_t* perm = (const GLOBAL uint32_t*)S->perm;
    const GArrays O = permuted ? gview(S->alt) : A;
    const int32_t cmask = (1 << cell_bits) - 1;
    const double tol = M.inside_tol;
    const double X = M.coord_bound;
    double ke = 0.0, pe = 0.0;
    constexpr bool use_lds = DMODE >= 2;
    if (threadIdx.x == 0) wcnt = 0;
    if (use_lds) lds_clear(lds_table());
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t last = e - 1;
    // Source index of a particle: through the pending cell order, if any.
    // The order is loaded one group ahead of the particle data it addresses.
    auto clampi = [&](int64_t j) { return j < e ? j : last; };
    PartRegs nxt;
    uint32_t src_next = 0;
    {
        const int64_t j0 = clampi(s + threadIdx.x);
        load_particle(A, permuted ? (int64_t)perm[j0] : j0, nxt);
        if (permuted) src_next = perm[clampi(s + threadIdx.x + blockDim.x)];
    }
    int grp = 0;  // groups of this chunk done
    for (int64_t base = s; base < e; base += blockDim.x) {
        const int64_t i = base + threadIdx.x;
        const bool active = i < e;
        const PartRegs p = nxt;
        const int sp = p.cs >> cell_bits;
        const int32_t c = p.cs & cmask;
        // Cell record and field copy of this particle's cell.
        const GLOBAL v2d* rp = (const GLOBAL v2d*)(M.rec + c);
        const GLOBAL v2d* ep = (const GLOBAL v2d*)(cellE + (int64_t)c * CELL_FIELD_STRIDE);
        double rb[12], ef[12];
        int32_t rd[8];  // dof[4], nbr[4] (the neighbours for a walker's first hop)
        for (int m = 0; m < 6; ++m) {
            const v2d t = rp[m];
            rb[2 * m] = t.x;
            rb[2 * m + 1] = t.y;
        }
        // The potential-energy variant loads no neighbour ids (no first hop
        // below): its four more field values would otherwise take it past
        // 128 VGPRs, three wavefronts per SIMD instead of four (measured:
        // 2.20 -> 2.03 ms per step on 64M particles with the hop dropped).
        for (int m = 0; m < (PE ? 1 : 2); ++m) {
            const v4i t = ((const GLOBAL v4i*)(rp + 6))[m];
            rd[4 * m] = t.x;
            rd[4 * m + 1] = t.y;
            rd[4 * m + 2] = t.z;
            rd[4 * m + 3] = t.w;
        }
        if (PE)
            for (int m = 4; m < 8; ++m) rd[m] = -1;
        for (int m = 0; m < 6; ++m) {
            const v2d t = ep[m];
            ef[2 * m] = t.x;
            ef[2 * m + 1] = t.y;
        }
        double ph[4];
        if (PE) {
            const v2d a = ep[6], b = ep[7];
            ph[0] = a.x;
            ph[1] = a.y;
            ph[2] = b.x;
            ph[3] = b.y;
        }
        __builtin_amdgcn_sched_barrier(0);
        // Prefetch: the order entries of the group after next, then the next
        // group (clamped index: lanes past the end reload the chunk's last
        // particle and discard it).  The entries go first so that taking
        // them over waits for nothing issued later.
        const int64_t inx = clampi(i + blockDim.x);
        const uint32_t src_after = permuted ? perm[clampi(i + 2 * (int64_t)blockDim.x)] : 0u;
        load_particle(A, permuted ? (int64_t)src_next : inx, nxt);
        src_next = src_after;
        __builtin_amdgcn_sched_barrier(0);
        double x = p.x, y = p.y, z = p.z, vx = p.vx, vy = p.vy, vz = p.vz;
        double l[4];
        bary4(rb, x, y, z, l);
        const SpCoef& K = coef[sp];
        if (PE && active) {
            double phi_x = 0.0;
            for (int k = 0; k < 4; ++k) phi_x += l[k] * ph[k];
            pe += 0.5 * K.q * phi_x;
        }
        double ke_l = 0.0;
        push_velocity<PUSH>(K, l, ef, vx, vy, vz, ke_l);
        // move / move_periodic (pusher.h:658, 676-677)
        x += dt * vx;
        y += dt * vy;
        z += dt * vz;
        if (MOVE == 2) {
            x -= Lx * floor(x / Lx);
            y -= Ly * floor(y / Ly);
            z -= Lz * floor(z / Lz);
        }
        if (active) {
            ke += ke_l;
            __builtin_nontemporal_store(mk2(x, y), &O.xy[i]);
            __builtin_nontemporal_store(mk2(z, vx), &O.zvx[i]);
            __builtin_nontemporal_store(mk2(vy, vz), &O.vyz[i]);
            if (permuted) __builtin_nontemporal_store(p.cs, &O.cs[i]);
        }
        double dl[4];
        bary4(rb, x, y, z, dl);
        const bool in_range = tol > 0.0 && fabs(x) <= X && fabs(y) <= X && fabs(z) <= X;
        int32_t dkey = -1;  // deposit cell (-1: nothing to deposit in this phase)
        int32_t ddof[4];
        bool walker = false;
        bool near = false;
        if (active) {
            if (in_range && certainly_inside(dl, tol)) {
                dkey = c;
                for (int k = 0; k < 4; ++k) ddof[k] = rd[k];
                near = near_face_test(M, c, x, y, z, dl, true);
            } else {
                walker = true;
            }
        }
        // A walker's first hop, decided here from the coordinates and the
        // neighbour ids already loaded, as walk_bary decides it at cell c
        // (the same lambda, the same margin, the first facet whose
        // coordinate is below -m): phase 2 then starts at that neighbour and
        // does not fetch this cell's record again.  Undecided cases (a
        // coordinate within the margin before the first clear exit, a point
        // outside the margin's coordinate range, an exit through the
        // boundary) start at c, and so does every walker of the
        // potential-energy variant.
        int32_t wstart = c;
        bool hopped = false;
        if (!PE && walker && in_range) {
            const double Sl = (fabs(dl[0]) + fabs(dl[1])) + (fabs(dl[2]) + fabs(dl[3]));
            const double m = tol * (Sl > 1.0 ? Sl : 1.0);
            int f = -1;
            for (int q = 0; q < 4; ++q) {
                if (dl[q] > m) continue;
                if (dl[q] < -m) f = q;
                break;  // decided exit, or within the margin (undecided)
            }
            if (f >= 0) {
                const int32_t nb = f == 0 ? rd[4] : f == 1 ? rd[5] : f == 2 ? rd[6] : rd[7];
                if (nb >= 0) {
                    wstart = nb;
                    hopped = true;
                }
            }
        }
        count_near(S, near);
        // Wavefront-aggregated append to this block's walk-list segment.
        const unsigned long long wm = __ballot(walker);
        if (wm) {
            const int leader = __ffsll((long long)wm) - 1;
            int wb = 0;
            if (lane == leader) wb = atomicAdd(&wcnt, __popcll(wm));
            wb = __shfl(wb, leader, 64);
            if (walker) {
                WalkRec wr;
                wr.x = x;
                wr.y = y;
                wr.z = z;
                wr.i = hopped ? ~(int32_t)i : (int32_t)i;  // negative: starts at the neighbour
                wr.cs = (sp << cell_bits) | wstart;
                const int wk = wb + __popcll(wm & ((1ull << lane) - 1ull));
                if (wk < wl_cap)
                    wl[wk] = wr;
                else
                    wrec[s + wk] = wr;
            }
        }
        deposit_cg1<DMODE>(M, rho_raw, dkey, 0.0, 0.0, 0.0, K.q, true, dl, ddof);
        // Table occupancy, every 8 groups before the chunk's last (all
        // block-uniform): a long chunk flushes a table that is 5/8 full
        // instead of sending its further adds to HBM after 16 failed probes.
        // A chunk of 8 groups or fewer never takes this branch.
        if (use_lds && (++grp & 7) == 0 && base + blockDim.x < e &&
            __syncthreads_or(threadIdx.x == 0 && *lds_table().count > LDS_SLOTS * 5 / 8))
            lds_flush(lds_table(), rho_raw);
    }
    const double t = block_sum(ke, sh);  // synchronizes: wcnt and the segment are complete
    if (threadIdx.x == 0) ke_part[blk] = t;
    if (PE) {
        const double tp = block_sum(pe, sh);
        if (threadIdx.x == 0) pe_part[blk] = tp;
    }
    // Phase 2: the walks of this block's list.  A.cs is rewritten only when
    // the cell changes.
    const int64_t we = s + wcnt;
    for (int64_t base = s; base < we; base += blockDim.x) {
        const int64_t j = base + threadIdx.x;
        int32_t dkey = -1;
        double dq = 0.0;
        double dl[4] = {0.0, 0.0, 0.0, 0.0};
        int32_t ddof[4] = {0, 0, 0, 0};
        int32_t leave_key = -1;
        int64_t i = 0;
        bool near = false;
        if (j < we) {
            const WalkRec r = j - s < wl_cap ? wl[j - s] : wrec[j];
            const bool hopped = r.i < 0;  // the walk's first hop was taken in phase 1
            i = hopped ? ~r.i : r.i;
            const int sp = r.cs >> cell_bits;
            const int32_t c = r.cs & cmask;  // where the walk starts
            const int32_t c2 = walk_bary(M, c, r.x, r.y, r.z, dl, ddof);
            if (c2 < 0) {
                leave_key = (c2 == WALK_FAILED) ? ABS_LOST : sp * MAX_BND + (-c2);
                O.cs[i] = c2;
            } else {
                if (hopped || c2 != c) O.cs[i] = (sp << cell_bits) | c2;
                dkey = c2;
                dq = coef[sp].q;
                const bool in_range = tol > 0.0 && fabs(r.x) <= X && fabs(r.y) <= X && fabs(r.z) <= X;
                near = near_face_test(M, c2, r.x, r.y, r.z, dl, in_range);
            }
        }
        record_leavers(S, holes, absorbed, leave_key, i);
        count_near(S, near);
        deposit_cg1<DMODE>(M, rho_raw, dkey, 0.0, 0.0, 0.0, dq, true, dl, ddof);
    }
    if (use_lds) lds_flush(lds_table(), rho_raw);
}

// After the fused step: a step that read through a pending cell order wrote
// the alternate buffers, which become current.
__global__ void k_step_end(DevState* __restrict__ S) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && S->perm_pending) {
        const PopArrays t = S->cur;
        S->cur = S->alt;
        S->alt = t;
        S->perm_pending = 0;
    }
}

__global__ void k_set_perm(DevState* __restrict__ S, const uint32_t* __restrict__ perm) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        S->perm = perm;
        S->perm_pending = 1;
    }
}

// E3: the CG1 field as the caller hands it over, 3 components per dof.
__global__ void k_cell_field(const DevMesh M, const double* __restrict__ E3, double* __restrict__ cellE) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < M.nc; c += (int64_t)gridDim.x * blockDim.x) {
        const CellRec& R = M.rec[c];
        double* o = cellE + c * CELL_FIELD_STRIDE;
        for (int k = 0; k < 4; ++k) {
            const double* v = E3 + 3 * (int64_t)R.dof[k];
            o[3 * k] = v[0];
            o[3 * k + 1] = v[1];
            o[3 * k + 2] = v[2];
        }
        // slots 12..15: phi at the four dofs (k_cell_potential), kept
    }
}

// phi at the cell's four dofs into slots 12..15 of its field row (the fused
// step's potential energy); phi == nullptr clears them.
__global__ void k_cell_potential(const DevMesh M, const double* __restrict__ phi, double* __restrict__ cellE) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < M.nc; c += (int64_t)gridDim.x * blockDim.x) {
        const CellRec& R = M.rec[c];
        double* o = cellE + c * CELL_FIELD_STRIDE + 12;
        for (int k = 0; k < 4; ++k) o[k] = phi ? phi[R.dof[k]] : 0.0;
    }
}

// Packed 3-component field -> one double4 per dof (w = 0).
__global__ void k_expand_field(const double* __restrict__ E3, double4* __restrict__ E, int64_t ndof) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < ndof; j += (int64_t)gridDim.x * blockDim.x)
        E[j] = make_double4(E3[3 * j], E3[3 * j + 1], E3[3 * j + 2], 0.0);
}

// Kinetic energy sum 0.5*m*v.v (reference diagnostics.h:248-265).
__global__ void __launch_bounds__(256)
k_kinetic(DevState* __restrict__ S, const SpCoef* __restrict__ coef, int cell_bits,
          double* __restrict__ ke_part) {
    __shared__ double sh[8];
    const GArrays A = gview(S->cur);
    int64_t s, e;
    chunk_range(S->n, s, e);
    double ke = 0.0;
    for (int64_t i = s + threadIdx.x; i < e; i += blockDim.x) {
        const int sp = A.cs[i] >> cell_bits;
        const double hm = coef[sp].half_m;
        double vx, vy, vz;
        ld_vel(A, i, vx, vy, vz);
        ke += (hm * vx) * vx;
        ke += (hm * vy) * vy;
        ke += (hm * vz) * vz;
    }
    const double t = block_sum(ke, sh);
    if (threadIdx.x == 0) ke_part[blockIdx.x] = t;
}

// Deterministic sum of np per-block partials by one block of 1024 threads
// (eight independent loads in flight per thread); the total in thread 0.
__device__ inline double sum_partials_1024(const double* __restrict__ part, int np, double* sh) {
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int k = threadIdx.x;
    for (; k + 7 * 1024 < np; k += 8 * 1024)
        for (int u = 0; u < 8; ++u) a[u] += part[k + u * 1024];
    for (; k < np; k += 1024) a[0] += part[k];
    const double v = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    return block_sum(v, sh);
}

__global__ void __launch_bounds__(1024) k_sum_partials(const double* __restrict__ part, int np,
                                                       double* __restrict__ out) {
    __shared__ double sh[16];
    const double t = sum_partials_1024(part, np, sh);
    if (threadIdx.x == 0) *out = t;
}

// rho_j = dv_inv_j * raw_j (distributor.h:205-208); raw is left zeroed for
// the next deposit.
__global__ void k_scale_rho(double* __restrict__ raw, const double* __restrict__ dv_inv,
                            double* __restrict__ rho, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        rho[i] = raw[i] * dv_inv[i];
        raw[i] = 0.0;
    }
}

// Last kernel of a fused step: every block scales its share of the deposit
// (and zeroes raw for the next step); block 0 also does the step's
// bookkeeping, which no other block reads: the swap to the buffers a
// permuted step wrote, the KE and PE sums of the per-block partials (the
// order of k_sum_partials), the count after removal and injection, and the
// step counter.  Blocks of 1024 threads.
__global__ void __launch_bounds__(1024)
k_scale_finish(double* __restrict__ raw, const double* __restrict__ dv_inv, double* __restrict__ rho, int64_t ndof,
               DevState* __restrict__ S, const StepClose C) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ndof;
         i += (int64_t)gridDim.x * blockDim.x) {
        rho[i] = raw[i] * dv_inv[i];
        raw[i] = 0.0;
    }
    if (blockIdx.x != 0) return;
    __shared__ double sh[16];
    const double ke = sum_partials_1024(C.ke_part, C.np, sh);
    const double pe = C.pe_part ? sum_partials_1024(C.pe_part, C.np, sh) : NAN;
    if (threadIdx.x == 0) {
        *C.ke_out = ke;
        *C.pe_out = pe;
        if (S->perm_pending) {
            const PopArrays t = S->cur;
            S->cur = S->alt;
            S->alt = t;
            S->perm_pending = 0;
        }
        int64_t n = S->n - S->n_holes;
        S->n_holes = 0;
        S->nH = 0;
        S->nS = 0;
        if (C.nkept) {  // injection (k_inject_finish's bookkeeping)
            int64_t kept = *C.nkept;
            const int64_t room = S->capacity - n;
            if (kept > room) kept = room;
            S->inj_kept = kept;
            n += kept;
        }
        S->n = n;
        if (C.advance) S->step += 1;
    }
}

__global__ void k_fill_f64(double* __restrict__ p, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}
__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// Removal of particles that left the mesh: in-place hole filling.  With L
// removed particles and n' = n - L survivors, each hole below n' receives one
// surviving particle from [n', n).  Mirrors the effect of the reference's
// swap-with-back deletion (population.h:686-692) on the particle set.
// ---------------------------------------------------------------------------
// The particle arrays a step's later kernels work on: the alternate buffers
// when the fused step read through a pending cell order (the swap itself is
// the step's last bookkeeping, k_scale_finish).
__device__ inline PopArrays step_arrays(const DevState* S) { return S->perm_pending ? S->alt : S->cur; }

// One pass: holes below n' and survivors in [n', n) each take a ticket from
// their own counter (the counts are equal); ticket k of both kinds meets in
// slot[k] by atomic exchange, and the second to arrive finds its partner's
// index there, moves the survivor into the hole and empties the slot (slots
// hold -1 between calls).  No block waits for another.  step: the arrays
// after a fused step (before its swap); the count is updated afterwards
// (k_compact_finish, or the step's k_scale_finish).
constexpr int COMPACT_BLOCK = 256;

__global__ void __launch_bounds__(COMPACT_BLOCK)
k_compact(DevState* __restrict__ S, const int32_t* __restrict__ holes, int32_t* __restrict__ slot, int step) {
    __shared__ unsigned int wc[COMPACT_BLOCK / 64];
    __shared__ unsigned long long bbase;
    const int64_t n = S->n, L = S->n_holes, nn = n - L;
    const GArrays A = gview(step ? step_arrays(S) : S->cur);
    const int64_t stride = (int64_t)gridDim.x * COMPACT_BLOCK;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // Tickets for the lanes that want one: one atomic per block (the
    // counters are single addresses, so per-wavefront atomics from a few
    // thousand wavefronts serialize).  Called by the whole block (the loops
    // below are block-uniform).
    auto meet = [&](unsigned long long* ctr, bool want, int32_t idx, bool is_hole) {
        const unsigned long long m = __ballot(want);
        if (lane == 0) wc[w] = (unsigned int)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long tot = 0;
            for (int v = 0; v < COMPACT_BLOCK / 64; ++v) tot += wc[v];
            bbase = tot ? atomicAdd(ctr, tot) : 0ull;
        }
        __syncthreads();
        unsigned long long base = bbase;
        for (int v = 0; v < w; ++v) base += wc[v];
        if (want) {
            const int64_t k = (int64_t)(base + __popcll(m & ((1ull << lane) - 1ull)));
            const int32_t other = atomicExch(&slot[k], idx);
            if (other >= 0) {  // second to arrive: move the survivor into the hole
                const int64_t d = is_hole ? idx : other, s = is_hole ? other : idx;
                copy_particle(A, d, A, s);
                slot[k] = -1;
            }
        }
        __syncthreads();  // wc and bbase are reused by the next call
    };
    for (int64_t b = blockIdx.x * (int64_t)COMPACT_BLOCK; b < L; b += stride) {
        const int64_t k = b + threadIdx.x;
        const int32_t h = k < L ? holes[k] : INT32_MAX;
        meet((unsigned long long*)&S->nH, h < nn, h, true);
    }
    for (int64_t b = nn + blockIdx.x * (int64_t)COMPACT_BLOCK; b < n; b += stride) {
        const int64_t j = b + threadIdx.x;
        meet((unsigned long long*)&S->nS, j < n && A.cs[j] >= 0, (int32_t)j, false);
    }
}

__global__ void k_compact_finish(DevState* __restrict__ S) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        S->n -= S->n_holes;
        S->n_holes = 0;
        S->nH = 0;
        S->nS = 0;
    }
}

// ---------------------------------------------------------------------------
// Counts per species (reference num_of_particles/positives/negatives,
// population.h:701-744, derive from these).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_count_species(DevState* __restrict__ S, int cell_bits, int nsp, unsigned long long* __restrict__ cnt) {
    __shared__ unsigned int c[256];
    for (int k = threadIdx.x; k < nsp; k += blockDim.x) c[k] = 0;
    __syncthreads();
    const GLOBAL int32_t* cs = (const GLOBAL int32_t*)S->cur.cs;
    const int64_t n = S->n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&c[cs[i] >> cell_bits], 1u);
    __syncthreads();
    for (int k = threadIdx.x; k < nsp; k += blockDim.x)
        if (c[k]) atomicAdd(&cnt[k], (unsigned long long)c[k]);
}

// ---------------------------------------------------------------------------
// Exclusive scan of int64 values (three phases, 1024 elements per block).
// ---------------------------------------------------------------------------
__device__ inline int64_t block_excl_scan_1024(int64_t v, int64_t* sh, int64_t* total) {
    // 256 threads x 4 elements handled by caller; here a generic 256-thread scan.
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        int64_t a = (t >= o) ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    const int64_t incl = sh[t];
    *total = sh[255];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(256)
k_scan_local(const int64_t* __restrict__ in, int64_t* __restrict__ out, int64_t len,
             int64_t* __restrict__ bsum) {
    __shared__ int64_t sh[256];
    const int64_t base = blockIdx.x * 1024ll + threadIdx.x * 4ll;
    int64_t v[4], s = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < len) ? in[base + k] : 0;
        s += v[k];
    }
    int64_t tot;
    int64_t ex = block_excl_scan_1024(s, sh, &tot);
    for (int k = 0; k < 4; ++k) {
        if (base + k < len) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256)
k_scan_bsums(int64_t* __restrict__ bsum, int64_t nb, int64_t* __restrict__ total) {
    __shared__ int64_t sh[256];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += 256) {
        const int64_t i = b0 + threadIdx.x;
        const int64_t v = i < nb ? bsum[i] : 0;
        int64_t tot;
        const int64_t ex = block_excl_scan_1024(v, sh, &tot);
        if (i < nb) bsum[i] = ex + carry;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void k_scan_add(int64_t* __restrict__ out, int64_t len, const int64_t* __restrict__ bsum) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < len) out[i] += bsum[i / 1024];
}

// ---------------------------------------------------------------------------
// Cell re-sort: least-significant-digit radix sort of (cell, index) pairs,
// RADIX_BITS per pass, then one gather of the particle data into the
// alternate buffers and a pointer swap.  No global atomics: every wavefront
// ranks one tile of consecutive keys (radix_tile: 1024 up to 16M particle
// slots, so that small populations still give every SIMD several
// wavefronts, 4096 beyond), stably, with ballot-matched
// lane groups and wavefront-private digit counters in LDS; the tile's digit
// counts, scanned digit-major over all tiles, give each (digit, tile) its
// output offset.  The result is a stable sort by cell (particles of one cell
// keep their previous relative order).
// ---------------------------------------------------------------------------
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
// Keys per wavefront tile for a population of `capacity` slots.
inline int radix_tile(int64_t capacity) { return capacity <= (16ll << 20) ? 1024 : 4096; }
constexpr int RADIX_TILE_MIN = 1024;

// Lanes of the wavefront whose digit equals this lane's.
__device__ inline unsigned long long match_digit(uint32_t dg) {
    unsigned long long m = ~0ull;
    for (int b = 0; b < RADIX_BITS; ++b) {
        const unsigned long long bb = __ballot((dg >> b) & 1u);
        m &= ((dg >> b) & 1u) ? bb : ~bb;
    }
    return m;
}

// Key of element i in pass `shift`: the first pass reads cells from the
// particle array, later passes the previous pass's keys.
// Pass elements after the first: (key << 32) | value, one 8-byte access each.
__device__ inline uint32_t radix_key(const GLOBAL int32_t* cs, const uint64_t* kv, int32_t cmask, int64_t i) {
    return cs ? (uint32_t)(cs[i] & cmask) : (uint32_t)(kv[i] >> 32);
}

// Keys (and values) are loaded RADIX_UNROLL wavefront steps at a time, all
// loads in flight together, then ranked step by step in order.
constexpr int RADIX_UNROLL = 8;
static_assert(RADIX_TILE_MIN % (64 * RADIX_UNROLL) == 0, "tiles hold whole unrolled groups");

// Digit counts of each tile: counts[digit * ntiles + tile].
__global__ void __launch_bounds__(256)
k_radix_count(const DevState* __restrict__ S, int cell_bits, const uint64_t* __restrict__ keys, int shift,
              int64_t ntiles, int tsize, int64_t* __restrict__ counts) {
    __shared__ uint32_t cnt[4][RADIX];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * 4 + w;
    const int64_t n = S->n;
    const GLOBAL int32_t* cs = keys ? nullptr : (const GLOBAL int32_t*)S->cur.cs;
    const int32_t cmask = (1 << cell_bits) - 1;
    for (int d = lane; d < RADIX; d += 64) cnt[w][d] = 0;
    if (tile < ntiles) {
        const int64_t t0 = tile * tsize;
        const int64_t t1 = t0 + tsize < n ? t0 + tsize : n;
        for (int64_t base = t0; base < t1; base += 64 * RADIX_UNROLL) {  // wavefront-uniform
            uint32_t dg[RADIX_UNROLL];
            for (int u = 0; u < RADIX_UNROLL; ++u) {
                const int64_t i = base + 64 * u + lane;
                dg[u] = i < t1 ? (radix_key(cs, keys, cmask, i) >> shift) & (RADIX - 1) : RADIX;
            }
            // Counting needs no order: each run of consecutive lanes with one
            // digit adds its length with one LDS atomic (the cell-sorted
            // input gives long runs), inactive lanes (digit RADIX) add nothing.
            for (int u = 0; u < RADIX_UNROLL; ++u) {
                const uint32_t d = dg[u];
                const uint32_t dprev = __shfl_up(d, 1, 64);
                const bool head = lane == 0 || dprev != d;
                const unsigned long long heads = __ballot(head);
                const unsigned long long upto = lane == 63 ? heads : (heads & ((2ull << lane) - 1ull));
                const int hl = 63 - __clzll((long long)upto);  // first lane of this lane's run
                const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
                if (tail && d < RADIX) atomicAdd(&cnt[w][d], (uint32_t)(lane - hl + 1));
            }
        }
        for (int d = lane; d < RADIX; d += 64) counts[(int64_t)d * ntiles + tile] = cnt[w][d];
    }
}

// Stable scatter of each tile's (key, value) pairs to offsets[digit * ntiles +
// tile] plus the rank among the tile's earlier keys of that digit; offsets
// holds the scan within blocks of 1024 entries, bsum the blocks' bases.  kv ==
// nullptr: the first pass, keys from the particles' cells and values the
// element indices.  kv_out == nullptr: the last pass, only the values are
// written (to vals_out).
__global__ void __launch_bounds__(256)
k_radix_scatter(const DevState* __restrict__ S, int cell_bits, const uint64_t* __restrict__ kv, int shift,
                int64_t ntiles, int tsize, const int64_t* __restrict__ offsets, const int64_t* __restrict__ bsum,
                uint64_t* __restrict__ kv_out, uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t run[4][RADIX];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * 4 + w;
    if (tile >= ntiles) return;  // wavefront-uniform, and no block barrier below
    const int64_t n = S->n;
    const GLOBAL int32_t* cs = kv ? nullptr : (const GLOBAL int32_t*)S->cur.cs;
    const int32_t cmask = (1 << cell_bits) - 1;
    for (int d = lane; d < RADIX; d += 64) {
        const int64_t k = (int64_t)d * ntiles + tile;
        run[w][d] = (uint32_t)(offsets[k] + bsum[k >> 10]);
    }
    const int64_t t0 = tile * tsize;
    const int64_t t1 = t0 + tsize < n ? t0 + tsize : n;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int64_t base = t0; base < t1; base += 64 * RADIX_UNROLL) {  // wavefront-uniform
        uint32_t kk[RADIX_UNROLL], vv[RADIX_UNROLL];
        for (int u = 0; u < RADIX_UNROLL; ++u) {
            const int64_t i = base + 64 * u + lane;
            const bool act = i < t1;
            if (kv) {
                const uint64_t e = act ? kv[i] : 0ull;
                kk[u] = (uint32_t)(e >> 32);
                vv[u] = (uint32_t)e;
            } else {
                kk[u] = act ? (uint32_t)(cs[i] & cmask) : 0u;
                vv[u] = (uint32_t)i;
            }
        }
        for (int u = 0; u < RADIX_UNROLL; ++u) {
            const bool act = base + 64 * u + lane < t1;
            const uint32_t dg = act ? (kk[u] >> shift) & (RADIX - 1) : RADIX;
            const unsigned long long m = match_digit(dg) & __ballot(act);
            uint32_t dst = 0;
            if (act) dst = run[w][dg] + (uint32_t)__popcll(m & below);
            // Every lane has read its group's counter before the group's lowest
            // lane advances it (one wavefront: LDS operations complete in order).
            if (act && lane == __ffsll((long long)m) - 1) run[w][dg] += (uint32_t)__popcll(m);
            if (act) {
                if (kv_out)
                    kv_out[dst] = ((uint64_t)kk[u] << 32) | vv[u];
                else
                    vals_out[dst] = vv[u];
            }
        }
    }
}

// Apply a pending cell order outside a fused step: alt[j] = cur[perm[j]]
// for all particle arrays (k_step_end then swaps the buffers).
__global__ void __launch_bounds__(256) k_sort_gather(DevState* __restrict__ S) {
    const int64_t n = S->n;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || !S->perm_pending) return;
    const GArrays A = gview(S->cur), B = gview(S->alt);
    const int64_t i = S->perm[j];
    __builtin_nontemporal_store(A.xy[i], &B.xy[j]);
    __builtin_nontemporal_store(A.zvx[i], &B.zvx[j]);
    __builtin_nontemporal_store(A.vyz[i], &B.vyz[j]);
    __builtin_nontemporal_store(A.cs[i], &B.cs[j]);
}


// ---------------------------------------------------------------------------
// Initial loading (reference load_particles, injector.h:244-278): positions by
// rejection in the box [cmin, cmax]^3 against the position pdf
// (inside(x) * (1 + amp*cos(kw*x[axis])), max 1+amp; amp = 0 is
// UniformPosition), velocities by the Maxwellian inverse CDF
// v = vd - sqrt(2)*vth*erfc_inv(2u) (distributions.cpp:41-50).  Slot k of
// species sp draws from Philox counter (k, attempt, sp, 0), key (seed, LOAD).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_load(const DevMesh M, DevState* __restrict__ S, int64_t base, int64_t N, int sp, int rng_sp,
       int32_t cell_lo, int32_t cell_hi, int cell_bits, uint64_t seed, double cmin, double cmax,
       double amp, double kw, int axis, VdfParams F, double vmax, unsigned long long* __restrict__ errs) {
    const GArrays A = gview(S->cur);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < N;
         k += (int64_t)gridDim.x * blockDim.x) {
        double p[3];
        int32_t c = -1;
        bool ok = false;
        for (uint64_t a = 0; a < 1000000ull; ++a) {
            U64x4 ctr = {{(uint64_t)k, a, (uint64_t)rng_sp, 0ull}};
            const U64x4 r = philox4x64_10(ctr, seed, TAG_LOAD);
            for (int d = 0; d < 3; ++d) p[d] = cmin + (cmax - cmin) * u01(r.v[d]);
            c = locate_in_grid(M.planes, M.bstart, M.bcells, M.gn, M.glo, M.gh, p[0], p[1], p[2],
                               cell_lo, cell_hi);
            const double pdf = (c >= 0) * (1.0 + amp * cos(kw * p[axis]));
            if (u01(r.v[3]) * (1.0 + amp) < pdf) {
                ok = true;
                break;
            }
        }
        if (!ok) {
            atomicAdd(errs, 1ull);
            c = 0;
        }
        double v[3];
        if (F.kind == VDF_MAXWELLIAN) {
            // inverse CDF per component (distributions.cpp:41-50)
            U64x4 ctr = {{(uint64_t)k, 0xFFFFFFFFull, (uint64_t)rng_sp, 0ull}};
            const U64x4 r = philox4x64_10(ctr, seed, TAG_LOAD);
            for (int d = 0; d < 3; ++d) v[d] = F.vd[d] - sqrt(2.0) * F.vth * erfc_inv_approx(2 * u01(r.v[d]));
        } else {
            // rejection against the pdf maximum in the box vd +- range*vth
            // (injector.cpp:25-49), attempts (k, a, species, 1)
            bool vok = false;
            double lo[3], hi[3];
            for (int d = 0; d < 3; ++d) {
                lo[d] = F.vd[d] - F.range * F.vth;
                hi[d] = F.vd[d] + F.range * F.vth;
            }
            for (uint64_t a = 0; a < 1000000ull && !vok; ++a) {
                U64x4 ctr = {{(uint64_t)k, a, (uint64_t)rng_sp, 1ull}};
                const U64x4 r = philox4x64_10(ctr, seed, TAG_LOAD);
                for (int d = 0; d < 3; ++d) v[d] = lo[d] + (hi[d] - lo[d]) * u01(r.v[d]);
                if (u01(r.v[3]) * vmax < vdf_eval(F, v[0], v[1], v[2])) vok = true;
            }
            if (!vok) atomicAdd(errs, 1ull);
        }
        const int64_t i = base + k;
        st_all(A, i, p[0], p[1], p[2], v[0], v[1], v[2]);
        A.cs[i] = (sp << cell_bits) | c;
    }
}

// ---------------------------------------------------------------------------
// Boundary injection (reference inject_particles, injector.h:157-232).
// ---------------------------------------------------------------------------
// Injection works on a table of (species, facet) pairs, species-major and
// facets ascending within a species, the order of the reference's loops
// (injector.h:168-205).  The step index comes from the device state so that a
// captured graph advances the stream on every replay.
//
// Candidates are processed INJECT_PER_WAVE per wavefront (lanes 0.. own one
// each).  The velocity rejection sampler (injector.cpp:51-76 against the flux
// pdf, distributions.h:93-102) runs cooperatively: for one candidate at a
// time the 64 lanes evaluate attempts 1+64r+lane of round r and the lowest
// accepting attempt wins - exactly the draw the sequential loop accepts,
// without a lane idling while the slowest candidate of its wavefront keeps
// rejecting (acceptance is about 1% for a non-drifting Maxwellian in the
// vd +- 5 vth box).  Few candidates per wavefront keep the per-wavefront
// chain of rounds short and spread a step's candidates over many
// wavefronts.  Then every owning lane finishes its candidate: facet point
// (injector.cpp:78-102), push by U*dt*v (injector.h:212-216), location by a
// walk from the facet's cell.
constexpr uint64_t INJECT_MAX_ATTEMPTS = 1000000ull;
constexpr int INJECT_PER_WAVE = 16;

// MAXW: every injected species is Maxwellian - the flux pdf without the
// other distributions' code (fewer registers, so more of its wavefronts run
// beside the particle kernel); the same draws as the general kernel.
template <bool MAXW>
__device__ __forceinline__ void inject_gen_body(const DevMesh& M, const DevInject& I, const DevState* __restrict__ S,
                                                const SpInject* __restrict__ spi, const int64_t* __restrict__ soff,
                                                const int32_t* __restrict__ spair, uint64_t seed, double dt,
                                                int cell_bits, int64_t cand_cap, double* __restrict__ cx,
                                                int32_t* __restrict__ ccs, int32_t* __restrict__ kc,
                                                unsigned long long* __restrict__ errs, int64_t blk, int64_t nblk) {
    const uint64_t step = S->step;
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blk * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = (nblk * (int64_t)blockDim.x) >> 6;
    for (int64_t g0 = wave * INJECT_PER_WAVE; g0 < cand_cap; g0 += nwaves * INJECT_PER_WAVE) {
        const int64_t k = g0 + lane;
        const bool owner = lane < INJECT_PER_WAVE;
        const bool in_cap = owner && k < cand_cap;
        // This lane's slot: its pair p (a table lookup) and its index within
        // the pair's floor(num)+1 slots.
        int64_t p = 0, s = 0, j = 0, slot = 0;
        if (in_cap) {
            p = spair[k];
            s = I.pair_sp[p];
            j = I.pair_f[p];
            slot = k - soff[p];
        }
        const SpInject P = spi[s];
        const uint64_t key = P.use_seed ? P.seed : seed;
        // The pair's candidate count this step, N = floor(num) + [U < frac]
        // (injector.h:186-191); the slot holds a candidate when slot < N.
        bool mine = false;
        if (in_cap) {
            const double num = I.num[p];
            int64_t N = (int64_t)num;
            U64x4 cc = {{step, (uint64_t)(j - P.facet_base), (uint64_t)P.rng_sp, 0ull}};
            const U64x4 rc = philox4x64_10(cc, key, TAG_INJECT_COUNT);
            if (u01(rc.v[0]) < (num - N)) N += 1;
            mine = slot < N;
        }
        const uint64_t sj = ((uint64_t)P.rng_sp << 32) | (uint64_t)(j - P.facet_base);
        const double* nrm = I.fnormal + 3 * j;
        const double n0 = nrm[0], n1 = nrm[1], n2 = nrm[2];
        const double pmax = I.pmax[p];
        double v[3] = {0.0, 0.0, 0.0};
        bool ok = false;
        const int64_t ng = cand_cap - g0 < INJECT_PER_WAVE ? cand_cap - g0 : INJECT_PER_WAVE;  // wavefront-uniform
        for (int c = 0; c < ng; ++c) {
            if (!__shfl((int)mine, c, 64)) continue;  // wavefront-uniform: slot c holds no candidate
            // broadcast candidate c's parameters
            const uint64_t key_c = __shfl(key, c, 64), sj_c = __shfl(sj, c, 64);
            const uint64_t slot_c = (uint64_t)__shfl(slot, c, 64);
            const double vth = __shfl(P.vth, c, 64), vrange = __shfl(P.vrange, c, 64);
            const double vd0 = __shfl(P.vd[0], c, 64), vd1 = __shfl(P.vd[1], c, 64), vd2 = __shfl(P.vd[2], c, 64);
            VdfParams F;
            F.kind = __shfl(P.kind, c, 64);
            F.factor = __shfl(P.factor, c, 64);
            F.vth2 = __shfl(P.vth2, c, 64);
            F.vth4 = F.vth2 * F.vth2;
            F.k = __shfl(P.k, c, 64);
            F.alpha = __shfl(P.alpha, c, 64);
            F.vd[0] = vd0;
            F.vd[1] = vd1;
            F.vd[2] = vd2;
            const double m0 = __shfl(n0, c, 64), m1 = __shfl(n1, c, 64), m2 = __shfl(n2, c, 64);
            const double pm = __shfl(pmax, c, 64);
            const double lo0 = vd0 - vrange * vth, hi0 = vd0 + vrange * vth;
            const double lo1 = vd1 - vrange * vth, hi1 = vd1 + vrange * vth;
            const double lo2 = vd2 - vrange * vth, hi2 = vd2 + vrange * vth;
            for (uint64_t a0 = 1; a0 < INJECT_MAX_ATTEMPTS; a0 += 64) {
                U64x4 ca = {{step, sj_c, slot_c, a0 + (uint64_t)lane}};
                const U64x4 ra = philox4x64_10(ca, key_c, TAG_INJECT);
                const double w0 = lo0 + (hi0 - lo0) * u01(ra.v[0]);
                const double w1 = lo1 + (hi1 - lo1) * u01(ra.v[1]);
                const double w2 = lo2 + (hi2 - lo2) * u01(ra.v[2]);
                const double f = MAXW ? vdf_flux_maxwellian(F.factor, F.vth2, F.vd, w0, w1, w2, m0, m1, m2)
                                      : vdf_flux(F, w0, w1, w2, m0, m1, m2);
                const unsigned long long acc = __ballot(u01(ra.v[3]) * pm < f);
                if (acc) {
                    const int win = __ffsll((long long)acc) - 1;
                    const double a = __shfl(w0, win, 64), b = __shfl(w1, win, 64), e = __shfl(w2, win, 64);
                    if (lane == c) {
                        v[0] = a;
                        v[1] = b;
                        v[2] = e;
                        ok = true;
                    }
                    break;
                }
            }
        }
        bool kept = false;
        if (mine) {
            if (!ok) atomicAdd(errs, 1ull);
            const double* V = I.fverts + 9 * j;
            U64x4 c0 = {{step, sj, (uint64_t)slot, 0ull}};
            const U64x4 r0 = philox4x64_10(c0, key, TAG_INJECT);
            double x[3];
            double r = u01(r0.v[0]);
            for (int d = 0; d < 3; ++d) x[d] = (1.0 - r) * V[d] + r * V[3 + d];
            r = sqrt(u01(r0.v[1]));
            for (int d = 0; d < 3; ++d) x[d] = r * x[d] + (1.0 - r) * V[6 + d];
            const double rp = u01(r0.v[2]);
            for (int d = 0; d < 3; ++d) x[d] = x[d] + rp * dt * v[d];
            const int32_t c = ok ? walk(M, I.fcell[j], x[0], x[1], x[2]) : -1;
            kept = c >= 0;
            cx[k] = x[0];
            cx[cand_cap + k] = x[1];
            cx[2 * cand_cap + k] = x[2];
            cx[3 * cand_cap + k] = v[0];
            cx[4 * cand_cap + k] = v[1];
            cx[5 * cand_cap + k] = v[2];
            ccs[k] = kept ? (int32_t)((s << cell_bits) | c) : -1;
        } else if (in_cap) {
            ccs[k] = -1;
        }
        // kept candidates of this group of INJECT_PER_WAVE slots (the append's
        // prefix sums read these instead of a scan)
        const unsigned long long km = __ballot(kept);
        if (lane == 0) kc[g0 / INJECT_PER_WAVE] = __popcll(km);
    }
}

// The general generator, and the Maxwellian-only one held to 128 VGPRs (4
// wavefronts per SIMD, no spills; the general one would spill there).
__global__ void __launch_bounds__(256)
k_inject_gen(const DevMesh M, const DevInject I, const DevState* __restrict__ S, const SpInject* __restrict__ spi,
             const int64_t* __restrict__ soff, const int32_t* __restrict__ spair, uint64_t seed, double dt,
             int cell_bits, int64_t cand_cap, double* __restrict__ cx, int32_t* __restrict__ ccs,
             int32_t* __restrict__ kc, unsigned long long* __restrict__ errs) {
    inject_gen_body<false>(M, I, S, spi, soff, spair, seed, dt, cell_bits, cand_cap, cx, ccs, kc, errs, blockIdx.x,
                           gridDim.x);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_inject_gen_maxwellian(const DevMesh M, const DevInject I, const DevState* __restrict__ S,
                        const SpInject* __restrict__ spi, const int64_t* __restrict__ soff,
                        const int32_t* __restrict__ spair, uint64_t seed, double dt, int cell_bits, int64_t cand_cap,
                        double* __restrict__ cx, int32_t* __restrict__ ccs, int32_t* __restrict__ kc,
                        unsigned long long* __restrict__ errs) {
    inject_gen_body<true>(M, I, S, spi, soff, spair, seed, dt, cell_bits, cand_cap, cx, ccs, kc, errs, blockIdx.x,
                          gridDim.x);
}

// Append kept candidates at the end of the population, in candidate (slot)
// order, and deposit their charge when rho_raw is given.  Block b (256
// threads) takes slots [256b, 256b + 256): its output base is the sum of the
// per-group kept counts before it (32 groups of INJECT_PER_WAVE per block),
// each slot's rank within the block comes from wavefront ballots.  The last
// block writes the total kept to *nkept.  Grid: ceil(cand_cap / 256).
constexpr int APPEND_BLOCK = 256;
static_assert(APPEND_BLOCK % INJECT_PER_WAVE == 0, "append blocks cover whole slot groups");

__global__ void __launch_bounds__(APPEND_BLOCK)
k_inject_append(const DevMesh M, DevState* __restrict__ S, const SpCoef* __restrict__ coef,
                int64_t cand_cap, const double* __restrict__ cx, const int32_t* __restrict__ ccs,
                const int32_t* __restrict__ kc, int64_t* __restrict__ nkept, int cell_bits,
                double* __restrict__ rho_raw) {
    __shared__ int64_t red[APPEND_BLOCK / 64];
    __shared__ int32_t wcnt[APPEND_BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // kept candidates in the slots before this block
    const int64_t g_end = (int64_t)blockIdx.x * (APPEND_BLOCK / INJECT_PER_WAVE);
    // eight independent loads in flight per thread (integer sums: any order)
    int64_t a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t g = threadIdx.x;
    for (; g + 7 * APPEND_BLOCK < g_end; g += 8 * APPEND_BLOCK)
        for (int u = 0; u < 8; ++u) a[u] += kc[g + u * APPEND_BLOCK];
    for (; g < g_end; g += APPEND_BLOCK) a[0] += kc[g];
    int64_t part = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
    const int64_t k = (int64_t)blockIdx.x * APPEND_BLOCK + threadIdx.x;
    const int32_t cs = k < cand_cap ? ccs[k] : -1;
    const bool kept = cs >= 0;
    const unsigned long long km = __ballot(kept);
    if (lane == 0) {
        red[w] = part;
        wcnt[w] = __popcll(km);
    }
    __syncthreads();
    int64_t base = 0, before = 0, in_block = 0;
    for (int v = 0; v < APPEND_BLOCK / 64; ++v) {
        base += red[v];
        if (v < w) before += wcnt[v];
        in_block += wcnt[v];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0 && nkept) *nkept = base + in_block;
    if (!kept) return;
    // After a fused step the removal's count update and the buffer swap are
    // still pending (k_scale_finish): the base is n - n_holes, the arrays
    // those the step wrote.  Otherwise n_holes is 0 and no order is pending.
    const GArrays A = gview(step_arrays(S));
    const int64_t n0 = S->n - S->n_holes, cap = S->capacity;
    const int32_t cmask = (1 << cell_bits) - 1;
    const int64_t d = n0 + base + before + __popcll(km & ((1ull << lane) - 1ull));
    if (d >= cap) {
        atomicAdd((unsigned long long*)&S->overflow, 1ull);
        return;
    }
    const double x = cx[k], y = cx[cand_cap + k], z = cx[2 * cand_cap + k];
    st_all(A, d, x, y, z, cx[3 * cand_cap + k], cx[4 * cand_cap + k], cx[5 * cand_cap + k]);
    A.cs[d] = cs;
    if (rho_raw) {
        const CellRec& R = M.rec[cs & cmask];
        double l[4];
        bary4(R.bary, x, y, z, l);
        const double q = coef[cs >> cell_bits].q;
        unsafeAtomicAdd(&rho_raw[R.dof[0]], q * l[0]);
        unsafeAtomicAdd(&rho_raw[R.dof[1]], q * l[1]);
        unsafeAtomicAdd(&rho_raw[R.dof[2]], q * l[2]);
        unsafeAtomicAdd(&rho_raw[R.dof[3]], q * l[3]);
    }
}

// Count update after an injection outside a fused step (room: particles
// past the capacity were dropped and counted as overflow by the append).
__global__ void k_inject_finish(DevState* __restrict__ S, const int64_t* __restrict__ nkept) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        int64_t kept = *nkept;
        const int64_t room = S->capacity - S->n;
        if (kept > room) kept = room;
        S->inj_kept = kept;
        S->n += kept;
    }
}

__global__ void k_set_step(DevState* __restrict__ S, uint64_t step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) S->step = step;
}

// ---------------------------------------------------------------------------
// Diagnostics (reference diagnostics.h, Population::statistics)
// ---------------------------------------------------------------------------
// Per-cell copy of a scalar CG1 function: cellPhi[c] = phi at the cell's
// four dofs in local vertex order (one 32-byte load per particle instead of a
// dependent gather through the cell's dof list).
__global__ void k_cell_scalar(const DevMesh M, const double* __restrict__ phi, double4* __restrict__ cellPhi) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < M.nc; c += (int64_t)gridDim.x * blockDim.x) {
        const CellRec& R = M.rec[c];
        cellPhi[c] = make_double4(phi[R.dof[0]], phi[R.dof[1]], phi[R.dof[2]], phi[R.dof[3]]);
    }
}

// particle_potential_energy_cg1 (diagnostics.h:379-412): per block the sum of
// 0.5 q phi(x) over its chunk, phi_x = sum_k lambda_k phi_k in the
// reference's order.

__global__ void __launch_bounds__(256)
k_potential_energy(const DevMesh M, DevState* __restrict__ S, const SpCoef* __restrict__ coef, int nsp,
                   int cell_bits, const double4* __restrict__ cellPhi, double* __restrict__ part) {
    // One particle per lane and group, pipelined as the fused step: the
    // group's cell rows (barycentric rows, phi at the four dofs) are
    // requested first, then the next group's x, y, z and cell ids, and only
    // then does the group wait for its cell data.  q per species from LDS.
    typedef double v2d __attribute__((ext_vector_type(2)));
    __shared__ double sh[8];
    __shared__ double q_of[MAX_SPECIES];
    for (int k = threadIdx.x; k < nsp; k += blockDim.x) q_of[k] = coef[k].q;
    __syncthreads();
    const GArrays A = gview(S->cur);
    int64_t s, e;
    chunk_range(S->n, s, e);
    const int32_t cmask = (1 << cell_bits) - 1;
    double pe = 0.0;
    if (s < e) {  // block-uniform
        const int64_t last = e - 1;
        auto clampi = [&](int64_t j) { return j < e ? j : last; };
        double nx, ny, nz;
        int32_t ncs;
        {
            const int64_t j0 = clampi(s + threadIdx.x);
            ld_pos(A, j0, nx, ny, nz);
            ncs = A.cs[j0];
        }
        for (int64_t base = s; base < e; base += blockDim.x) {  // block-uniform trip count
            const bool active = base + threadIdx.x < e;
            const double x = nx, y = ny, z = nz;
            const int32_t cs = ncs;
            const int32_t c = cs & cmask;
            const GLOBAL v2d* rp = (const GLOBAL v2d*)(M.rec + c);
            const GLOBAL v2d* fp = (const GLOBAL v2d*)(cellPhi + c);
            double rb[12];
            for (int m = 0; m < 6; ++m) {
                const v2d t = rp[m];
                rb[2 * m] = t.x;
                rb[2 * m + 1] = t.y;
            }
            const v2d f01 = fp[0], f23 = fp[1];
            __builtin_amdgcn_sched_barrier(0);
            {
                const int64_t j = clampi(base + threadIdx.x + blockDim.x);
                ld_pos(A, j, nx, ny, nz);
                ncs = A.cs[j];
            }
            __builtin_amdgcn_sched_barrier(0);
            double l[4];
            bary4(rb, x, y, z, l);
            double phi_x = 0.0;
            phi_x += l[0] * f01.x;
            phi_x += l[1] * f01.y;
            phi_x += l[2] * f23.x;
            phi_x += l[3] * f23.y;
            if (active) pe += 0.5 * q_of[cs >> cell_bits] * phi_x;
        }
    }
    const double t = block_sum(pe, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// DG0 counts (density_dg0, diagnostics.h:432-465) into raw[2 * cell + class]
// (class 1 for q > 0).  The pass reads 4 B per particle; a block iteration
// covers DG0_GROUPS groups of blockDim particles, whose cell ids (one per
// lane and group, coalesced) are loaded one iteration ahead.  Per group,
// runs of consecutive lanes with the same (cell, class) are counted with a
// ballot and the run's last lane adds the count.  A block's chunk of the
// cell-sorted population covers a compact range of cells, so the counts go
// to a direct-mapped window of DG0_WINDOW cells in LDS (an integer LDS add,
// no hashing) that starts a little below the chunk's first cell; only
// particles outside it (moved far, or placed by compaction) use the block's
// hash table.  The counts are integers, so the sums are exact in any order.
constexpr int DG0_GROUPS = 8;
constexpr int DG0_WINDOW = 1024;  // cells; 2 int32 counters each
constexpr int DG0_SLACK = 128;    // window cells below the chunk's first cell

__global__ void __launch_bounds__(256)
k_density_dg0(DevState* __restrict__ S, const SpCoef* __restrict__ coef, int nsp, int cell_bits,
              double* __restrict__ raw) {
    __shared__ int32_t cls_of[MAX_SPECIES];
    __shared__ uint32_t win[2 * DG0_WINDOW];
    for (int k = threadIdx.x; k < nsp; k += blockDim.x) cls_of[k] = coef[k].q > 0 ? 1 : 0;
    for (int k = threadIdx.x; k < 2 * DG0_WINDOW; k += blockDim.x) win[k] = 0u;
    const GLOBAL int32_t* cs = (const GLOBAL int32_t*)S->cur.cs;
    int64_t s, e;
    chunk_range(S->n, s, e);
    if (s >= e) return;  // block-uniform
    const int32_t cmask = (1 << cell_bits) - 1;
    const int32_t c0 = cs[s] & cmask;  // the same address for every lane: one broadcast load
    const int32_t wbase = c0 > DG0_SLACK ? c0 - DG0_SLACK : 0;
    const LdsTable t = lds_table();
    lds_clear(t);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t per_iter = (int64_t)blockDim.x * DG0_GROUPS;
    int32_t nxt[DG0_GROUPS];
    for (int g = 0; g < DG0_GROUPS; ++g) {
        const int64_t i = s + (int64_t)g * blockDim.x + threadIdx.x;
        nxt[g] = i < e ? cs[i] : -1;
    }
    int iter = 0;
    for (int64_t b = s; b < e; b += per_iter) {  // block-uniform trip count
        int32_t cur[DG0_GROUPS];
        for (int g = 0; g < DG0_GROUPS; ++g) cur[g] = nxt[g];
        for (int g = 0; g < DG0_GROUPS; ++g) {
            const int64_t i = b + per_iter + (int64_t)g * blockDim.x + threadIdx.x;
            nxt[g] = i < e ? cs[i] : -1;
        }
        for (int g = 0; g < DG0_GROUPS; ++g) {
            const int32_t x = cur[g];
            const int32_t key = x >= 0 ? 2 * (x & cmask) + cls_of[x >> cell_bits] : -1;
            const int32_t kprev = __shfl_up(key, 1, 64);
            const bool head = lane == 0 || kprev != key;
            const unsigned long long heads = __ballot(head);
            const unsigned long long upto = lane == 63 ? heads : (heads & ((2ull << lane) - 1ull));
            const int hl = 63 - __clzll((long long)upto);  // first lane of this lane's run
            const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
            if (key >= 0 && tail) {
                const uint32_t w = (uint32_t)(key - 2 * wbase);
                if (w < 2u * DG0_WINDOW)
                    atomicAdd(&win[w], (uint32_t)(lane - hl + 1));
                else
                    lds_add(t, raw, key, (double)(lane - hl + 1));
            }
        }
        // every 2 iterations (16 groups), one decision for the whole block
        if ((++iter & 1) == 0 && __syncthreads_or(threadIdx.x == 0 && *t.count > LDS_SLOTS * 5 / 8))
            lds_flush(t, raw);
    }
    lds_flush(t, raw);  // synchronizes: the window is complete
    const int64_t wend = (int64_t)2 * (cmask + 1);
    for (int k = threadIdx.x; k < 2 * DG0_WINDOW; k += blockDim.x) {
        const uint32_t c = win[k];
        if (c != 0u && 2 * (int64_t)wbase + k < wend) unsafeAtomicAdd(&raw[2 * (int64_t)wbase + k], (double)c);
    }
}

// CG1 number densities (density_cg1, diagnostics.h:492-555) into
// raw[2 * dof + class] (class 0 for q < 0, else 1), with the fused step's
// pipeline and deposit: a group's cell records are requested first, then the
// next group's x, y, z and cell ids, and only then does the group wait for
// its cell data; the lanes of a quad with the same (cell, class) are summed
// with DPP before the block's LDS table takes their lambda_k (deposit mode 4,
// q = 1: the added values are lambda_k exactly).
__global__ void __launch_bounds__(256)
k_density_cg1(const DevMesh M, DevState* __restrict__ S, const SpCoef* __restrict__ coef, int nsp, int cell_bits,
              double* __restrict__ raw) {
    typedef double v2d __attribute__((ext_vector_type(2)));
    typedef int32_t v4i __attribute__((ext_vector_type(4)));
    __shared__ int32_t cls_of[MAX_SPECIES];
    for (int k = threadIdx.x; k < nsp; k += blockDim.x) cls_of[k] = coef[k].q < 0 ? 0 : 1;
    const GArrays A = gview(S->cur);
    int64_t s, e;
    chunk_range(S->n, s, e);
    if (s >= e) return;  // block-uniform
    const int32_t cmask = (1 << cell_bits) - 1;
    lds_clear(lds_table());
    __syncthreads();
    const int64_t last = e - 1;
    auto clampi = [&](int64_t j) { return j < e ? j : last; };
    double nx, ny, nz;
    int32_t ncs;
    {
        const int64_t j0 = clampi(s + threadIdx.x);
        ld_pos(A, j0, nx, ny, nz);
        ncs = A.cs[j0];
    }
    int grp = 0;
    for (int64_t base = s; base < e; base += blockDim.x) {  // block-uniform trip count
        const int64_t i = base + threadIdx.x;
        const bool active = i < e;
        const double x = nx, y = ny, z = nz;
        const int32_t cs = ncs;
        const int32_t c = cs & cmask;
        const GLOBAL v2d* rp = (const GLOBAL v2d*)(M.rec + c);
        double rb[12];
        for (int m = 0; m < 6; ++m) {
            const v2d t = rp[m];
            rb[2 * m] = t.x;
            rb[2 * m + 1] = t.y;
        }
        const v4i dv = *(const GLOBAL v4i*)(rp + 6);
        __builtin_amdgcn_sched_barrier(0);
        {
            const int64_t j = clampi(i + blockDim.x);
            ld_pos(A, j, nx, ny, nz);
            ncs = A.cs[j];
        }
        __builtin_amdgcn_sched_barrier(0);
        const int cls = cls_of[cs >> cell_bits];
        double l[4];
        bary4(rb, x, y, z, l);
        const int32_t ddof[4] = {2 * dv.x + cls, 2 * dv.y + cls, 2 * dv.z + cls, 2 * dv.w + cls};
        deposit_cg1<4>(M, raw, active ? 2 * c + cls : -1, 0.0, 0.0, 0.0, 1.0, true, l, ddof);
        // every 8 groups, one decision for the whole block
        if ((++grp & 7) == 0 && base + blockDim.x < e &&
            __syncthreads_or(threadIdx.x == 0 && *lds_table().count > LDS_SLOTS * 5 / 8))
            lds_flush(lds_table(), raw);
    }
    lds_flush(lds_table(), raw);
}

// CG1: n_j = raw_j * (dv_inv_j * w) (diagnostics.h:549-553).  DG0: n_k =
// (w + w + ... count_k times) / volume_k, the weights summed one particle at a
// time as diagnostics.h:452-459 does (count_k * w can round differently).
__device__ inline double repeat_sum(double count, double w) {
    double a = 0.0;
    for (int64_t k = (int64_t)count; k > 0; --k) a += w;
    return a;
}

template <bool DG0>
__global__ void k_density_scale(const double* __restrict__ raw, const double* __restrict__ scale, double we,
                                double wi, double* __restrict__ ne, double* __restrict__ ni, int64_t n) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        if (DG0) {
            ne[j] = repeat_sum(raw[2 * j], we) / scale[j];
            ni[j] = repeat_sum(raw[2 * j + 1], wi) / scale[j];
        } else {
            ne[j] = raw[2 * j] * (scale[j] * we);
            ni[j] = raw[2 * j + 1] * (scale[j] * wi);
        }
    }
}

// Population::statistics (population.h:746-805) in parallel: every thread
// runs the reference's running update (Welford) over its particles, then
// the partial (count, mean, M2) triples are merged pairwise (Chan, Golub and
// LeVeque: mean += delta*nb/n, M2 += M2b + delta^2*na*nb/n).  Class 0:
// q < 0, class 1: q > 0.  part: per block 6 doubles (n, mean, M2 per class).
struct Welford {
    double n, mean, m2;
};
__device__ inline void welford_merge(Welford& a, const Welford& b) {
    if (b.n == 0.0) return;
    if (a.n == 0.0) {
        a = b;
        return;
    }
    const double n = a.n + b.n;
    const double delta = b.mean - a.mean;
    a.mean = a.mean + delta * (b.n / n);
    a.m2 = (a.m2 + b.m2) + delta * delta * (a.n * b.n / n);
    a.n = n;
}

__global__ void __launch_bounds__(256)
k_statistics(DevState* __restrict__ S, const SpCoef* __restrict__ coef, int nsp, int cell_bits,
             double* __restrict__ part) {
    __shared__ Welford sh[2][256];
    __shared__ double qs[MAX_SPECIES];
    for (int k = threadIdx.x; k < nsp; k += blockDim.x) qs[k] = coef[k].q;
    __syncthreads();
    const GArrays A = gview(S->cur);
    int64_t s, e;
    chunk_range(S->n, s, e);
    Welford w[2] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    // STATS_UNROLL particles' loads in flight at once, then the running
    // updates in the thread's particle order (the same order as one at a time)
    constexpr int STATS_UNROLL = 4;
    for (int64_t i0 = s + threadIdx.x; i0 < e; i0 += (int64_t)STATS_UNROLL * blockDim.x) {
        int32_t cs[STATS_UNROLL];
        double vx[STATS_UNROLL], vy[STATS_UNROLL], vz[STATS_UNROLL];
        for (int u = 0; u < STATS_UNROLL; ++u) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            cs[u] = -1;
            if (i < e) {
                cs[u] = A.cs[i];
                ld_vel(A, i, vx[u], vy[u], vz[u]);
            }
        }
        for (int u = 0; u < STATS_UNROLL; ++u) {
            if (cs[u] < 0) continue;
            const double q = qs[cs[u] >> cell_bits];
            if (q == 0.0) continue;
            double v = 0;
            v += vx[u] * vx[u];
            v += vy[u] * vy[u];
            v += vz[u] * vz[u];
            v = sqrt(v);
            Welford& W = w[q < 0 ? 0 : 1];
            W.n += 1.0;
            const double old = W.mean;
            W.mean = old + (v - old) / W.n;
            W.m2 += (v - old) * (v - W.mean);
        }
    }
    for (int k = 0; k < 2; ++k) sh[k][threadIdx.x] = w[k];
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 2; ++k) welford_merge(sh[k][threadIdx.x], sh[k][threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x < 2) {
        const Welford& r = sh[threadIdx.x][0];
        part[6 * blockIdx.x + 3 * threadIdx.x] = r.n;
        part[6 * blockIdx.x + 3 * threadIdx.x + 1] = r.mean;
        part[6 * blockIdx.x + 3 * threadIdx.x + 2] = r.m2;
    }
}

// Merge of the per-block triples (one block of 256): each thread merges a
// contiguous range of blocks, then a tree over the threads.
__global__ void __launch_bounds__(256) k_statistics_merge(const double* __restrict__ part, int np,
                                                          double* __restrict__ out) {
    __shared__ Welford sh[2][256];
    const int per = (np + blockDim.x - 1) / blockDim.x;
    const int b0 = threadIdx.x * per, b1 = min(np, b0 + per);
    for (int k = 0; k < 2; ++k) {
        Welford a = {0.0, 0.0, 0.0};
        for (int b = b0; b < b1; ++b) {
            const Welford w = {part[6 * b + 3 * k], part[6 * b + 3 * k + 1], part[6 * b + 3 * k + 2]};
            welford_merge(a, w);
        }
        sh[k][threadIdx.x] = a;
    }
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 2; ++k) welford_merge(sh[k][threadIdx.x], sh[k][threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x < 2) {
        out[3 * threadIdx.x] = sh[threadIdx.x][0].n;
        out[3 * threadIdx.x + 1] = sh[threadIdx.x][0].mean;
        out[3 * threadIdx.x + 2] = sh[threadIdx.x][0].m2;
    }
}

// ema (diagnostics.h:573-588): g = w f + (1 - w) g.
__global__ void k_ema(const double* __restrict__ f, double* __restrict__ g, double w, int64_t n) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        g[j] = w * f[j] + (1.0 - w) * g[j];
}

// ---------------------------------------------------------------------------
// Order-independent digest of a population: per species the wrapping sum of a
// 64-bit hash of each particle's bit patterns (x, y, z, vx, vy, vz and the
// cell id relative to c0[species]) and the particle count.  Two populations
// hold the same particles (as sets, bit for bit) exactly when their
// digests agree, up to hash collisions.
// ---------------------------------------------------------------------------
__device__ inline uint64_t mix64(uint64_t z) {  // the splitmix64 finalizer
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256)
k_pop_digest(DevState* __restrict__ S, int cell_bits, const int32_t* __restrict__ c0, int nsp,
             unsigned long long* __restrict__ out) {
    __shared__ unsigned long long sh[2 * MAX_SPECIES];
    for (int k = threadIdx.x; k < 2 * nsp; k += blockDim.x) sh[k] = 0ull;
    __syncthreads();
    const GArrays A = gview(S->cur);
    const int64_t n = S->n;
    const int32_t cmask = (1 << cell_bits) - 1;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t cs = A.cs[i];
        const int sp = cs >> cell_bits;
        const uint32_t cell = (uint32_t)((cs & cmask) - c0[sp]);
        double px, py, pz, qx, qy, qz;
        ld_all(A, i, px, py, pz, qx, qy, qz);
        uint64_t h = mix64((uint64_t)__double_as_longlong(px) ^ 0x9e3779b97f4a7c15ull);
        h = mix64(h ^ (uint64_t)__double_as_longlong(py));
        h = mix64(h ^ (uint64_t)__double_as_longlong(pz));
        h = mix64(h ^ (uint64_t)__double_as_longlong(qx));
        h = mix64(h ^ (uint64_t)__double_as_longlong(qy));
        h = mix64(h ^ (uint64_t)__double_as_longlong(qz));
        h = mix64(h ^ (uint64_t)cell);
        atomicAdd(&sh[sp], (unsigned long long)h);
        atomicAdd(&sh[nsp + sp], 1ull);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * nsp; k += blockDim.x)
        if (sh[k]) atomicAdd(&out[k], sh[k]);
}

// ---------------------------------------------------------------------------
// Flux tables by Monte Carlo (create_flux, injector.cpp:127-168) for the
// distributions without a closed form: per exterior facet, n_iter uniform
// velocities in the box vd +- range*vth give the flux integral
// area * sum * volume / n_iter, and n_iter more give the flux-pdf maximum
// (times 1.01).  One block per facet; sample s of facet j uses Philox
// counters (species, j, s, 0) for the integral and (species, j, s, 1) for the
// maximum, key (seed, FLUX).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_flux_mc(const double* __restrict__ fnormal, const double* __restrict__ farea, int64_t nf, VdfParams F,
          uint64_t seed, int rng_sp, int64_t n_iter, int do_num, int do_max, double* __restrict__ num,
          double* __restrict__ pmax) {
    __shared__ double sh[8];
    __shared__ double shm[256];
    const int dim = 3;
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = F.vd[d] - F.range * F.vth;  // Pdf::update_domain (distributions.h:139-146)
        hi[d] = F.vd[d] + F.range * F.vth;
    }
    double volume = 1.0;
    for (int d = 0; d < dim; ++d) volume *= hi[d] - lo[d];
    for (int64_t j = blockIdx.x; j < nf; j += gridDim.x) {  // block-uniform
        const double n0 = fnormal[3 * j], n1 = fnormal[3 * j + 1], n2 = fnormal[3 * j + 2];
        double sum = 0.0, mx = 0.0;
        for (int64_t s = threadIdx.x; s < n_iter; s += blockDim.x) {
            if (do_num) {
                U64x4 c = {{(uint64_t)rng_sp, (uint64_t)j, (uint64_t)s, 0ull}};
                const U64x4 r = philox4x64_10(c, seed, TAG_FLUX);
                double x[3];
                for (int d = 0; d < 3; ++d) x[d] = lo[d] + u01(r.v[d]) * (hi[d] - lo[d]);
                sum += vdf_flux(F, x[0], x[1], x[2], n0, n1, n2);
            }
            if (do_max) {
                U64x4 c = {{(uint64_t)rng_sp, (uint64_t)j, (uint64_t)s, 1ull}};
                const U64x4 r = philox4x64_10(c, seed, TAG_FLUX);
                double x[3];
                for (int d = 0; d < 3; ++d) x[d] = lo[d] + u01(r.v[d]) * (hi[d] - lo[d]);
                const double p = vdf_flux(F, x[0], x[1], x[2], n0, n1, n2);
                mx = mx > p ? mx : p;
            }
        }
        const double tot = block_sum(sum, sh);
        shm[threadIdx.x] = mx;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) shm[threadIdx.x] = shm[threadIdx.x] > shm[threadIdx.x + o] ? shm[threadIdx.x]
                                                                                                 : shm[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (do_num) num[j] = farea[j] * tot * volume / n_iter;
            if (do_max) pmax[j] = shm[0] * 1.01;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Population files (Population::save_file / load_file, population.h:807-880):
// one record per particle, the reference's Particle<3> layout of 8 doubles
// x[3], v[3], q, m (64 bytes, no padding).
// ---------------------------------------------------------------------------
__global__ void k_pack_records(const DevState* __restrict__ S, const SpCoef* __restrict__ coef, int cell_bits,
                               int64_t first, int64_t count, double* __restrict__ rec) {
    const GArrays A = gview(S->cur);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = first + k;
        const SpCoef& C = coef[A.cs[i] >> cell_bits];
        double2* r = (double2*)(rec + 8 * k);
        const v2d a = A.xy[i], b = A.zvx[i], c = A.vyz[i];
        r[0] = make_double2(a.x, a.y);
        r[1] = make_double2(b.x, b.y);
        r[2] = make_double2(c.x, c.y);
        r[3] = make_double2(C.q, C.m);
    }
}

// Records -> located candidates: species by exact (q, m) match (the host has
// registered every pair first), cell by the grid point location; points
// outside the mesh are dropped (keep 0).
__global__ void k_unpack_locate(const DevMesh M, const double* __restrict__ rec, int64_t count,
                                const double* __restrict__ spq, const double* __restrict__ spm, int nsp,
                                int cell_bits, int64_t cap, double* __restrict__ cx, int32_t* __restrict__ ccs,
                                int64_t* __restrict__ keep, unsigned long long* __restrict__ errs) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x) {
        const double* r = rec + 8 * k;
        int s = -1;
        for (int j = 0; j < nsp && s < 0; ++j)
            if (spq[j] == r[6] && spm[j] == r[7]) s = j;
        int32_t c = -1;
        if (s >= 0)
            c = locate_in_grid(M.planes, M.bstart, M.bcells, M.gn, M.glo, M.gh, r[0], r[1], r[2], 0, (int32_t)M.nc);
        else
            atomicAdd(errs, 1ull);
        for (int d = 0; d < 6; ++d) cx[d * cap + k] = r[d];
        ccs[k] = c >= 0 ? ((s << cell_bits) | c) : -1;
        keep[k] = c >= 0 ? 1 : 0;
    }
}

// Appends the kept records at n + pos[k] in record order, then n += kept.
__global__ void k_file_append(DevState* __restrict__ S, int64_t count, int64_t cap, const double* __restrict__ cx,
                              const int32_t* __restrict__ ccs, const int64_t* __restrict__ pos) {
    const GArrays A = gview(S->cur);
    const int64_t n0 = S->n, pcap = S->capacity;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t cs = ccs[k];
        if (cs < 0) continue;
        const int64_t d = n0 + pos[k];
        if (d >= pcap) continue;  // counted by k_file_finish
        st_all(A, d, cx[k], cx[cap + k], cx[2 * cap + k], cx[3 * cap + k], cx[4 * cap + k], cx[5 * cap + k]);
        A.cs[d] = cs;
    }
}

__global__ void k_file_finish(DevState* __restrict__ S, const int64_t* __restrict__ kept) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t room = S->capacity - S->n;
        const int64_t k = *kept;
        if (k > room) {
            S->overflow += k - room;
            S->n = S->capacity;
        } else {
            S->n += k;
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
#define LAUNCH_CHECK() \
    do { hipError_t e_ = hipGetLastError(); if (e_ != hipSuccess) return e_; } while (0)

template <int PUSH, int MOVE, bool LOCATE, bool DEPOSIT, int DMODE>
static hipError_t launch_particles_t(const ParticleLaunch& L, hipStream_t st) {
    const size_t shmem = (DEPOSIT && DMODE >= 2) ? LDS_TABLE_BYTES : 0;
    hipLaunchKernelGGL((k_particles<PUSH, MOVE, LOCATE, DEPOSIT, DMODE>), dim3(L.grid), dim3(L.block), shmem,
                       st, L.mesh, L.state, L.coef, L.cellE, L.dt, L.Ld[0], L.Ld[1], L.Ld[2], L.cell_bits,
                       L.rho_raw, L.ke_part, L.holes, L.absorbed);
    LAUNCH_CHECK();
    return hipSuccess;
}

template <int PUSH, int MOVE, bool LOCATE>
static hipError_t launch_pml(const ParticleLaunch& L, hipStream_t st) {
    if (!L.deposit) return launch_particles_t<PUSH, MOVE, LOCATE, false, 0>(L, st);
    switch (L.dep_mode) {
        case 0: return launch_particles_t<PUSH, MOVE, LOCATE, true, 0>(L, st);
        case 1: return launch_particles_t<PUSH, MOVE, LOCATE, true, 1>(L, st);
        case 2: return launch_particles_t<PUSH, MOVE, LOCATE, true, 2>(L, st);
        case 3: return launch_particles_t<PUSH, MOVE, LOCATE, true, 3>(L, st);
        case 4: return launch_particles_t<PUSH, MOVE, LOCATE, true, 4>(L, st);
    }
    return hipErrorInvalidValue;
}

template <int PUSH, int MOVE>
static hipError_t launch_pm(const ParticleLaunch& L, hipStream_t st) {
    if (L.locate) return launch_pml<PUSH, MOVE, true>(L, st);
    return launch_pml<PUSH, MOVE, false>(L, st);
}

static int fill_grid(int64_t n);

// Blocks of the injection generator: INJECT_PER_WAVE slots per wavefront, 4
// wavefronts per block of 256 (a grid-stride loop covers any shortfall).
static int inject_gen_blocks(int64_t cand_cap) {
    const int64_t b = (cand_cap + 4 * INJECT_PER_WAVE - 1) / (4 * INJECT_PER_WAVE);
    return (int)std::max<int64_t>(std::min<int64_t>(b, 16384), 1);
}

template <int PUSH, int MOVE, int DMODE>
static hipError_t launch_step_t(const ParticleLaunch& L, hipStream_t st) {
    if (L.nsp < 1 || L.nsp > MAX_SPECIES) return hipErrorInvalidValue;
    size_t shmem = (DMODE >= 2 ? LDS_TABLE_BYTES : 0) + (size_t)L.nsp * sizeof(SpCoef);
    // LDS walk records: what fits beside the table and the coefficients in a
    // quarter of the CU's 160 KB (4 blocks of 256 per CU, the occupancy the
    // kernel's registers allow), at most STEP_WALK_LDS.
    const size_t budget = 40 * 1024 - 1024;  // less the static LDS and allocation rounding
    const int wl_cap = shmem < budget ? (int)std::min<size_t>(STEP_WALK_LDS, (budget - shmem) / sizeof(WalkRec)) : 0;
    shmem += (size_t)wl_cap * sizeof(WalkRec);
    GenArgs G{};  // no generator blocks
    int grid = L.grid;
    if (L.gen) {
        const InjectLaunch& J = *L.gen;
        if (!J.all_maxwellian) return hipErrorInvalidValue;
        G.I = J.inj;
        G.spi = J.spi;
        G.soff = J.soff;
        G.spair = J.spair;
        G.seed = J.seed;
        G.dt = J.dt;
        G.cand_cap = J.cand_cap;
        G.cx = J.cx;
        G.ccs = J.ccs;
        G.kc = J.kc;
        G.errs = J.errs;
        G.blocks = inject_gen_blocks(J.cand_cap);
        grid += G.blocks;
    }
    if (L.pe_part)
        hipLaunchKernelGGL((k_step_main<PUSH, MOVE, DMODE, true>), dim3(grid), dim3(L.block), shmem, st, L.mesh,
                           L.state, L.coef, L.nsp, L.cellE, L.dt, L.Ld[0], L.Ld[1], L.Ld[2], L.cell_bits, L.rho_raw,
                           L.ke_part, L.holes, L.absorbed, (WalkRec*)L.walk, L.pe_part, wl_cap, G);
    else
        hipLaunchKernelGGL((k_step_main<PUSH, MOVE, DMODE, false>), dim3(grid), dim3(L.block), shmem, st, L.mesh,
                           L.state, L.coef, L.nsp, L.cellE, L.dt, L.Ld[0], L.Ld[1], L.Ld[2], L.cell_bits, L.rho_raw,
                           L.ke_part, L.holes, L.absorbed, (WalkRec*)L.walk, nullptr, wl_cap, G);
    LAUNCH_CHECK();
    return hipSuccess;
}

template <int PUSH, int MOVE>
static hipError_t launch_step_pm(const ParticleLaunch& L, hipStream_t st) {
    switch (L.dep_mode) {
        case 0: return launch_step_t<PUSH, MOVE, 0>(L, st);
        case 1: return launch_step_t<PUSH, MOVE, 1>(L, st);
        case 2: return launch_step_t<PUSH, MOVE, 2>(L, st);
        case 3: return launch_step_t<PUSH, MOVE, 3>(L, st);
        case 4: return launch_step_t<PUSH, MOVE, 4>(L, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_step_passes(const ParticleLaunch& L, hipStream_t st) {
    if (!L.locate || !L.deposit || !L.walk) return hipErrorInvalidValue;
    switch (L.push * 3 + L.move) {
        case 4: return launch_step_pm<1, 1>(L, st);
        case 5: return launch_step_pm<1, 2>(L, st);
        case 7: return launch_step_pm<2, 1>(L, st);
        case 8: return launch_step_pm<2, 2>(L, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_cell_field(const DevMesh& M, const double* E3, double* cellE, hipStream_t st) {
    hipLaunchKernelGGL(k_cell_field, dim3(fill_grid(M.nc)), dim3(256), 0, st, M, E3, cellE);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_cell_potential(const DevMesh& M, const double* phi, double* cellE, hipStream_t st) {
    hipLaunchKernelGGL(k_cell_potential, dim3(fill_grid(M.nc)), dim3(256), 0, st, M, phi, cellE);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_expand_field(const double* E3, double4* E, int64_t ndof, hipStream_t st) {
    if (ndof <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_expand_field, dim3(fill_grid(ndof)), dim3(256), 0, st, E3, E, ndof);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_particles(const ParticleLaunch& L, hipStream_t st) {
    switch (L.push * 3 + L.move) {
        case 0: return launch_pm<0, 0>(L, st);
        case 1: return launch_pm<0, 1>(L, st);
        case 2: return launch_pm<0, 2>(L, st);
        case 3: return launch_pm<1, 0>(L, st);
        case 4: return launch_pm<1, 1>(L, st);
        case 5: return launch_pm<1, 2>(L, st);
        case 6: return launch_pm<2, 0>(L, st);
        case 7: return launch_pm<2, 1>(L, st);
        case 8: return launch_pm<2, 2>(L, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_kinetic(DevState* S, const SpCoef* coef, int cell_bits, double* ke_part, int grid,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_kinetic, dim3(grid), dim3(256), 0, st, S, coef, cell_bits, ke_part);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_sum_partials(const double* part, int np, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, st, part, np, out);
    LAUNCH_CHECK();
    return hipSuccess;
}

static int fill_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return (int)g;
}

hipError_t launch_fill_f64(double* p, int64_t n, double v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3(fill_grid(n)), dim3(256), 0, st, p, n, v);
    LAUNCH_CHECK();
    return hipSuccess;
}
hipError_t launch_fill_u64(unsigned long long* p, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(fill_grid(n)), dim3(256), 0, st, p, n);
    LAUNCH_CHECK();
    return hipSuccess;
}
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(n)), dim3(256), 0, st, p, n, v);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_scale_rho(double* raw, const double* dv_inv, double* rho, int64_t n,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_rho, dim3(fill_grid(n)), dim3(256), 0, st, raw, dv_inv, rho, n);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_compact(DevState* S, const int32_t* holes, int32_t* slot, int grid, bool after_step,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_compact, dim3(grid), dim3(COMPACT_BLOCK), 0, st, S, holes, slot, after_step ? 1 : 0);
    LAUNCH_CHECK();
    if (!after_step) {
        hipLaunchKernelGGL(k_compact_finish, dim3(1), dim3(64), 0, st, S);
        LAUNCH_CHECK();
    }
    return hipSuccess;
}

hipError_t launch_scale_finish(double* raw, const double* dv_inv, double* rho, int64_t ndof, DevState* S,
                               const StepClose& C, hipStream_t st) {
    int64_t g = (ndof + 1023) / 1024;
    g = g < 1 ? 1 : (g > 1024 ? 1024 : g);
    hipLaunchKernelGGL(k_scale_finish, dim3((unsigned)g), dim3(1024), 0, st, raw, dv_inv, rho, ndof, S, C);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_count_species(DevState* S, int cell_bits, int nsp, unsigned long long* cnt,
                                int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_count_species, dim3(grid), dim3(256), 0, st, S, cell_bits, nsp, cnt);
    LAUNCH_CHECK();
    return hipSuccess;
}

// The first two kernels of launch_scan: out[i] + bsum[i / 1024] is then the
// exclusive scan (a consumer that adds the block sum itself saves the third).
static hipError_t launch_scan_blocks(const int64_t* in, int64_t* out, int64_t len, int64_t* bsum,
                                     int64_t* total, hipStream_t st) {
    const int64_t nb = (len + 1023) / 1024;
    if (nb == 0) {
        return launch_fill_u64((unsigned long long*)total, 1, st);
    }
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(256), 0, st, in, out, len, bsum);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(256), 0, st, bsum, nb, total);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_scan(const int64_t* in, int64_t* out, int64_t len, int64_t* bsum, int64_t* total,
                       hipStream_t st) {
    hipError_t e = launch_scan_blocks(in, out, len, bsum, total, st);
    if (e != hipSuccess || len <= 0) return e;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, out, len, bsum);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_sort(const SortLaunch& L, hipStream_t st) {
    const int tsize = radix_tile(L.capacity);
    const int64_t ntiles = std::max<int64_t>((L.capacity + tsize - 1) / tsize, 1);
    const unsigned blocks = (unsigned)((ntiles + 3) / 4);
    const int passes = (L.cell_bits + RADIX_BITS - 1) / RADIX_BITS;
    const int64_t ncount = ntiles * RADIX;
    uint64_t* kvbuf[2] = {L.kv0, L.kv1};
    const uint64_t* kin = nullptr;
    for (int ps = 0; ps < passes; ++ps) {
        const int shift = ps * RADIX_BITS;
        hipLaunchKernelGGL(k_radix_count, dim3(blocks), dim3(256), 0, st, L.state, L.cell_bits, kin, shift, ntiles,
                           tsize, L.counts);
        LAUNCH_CHECK();
        hipError_t e = launch_scan_blocks(L.counts, L.offsets, ncount, L.bsum, L.total, st);
        if (e != hipSuccess) return e;
        uint64_t* kout = ps + 1 < passes ? kvbuf[ps & 1] : nullptr;
        hipLaunchKernelGGL(k_radix_scatter, dim3(blocks), dim3(256), 0, st, L.state, L.cell_bits, kin, shift, ntiles,
                           tsize, L.offsets, L.bsum, kout, L.perm);
        LAUNCH_CHECK();
        kin = kout;
    }
    hipLaunchKernelGGL(k_set_perm, dim3(1), dim3(64), 0, st, L.state, L.perm);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_apply_order(DevState* S, int64_t capacity, hipStream_t st) {
    const int64_t gblocks = std::max<int64_t>((capacity + 255) / 256, 1);
    hipLaunchKernelGGL(k_sort_gather, dim3((unsigned)gblocks), dim3(256), 0, st, S);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, st, S);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_step_end(DevState* S, hipStream_t st) {
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, st, S);
    LAUNCH_CHECK();
    return hipSuccess;
}

int64_t sort_count_entries(int64_t capacity) {
    const int tsize = radix_tile(capacity);
    return std::max<int64_t>((capacity + tsize - 1) / tsize, 1) * RADIX;
}

hipError_t launch_load(const DevMesh& M, DevState* S, int64_t base, int64_t N, int sp, int rng_sp,
                       int32_t cell_lo, int32_t cell_hi, int cell_bits, uint64_t seed, double cmin,
                       double cmax, double amp, double kw, int axis, const VdfParams& F, double vmax,
                       unsigned long long* errs, hipStream_t st) {
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_load, dim3(fill_grid(N)), dim3(256), 0, st, M, S, base, N, sp, rng_sp, cell_lo,
                       cell_hi, cell_bits, seed, cmin, cmax, amp, kw, axis, F, vmax, errs);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_inject_gen(const InjectLaunch& L, hipStream_t st) {
    const int gen_blocks = inject_gen_blocks(L.cand_cap);
    if (L.all_maxwellian)
        hipLaunchKernelGGL(k_inject_gen_maxwellian, dim3((unsigned)gen_blocks), dim3(256), 0, st,
                           L.mesh, L.inj, L.state, L.spi, L.soff, L.spair, L.seed, L.dt, L.cell_bits, L.cand_cap, L.cx,
                           L.ccs, L.kc, L.errs);
    else
        hipLaunchKernelGGL(k_inject_gen, dim3((unsigned)gen_blocks), dim3(256), 0, st,
                           L.mesh, L.inj, L.state, L.spi, L.soff, L.spair, L.seed, L.dt, L.cell_bits, L.cand_cap, L.cx,
                           L.ccs, L.kc, L.errs);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_inject_append(const InjectLaunch& L, hipStream_t st, bool finish) {
    const int64_t blocks = std::max<int64_t>((L.cand_cap + APPEND_BLOCK - 1) / APPEND_BLOCK, 1);
    hipLaunchKernelGGL(k_inject_append, dim3((unsigned)blocks), dim3(APPEND_BLOCK), 0, st, L.mesh, L.state,
                       L.coef, L.cand_cap, L.cx, L.ccs, L.kc, L.nkept, L.cell_bits, L.rho_raw);
    LAUNCH_CHECK();
    if (finish) {
        hipLaunchKernelGGL(k_inject_finish, dim3(1), dim3(64), 0, st, L.state, L.nkept);
        LAUNCH_CHECK();
    }
    return hipSuccess;
}

int64_t inject_group_count(int64_t cand_cap) { return (cand_cap + INJECT_PER_WAVE - 1) / INJECT_PER_WAVE; }

hipError_t launch_inject(const InjectLaunch& L, hipStream_t st) {
    hipError_t e = launch_inject_gen(L, st);
    if (e != hipSuccess) return e;
    return launch_inject_append(L, st);
}

}  // namespace punc_gpu

namespace punc_gpu {
hipError_t launch_set_step(DevState* S, uint64_t step, hipStream_t st) {
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, st, S, step);
    LAUNCH_CHECK();
    return hipSuccess;
}
}  // namespace punc_gpu

// ---------------------------------------------------------------------------
// Streaming copy of 16-byte vectors: the reference point for achievable HBM
// bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE counters.
// ---------------------------------------------------------------------------
namespace punc_gpu {
__global__ void __launch_bounds__(256) k_copy16(const double2* __restrict__ a, double2* __restrict__ b, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}
hipError_t launch_copy16(const void* a, void* b, int64_t n16, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, (const double2*)a, (double2*)b, n16);
    LAUNCH_CHECK();
    return hipSuccess;
}
}  // namespace punc_gpu

namespace punc_gpu {
hipError_t launch_potential_energy(const DevMesh& M, DevState* S, const SpCoef* coef, int nsp, int cell_bits,
                                   const double* phi, double4* cellPhi, double* part, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_cell_scalar, dim3(fill_grid(M.nc)), dim3(256), 0, st, M, phi, cellPhi);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_potential_energy, dim3(grid), dim3(256), 0, st, M, S, coef, nsp, cell_bits, cellPhi, part);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_density(const DevMesh& M, DevState* S, const SpCoef* coef, int nsp, int cell_bits, bool dg0,
                          double* raw, int grid, hipStream_t st) {
    if (nsp < 1 || nsp > MAX_SPECIES) return hipErrorInvalidValue;
    if (dg0)
        hipLaunchKernelGGL(k_density_dg0, dim3(grid), dim3(256), LDS_TABLE_BYTES, st, S, coef, nsp, cell_bits, raw);
    else
        hipLaunchKernelGGL(k_density_cg1, dim3(grid), dim3(256), LDS_TABLE_BYTES, st, M, S, coef, nsp, cell_bits, raw);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_density_scale(bool dg0, const double* raw, const double* scale, double we, double wi,
                                double* ne, double* ni, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (dg0)
        hipLaunchKernelGGL(k_density_scale<true>, dim3(fill_grid(n)), dim3(256), 0, st, raw, scale, we, wi, ne, ni, n);
    else
        hipLaunchKernelGGL(k_density_scale<false>, dim3(fill_grid(n)), dim3(256), 0, st, raw, scale, we, wi, ne, ni, n);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_statistics(DevState* S, const SpCoef* coef, int nsp, int cell_bits, double* part, double* out,
                             int grid, hipStream_t st) {
    if (nsp < 1 || nsp > MAX_SPECIES) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_statistics, dim3(grid), dim3(256), 0, st, S, coef, nsp, cell_bits, part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_statistics_merge, dim3(1), dim3(256), 0, st, part, grid, out);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_ema(const double* f, double* g, double w, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ema, dim3(fill_grid(n)), dim3(256), 0, st, f, g, w, n);
    LAUNCH_CHECK();
    return hipSuccess;
}
}  // namespace punc_gpu

namespace punc_gpu {
hipError_t launch_pop_digest(DevState* S, int cell_bits, const int32_t* c0, int nsp, unsigned long long* out,
                             int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_pop_digest, dim3(grid), dim3(256), 0, st, S, cell_bits, c0, nsp, out);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_flux_mc(const double* fnormal, const double* farea, int64_t nf, const VdfParams& F, uint64_t seed,
                          int rng_sp, int64_t n_iter, bool do_num, bool do_max, double* num, double* pmax,
                          hipStream_t st) {
    if (nf <= 0 || (!do_num && !do_max)) return hipSuccess;
    const int grid = (int)std::min<int64_t>(nf, 65536);
    hipLaunchKernelGGL(k_flux_mc, dim3(grid), dim3(256), 0, st, fnormal, farea, nf, F, seed, rng_sp, n_iter,
                       do_num ? 1 : 0, do_max ? 1 : 0, num, pmax);
    LAUNCH_CHECK();
    return hipSuccess;
}
}  // namespace punc_gpu

namespace punc_gpu {
hipError_t launch_pack_records(const DevState* S, const SpCoef* coef, int cell_bits, int64_t first, int64_t count,
                               double* rec, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_records, dim3(fill_grid(count)), dim3(256), 0, st, S, coef, cell_bits, first, count,
                       rec);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_unpack_append(const DevMesh& M, DevState* S, const double* rec, int64_t count, const double* spq,
                                const double* spm, int nsp, int cell_bits, double* cx, int32_t* ccs, int64_t* keep,
                                int64_t* pos, int64_t* bsum, int64_t* total, unsigned long long* errs,
                                hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack_locate, dim3(fill_grid(count)), dim3(256), 0, st, M, rec, count, spq, spm, nsp,
                       cell_bits, count, cx, ccs, keep, errs);
    LAUNCH_CHECK();
    hipError_t e = launch_scan(keep, pos, count, bsum, total, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_file_append, dim3(fill_grid(count)), dim3(256), 0, st, S, count, count, cx, ccs, pos);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_file_finish, dim3(1), dim3(64), 0, st, S, total);
    LAUNCH_CHECK();
    return hipSuccess;
}
}  // namespace punc_gpu

// Device-side state shared by the kernels (kernels.hip) and the C ABI (capi.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "punc_types.h"

namespace punc_gpu {

constexpr int MAX_BND = 64;       // boundary ids 1..63
constexpr int MAX_SPECIES = 128;  // species per population

// Particle storage: FP64 structure of arrays plus one packed int32 cell field.
// Particle state: three arrays of 16-byte pairs, (x, y), (z, vx) and
// (vy, vz) - 2 doubles per particle each - and the packed species/cell.
// A pass over all six values is three 16-byte loads (or stores) per lane
// instead of six 8-byte ones; the particle kernels are bound by how many
// vector-memory instructions the texture units process, not by bytes.
struct PopArrays {
    double* xy;
    double* zvx;
    double* vyz;
    int32_t* cs;
};

// Population state in device memory.  Kernels read the array pointers and the
// counters through a pointer, so the cell re-sort can swap buffers and the
// counts can change inside a captured graph without host involvement.
struct DevState {
    PopArrays cur;
    PopArrays alt;
    int64_t n;         // live particles
    int64_t capacity;  // allocated slots per buffer
    int64_t n_holes;   // particles removed by the last relocation (hole list length)
    int64_t n_lost;    // relocation walks that exceeded the hop cap
    int64_t nH;        // compaction: holes below the new count
    int64_t nS;        // compaction: survivors above the new count
    int64_t inj_cand;  // injection: candidates this call
    int64_t inj_kept;  // injection: kept candidates this call
    int64_t overflow;  // particles dropped because capacity was full
    uint64_t step;     // injection stream step index
    // Pending cell order from the last re-sort: cur[perm[j]] is the j-th
    // particle in cell order.  The next fused step reads through it and writes
    // the alternate buffers; other operations apply it first (gather + swap).
    const uint32_t* perm;
    int64_t perm_pending;
    int64_t n_near;    // relocated particles near a facet of their cell (NEAR_FACE_DIST), since last read
    int64_t pad[3];
};

// Per-species coefficients for one call (depends on dt and B).
struct SpCoef {
    double q, m;
    double qm;      // q / m
    double half_m;  // 0.5 * m
    double a;       // accel: dt*(q/m); boris: (0.5*dt)*(q/m)
    double t[3];    // boris: tan(((dt*q)/(2.0*m))*B_i)
    double s[3];    // boris: 2*t_i/(1+|t|^2)
    double pad;
};

// Mesh data on the device.
struct DevMesh {
    const CellRec* rec;
    const CellPlanes* planes;
    const int32_t* bstart;
    const int32_t* bcells;
    int64_t nc;
    int64_t ndof;
    double inside_tol;   // 0: no "certainly inside" shortcut
    double coord_bound;  // margin valid for |x|, |y|, |z| <= coord_bound
    double near_lam;     // below it a barycentric coordinate may be near a facet (MeshTopo::near_lam)
    int gn[3];
    double glo[3];
    double gh[3];
};

// Exterior-facet injection table for one species set.
struct DevInject {
    const double* fverts;   // nf*9 facet vertices
    const double* fnormal;  // nf*3 inward normals
    const int32_t* fcell;   // nf owning cell
    const int32_t* pair_sp; // npairs species of each (species, facet) pair
    const int32_t* pair_f;  // npairs exterior facet of each pair
    const double* num;      // npairs expected count per step (flux*n*dt)
    const double* pmax;     // npairs flux pdf maximum
    int64_t npairs;
};

struct SpInject {
    double vth;
    double vd[3];
    double factor;  // the distribution's normalisation (VdfParams::factor)
    double vth2;
    double vrange;  // rejection box half width in units of vth
    double q, m;
    double density;  // Species::n
    // Random stream identity: draws use key (seed, tag) and the species id
    // rng_sp, and facet ids relative to facet_base, so an instance of a batched
    // mesh draws exactly what it draws as a mesh of its own.
    uint64_t seed;
    int32_t use_seed;  // 0: the seed given to the call
    int32_t rng_sp;
    int64_t facet_base;
    int32_t kind;  // VdfKind of the injected velocity distribution
    int32_t pad0;
    double k;      // kappa (Kappa, KappaCairns)
    double alpha;  // alpha (Cairns, KappaCairns)
};
static_assert(sizeof(SpInject) == 128, "SpInject layout");

}  // namespace punc_gpu

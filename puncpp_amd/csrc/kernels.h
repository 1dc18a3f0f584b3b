// Launch wrappers for the kernels in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "device.h"

namespace punc_gpu {

struct InjectLaunch;

struct ParticleLaunch {
    DevMesh mesh;
    DevState* state;
    const SpCoef* coef;
    int nsp;              // entries of coef
    const double* cellE;  // per-cell field copy (k_cell_field)
    double dt;
    double Ld[3];
    int cell_bits;
    double* rho_raw;
    double* ke_part;
    int32_t* holes;
    unsigned long long* absorbed;
    void* walk;           // fused step: walk list, capacity records of 32 bytes
    int push;  // 0 none, 1 accel, 2 boris
    int move;  // 0 none, 1 plain, 2 periodic
    bool locate;
    bool deposit;
    int dep_mode;  // 0: per-particle atomics, 1: wavefront run reduction, 2: + LDS table
    double* pe_part;  // fused step: per-block potential energy (nullptr: not computed)
    int grid;
    int block;
    // Fused step: the Maxwellian injection generator's blocks run first in
    // the same grid (nullptr: no generation in the step kernel).
    const InjectLaunch* gen;
};

struct InjectLaunch {
    DevMesh mesh;
    DevInject inj;
    DevState* state;
    const SpInject* spi;
    const SpCoef* coef;
    uint64_t seed;
    double dt;
    int cell_bits;
    int64_t cand_cap;     // candidate slots: floor(num)+1 per (species, facet) pair
    const int64_t* soff;  // npairs: first slot of each pair (prefix sums of floor(num)+1)
    const int32_t* spair; // cand_cap: the pair of each slot
    int64_t* nkept;       // kept candidates this call (written by the append)
    double* cx;           // 6 * cand_cap: candidate positions and velocities
    int32_t* ccs;         // cand_cap: packed cell of a kept candidate, else -1
    int32_t* kc;          // inject_group_count(cand_cap): kept per group of slots
    double* rho_raw;  // nullable: deposit the kept particles
    unsigned long long* errs;
    bool all_maxwellian;  // every species with injection pairs draws from a Maxwellian
};

hipError_t launch_particles(const ParticleLaunch& L, hipStream_t st);
// Fused step (push + move + in-cell relocation + deposit, then the
// relocation walk + deposit for the particles that left their cell).
// Requires L.push != 0, L.move != 0, L.locate and L.deposit.
hipError_t launch_step_passes(const ParticleLaunch& L, hipStream_t st);
// Per-cell field copy: cellE[c*16 + 3k + j] = E[dof_k(c)].j.
// E3: 3 FP64 components per dof, packed.
hipError_t launch_cell_field(const DevMesh& M, const double* E3, double* cellE, hipStream_t st);
hipError_t launch_expand_field(const double* E3, double4* E, int64_t ndof, hipStream_t st);
// phi (ndof, device) at each cell's dofs into slots 12..15 of the cell's field
// row, read by the fused step's potential energy; nullptr clears them.
hipError_t launch_cell_potential(const DevMesh& M, const double* phi, double* cellE, hipStream_t st);
hipError_t launch_kinetic(DevState* S, const SpCoef* coef, int cell_bits, double* ke_part, int grid,
                          hipStream_t st);
hipError_t launch_sum_partials(const double* part, int np, double* out, hipStream_t st);
hipError_t launch_fill_f64(double* p, int64_t n, double v, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, int64_t n, hipStream_t st);
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);
// rho = raw * dv_inv, then raw = 0 (raw is all zero outside a deposit).
hipError_t launch_scale_rho(double* raw, const double* dv_inv, double* rho, int64_t n,
                            hipStream_t st);
// Removal of the particles recorded in holes (DevState::n_holes of them):
// survivors from above the new count fill the holes below it.  slot:
// capacity int32 entries, all -1 (left so).  after_step: the fused step's
// arrays, count update left to launch_scale_finish; otherwise the count is
// updated here.
hipError_t launch_compact(DevState* S, const int32_t* holes, int32_t* slot, int grid, bool after_step,
                          hipStream_t st);
// Closing bookkeeping of a fused step, done by block 0 of its last kernel.
struct StepClose {
    const double* ke_part;  // np per-block partials of the fused kernel
    const double* pe_part;  // nullptr: no potential energy (pe_out = NaN)
    int np;
    double* ke_out;
    double* pe_out;
    const int64_t* nkept;   // injected particles kept (nullptr: no injection)
    int advance;            // step counter += 1 (a captured step)
};
// Last kernel of a fused step: rho = raw * dv_inv, raw = 0, and the closing
// bookkeeping (buffer swap, KE/PE sums, count after removal and injection,
// step counter).
hipError_t launch_scale_finish(double* raw, const double* dv_inv, double* rho, int64_t ndof, DevState* S,
                               const StepClose& C, hipStream_t st);
hipError_t launch_count_species(DevState* S, int cell_bits, int nsp, unsigned long long* cnt,
                                int grid, hipStream_t st);
hipError_t launch_scan(const int64_t* in, int64_t* out, int64_t len, int64_t* bsum, int64_t* total,
                       hipStream_t st);
// Cell re-sort (LSD radix sort of (cell, index) pairs, then a gather of the
// particle data into the alternate buffers and a buffer swap).
struct SortLaunch {
    DevState* state;
    int cell_bits;
    int64_t capacity;   // particle slots per buffer
    uint64_t* kv0;      // capacity each: pass scratch ((key << 32) | value), free once
    uint64_t* kv1;      // the sort has run
    uint32_t* perm;     // capacity: the resulting order; must stay untouched until
                        // the next fused step has read it (that step writes its
                        // walk list, so not in the walk-list buffer)
    int64_t* counts;    // sort_count_entries(capacity) each
    int64_t* offsets;
    int64_t* bsum;      // scan scratch: sort_count_entries / 1024 + 2
    int64_t* total;
};
// Computes the cell order and leaves it pending (DevState::perm): the next
// fused step reads through it; launch_apply_order applies it otherwise.
hipError_t launch_sort(const SortLaunch& L, hipStream_t st);
// Gather + swap of a pending cell order (no-op on the device when none is).
hipError_t launch_apply_order(DevState* S, int64_t capacity, hipStream_t st);
// After a fused step: swap in the buffers a permuted step wrote.
hipError_t launch_step_end(DevState* S, hipStream_t st);
int64_t sort_count_entries(int64_t capacity);
hipError_t launch_load(const DevMesh& M, DevState* S, int64_t base, int64_t N, int sp, int rng_sp,
                       int32_t cell_lo, int32_t cell_hi, int cell_bits, uint64_t seed, double cmin,
                       double cmax, double amp, double kw, int axis, const VdfParams& F, double vmax,
                       unsigned long long* errs, hipStream_t st);
// Flux tables by Monte Carlo (create_flux) over nf facets: num[j] when
// do_num, pmax[j] when do_max (fnormal: 3 per facet, farea: 1 per facet).
hipError_t launch_flux_mc(const double* fnormal, const double* farea, int64_t nf, const VdfParams& F, uint64_t seed,
                          int rng_sp, int64_t n_iter, bool do_num, bool do_max, double* num, double* pmax,
                          hipStream_t st);
// Injection in two halves: generation (per-facet counts, candidates,
// rejection-sampled velocities, location; reads only the mesh, the tables and
// the step counter, so it may run beside a particle kernel) and the append of
// the kept candidates at the end of the population.  launch_inject = both.
hipError_t launch_inject_gen(const InjectLaunch& L, hipStream_t st);
// finish = false inside a fused step: the count update is left to its
// launch_scale_finish.
hipError_t launch_inject_append(const InjectLaunch& L, hipStream_t st, bool finish = true);
hipError_t launch_inject(const InjectLaunch& L, hipStream_t st);
// Entries of InjectLaunch::kc for cand_cap slots.
int64_t inject_group_count(int64_t cand_cap);
hipError_t launch_set_step(DevState* S, uint64_t step, hipStream_t st);
hipError_t launch_copy16(const void* a, void* b, int64_t n16, int grid, hipStream_t st);

// Diagnostics.  potential energy: per-cell copy of phi into cellPhi (nc),
// then per-block partial sums into part (grid entries); density: raw sums (2 per dof for CG1, 2 per cell for DG0) into
// raw, which must be zeroed first; density_scale: CG1 n = raw * (dv_inv * w),
// DG0 n = (raw * w) / volume; statistics: per-block (n, mean, M2) per class
// into part (6 * grid), merged into out (6).
hipError_t launch_potential_energy(const DevMesh& M, DevState* S, const SpCoef* coef, int nsp, int cell_bits,
                                   const double* phi, double4* cellPhi, double* part, int grid, hipStream_t st);
hipError_t launch_density(const DevMesh& M, DevState* S, const SpCoef* coef, int nsp, int cell_bits, bool dg0,
                          double* raw, int grid, hipStream_t st);
hipError_t launch_density_scale(bool dg0, const double* raw, const double* scale, double we, double wi,
                                double* ne, double* ni, int64_t n, hipStream_t st);
hipError_t launch_statistics(DevState* S, const SpCoef* coef, int nsp, int cell_bits, double* part, double* out, int grid,
                             hipStream_t st);
hipError_t launch_ema(const double* f, double* g, double w, int64_t n, hipStream_t st);
// Per species (nsp <= MAX_SPECIES): out[s] += sum of particle hashes (cell ids
// relative to c0[s]), out[nsp + s] += count; out must be zeroed first.
hipError_t launch_pop_digest(DevState* S, int cell_bits, const int32_t* c0, int nsp, unsigned long long* out,
                             int grid, hipStream_t st);

// Population files: pack particles [first, first+count) as 8-double records
// (x, y, z, vx, vy, vz, q, m); unpack `count` records, locate them and append
// the ones inside the mesh in record order (cx: 6*count, ccs, keep, pos:
// count each, bsum: count/1024 + 2, total: 1; errs counts records whose
// (q, m) matches no species).
hipError_t launch_pack_records(const DevState* S, const SpCoef* coef, int cell_bits, int64_t first, int64_t count,
                               double* rec, hipStream_t st);
hipError_t launch_unpack_append(const DevMesh& M, DevState* S, const double* rec, int64_t count, const double* spq,
                                const double* spm, int nsp, int cell_bits, double* cx, int32_t* ccs, int64_t* keep,
                                int64_t* pos, int64_t* bsum, int64_t* total, unsigned long long* errs,
                                hipStream_t st);

}  // namespace punc_gpu

/*
 * C ABI of the MI355X particle path for PUNC++ (libpunc_gpu.so).
 *
 * The reference (PUNC++) is a header-only C++ template library over DOLFIN
 * (punc/include/punc/{population,pusher,distributor,injector}.h).  This ABI is
 * what its particle operators bind to when they run on the GPU: plain
 * pointers and sizes, host arrays in and out, device state behind opaque
 * handles.  The C++ drop-in API in include/punc/ sits on top of it.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative value on error; the
 *    message is available from punc_gpu_last_error() (thread-local).  Nothing
 *    throws across the ABI.
 *  - Vectors of points are row-major n x 3 doubles.
 *  - Cells are tetrahedra with ascending vertex ids; local facet i is opposite
 *    local vertex i.  A facet neighbour < 0 is -(boundary id): 1 = exterior,
 *    2.. = objects (reference mesh.h:61-71).
 *  - A population handle is bound to one HIP stream; calls on one handle are
 *    not thread-safe, calls on different handles are.
 */
#ifndef PUNC_GPU_H
#define PUNC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct punc_hostmesh_s punc_hostmesh_t;
typedef struct punc_mesh_s punc_mesh_t;
typedef struct punc_pop_s punc_pop_t;

#define PUNC_GPU_MAX_BND 64

/* ---- errors ------------------------------------------------------------ */
/* Last error message of the calling thread ("" if none). */
const char* punc_gpu_last_error(void);
/* ABI version (increments on incompatible changes). */
int punc_gpu_abi_version(void);

/* ---- synthetic meshes (host) ------------------------------------------- */
/* Box [lo,hi]^3 of n^3 cubes x 6 Kuhn tetrahedra; cells with centroid inside
 * radius sphere_r (>0) are removed and form object boundary 2; exterior is 1.
 * periodic != 0 maps opposite vertices to one CG1 dof.  Stands in for the
 * reference's Mesh(fname) file loading (mesh.cpp:41-92), which needs DOLFIN. */
int punc_gpu_meshgen_box(int n, const double lo[3], const double hi[3], double sphere_r,
                         int periodic, punc_hostmesh_t** out);
int punc_gpu_hostmesh_sizes(const punc_hostmesh_t* m, int64_t* nv, int64_t* nc, int64_t* ndof);
int punc_gpu_hostmesh_copy(const punc_hostmesh_t* m, double* xyz, int32_t* cells,
                           int32_t* facet_bnd, int32_t* vdof);
void punc_gpu_hostmesh_free(punc_hostmesh_t* m);

/* ---- mesh / localizer -------------------------------------------------- */
/* Builds the localizer (facet neighbours, facet planes, barycentric matrices)
 * and uploads the packed cell records.  Replaces Population<3>::Population and
 * init_localizer (population.h:310-389) and Cell<3>::init_barycentric_matrix
 * (population.cpp:181-240).  facet_bnd[c*4+i] gives the boundary id of local
 * facet i when that facet has one adjacent cell.  vdof maps vertex -> CG1 dof. */
int punc_gpu_mesh_create(int64_t nv, const double* xyz, int64_t nc, const int32_t* cells,
                         const int32_t* facet_bnd, const int32_t* vdof, int64_t ndof,
                         punc_mesh_t** out);
void punc_gpu_mesh_destroy(punc_mesh_t* m);
/* Sizes, number of exterior facets and the mesh volume (Mesh::volume, mesh.cpp:147-164). */
int punc_gpu_mesh_info(const punc_mesh_t* m, int64_t* nv, int64_t* nc, int64_t* ndof,
                       int64_t* n_exterior, double* volume);
/* Host copies of the localizer: adj nc*4 (Cell::facet_adjacents), planes nc*16
 * (Cell::facet_plane_coeffs), bary nc*12 (barycentric_matrix rows 0..2),
 * vcoords nc*12 (Cell::vertex_coordinates), dofs nc*4.  Any pointer may be NULL. */
int punc_gpu_mesh_localizer(const punc_mesh_t* m, int64_t* adj, double* planes, double* bary,
                            double* vcoords, int32_t* dofs);
/* element_volume(V, voronoi) (distributor.cpp:35-69): per dof (D+1)/patch volume
 * (voronoi != 0) or 1/patch volume. */
int punc_gpu_element_volume(const punc_mesh_t* m, int voronoi, double* dv_inv);
/* Exterior facets (Mesh::exterior_boundaries, mesh.cpp:166-236): area, 9
 * vertex coordinates, inward unit normal, 9-entry basis, and the owning cell. */
int punc_gpu_exterior_facets(const punc_mesh_t* m, double* area, double* vertices,
                             double* normal, double* basis, int32_t* cell);
/* Point location (punc::locate, population.cpp:24-39): lowest-id cell whose
 * four facet planes all give proj <= 0, else -1. */
int punc_gpu_mesh_locate(const punc_mesh_t* m, int64_t n, const double* xs, int64_t* cells);

/* Maxwellian flux tables for create_flux (injector.cpp:104-171 with
 * Maxwellian::flux_num_particles / flux_max, distributions.cpp:52-72):
 * num[j] = expected particles per unit density per unit time through exterior
 * facet j, pdf_max[j] = maximum of the flux pdf for facet j's normal. */
int punc_gpu_maxwellian_flux(const punc_mesh_t* m, double vth, const double vd[3], double* num,
                             double* pdf_max);

/* Velocity distributions (distributions.h:233-360, distributions.cpp:23-245).
 * kind: 0 Maxwellian, 1 Kappa (kappa > 3/2), 2 Cairns (alpha), 3 KappaCairns
 * (kappa, alpha); vdf_range: half width of the sampling box in units of vth
 * (the reference's defaults: 5, 7, 7, 15). */
#define PUNC_VDF_MAXWELLIAN 0
#define PUNC_VDF_KAPPA 1
#define PUNC_VDF_CAIRNS 2
#define PUNC_VDF_KAPPA_CAIRNS 3
/* The constructor's normalisation factor, Pdf::max(), and the create_flux
 * flags has_flux_number / has_flux_max (any output may be NULL). */
int punc_gpu_vdf_info(int kind, double vth, const double vd[3], double kappa, double alpha, double vdf_range,
                      double* factor, double* pdf_max, int* has_flux_num, int* has_flux_max);
/* create_flux (injector.cpp:104-171) for one species over the exterior
 * facets: closed forms where the distribution has them, otherwise n_iter
 * (the reference: 500000) Monte Carlo samples per facet on the device from a
 * Philox stream keyed by seed, with rng_species naming the species' stream.
 * num[j], pdf_max[j]: per exterior facet. */
int punc_gpu_vdf_flux(punc_mesh_t* m, int kind, double vth, const double vd[3], double kappa, double alpha,
                      double vdf_range, uint64_t seed, int rng_species, int64_t n_iter, double* num,
                      double* pdf_max);

/* create_flux(species, facets) (injector.cpp:104-171) for one species over
 * nf facets given by their inward unit normals (nf*3) and areas (nf), as the
 * reference's form takes the facet list rather than a mesh; otherwise as
 * punc_gpu_vdf_flux (the Monte Carlo runs on the current device). */
int punc_gpu_vdf_flux_facets(int64_t nf, const double* normal, const double* area, int kind, double vth,
                             const double vd[3], double kappa, double alpha, double vdf_range, uint64_t seed,
                             int rng_species, int64_t n_iter, double* num, double* pdf_max);

/* punc_gpu_vdf_flux_facets without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * facets and the tables are staged through a fixed device array that each
 * GPU holds from the first call on it until the process ends, at most
 * 131072 facets per call (larger nf is rejected with an error).  Calls that
 * overlap on one GPU share that array and race.  Same arguments and results
 * as punc_gpu_vdf_flux_facets, bit for bit. */
int punc_gpu_vdf_flux_facets_noalloc(int64_t nf, const double* normal, const double* area, int kind, double vth,
                                     const double vd[3], double kappa, double alpha, double vdf_range,
                                     uint64_t seed, int rng_species, int64_t n_iter, double* num, double* pdf_max);

/* ---- population -------------------------------------------------------- */
/* A population of nspecies species with weighted charge q[s] and mass m[s]
 * (Species::q, Species::m, population.cpp:61-64), room for capacity particles.
 * nspecies may be 0; species can be added later (up to 128).
 * stream: a hipStream_t, or NULL to let the library create one. */
int punc_gpu_pop_create(punc_mesh_t* m, int nspecies, const double* q, const double* mass,
                        int64_t capacity, void* stream, punc_pop_t** out);
void punc_gpu_pop_destroy(punc_pop_t* p);
/* Registers species (q, mass) and returns its index (the existing index when
 * already registered).  Population<3> registers species on first use, as the
 * reference's particles carry q and m (population.h:74-94). */
int punc_gpu_pop_add_species(punc_pop_t* p, double q, double mass, int* index);
/* The population's species: their number, and (when q / mass are non-NULL,
 * nspecies entries each) charge and mass in index order. */
int punc_gpu_pop_species(punc_pop_t* p, int* nspecies, double* q, double* mass);
/* add_particles(xs, vs, q, m) (population.h:411-435): locates each point and
 * appends it as species s; points outside the mesh are skipped (the reference
 * indexes out of bounds there) and *n_added tells how many were kept. */
int punc_gpu_pop_add(punc_pop_t* p, int species, int64_t n, const double* xs, const double* vs,
                     int64_t* n_added);
/* Appends particles with known species and cells (restart from a saved
 * state); cells must be valid cell ids. */
int punc_gpu_pop_add_located(punc_pop_t* p, int64_t n, const double* xs, const double* vs,
                             const int32_t* species, const int64_t* cells);
/* load_particles (injector.h:244-278) on the device for species s: N particles
 * with positions by rejection against inside(x)*(1+amp*cos(kw*x[axis]))
 * (amp = 0: UniformPosition) and Maxwellian velocities by inverse CDF. */
int punc_gpu_pop_load(punc_pop_t* p, int species, int64_t N, uint64_t seed, double amp, double kw,
                      int axis, double vth, const double vd[3]);
/* load_particles with any velocity distribution: Maxwellian velocities by
 * inverse CDF (as punc_gpu_pop_load), the others by rejection against
 * Pdf::max() in the box vd +- vdf_range*vth (injector.cpp:25-49). */
int punc_gpu_pop_load_vdf(punc_pop_t* p, int species, int64_t N, uint64_t seed, double amp, double kw, int axis,
                          int kind, double vth, const double vd[3], double kappa, double alpha, double vdf_range);
int punc_gpu_pop_size(punc_pop_t* p, int64_t* n);
/* Particles per species (num_of_particles/positives/negatives derive from it,
 * population.h:701-744). */
int punc_gpu_pop_counts(punc_pop_t* p, int64_t* per_species);
/* Download all particles: xs, vs n*3, species n, cells n (any may be NULL). */
int punc_gpu_pop_get(punc_pop_t* p, double* xs, double* vs, int32_t* species, int64_t* cells);
/* Population::save_file (population.h:807-839): every particle as the
 * reference's 64-byte record x[3], v[3], q, m (binary != 0) or as the text line
 * "%.17g\t" x 6, "%.17g\t %.17g\t", "\n"; in cell order (the population is
 * re-sorted by cell first, as the reference writes cell by cell). */
int punc_gpu_pop_save_file(punc_pop_t* p, const char* fname, int binary);
/* Population::load_file (population.h:841-880): appends the records of a
 * file in that format (text: the numbers of each line fill x, v, q, m in
 * turn; a line that gives fewer keeps the previous line's values), locating
 * each point and skipping those outside the mesh.  Every new (q, m) pair
 * becomes a species (punc_gpu_pop_add_species).  n_added may be NULL.  On
 * an error (species limit, capacity, HIP failure) the population keeps its
 * particles and species from before the call and *n_added is not set. */
int punc_gpu_pop_load_file(punc_pop_t* p, const char* fname, int binary, int64_t* n_added);

/* ---- the particle operators -------------------------------------------- */
/* Upload the CG1 electric field, ndof*3 doubles (E_x,E_y,E_z per dof).
 * From memory of punc_gpu_host_alloc the copy is one DMA transfer; from
 * pageable memory the runtime stages it. */
int punc_gpu_set_field(punc_pop_t* p, const double* E);
/* Page-locked host memory (hipHostMalloc) for arrays that cross the host
 * boundary every step: the drop-in layer's Function values live in it.  No
 * reference counterpart (the reference's field dofs never leave the host).
 * bytes 0 allocates a minimal block; free with punc_gpu_host_free. */
int punc_gpu_host_alloc(int64_t bytes, void** ptr);
int punc_gpu_host_free(void* ptr);
/* accel_cg1(pop, E, dt) (pusher.h:154-201); returns the mid-step KE. */
int punc_gpu_accel(punc_pop_t* p, double dt, double* KE);
/* boris_cg1(pop, E, B, dt) (pusher.h:282-372); returns 0.5*m*|v-|^2 summed. */
int punc_gpu_boris(punc_pop_t* p, const double B[3], double dt, double* KE);
/* move(pop, dt) (pusher.h:648-662). */
int punc_gpu_move(punc_pop_t* p, double dt);
/* move_periodic(pop, dt, Ld) (pusher.h:666-681). */
int punc_gpu_move_periodic(punc_pop_t* p, double dt, const double Ld[3]);
/* Population::update (population.h:641-699): relocation walk, removal of
 * particles that left the mesh.  absorbed[s*PUNC_GPU_MAX_BND + b] receives the
 * number of species-s particles that left through boundary b this call
 * (may be NULL).  Object charge/current bookkeeping is done by the caller. */
int punc_gpu_update(punc_pop_t* p, int64_t* absorbed);
/* distribute_cg1(pop, rho, dv_inv) (distributor.h:168-210).  dv_inv (ndof) is
 * uploaded when non-NULL and kept for later calls; rho (ndof) may be NULL to
 * keep the result on the device. */
int punc_gpu_distribute(punc_pop_t* p, const double* dv_inv, double* rho);
/* kinetic_energy(pop) (diagnostics.h:248-265). */
int punc_gpu_kinetic_energy(punc_pop_t* p, double* KE);
/* particle_potential_energy_cg1(pop, phi) (diagnostics.h:379-412): sum over
 * the particles of 0.5 q phi(x), phi given by its CG1 dof values (ndof). */
int punc_gpu_potential_energy(punc_pop_t* p, const double* phi, double* PE);
/* density_cg1(V, pop, species, ne, ni, dv_inv) (diagnostics.h:492-555):
 * particles with q < 0 count as electrons, the others as ions; weight_e and
 * weight_i are species[0].weight and species[1].weight.  dv_inv as for
 * punc_gpu_distribute (NULL: the one already uploaded, else the Voronoi
 * volumes).  ne, ni (ndof each) may be NULL to keep the result on the
 * device for punc_gpu_density_ema. */
int punc_gpu_density_cg1(punc_pop_t* p, const double* dv_inv, double weight_e, double weight_i, double* ne,
                         double* ni);
/* density_dg0(Q, pop, species, ne, ni) (diagnostics.h:432-465): q > 0 counts
 * as ions, the others as electrons; one value per cell (nc each, indexed by
 * cell id), weights summed and divided by the cell volume. */
int punc_gpu_density_dg0(punc_pop_t* p, double weight_e, double weight_i, double* ne, double* ni);
/* Population::statistics(stats) (population.h:746-805): stats[0], stats[1]
 * mean and sample standard deviation of the speed of the q < 0 particles,
 * stats[2], stats[3] of the q > 0 ones; a class without particles leaves its
 * entries unchanged. */
int punc_gpu_statistics(punc_pop_t* p, double* stats);
/* ema(rho, rho_ema, dt, tau) (diagnostics.h:573-588) with rho the
 * population's last deposited charge density and rho_ema kept on the device
 * (zero when first used; reset != 0 zeroes it first).  g (ndof) receives
 * rho_ema when non-NULL. */
int punc_gpu_rho_ema(punc_pop_t* p, double dt, double tau, int reset, double* g);
/* ema(f, g, dt, tau) (diagnostics.h:573-588) over caller arrays of n values:
 * g = w f + (1 - w) g with w = 1 - exp(-dt/tau), computed on the current
 * device (the reference's form, for Functions that live on the host). */
int punc_gpu_ema(int64_t n, const double* f, double* g, double dt, double tau);
/* ema(ne, ne_ema, dt, tau) and ema(ni, ni_ema, dt, tau) over the last density
 * computed by punc_gpu_density_cg1/dg0, the averages kept on the device
 * (reset != 0, needed when the density's length changed, zeroes them). */
int punc_gpu_density_ema(punc_pop_t* p, double dt, double tau, int reset, double* ne_g, double* ni_g);
/* Re-sort the particles by cell (counting sort, order within a cell
 * arbitrary).  Enqueued on the population's stream; returns without waiting. */
int punc_gpu_sort(punc_pop_t* p);

/* ---- field solve -------------------------------------------------------- */
/* PoissonSolver (poisson.cpp:260-376, Potential3D.ufl: eps0 grad(phi).grad(v)
 * dx = rho v dx) and ESolver (efield.cpp:41-75, EField3D.ufl: the vector-CG1
 * L2 projection of -grad(phi)) on the device.  Dirichlet conditions: the
 * exterior boundary (boundary id 1) when dirichlet_exterior, and objects
 * 0..n_objects-1 (boundary ids 2..) at given potentials; remove_null_space
 * for problems without Dirichlet conditions (periodic boxes): the
 * right-hand side is made orthogonal to the constant vector and the
 * potential has zero mean.  Jacobi-preconditioned conjugate gradients to the
 * reference's tolerances (relative 1e-12, absolute 1e-14). */
typedef struct punc_field_s punc_field_t;
int punc_gpu_field_create(punc_mesh_t* m, double eps0, int dirichlet_exterior, int n_objects, int remove_null_space,
                          punc_field_t** out);
void punc_gpu_field_destroy(punc_field_t* f);
/* Iteration limit per solve (default 20000; failing to converge is an error). */
int punc_gpu_field_set_max_iterations(punc_field_t* f, int maxit);
/* warm != 0: each solve starts from the previous solution (the reference's
 * solvers start from zero; the result agrees to the solver tolerance). */
int punc_gpu_field_set_warm_start(punc_field_t* f, int warm);
/* Preconditioner of the Poisson solves: 1 smoothed-aggregation algebraic
 * multigrid, the default whenever its hierarchy can be built (the reference's
 * PETSc solver takes hypre's AMG, poisson.cpp:299-304); 0 Jacobi.  Setting 1
 * fails when there is no hierarchy (operators of at most 256 dofs, or
 * coarsening that stalls).  The E projection is always Jacobi-preconditioned
 * (a mass matrix).  No counterpart in the reference's API beyond that choice
 * of PETSc preconditioner. */
int punc_gpu_field_set_preconditioner(punc_field_t* f, int kind);
/* kind: current preconditioner; levels: multigrid levels above the dense
 * coarsest solve; coarse: that solve's size (any pointer may be NULL). */
int punc_gpu_field_get_preconditioner(punc_field_t* f, int* kind, int* levels, int64_t* coarse);
/* rho (ndof) -> phi (ndof).  exterior_values: one value per dof, read at the
 * exterior Dirichlet dofs (NULL: 0); object_potentials: n_objects values
 * (NULL: 0).  iterations may be NULL. */
int punc_gpu_poisson_solve(punc_field_t* f, const double* rho, const double* exterior_values,
                           const double* object_potentials, double* phi, int* iterations);
/* phi (ndof) -> E (3*ndof, dof-major x, y, z). */
int punc_gpu_efield_solve(punc_field_t* f, const double* phi, double* E, int* iterations);
/* Reverse of punc_gpu_efield_solve: given gE = dJ/dE (3*ndof, dof-major),
 * gphi = dJ/dphi (ndof) = B^T M^-1 gE, with M the vector-CG1 mass matrix and
 * B the gradient form of the projection (EField3D.ufl, efield.cpp:70-75;
 * E = M^-1 B phi, M symmetric).  One mass-matrix CG to the solver tolerance.
 * No counterpart in the reference (used by the adjoint). */
int punc_gpu_efield_adjoint(punc_field_t* f, const double* gE, double* gphi, int* iterations);
/* Charge on each object of the potential phi (ndof): the reference's Charge
 * form eps0 grad(phi).n ds over the object's facets (Charge.ufl), n_objects
 * values. */
int punc_gpu_object_charges(punc_field_t* f, const double* phi, double* charges);
/* Capacitance-matrix circuit (CircuitCM, object_CM.cpp:53-357; charge-sharing
 * sets object.cpp:68-135): n_objects Laplace solves give the inverse
 * capacitance matrix (inv_capacitance: n_objects^2 values, row-major, may be
 * NULL); voltage sources (node_a -1 for ground) and current sources between
 * objects, dt for the current sources. */
int punc_gpu_circuit_set(punc_field_t* f, double dt, int n_vsources, const int* v_a, const int* v_b,
                         const double* v_value, int n_isources, const int* i_a, const int* i_b, const double* i_value,
                         double* inv_capacitance);
/* PoissonSolver::solve_circuit (poisson.cpp:364-376): charges (n_objects,
 * in/out) are the objects' collected charges; potentials (out) the object
 * potentials the circuit sets; phi (ndof, may be NULL). */
int punc_gpu_poisson_solve_circuit(punc_field_t* f, const double* rho, double* charges, double* potentials,
                                   double* phi, int* iterations);
/* The population's last deposited rho -> phi -> E, written straight into the
 * population's field for the next push (no host round trip); phi and E are
 * copied out when non-NULL; iterations[0..1] (may be NULL): Poisson and E. */
int punc_gpu_field_solve_pop(punc_field_t* f, punc_pop_t* p, const double* exterior_values,
                             const double* object_potentials, double* phi, double* E, int* iterations);
/* As punc_gpu_field_solve_pop with the circuit's solve (charges in/out,
 * potentials out, as punc_gpu_poisson_solve_circuit). */
int punc_gpu_field_solve_pop_circuit(punc_field_t* f, punc_pop_t* p, double* charges, double* potentials, double* phi,
                                     double* E, int* iterations);

/* ---- injection --------------------------------------------------------- */
/* Per species Maxwellian flux source: thermal speed, drift, density n
 * (Species::n) and rejection box half width in units of vth (Pdf::vdf_range). */
int punc_gpu_injector_set(punc_pop_t* p, int species, double vth, const double vd[3],
                          double density, double vrange);
/* Flux source with any velocity distribution (kind, kappa, alpha as for
 * punc_gpu_vdf_info): num / pdf_max are the species' create_flux tables (one
 * entry per exterior facet, e.g. from punc_gpu_vdf_flux), or both NULL to
 * compute them with seed 0 and n_iter 500000.  Registering the same species
 * with unchanged parameters and tables keeps the built injection tables. */
int punc_gpu_injector_set_vdf(punc_pop_t* p, int species, int kind, double vth, const double vd[3], double kappa,
                              double alpha, double density, double vdf_range, const double* num,
                              const double* pdf_max);
/* Batched instances.  Binds species s to one instance of a concatenated mesh:
 * its particles live in cells [cell_begin, cell_end) (loading and host
 * location stay there), it injects through exterior facets [facet_begin,
 * facet_end), and its random streams use `seed`, species id rng_species and
 * facet ids counted from facet_begin - so the instance draws exactly what it
 * draws as a mesh of its own with that seed and species index.  No
 * counterpart in the reference (one instance per process there). */
int punc_gpu_species_instance(punc_pop_t* p, int species, int rng_species, uint64_t seed,
                              int64_t cell_begin, int64_t cell_end, int64_t facet_begin,
                              int64_t facet_end);
/* inject_particles(pop, species, facets, dt) (injector.h:157-232) with a
 * counter-based stream keyed by (seed, step).  n_injected may be NULL. */
int punc_gpu_inject(punc_pop_t* p, double dt, uint64_t seed, uint64_t step, int64_t* n_injected);

/* ---- fused step -------------------------------------------------------- */
/* One particle half-step as the reference time loop orders it
 * (interaction.cpp:384-422): push (pusher 1 = accel_cg1, 2 = boris_cg1, with
 * push_dt), move by dt (periodic when Ld != NULL), update, inject (when
 * inject != 0), and the CG1 deposit of the resulting population for the next
 * distribute_cg1, all on the device.  Results stay on the device until read
 * by punc_gpu_step_results / punc_gpu_get_rho. */
int punc_gpu_step(punc_pop_t* p, int pusher, double push_dt, const double B[3], double dt,
                  const double* Ld, int inject, uint64_t seed, uint64_t step);
/* Captures one punc_gpu_step (same arguments) into a HIP graph; each replay
 * advances the injection step index, starting at first_step.  All buffers are
 * prepared before the capture; the graph holds kernel nodes only. */
int punc_gpu_step_graph(punc_pop_t* p, int pusher, double push_dt, const double B[3], double dt,
                        const double* Ld, int inject, uint64_t seed, uint64_t first_step);
/* Replays the captured step count times on the population's stream. */
int punc_gpu_graph_launch(punc_pop_t* p, int count);
/* KE of the last push, absorbed counts since the last read (nspecies*64), and
 * particles injected by the last step.  Any pointer may be NULL. */
int punc_gpu_step_results(punc_pop_t* p, double* KE, int64_t* absorbed, int64_t* n_injected);
/* Particles that a relocation (punc_gpu_update or a fused step, eager or
 * replayed) left within 1e-12 of a facet plane of their cell, |proj| as
 * relocate_fast evaluates it (population.h:547-565), summed since the last
 * call (which resets the count).  Where |proj| is that small the walk's
 * decision rests on rounding: these are the particles whose cell another
 * evaluation order could place differently.  Reporting only; the relocation
 * itself is unchanged. */
int punc_gpu_near_face(punc_pop_t* p, int64_t* n_near);
/* Order-independent digest of the population, per species (nspecies entries
 * each): digest[s] the wrapping 64-bit sum of a hash of every species-s
 * particle's x, y, z, vx, vy, vz bit patterns and its cell id (relative to
 * the species' instance cell range, punc_gpu_species_instance), counts[s]
 * (may be NULL) the number of such particles.  Equal particle sets give equal
 * digests whatever the order, so a batched instance can be compared with its
 * single-instance run without downloading either.  No counterpart in the
 * reference. */
int punc_gpu_pop_digest(punc_pop_t* p, uint64_t* digest, int64_t* counts);
/* rho (ndof) from the last deposit. */
int punc_gpu_get_rho(punc_pop_t* p, double* rho);
/* The potential phi (ndof) at the particles' positions at the start of each
 * fused step: with it set, every punc_gpu_step / graph replay also computes
 * particle_potential_energy_cg1 (diagnostics.h:379-412) of the positions it
 * starts from - the pair the reference's loop evaluates before the push
 * (interaction.cpp:368-371) - inside the particle kernel, read back with
 * punc_gpu_step_potential_energy.  phi == NULL switches it off.
 * punc_gpu_field_solve_pop sets it from the solved potential. */
int punc_gpu_set_potential(punc_pop_t* p, const double* phi);
/* Potential energy computed by the last fused step (NaN when no potential
 * was set). */
int punc_gpu_step_potential_energy(punc_pop_t* p, double* PE);

/* ---- adjoint ----------------------------------------------------------- */
/* Reverse mode through K steps of boris_cg1 + move (+ the relocation walk)
 * and the CG1 deposit of the final positions, along the forward run's cell
 * path, with step checkpoints.  No counterpart in the reference.
 * n particles with per-particle species (q, mass per species). */
typedef struct punc_adj_s punc_adj_t;
const char* punc_gpu_adjoint_last_error(void);
int punc_gpu_adjoint_create(punc_mesh_t* m, int nspecies, const double* q, const double* mass, int64_t n,
                            const int32_t* species, punc_adj_t** out);
void punc_gpu_adjoint_destroy(punc_adj_t* a);
/* Forward K steps from (x0, cell0, v0) in the fixed field E (ndof*3), with B,
 * dt and the periodic box Ld (NULL: not periodic), keeping a checkpoint every
 * `every` steps.  w (ndof, may be NULL) receives w_j = sum_p q_p lambda_j(x_K)
 * over particles still in the mesh; xK, vK, cellK (may be NULL) the final state. */
int punc_gpu_adjoint_forward(punc_adj_t* a, const double* x0, const int64_t* cell0, const double* v0,
                             const double* E, const double B[3], double dt, const double* Ld, int K,
                             int every, double* w, double* xK, double* vK, int64_t* cellK);
/* Given g = dJ/dw (ndof), returns dJ/dE (ndof*3), dJ/dv0 and dJ/dx0 (n*3);
 * any output may be NULL. */
int punc_gpu_adjoint_backward(punc_adj_t* a, const double* g, double* dJdE, double* dJdv0, double* dJdx0);
/* The collected charge of the last forward: per particle, weight the weight
 * (K - n - s)/K of its charge in Q_b, the time average over the K steps of
 * the charge boundary b has collected (n: the step in which it left, s in
 * [0, 1): the fraction of that step at which its path crosses the plane of
 * the facet the walk left through); bnd the boundary it left through and
 * step the step (0 and -1 while in the mesh).  Q_b = sum_p q_p weight_p over
 * bnd_p == b is continuous in the inputs along a fixed crossing path, unlike
 * the collected charge itself.  irregular: exits where the walk left through
 * a plane the step's segment does not cross (weight uses s clamped to 0 or
 * 1; no sensitivity).  Any pointer may be NULL. */
int punc_gpu_adjoint_collected(punc_adj_t* a, double* weight, int32_t* bnd, int32_t* step, int64_t* irregular);
/* As punc_gpu_adjoint_backward with also gq[b] = dJ/dQ_b (64 values, index =
 * boundary id; NULL: 0), the derivative of J with respect to the collected
 * charges above. */
int punc_gpu_adjoint_backward_q(punc_adj_t* a, const double* g, const double* gq, double* dJdE, double* dJdv0,
                                double* dJdx0);

/* ---- launch control, timing -------------------------------------------- */
/* Launch shape of the particle kernels: grid size (each block takes one
 * contiguous chunk of particles), block size (64, 128 or 256) and deposit
 * mode (0: four FP64 atomics per particle; 1: wavefront runs of equal cells
 * are summed across lanes first; 2: run sums go into a per-block LDS table
 * keyed by dof, flushed with one atomic per occupied slot; 3: every lane
 * adds its four contributions to that LDS table directly; 4: lanes of a
 * quad with the same cell are summed first, then added as in 3).  Default:
 * 128 x CUs, 256, mode 3. */
int punc_gpu_set_launch(punc_pop_t* p, int grid, int block, int deposit_mode);
int punc_gpu_get_launch(punc_pop_t* p, int* grid, int* block, int* deposit_mode);
/* Where a fused step with Maxwellian-only injection generates its injected
 * particles: 1 - the generator's blocks run first in the step kernel's own
 * grid; 0 (default) - a generator kernel on a second stream beside the step
 * kernel, joined before the append.  The particles drawn are the same either
 * way.  A change drops a captured step graph (its launch would then fail
 * until the step is captured again).  No counterpart in the reference. */
int punc_gpu_set_inject_in_step(punc_pop_t* p, int on);
int punc_gpu_get_inject_in_step(punc_pop_t* p, int* on);
/* Autotuner: times `reps` fused steps (arguments as punc_gpu_step) for every
 * launch shape (grid 4..128 x CUs, block 128/256, deposit mode 0..4) on a
 * snapshot of the population, restores the snapshot (particles, counts,
 * absorbed counters, rho, KE), and keeps the fastest shape.  With injection
 * from Maxwellians only, the fastest shape is also timed with the injection
 * generator's blocks inside the step kernel's grid instead of in a kernel on
 * a second stream, and the faster of the two is kept (it pays on small
 * populations, where the join of the streams is a visible part of a step).
 * A captured step graph is dropped.  No counterpart in the reference. */
int punc_gpu_autotune(punc_pop_t* p, int pusher, double push_dt, const double B[3], double dt,
                      const double* Ld, int inject, uint64_t seed, int reps, int* best_grid,
                      int* best_block, int* best_mode, double* best_ms);
/* Record HIP events around every launch of the main particle kernel. */
int punc_gpu_timing(punc_pop_t* p, int enable);
int punc_gpu_timing_read(punc_pop_t* p, double* ms_total, int64_t* launches);
int punc_gpu_synchronize(punc_pop_t* p);
/* Streaming-copy reference kernel: copies nbytes between two fresh buffers
 * reps times and returns the mean time of one copy (achievable bandwidth,
 * and the scale for HBM byte counters).  No counterpart in the reference. */
int punc_gpu_copy_probe(int64_t nbytes, int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* PUNC_GPU_H */

// Field solve on the device: the CG1 Poisson problem of PoissonSolver
// (reference poisson.cpp:260-376, Potential3D.ufl) and the vector-CG1 L2
// projection E = -grad(phi) of ESolver (efield.cpp:41-75, EField3D.ufl).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mesh_host.h"

namespace punc_gpu {

struct FieldSolver;

// A voltage or current source between objects node_a (-1: ground) and
// node_b (reference object.h:42-70).
struct CircuitSource {
    int node_a, node_b;
    double value;
};

// Dirichlet conditions: the exterior boundary (boundary id 1) when
// dirichlet_exterior, and the boundaries of objects 0..n_objects-1 (ids 2..).
// remove_null_space: the constant vector is the null space (no Dirichlet
// condition at all, e.g. a periodic box); the right-hand side is
// orthogonalised against it and the solution has zero mean.
FieldSolver* field_create(const HostMesh& m, const MeshTopo& t, double eps0, bool dirichlet_exterior, int n_objects,
                          bool remove_null_space, int device, hipStream_t stream, std::string& err);
void field_destroy(FieldSolver* F);

int64_t field_ndof(const FieldSolver* F);
int field_num_objects(const FieldSolver* F);
double* field_phi(FieldSolver* F);  // device, ndof
double* field_E(FieldSolver* F);    // device, 3*ndof, dof-major (x, y, z)

// Poisson: d_rho (device, ndof) -> field_phi.  ext_values: host, one value
// per dof, read at the exterior Dirichlet dofs (NULL: 0); object_potentials:
// host, n_objects values (NULL: 0).  Solves to |r| <= max(rtol |b|, atol)
// with the reference's tolerances (rtol 1e-12, atol 1e-14).
bool field_poisson(FieldSolver* F, const double* d_rho, const double* ext_values, const double* object_potentials,
                   int* iterations, std::string& err);
// E projection of d_phi (device, ndof) into field_E.
bool field_efield(FieldSolver* F, const double* d_phi, int* iterations, std::string& err);
// Reverse of the E projection: d_gphi (device, ndof) = B^T M^-1 d_gE (device,
// 3*ndof dof-major), the derivative of J(E(phi)) in phi given dJ/dE.
bool field_efield_adjoint(FieldSolver* F, const double* d_gE, double* d_gphi, int* iterations, std::string& err);
// Charges on the objects (the Charge form, Charge.ufl) of the potential d_phi
// (device, ndof): out[j] for j < n_objects (host).
bool field_object_charges(FieldSolver* F, const double* d_phi, double* out, std::string& err);
// The capacitance-matrix circuit (CircuitCM, object_CM.cpp:53-357, with the
// charge-sharing sets of object.cpp:68-135): the inverse capacitance matrix
// from n_objects Laplace solves (inv_capacitance: n x n, row-major, may be
// NULL), then the circuit matrix of the sources.
bool field_set_circuit(FieldSolver* F, double dt, const std::vector<CircuitSource>& vsources,
                       const std::vector<CircuitSource>& isources, double* inv_capacitance, std::string& err);
// PoissonSolver::solve_circuit (poisson.cpp:364-376): current sources added
// to the object charges, a solve with the objects grounded, the image
// charges and the circuit giving the object potentials and charges, then the
// corrected solve.  charges: in/out (n_objects); potentials: out.
bool field_poisson_circuit(FieldSolver* F, const double* d_rho, double* charges, double* potentials, int* iterations,
                           std::string& err);
// Iteration limit of every solve (default 20000).
void field_set_max_iterations(FieldSolver* F, int maxit);
// Start each solve from the previous solution instead of zero (default off,
// as the reference's solvers start from zero); the answer is the same to the
// solver tolerance, reached in fewer iterations when the field changes little.
void field_set_warm_start(FieldSolver* F, bool warm);
// Preconditioner of the Poisson (stiffness) solves: 1 smoothed-aggregation
// multigrid (the default when its hierarchy could be built), 0 Jacobi.  The E
// projection (mass matrix) is always Jacobi-preconditioned.
bool field_set_preconditioner(FieldSolver* F, int kind, std::string& err);
// The current preconditioner (0/1); levels: multigrid levels above the dense
// coarsest solve, coarse: its size (0 without a hierarchy).
int field_preconditioner(const FieldSolver* F, int* levels, int64_t* coarse);

}  // namespace punc_gpu

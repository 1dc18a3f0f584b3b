// Drop-in for PoissonSolver of punc/include/punc/poisson.h (PUNC++) on the
// device: the CG1 Poisson problem eps0 grad(phi).grad(v) dx = rho v dx with a
// Dirichlet exterior boundary and objects at given potentials, or a periodic
// box with the null space removed (poisson.cpp:260-376).
#ifndef PUNC_GPU_POISSON_H
#define PUNC_GPU_POISSON_H

#include <punc/mesh.h>
#include <punc/object.h>
#include <punc/population.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace punc {

class PoissonSolver {
public:
    //! exterior_dirichlet: phi fixed on the exterior boundary (the
    //! reference's ext_bc); objects: held at their potentials (Object::potential);
    //! remove_null_space: no Dirichlet condition (periodic box).
    PoissonSolver(const Mesh& mesh, const ObjectVector& objects, bool exterior_dirichlet = true, double eps0 = 1.0,
                  bool remove_null_space = false)
        : mesh_(mesh), nobj_(objects.size()) {
        punc_field_t* h = nullptr;
        check(punc_gpu_field_create(mesh.handle(), eps0, exterior_dirichlet ? 1 : 0, (int)objects.size(),
                                    remove_null_space ? 1 : 0, &h));
        handle_.reset(h, punc_gpu_field_destroy);
    }

    //! The preconditioner of the solves, by the reference's PETSc names
    //! (poisson.cpp:299-304): "hypre_amg" or "amg" (smoothed-aggregation
    //! multigrid, the default where a hierarchy can be built) or "jacobi".
    void set_preconditioner(const std::string& name) {
        int kind = -1;
        if (name == "hypre_amg" || name == "amg") kind = 1;
        if (name == "jacobi") kind = 0;
        if (kind < 0) throw std::invalid_argument("preconditioner: hypre_amg, amg or jacobi");
        check(punc_gpu_field_set_preconditioner(handle_.get(), kind));
    }

    //! solve (poisson.cpp:349-362): phi from rho, the objects at their
    //! potentials; exterior_values (one per dof) for a non-zero exterior.
    void solve(Function& phi, const Function& rho, const ObjectVector& objects,
               const std::vector<double>& exterior_values = {}) {
        std::vector<double> pot = potentials(objects);
        check(punc_gpu_poisson_solve(handle_.get(), rho.values.data(),
                                     exterior_values.empty() ? nullptr : exterior_values.data(),
                                     pot.empty() ? nullptr : pot.data(), phi.values.data(), &iterations));
    }

    //! The population's deposited rho -> phi -> E, left on the device as the
    //! field of the population's next push (no host round trip); phi and E
    //! (when given) receive copies.
    template <typename PopulationType>
    void solve_population(PopulationType& pop, const ObjectVector& objects, Function* phi = nullptr,
                          Function* E = nullptr) {
        std::vector<double> pot = potentials(objects);
        int its[2] = {0, 0};
        check(punc_gpu_field_solve_pop(handle_.get(), pop.handle(), nullptr, pot.empty() ? nullptr : pot.data(),
                                       phi ? phi->values.data() : nullptr, E ? E->values.data() : nullptr, its));
        iterations = its[0];
    }

    //! solve_circuit (poisson.cpp:364-376) with the circuit set by CircuitCM:
    //! the objects' charges are read and, with their potentials, updated.
    void solve_circuit(Function& phi, const Function& rho, const ObjectVector& objects) {
        std::vector<double> q(nobj_), v(nobj_);
        for (size_t j = 0; j < nobj_; ++j) q[j] = objects[j]->charge;
        check(punc_gpu_poisson_solve_circuit(handle_.get(), rho.values.data(), q.data(), v.data(),
                                             phi.values.data(), &iterations));
        for (size_t j = 0; j < nobj_; ++j) {
            objects[j]->charge = q[j];
            objects[j]->set_potential(v[j]);
        }
    }

    punc_field_t* handle() const { return handle_.get(); }
    size_t num_objects() const { return nobj_; }
    int iterations = 0;  ///< CG iterations of the last solve

private:
    std::vector<double> potentials(const ObjectVector& objects) const {
        std::vector<double> pot(nobj_, 0.0);
        for (size_t j = 0; j < nobj_ && j < objects.size(); ++j) pot[j] = objects[j]->get_potential();
        return pot;
    }
    const Mesh& mesh_;
    size_t nobj_;
    std::shared_ptr<punc_field_t> handle_;
};

//! CircuitCM (object_CM.cpp:96-357): the capacitance-matrix circuit of a
//! PoissonSolver's objects with voltage and current sources.
class CircuitCM {
public:
    CircuitCM(PoissonSolver& poisson, const VSourceVector& vsources, const ISourceVector& isources, double dt)
        : inv_capacitance(poisson.num_objects() * poisson.num_objects()) {
        std::vector<int> va, vb, ia, ib;
        std::vector<double> vv, iv;
        for (const auto& s : vsources) {
            va.push_back(s.node_a);
            vb.push_back(s.node_b);
            vv.push_back(s.value);
        }
        for (const auto& s : isources) {
            ia.push_back(s.node_a);
            ib.push_back(s.node_b);
            iv.push_back(s.value);
        }
        check(punc_gpu_circuit_set(poisson.handle(), dt, (int)va.size(), va.data(), vb.data(), vv.data(),
                                   (int)ia.size(), ia.data(), ib.data(), iv.data(), inv_capacitance.data()));
    }
    std::vector<double> inv_capacitance;  ///< row-major, n_objects^2
};

}  // namespace punc

#endif

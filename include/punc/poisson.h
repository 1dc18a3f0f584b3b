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

    punc_field_t* handle() const { return handle_.get(); }
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

}  // namespace punc

#endif

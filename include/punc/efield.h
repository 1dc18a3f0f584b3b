// Drop-in for ESolver of punc/include/punc/efield.h (PUNC++): the vector-CG1
// L2 projection E = -grad(phi) (efield.cpp:41-75), on the device.
#ifndef PUNC_GPU_EFIELD_H
#define PUNC_GPU_EFIELD_H

#include <punc/poisson.h>

namespace punc {

class ESolver {
public:
    //! Shares the device matrices of a PoissonSolver on the same mesh.
    explicit ESolver(PoissonSolver& poisson) : field_(poisson.handle()) {}

    //! solve (efield.cpp:59-64): E (3 components per dof) from phi.
    void solve(Function& E, const Function& phi) {
        check(punc_gpu_efield_solve(field_, phi.values.data(), E.values.data(), &iterations));
    }

    int iterations = 0;

private:
    punc_field_t* field_;
};

}  // namespace punc

#endif

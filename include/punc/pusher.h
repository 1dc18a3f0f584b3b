// Drop-in for punc/include/punc/pusher.h (PUNC++): accel_cg1, boris_cg1,
// move, move_periodic on the GPU.  E is a 3-component CG1 Function.
#ifndef PUNC_GPU_PUSHER_H
#define PUNC_GPU_PUSHER_H

#include <punc/population.h>

#include <cassert>
#include <vector>

namespace punc {

//! Uploads the CG1 field dofs (the host -> device field transfer).
template <typename PopulationType>
void set_field(PopulationType& pop, const Function& E) {
    assert(E.value_dim == 3);
    check(punc_gpu_set_field(pop.handle(), E.values.data()));
}

//! accel_cg1 (pusher.h:154-201); returns the mid-step kinetic energy.
template <typename PopulationType>
double accel_cg1(PopulationType& pop, const Function& E, double dt) {
    set_field(pop, E);
    double KE = 0.0;
    check(punc_gpu_accel(pop.handle(), dt, &KE));
    return KE;
}

//! boris_cg1 (pusher.h:282-372); B uniform, 3 components.
template <typename PopulationType>
double boris_cg1(PopulationType& pop, const Function& E, const std::vector<double>& B, double dt) {
    assert(B.size() == 3 && "The algorithm is only valid for 3D.");
    set_field(pop, E);
    double KE = 0.0;
    check(punc_gpu_boris(pop.handle(), B.data(), dt, &KE));
    return KE;
}

//! move (pusher.h:648-662).
template <typename PopulationType>
void move(PopulationType& pop, double dt) {
    check(punc_gpu_move(pop.handle(), dt));
}

//! move_periodic (pusher.h:666-681); the box corner is at the origin.
template <typename PopulationType>
void move_periodic(PopulationType& pop, const double dt, const std::vector<double>& Ld) {
    check(punc_gpu_move_periodic(pop.handle(), dt, Ld.data()));
}

}  // namespace punc

#endif

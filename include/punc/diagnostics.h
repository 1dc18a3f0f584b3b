// Drop-in for kinetic_energy of punc/include/punc/diagnostics.h (PUNC++).
#ifndef PUNC_GPU_DIAGNOSTICS_H
#define PUNC_GPU_DIAGNOSTICS_H

#include <punc/population.h>

namespace punc {

//! kinetic_energy (diagnostics.h:248-265).
template <typename PopulationType>
double kinetic_energy(PopulationType& pop) {
    double KE = 0.0;
    check(punc_gpu_kinetic_energy(pop.handle(), &KE));
    return KE;
}

}  // namespace punc

#endif

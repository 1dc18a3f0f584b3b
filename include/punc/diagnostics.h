// Drop-in for the particle diagnostics of punc/include/punc/diagnostics.h
// (PUNC++): kinetic and potential energy, CG1/DG0 number densities and the
// exponential moving averages of the deposited densities.
#ifndef PUNC_GPU_DIAGNOSTICS_H
#define PUNC_GPU_DIAGNOSTICS_H

#include <punc/population.h>

#include <stdexcept>
#include <vector>

namespace punc {

//! kinetic_energy (diagnostics.h:248-265).
template <typename PopulationType>
double kinetic_energy(PopulationType& pop) {
    double KE = 0.0;
    check(punc_gpu_kinetic_energy(pop.handle(), &KE));
    return KE;
}

//! particle_potential_energy_cg1 (diagnostics.h:379-412); phi is a scalar CG1 Function.
template <typename PopulationType>
double particle_potential_energy_cg1(PopulationType& pop, const Function& phi) {
    double PE = 0.0;
    check(punc_gpu_potential_energy(pop.handle(), phi.values.data(), &PE));
    return PE;
}

//! density_cg1 (diagnostics.h:492-555).  The function space argument is
//! accepted for the reference's signature; the dofs are the mesh's CG1 dofs.
template <typename Space, typename PopulationType>
void density_cg1(const Space&, PopulationType& pop, const std::vector<Species>& species, Function& ne,
                 Function& ni, const std::vector<double>& dv_inv) {
    check(punc_gpu_density_cg1(pop.handle(), dv_inv.data(), species[0].weight, species[1].weight,
                               ne.values.data(), ni.values.data()));
}

//! density_dg0 (diagnostics.h:432-465); ne and ni are DG0 Functions (one
//! value per cell, Function(mesh, 1, true)).
template <typename Space, typename PopulationType>
void density_dg0(const Space&, PopulationType& pop, const std::vector<Species>& species, Function& ne,
                 Function& ni) {
    if (!ne.dg0 || !ni.dg0) throw std::runtime_error("density_dg0: ne and ni must be DG0 Functions");
    check(punc_gpu_density_dg0(pop.handle(), species[0].weight, species[1].weight, ne.values.data(),
                               ni.values.data()));
}

//! ema(rho, rho_ema, dt, tau) (diagnostics.h:573-588) for the population's
//! deposited charge density: the average lives on the device and is copied
//! into g.  reset zeroes it first (a freshly constructed rho_ema).
template <typename PopulationType>
void rho_ema(PopulationType& pop, Function& g, double dt, double tau, bool reset = false) {
    check(punc_gpu_rho_ema(pop.handle(), dt, tau, reset ? 1 : 0, g.values.data()));
}

//! ema(ne, ne_ema, dt, tau) and ema(ni, ni_ema, dt, tau) over the last
//! density_cg1 / density_dg0 result, averaged on the device.
template <typename PopulationType>
void density_ema(PopulationType& pop, Function& ne_g, Function& ni_g, double dt, double tau, bool reset = false) {
    check(punc_gpu_density_ema(pop.handle(), dt, tau, reset ? 1 : 0, ne_g.values.data(), ni_g.values.data()));
}

}  // namespace punc

#endif

// Drop-in for punc/include/punc/injector.h (PUNC++): flux tables, boundary
// injection and initial loading, sampled on the GPU from a counter-based
// Philox stream (seeded through Population::set_seed).
#ifndef PUNC_GPU_INJECTOR_H
#define PUNC_GPU_INJECTOR_H

#include <punc/population.h>

#include <vector>

namespace punc {

//! create_flux (injector.cpp:104-171) for Maxwellian species.
inline void create_flux(const Mesh& mesh, std::vector<Species>& species) {
    const size_t nf = mesh.exterior_facets.size();
    for (auto& s : species) {
        s.vdf->num_particles.assign(nf, 0.0);
        s.vdf->pdf_max.assign(nf, 0.0);
        std::vector<double> vd = s.vdf->vd();
        check(punc_gpu_maxwellian_flux(mesh.handle(), s.vdf->vth(), vd.data(), s.vdf->num_particles.data(),
                                       s.vdf->pdf_max.data()));
    }
}

//! inject_particles (injector.h:157-232).
template <typename PopulationType>
void inject_particles(PopulationType& pop, std::vector<Species>& species,
                      std::vector<ExteriorFacet>& facets, double dt) {
    (void)facets;  // the population's mesh holds the exterior facets
    for (auto& s : species) {
        int idx = pop.species_index(s.q, s.m);
        std::vector<double> vd = s.vdf->vd();
        check(punc_gpu_injector_set(pop.handle(), idx, s.vdf->vth(), vd.data(), s.n, s.vdf->vdf_range));
    }
    int64_t k = 0;
    check(punc_gpu_inject(pop.handle(), dt, pop.seed(), pop.next_injection_step(), &k));
}

//! load_particles (injector.h:244-278).
template <typename PopulationType>
void load_particles(PopulationType& pop, std::vector<Species>& species) {
    for (auto& s : species) {
        int idx = pop.species_index(s.q, s.m);
        std::vector<double> vd = s.vdf->vd();
        check(punc_gpu_pop_load(pop.handle(), idx, (int64_t)s.num, pop.seed(), s.pdf->amp, s.pdf->kw,
                                s.pdf->axis, s.vdf->vth(), vd.data()));
    }
}

}  // namespace punc

#endif

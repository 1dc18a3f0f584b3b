// Drop-in for punc/include/punc/injector.h (PUNC++): flux tables, boundary
// injection and initial loading, sampled on the GPU from a counter-based
// Philox stream (seeded through Population::set_seed).
#ifndef PUNC_GPU_INJECTOR_H
#define PUNC_GPU_INJECTOR_H

#include <punc/population.h>

#include <stdexcept>
#include <vector>

namespace punc {

//! The device description of a species' velocity distribution (any Pdf
//! other than the library's velocity pdfs is rejected).
inline const VelocityPdf& velocity_pdf(const Species& s) {
    auto* v = dynamic_cast<const VelocityPdf*>(s.vdf.get());
    if (!v) throw std::runtime_error("the GPU samplers take Maxwellian, Kappa, Cairns or KappaCairns velocity pdfs");
    return *v;
}

//! create_flux (injector.cpp:104-171): closed forms where the distribution
//! has them, the device Monte Carlo (500000 samples per facet, Philox stream
//! keyed by `seed`; species i uses substream i) otherwise.
inline void create_flux(const Mesh& mesh, std::vector<Species>& species, uint64_t seed = 0,
                        int64_t n_iter = 500000) {
    const size_t nf = mesh.exterior_facets.size();
    for (size_t i = 0; i < species.size(); ++i) {
        auto& s = species[i];
        const VelocityPdf& v = velocity_pdf(s);
        s.vdf->num_particles.assign(nf, 0.0);
        s.vdf->pdf_max.assign(nf, 0.0);
        std::vector<double> vd = s.vdf->vd();
        check(punc_gpu_vdf_flux(mesh.handle(), v.kind, s.vdf->vth(), vd.data(), v.kappa, v.alpha, v.vdf_range, seed,
                                (int)i, n_iter, s.vdf->num_particles.data(), s.vdf->pdf_max.data()));
    }
}

//! create_flux(species, facets) in the reference's form (injector.h:144,
//! called at interaction.cpp:119): the tables over the given exterior facets,
//! Monte Carlo where needed from the Philox stream with seed 0 (species i
//! uses substream i; the reference's stream is unseeded).
inline void create_flux(std::vector<Species>& species, std::vector<ExteriorFacet>& facets) {
    const size_t nf = facets.size();
    std::vector<double> normal(3 * nf), area(nf);
    for (size_t j = 0; j < nf; ++j) {
        area[j] = facets[j].area;
        for (int k = 0; k < 3; ++k) normal[3 * j + k] = facets[j].normal[k];
    }
    for (size_t i = 0; i < species.size(); ++i) {
        auto& s = species[i];
        const VelocityPdf& v = velocity_pdf(s);
        s.vdf->num_particles.assign(nf, 0.0);
        s.vdf->pdf_max.assign(nf, 0.0);
        std::vector<double> vd = s.vdf->vd();
        check(punc_gpu_vdf_flux_facets((int64_t)nf, normal.data(), area.data(), v.kind, s.vdf->vth(), vd.data(),
                                       v.kappa, v.alpha, v.vdf_range, 0, (int)i, 500000,
                                       s.vdf->num_particles.data(), s.vdf->pdf_max.data()));
    }
}

//! inject_particles (injector.h:157-232) with the create_flux tables.
template <typename PopulationType>
void inject_particles(PopulationType& pop, std::vector<Species>& species,
                      std::vector<ExteriorFacet>& facets, double dt) {
    for (auto& s : species) {
        int idx = pop.species_index(s.q, s.m);
        const VelocityPdf& v = velocity_pdf(s);
        std::vector<double> vd = s.vdf->vd();
        const bool tables = s.vdf->num_particles.size() == facets.size() && s.vdf->pdf_max.size() == facets.size();
        check(punc_gpu_injector_set_vdf(pop.handle(), idx, v.kind, s.vdf->vth(), vd.data(), v.kappa, v.alpha, s.n,
                                        v.vdf_range, tables ? s.vdf->num_particles.data() : nullptr,
                                        tables ? s.vdf->pdf_max.data() : nullptr));
    }
    int64_t k = 0;
    check(punc_gpu_inject(pop.handle(), dt, pop.seed(), pop.next_injection_step(), &k));
}

//! load_particles (injector.h:244-278).
template <typename PopulationType>
void load_particles(PopulationType& pop, std::vector<Species>& species) {
    for (auto& s : species) {
        int idx = pop.species_index(s.q, s.m);
        const VelocityPdf& v = velocity_pdf(s);
        std::vector<double> vd = s.vdf->vd();
        check(punc_gpu_pop_load_vdf(pop.handle(), idx, (int64_t)s.num, pop.seed(), s.pdf->amp, s.pdf->kw, s.pdf->axis,
                                    v.kind, s.vdf->vth(), vd.data(), v.kappa, v.alpha, v.vdf_range));
    }
}

}  // namespace punc

#endif

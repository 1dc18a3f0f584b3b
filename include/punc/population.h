// Drop-in for punc/include/punc/population.h (PUNC++) on the MI355X particle
// path.  Population<3> keeps its particles on the GPU (FP64 structure of
// arrays, species-tagged cell field); the member functions keep the
// reference's names and meaning (population.h:236-308).
#ifndef PUNC_GPU_POPULATION_H
#define PUNC_GPU_POPULATION_H

#include <punc/distributions.h>
#include <punc/mesh.h>
#include <punc/object.h>
#include <punc_gpu.h>

#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace punc {

//! A simulation particle (reference population.h:74-94).
template <std::size_t len = 3>
struct Particle {
    double x[len];
    double v[len];
    double q;
    double m;
    Particle() {}
    Particle(const double* x_, const double* v_, double q, double m) : q(q), m(m) {
        for (std::size_t i = 0; i < len; i++) {
            x[i] = x_[i];
            v[i] = v_[i];
        }
    }
};

enum class ParticleAmountType { per_cell, in_total, per_volume, phys_per_sim };

//! Species (reference population.h:99-125, population.cpp:41-70).
class Species {
public:
    double q;       ///< Charge of a simulation particle (physical charge x weight)
    double m;       ///< Mass of a simulation particle
    double n;       ///< Density of simulation particles
    size_t num;     ///< Initial number of simulation particles
    std::shared_ptr<Pdf> pdf;
    std::shared_ptr<Pdf> vdf;
    double debye;
    double weight;

    Species(double charge, double mass, double density, double amount, ParticleAmountType type,
            const Mesh& mesh, std::shared_ptr<Pdf> pdf, std::shared_ptr<Pdf> vdf, double eps0)
        : pdf(pdf), vdf(vdf) {
        switch (type) {  // deliberate fall-through, as in the reference
            case ParticleAmountType::per_cell:
                amount *= mesh.num_cells();
                [[fallthrough]];
            case ParticleAmountType::in_total:
                amount /= mesh.volume();
                [[fallthrough]];
            case ParticleAmountType::per_volume:
                amount = density / amount;
                [[fallthrough]];
            case ParticleAmountType::phys_per_sim:
                break;
        }
        weight = amount;
        q = charge * amount;
        m = mass * amount;
        n = density / amount;
        num = n * mesh.volume();
        debye = vdf->debye(m, q, n, eps0);
    }
};

template <std::size_t len>
class Population;

//! Device-resident population on a 3D tetrahedral mesh.
template <>
class Population<3> {
public:
    const std::size_t g_dim = 3;
    const Mesh& mesh;

    //! capacity: maximum number of particles held at once.
    explicit Population(const Mesh& mesh, std::size_t capacity = std::size_t(1) << 24,
                        void* hip_stream = nullptr)
        : mesh(mesh) {
        punc_pop_t* h = nullptr;
        check(punc_gpu_pop_create(mesh.handle(), 0, nullptr, nullptr, (int64_t)capacity, hip_stream, &h));
        handle_.reset(h, punc_gpu_pop_destroy);
    }

    //! Species index of (q, m), registering it on first use.
    int species_index(double q, double m) {
        for (size_t s = 0; s < q_.size(); ++s)
            if (q_[s] == q && m_[s] == m) return (int)s;
        int idx = -1;
        check(punc_gpu_pop_add_species(handle(), q, m, &idx));
        q_.push_back(q);
        m_.push_back(m);
        return idx;
    }

    //! add_particles (population.h:411-435); points outside the mesh are skipped.
    void add_particles(const std::vector<double>& xs, const std::vector<double>& vs, double q, double m) {
        int s = species_index(q, m);
        int64_t added = 0;
        check(punc_gpu_pop_add(handle(), s, (int64_t)xs.size() / 3, xs.data(), vs.data(), &added));
    }

    void add_particles(const std::vector<Particle<3>>& ps) {
        for (const auto& p : ps)
            add_particles(std::vector<double>(p.x, p.x + 3), std::vector<double>(p.v, p.v + 3), p.q, p.m);
    }

    //! locate (population.h:448-452): cell id or -1.
    signed long int locate(const double* p) {
        int64_t c = -1;
        check(punc_gpu_mesh_locate(mesh.handle(), 1, p, &c));
        return c;
    }

    //! update (population.h:641-699): relocate, remove particles that left,
    //! and credit absorbed charge to the objects whose boundary they crossed.
    void update(ObjectVector objects, double dt) {
        std::vector<int64_t> ab(std::max<size_t>(q_.size(), 1) * PUNC_GPU_MAX_BND, 0);
        check(punc_gpu_update(handle(), ab.data()));
        credit_objects(objects, ab, dt);
    }

    std::size_t num_of_particles() {
        std::size_t n = 0;
        for (auto c : counts()) n += c;
        return n;
    }
    std::size_t num_of_positives() {
        std::size_t n = 0;
        auto c = counts();
        for (size_t s = 0; s < c.size(); ++s)
            if (q_[s] > 0) n += c[s];
        return n;
    }
    std::size_t num_of_negatives() {
        std::size_t n = 0;
        auto c = counts();
        for (size_t s = 0; s < c.size(); ++s)
            if (q_[s] < 0) n += c[s];
        return n;
    }

    //! All particles, downloaded (order: device storage order).
    std::vector<Particle<3>> particles() {
        int64_t n = 0;
        check(punc_gpu_pop_size(handle(), &n));
        std::vector<double> xs(n * 3), vs(n * 3);
        std::vector<int32_t> sp(n);
        check(punc_gpu_pop_get(handle(), xs.data(), vs.data(), sp.data(), nullptr));
        std::vector<Particle<3>> out(n);
        for (int64_t i = 0; i < n; ++i) out[i] = Particle<3>(&xs[3 * i], &vs[3 * i], q_[sp[i]], m_[sp[i]]);
        return out;
    }

    //! Particles relocated within 1e-12 of a facet plane of their cell since
    //! the last call (reporting only; the walk is relocate_fast's).
    std::size_t num_near_face() {
        int64_t k = 0;
        check(punc_gpu_near_face(handle(), &k));
        return (std::size_t)k;
    }

    //! Re-sort particles by cell (keeps per-cell data reads coalesced).
    void sort() { check(punc_gpu_sort(handle())); }

    //! save_file (population.h:807-839): 64-byte records or "%.17g" text lines,
    //! in cell order (the particles are re-sorted by cell first).
    void save_file(const std::string& fname, bool binary) {
        check(punc_gpu_pop_save_file(handle(), fname.c_str(), binary ? 1 : 0));
    }

    //! load_file (population.h:841-880): appends the file's particles; new
    //! (q, m) pairs become species.
    void load_file(const std::string& fname, bool binary) {
        int64_t added = 0;
        check(punc_gpu_pop_load_file(handle(), fname.c_str(), binary ? 1 : 0, &added));
        int n = 0;
        check(punc_gpu_pop_species(handle(), &n, nullptr, nullptr));
        q_.assign((size_t)n, 0.0);
        m_.assign((size_t)n, 0.0);
        if (n > 0) check(punc_gpu_pop_species(handle(), &n, q_.data(), m_.data()));
    }

    //! statistics (population.h:746-805): mean and standard deviation of the
    //! speed of the q < 0 (stats[0..1]) and q > 0 (stats[2..3]) particles.
    void statistics(double* stats) { check(punc_gpu_statistics(handle(), stats)); }

    //! Seed of the injection stream (the reference's RNG is unseeded).
    void set_seed(uint64_t seed) { seed_ = seed; }
    uint64_t seed() const { return seed_; }
    uint64_t next_injection_step() { return inj_step_++; }

    punc_pop_t* handle() const { return handle_.get(); }
    const std::vector<double>& species_q() const { return q_; }

    //! Per object: current = sum of absorbed q; charge += current; current /= dt.
    void credit_objects(ObjectVector& objects, const std::vector<int64_t>& ab, double dt) {
        for (auto& o : objects) o->current = 0;
        for (auto& o : objects) {
            if (o->bnd_id >= PUNC_GPU_MAX_BND) continue;
            for (size_t s = 0; s < q_.size(); ++s) o->current += (double)ab[s * PUNC_GPU_MAX_BND + o->bnd_id] * q_[s];
        }
        for (auto& o : objects) {
            o->charge += o->current;
            o->current /= dt;
        }
    }

private:
    std::vector<int64_t> counts() {
        std::vector<int64_t> c(q_.size(), 0);
        if (!c.empty()) check(punc_gpu_pop_counts(handle(), c.data()));
        return c;
    }

    std::shared_ptr<punc_pop_t> handle_;
    std::vector<double> q_, m_;
    uint64_t seed_ = 0;
    uint64_t inj_step_ = 0;
};

}  // namespace punc

#endif

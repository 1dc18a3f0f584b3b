// The reference time loop (interaction/interaction.cpp:342-423) written
// against the drop-in C++ API (include/punc.h): distribute, field solve on
// the device (Poisson with a grounded exterior and a floating object through
// the capacitance-matrix circuit, then the E projection), count, push (half
// step at n == 0), move, update, inject.  Prints one line per step and checks particle-number
// bookkeeping: N(n+1) = N(n) - absorbed + injected, and the diagnostics'
// conservation properties.
#include <punc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

using namespace punc;
namespace df = punc;  // the reference's df::Function / df::FunctionSpace lines compile as written

int main(int argc, char** argv) {
    int steps = argc > 1 ? atoi(argv[1]) : 5;
    Mesh mesh = Mesh::box(8, {-4.0, -4.0, -4.0}, {4.0, 4.0, 4.0}, 1.0);
    std::vector<double> vd{0.0, 0.0, 0.5};
    auto pdf = std::make_shared<UniformPosition>(mesh);
    auto make_species = [&]() {
        std::vector<Species> sp;
        sp.emplace_back(-1.0, 1.0, 20.0, 1.0, ParticleAmountType::phys_per_sim, mesh, pdf,
                        std::make_shared<Maxwellian>(1.0, vd), 1.0);
        sp.emplace_back(1.0, 1836.0, 20.0, 1.0, ParticleAmountType::phys_per_sim, mesh, pdf,
                        std::make_shared<Maxwellian>(1.0 / std::sqrt(1836.0), vd), 1.0);
        return sp;
    };
    std::vector<Species> species = make_species();
    // interaction.cpp:90-94, 97-108 and 119, as written there
    auto V = CG1_space(mesh);
    auto W = CG1_vector_space(mesh);
    auto Q = DG0_space(mesh);
    auto dv_inv = element_volume(V);
    df::Function rho(std::make_shared<const df::FunctionSpace>(V));
    df::Function E(std::make_shared<const df::FunctionSpace>(W));
    df::Function phi(std::make_shared<const df::FunctionSpace>(V));
    df::Function rho_ema(std::make_shared<const df::FunctionSpace>(V));
    df::Function ne(std::make_shared<const df::FunctionSpace>(V));
    df::Function ni(std::make_shared<const df::FunctionSpace>(V));
    df::Function ne_ema(std::make_shared<const df::FunctionSpace>(V));
    df::Function ni_ema(std::make_shared<const df::FunctionSpace>(V));
    create_flux(species, mesh.exterior_facets);
    // the facet-list form gives the mesh form's tables (same stream, seed 0)
    std::vector<Species> species_m = make_species();  // own Pdf objects
    create_flux(mesh, species_m);
    int bad = 0;
    for (size_t i = 0; i < species.size(); ++i)
        if (species[i].vdf->num_particles != species_m[i].vdf->num_particles ||
            species[i].vdf->pdf_max != species_m[i].vdf->pdf_max)
            bad++;
    Population<3> pop(mesh, 1 << 20);
    pop.set_seed(42);
    load_particles(pop, species);
    Function phi_one(mesh, 1), rho_avg(mesh, 1);
    Function ne0(Q), ni0(Q);
    if (dv_inv != element_volume(mesh) || ne0.values.size() != (size_t)mesh.num_cells()) bad++;
    for (double& v : phi_one.values) v = 1.0;
    ObjectVector objects{std::make_shared<Object>(2)};
    PoissonSolver poisson(mesh, objects, true, 1.0);
    ESolver esolver(poisson);
    CircuitCM circuit(poisson, VSourceVector{}, ISourceVector{}, 0.05);  // a floating object
    std::vector<double> B{0.0, 0.0, 1.0};
    const double dt = 0.05;
    size_t n_prev = pop.num_of_particles();
    const double tau_rho = 10 * dt, tau_n = 5 * dt;
    const bool filter_rho = true, filter_n = true;
    if (!(circuit.inv_capacitance[0] > 0)) bad++;
    for (int n = 0; n <= steps; ++n) {
        distribute_cg1(pop, rho, dv_inv);
        poisson.solve_circuit(phi, rho, objects);
        esolver.solve(E, phi);
        if (n == 0) {
            // the device-resident path gives the same field
            Function phi2(mesh, 1), E2(mesh, 3);
            poisson.solve_population(pop, objects, &phi2, &E2);
            for (int64_t d = 0; d < 3 * mesh.ndof; ++d)
                if (std::fabs(E2.values[d] - E.values[d]) > 1e-9 * (1.0 + std::fabs(E.values[d]))) bad++;
            // and so does the Jacobi-preconditioned solve
            Function phi3(mesh, 1);
            poisson.set_preconditioner("jacobi");
            poisson.solve(phi3, rho, objects);
            poisson.set_preconditioner("hypre_amg");
            double pmax = 0.0;
            for (int64_t d = 0; d < mesh.ndof; ++d) pmax = std::max(pmax, std::fabs(phi2.values[d]));
            for (int64_t d = 0; d < mesh.ndof; ++d)
                if (std::fabs(phi3.values[d] - phi2.values[d]) > 1e-9 * pmax) bad++;
        }
        double qsum = 0.0;
        for (int64_t d = 0; d < mesh.ndof; ++d) qsum += rho.values[d] / dv_inv[d];
        size_t num_e = pop.num_of_negatives(), num_i = pop.num_of_positives();
        // diagnostics: phi = 1 gives PE = sum q / 2; CG1 densities keep
        // sum_j n_j / dv_inv_j = weight * count; rho_ema after a reset is w rho
        double PE = particle_potential_energy_cg1(pop, phi_one);
        density_cg1(V, pop, species, ne, ni, dv_inv);
        density_dg0(Q, pop, species, ne0, ni0);
        double ne_sum = 0.0, ni_sum = 0.0, ne0_max = 0.0;
        for (int64_t d = 0; d < mesh.ndof; ++d) {
            ne_sum += ne.values[d] / dv_inv[d];
            ni_sum += ni.values[d] / dv_inv[d];
        }
        for (double v : ne0.values) ne0_max = std::max(ne0_max, v);
        double stats[4] = {0, 0, 0, 0};
        pop.statistics(stats);
        // interaction.cpp:432 and 458-459 as written there; the device-resident
        // average of the same deposits must agree
        if (filter_rho) ema(rho, rho_ema, dt, tau_rho);
        if (filter_n) {
            ema(ne, ne_ema, dt, tau_n);
            ema(ni, ni_ema, dt, tau_n);
        }
        punc::rho_ema(pop, rho_avg, dt, tau_rho, n == 0);
        const double w = 1.0 - std::exp(-dt / tau_rho);
        for (int64_t d = 0; d < mesh.ndof; ++d) {
            if (std::fabs(rho_avg.values[d] - rho_ema.values[d]) > 1e-12 * (1 + std::fabs(rho.values[d]))) bad++;
            if (n == 0 && std::fabs(rho_ema.values[d] - w * rho.values[d]) > 1e-12 * (1 + std::fabs(rho.values[d])))
                bad++;
        }
        const double wn = 1.0 - std::exp(-dt / tau_n);
        if (n == 0)
            for (int64_t d = 0; d < mesh.ndof; ++d)
                if (ne_ema.values[d] != wn * ne.values[d] || ni_ema.values[d] != wn * ni.values[d]) bad++;
        if (std::fabs(PE - 0.5 * ((double)num_i - (double)num_e)) > 1e-8 * (num_e + num_i)) bad++;
        if (std::fabs(ne_sum - species[0].weight * num_e) > 1e-8 * num_e) bad++;
        if (std::fabs(ni_sum - species[1].weight * num_i) > 1e-8 * num_i) bad++;
        if (!(ne0_max > 0) || !(stats[0] > 0) || !(stats[1] > 0) || !(stats[2] > 0) || !(stats[3] > 0)) bad++;
        double KE = boris_cg1(pop, E, B, (1.0 - 0.5 * (n == 0)) * dt);
        if (n == 0) KE = kinetic_energy(pop);
        move(pop, dt);
        size_t before = pop.num_of_particles();
        double q_before = objects[0]->charge;
        pop.update(objects, dt);
        (void)pop.num_near_face();
        size_t after_update = pop.num_of_particles();
        inject_particles(pop, species, mesh.exterior_facets, dt);
        size_t after = pop.num_of_particles();
        double qexp = (double)num_i - (double)num_e;
        printf("step %d ne %zu ni %zu KE %.6e Q_rho %.6e (expect %.0f) removed %zu injected %zu obj_charge %.1f\n",
               n, num_e, num_i, KE, qsum, qexp, before - after_update, after - after_update,
               objects[0]->charge);
        if (std::fabs(qsum - qexp) > 1e-8 * (num_e + num_i)) bad++;
        if (num_e + num_i != n_prev) bad++;
        if (objects[0]->charge - q_before > (double)(before - after_update) + 1e-9) bad++;
        n_prev = after;
    }
    printf(bad ? "FAIL\n" : "OK\n");
    return bad ? 1 : 0;
}

// Drop-in for the particle diagnostics of punc/include/punc/diagnostics.h
// (PUNC++): kinetic and potential energy, CG1/DG0 number densities and the
// exponential moving averages of the deposited densities.
#ifndef PUNC_GPU_DIAGNOSTICS_H
#define PUNC_GPU_DIAGNOSTICS_H

#include <punc/population.h>

#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <limits>
#include <stdexcept>
#include <string>
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

//! ema(f, g, dt, tau) (diagnostics.h:573-588) in the reference's form:
//! g = w f + (1 - w) g, w = 1 - exp(-dt/tau), over the values of two
//! Functions of one space (computed on the GPU).
inline void ema(const Function& f, Function& g, double dt, double tau) {
    if (f.values.size() != g.values.size()) throw std::runtime_error("ema: f and g differ in length");
    check(punc_gpu_ema((int64_t)g.values.size(), f.values.data(), g.values.data(), dt, tau));
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

//! Run state for restarts (diagnostics.h:39-71, diagnostics.cpp:35-68).
class State {
public:
    std::string fname;
    explicit State(std::string fname) : fname(std::move(fname)) {}

    //! First line: n, t, then per object charge and current (objects after
    //! the first read 0, their values being on later lines).
    void load(std::size_t& n, double& t, ObjectVector objects) {
        std::ifstream ifile(fname);
        std::string line;
        std::getline(ifile, line);
        char* s = (char*)line.c_str();
        n = strtol(s, &s, 10);
        t = strtod(s, &s);
        for (auto& o : objects) {
            o->charge = strtod(s, &s);
            o->current = strtod(s, &s);
        }
    }

    //! "n+1\tt\t", then "charge\tcurrent\n" per object (default stream format).
    void save(std::size_t n, double t, const ObjectVector& objects) {
        std::ofstream ofile(fname, std::ofstream::out);
        ofile << n + 1 << "\t" << t << "\t";
        for (auto& o : objects) {
            ofile << o->charge << "\t";
            ofile << o->current << "\n";
        }
    }
};

//! Time-series file (diagnostics.h:73-170, diagnostics.cpp:70-126).
class History {
public:
    std::ofstream ofile;
    std::size_t dim;
    bool stats;
    bool hex_output;  ///< accepted; decimal output either way, as in the reference

    History(const std::string& fname, const ObjectVector& objects, std::size_t dim = 3, bool stats = false,
            bool continue_simulation = false, bool hex_output = false)
        : dim(dim), stats(stats), hex_output(hex_output) {
        if (continue_simulation) {
            ofile.open(fname, std::ofstream::out | std::ofstream::app);
            return;
        }
        ofile.open(fname, std::ofstream::out);
        ofile << "#:xaxis\tt\n";
        ofile << "#:name\tn\tt\tne\tni\tKE\tPE";
        for (std::size_t i = 0; i < objects.size(); ++i) ofile << "\tV[" << i << "]\tI[" << i << "]\tQ[" << i << "]";
        if (stats) ofile << "\tmean_e\tstdev_e\tmean_i\tstdev_i";
        ofile << "\n";
        ofile << "#:long\ttimestep\ttime\t\"number of electrons\"\t";
        ofile << "\"number of ions\"\t\"kinetic energy\"\t";
        ofile << "\"potential energy\"";
        for (std::size_t i = 0; i < objects.size(); ++i) ofile << "\tvoltage\tcurrent\tcharge";
        if (stats)
            ofile << "\tmean velocity of electrons\tstandard deviation of electron velocities"
                  << "\tmean velocity of ions\tstandard deviation of ion velocities";
        ofile << "\n";
        ofile << "#:units\t1\ts\tm**(-3)\tm**(-3)\tJ\tJ";
        for (std::size_t i = 0; i < objects.size(); ++i) {
            ofile << "\tV";
            if (dim == 1) ofile << "\tA/m**2";
            else if (dim == 2) ofile << "\tA/m";
            else if (dim == 3) ofile << "\tA";
            ofile << "\tC";
        }
        if (stats) ofile << "\tm/s\tm/s\tm/s\tm/s";
        ofile << "\n";
    }

    //! One line: n, then t, ne, ni, KE, PE, per object V, I, Q, and with
    //! stats the population's speed statistics, all "%.16e".
    template <typename PopulationType>
    void save(std::size_t n, double t, double num_e, double num_i, double KE, double PE, const ObjectVector& objects,
              PopulationType& pop) {
        ofile << n << "\t";
        ofile << std::setprecision(std::numeric_limits<double>::digits10 + 1);
        ofile << std::scientific;
        ofile << t << "\t" << num_e << "\t" << num_i << "\t" << KE << "\t" << PE;
        for (auto& o : objects) ofile << "\t" << o->get_potential() << "\t" << o->current << "\t" << o->charge;
        if (stats) {
            double statistics[4] = {0, 0, 0, 0};
            pop.statistics(statistics);
            for (int i = 0; i < 4; ++i) ofile << "\t" << statistics[i];
        }
        ofile << std::endl;
    }
};

}  // namespace punc

#endif

// Drop-in for the Pdf / UniformPosition / Maxwellian part of
// punc/include/punc/distributions.h (PUNC++).  On the GPU path a position pdf
// is described by its parameters (inside(x) * (1 + amp*cos(kw*x[axis]))) and a
// velocity pdf by (vth, vd): the samplers run on the device.
#ifndef PUNC_GPU_DISTRIBUTIONS_H
#define PUNC_GPU_DISTRIBUTIONS_H

#include <cmath>
#include <limits>
#include <vector>

namespace punc {

class Pdf {
protected:
    double _vth = 0.0;
    std::vector<double> _vd;

public:
    int dim = 3;
    bool has_icdf = false, has_flux_number = false, has_flux_max = false;
    double vdf_range = 5.0;
    std::vector<double> domain;
    std::vector<double> pdf_max, num_particles;  ///< per exterior facet (create_flux)
    // Position pdf parameters: inside(x) * (1 + amp*cos(kw*x[axis])).
    double amp = 0.0, kw = 0.0;
    int axis = 0;

    virtual ~Pdf() {}
    virtual double max() { return 1.0 + amp; }
    virtual double vth() { return _vth; }
    virtual std::vector<double> vd() { return _vd; }
    virtual double debye(double m, double q, double n, double eps0) { return 0.0; }
};

//! UniformPosition (distributions.h:186-221).
class UniformPosition : public Pdf {
public:
    UniformPosition() {}
    template <typename MeshT>
    explicit UniformPosition(const MeshT&) {}
};

//! Uniform position with a cosine density perturbation, the pattern of the
//! Langmuir-wave demo (demo/langmuir/langmuir.cpp:17-54).
class CosinePosition : public Pdf {
public:
    CosinePosition(double amplitude, double wavenumber, int ax = 0) {
        amp = amplitude;
        kw = wavenumber;
        axis = ax;
    }
};

//! Maxwellian (distributions.h:233-249, distributions.cpp:23-77).
class Maxwellian : public Pdf {
public:
    Maxwellian(double vth, std::vector<double> vd, double range = 5.0) {
        _vth = vth;
        _vd = vd;
        dim = (int)vd.size();
        has_icdf = has_flux_number = has_flux_max = true;
        vdf_range = range;
        domain.resize(2 * dim);
        for (int i = 0; i < dim; ++i) {
            domain[i] = _vd[i] - vdf_range * _vth;
            domain[i + dim] = _vd[i] + vdf_range * _vth;
        }
    }
    double debye(double m, double q, double n, double eps0) override {
        return std::sqrt(eps0 * m / (n * q * q)) * _vth;
    }
};

}  // namespace punc

#endif

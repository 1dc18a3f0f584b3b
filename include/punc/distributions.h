// Drop-in for punc/include/punc/distributions.h (PUNC++): Pdf,
// UniformPosition and the velocity distributions Maxwellian, Kappa, Cairns and
// KappaCairns.  On the GPU path a position pdf is described by its parameters
// (inside(x) * (1 + amp*cos(kw*x[axis]))) and a velocity pdf by its kind and
// (vth, vd, kappa, alpha, range): the samplers run on the device.
#ifndef PUNC_GPU_DISTRIBUTIONS_H
#define PUNC_GPU_DISTRIBUTIONS_H

#include <punc/mesh.h>
#include <punc_gpu.h>

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
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

//! A velocity distribution sampled on the device: kind (PUNC_VDF_*), vth, vd,
//! kappa, alpha and the sampling box.  Normalisation, max() and the
//! create_flux flags come from the library (punc_gpu_vdf_info).
class VelocityPdf : public Pdf {
public:
    int kind = PUNC_VDF_MAXWELLIAN;
    double kappa = 0.0, alpha = 0.0;
    double factor = 0.0;

    double max() override { return max_; }

protected:
    double max_ = 0.0;
    void init(int k, double vth, const std::vector<double>& vd, double kap, double alp, double range) {
        kind = k;
        _vth = vth;
        _vd = vd;
        kappa = kap;
        alpha = alp;
        dim = (int)vd.size();
        vdf_range = range;
        domain.resize(2 * dim);
        for (int i = 0; i < dim; ++i) {  // Pdf::update_domain (distributions.h:139-146)
            domain[i] = _vd[i] - vdf_range * _vth;
            domain[i + dim] = _vd[i] + vdf_range * _vth;
        }
        int fn = 0, fm = 0;
        check(punc_gpu_vdf_info(kind, vth, _vd.data(), kappa, alpha, vdf_range, &factor, &max_, &fn, &fm));
        has_icdf = kind == PUNC_VDF_MAXWELLIAN;
        has_flux_number = fn != 0;
        has_flux_max = fm != 0;
    }
};

//! Maxwellian (distributions.h:233-249, distributions.cpp:23-77).
class Maxwellian : public VelocityPdf {
public:
    Maxwellian(double vth, std::vector<double> vd, double range = 5.0) {
        init(PUNC_VDF_MAXWELLIAN, vth, vd, 0.0, 0.0, range);
    }
    double debye(double m, double q, double n, double eps0) override {
        return std::sqrt(eps0 * m / (n * q * q)) * _vth;
    }
};

//! Kappa (distributions.h:265-297, distributions.cpp:79-134), kappa > 3/2.
class Kappa : public VelocityPdf {
public:
    double k;
    Kappa(double vth, std::vector<double> vd, double k, double range = 7.0) : k(k) {
        init(PUNC_VDF_KAPPA, vth, vd, k, 0.0, range);
    }
    double debye(double m, double q, double n, double eps0) override {
        double B = (k - 0.5) / (k - 1.5);
        return std::sqrt(eps0 * m / (n * q * q * B)) * _vth;
    }
};

//! Cairns (distributions.h:299-333, distributions.cpp:136-191).
class Cairns : public VelocityPdf {
public:
    Cairns(double vth, std::vector<double> vd, double alpha_, double range = 7.0) {
        init(PUNC_VDF_CAIRNS, vth, vd, 0.0, alpha_, range);
    }
    double debye(double m, double q, double n, double eps0) override {
        double B = (1. + 3. * alpha) / (1. + 15 * alpha);
        return std::sqrt(eps0 * m / (n * q * q * B)) * _vth;
    }
};

//! KappaCairns (distributions.h:335-360, distributions.cpp:193-243).
class KappaCairns : public VelocityPdf {
public:
    double k;
    KappaCairns(double vth, std::vector<double> vd, double k, double alpha_, double range = 15.0) : k(k) {
        init(PUNC_VDF_KAPPA_CAIRNS, vth, vd, k, alpha_, range);
    }
    double debye(double m, double q, double n, double eps0) override {
        double B = ((k - 0.5) / (k - 1.5)) *
                   ((1. + 3. * alpha * ((k - 1.5) / (k - 0.5))) / (1. + 15 * alpha * ((k - 1.5) / (k - 2.5))));
        return std::sqrt(eps0 * m / (n * q * q * B)) * _vth;
    }
};

}  // namespace punc

#endif

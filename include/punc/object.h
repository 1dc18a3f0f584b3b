// Drop-in for the Object interface of punc/include/punc/object.h (PUNC++):
// the particle path credits absorbed charge to charge/current; the
// capacitance-matrix circuit (include/punc/poisson.h) reads and sets charge
// and potential.
#ifndef PUNC_GPU_OBJECT_H
#define PUNC_GPU_OBJECT_H

#include <memory>
#include <vector>

namespace punc {

class Object {
public:
    using size_t = std::size_t;
    explicit Object(size_t bnd_id) : bnd_id(bnd_id) {}
    virtual ~Object() {}
    virtual void set_potential(double p) { potential = p; }
    double get_potential() { return potential; }
    double charge = 0;   ///< incremented by Population::update
    double current = 0;  ///< current collected during the last Population::update
    size_t bnd_id;

protected:
    double potential = 0;
};

using ObjectVector = std::vector<std::shared_ptr<Object>>;

//! A two-port between objects node_a (-1: ground, minus side) and node_b
//! (plus side) with a value (reference object.h:42-70).
struct Source {
    Source(int node_a, int node_b, double value) : node_a(node_a), node_b(node_b), value(value) {}
    int node_a;
    int node_b;
    double value;
};
struct VSource : Source {
    VSource(int a, int b, double v) : Source(a, b, v) {}
};
struct ISource : Source {
    ISource(int a, int b, double v) : Source(a, b, v) {}
};
using VSourceVector = std::vector<VSource>;
using ISourceVector = std::vector<ISource>;

}  // namespace punc

#endif

// Drop-in for the Object interface of punc/include/punc/object.h (PUNC++):
// the particle path credits absorbed charge to charge/current; the circuit
// solve that reads them stays on the host.
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

}  // namespace punc

#endif

// Drop-in for punc/include/punc/distributor.h (PUNC++): CG1 charge deposit
// and patch volumes.
#ifndef PUNC_GPU_DISTRIBUTOR_H
#define PUNC_GPU_DISTRIBUTOR_H

#include <punc/population.h>

#include <stdexcept>
#include <vector>

namespace punc {

//! element_volume (distributor.cpp:35-69): (D+1)/patch volume per dof.
inline std::vector<double> element_volume(const Mesh& mesh, bool voronoi = true) {
    std::vector<double> dv((size_t)mesh.ndof);
    check(punc_gpu_element_volume(mesh.handle(), voronoi ? 1 : 0, dv.data()));
    return dv;
}

//! element_volume(V, voronoi) in the reference's form (distributor.h:52),
//! V a scalar CG1 space (CG1_space(mesh)).
inline std::vector<double> element_volume(const FunctionSpace& V, bool voronoi = true) {
    if (V.dg0() || V.value_dim() != 1) throw std::runtime_error("element_volume: V must be a scalar CG1 space");
    return element_volume(V.mesh(), voronoi);
}

//! distribute_cg1 (distributor.h:168-210): rho_j = dv_inv_j * sum_p q_p lambda_j(x_p).
template <typename PopulationType>
void distribute_cg1(PopulationType& pop, Function& rho, const std::vector<double>& dv_inv) {
    check(punc_gpu_distribute(pop.handle(), dv_inv.data(), rho.values.data()));
}

}  // namespace punc

#endif

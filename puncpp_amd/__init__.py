"""MI355X-native particle half of the PUNC++ PIC step.

libpunc_gpu.so (HIP kernels for gfx950 behind a C ABI, include/punc_gpu.h)
holds the product; this package is its Python mirror of the reference's
operator API (puncpp_amd.punc) and the build recipe (puncpp_amd.build).
"""
from .punc import (Cairns, CosinePosition, ExteriorFacet, FieldSolver, History, HostMesh, Kappa,  # noqa: F401
                   KappaCairns,
                   Maxwellian, Mesh, Object, State,
                   ParticleAmountType, Population, Species, UniformPosition, accel_cg1, boris_cg1,
                   box_mesh, create_flux, density_cg1, density_dg0, density_ema, distribute_cg1,
                   element_volume, inject_particles, kinetic_energy, load_particles, move, move_periodic,
                   particle_potential_energy_cg1, rho_ema)

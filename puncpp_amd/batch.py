"""Batched instances: many independent PIC instances (for example a probe-bias
sweep with uneven meshes and particle counts) advanced by one launch per
phase.

The instances' meshes are concatenated into one disconnected mesh (vertex,
cell and dof ids offset per instance), so every per-particle kernel works on
global cell and dof ids with no per-particle instance lookup.  Each instance
owns its own species (global species i*ns + s), which keeps the absorbed
charge per instance, and every species is bound to its instance
(punc_gpu_species_instance): its cell range, its contiguous range of
exterior facets (facets are ordered by cell), and the random streams the
instance would use as a mesh of its own.  An instance therefore evolves
exactly as the plain single-instance path does with the same seed.
"""
import numpy as np

from . import _native as N
from .punc import HostMesh, Mesh, Population
from ._native import check, dptr


class InstanceTable:
    """Offsets of each instance in the concatenated arrays."""

    def __init__(self, meshes):
        self.k = len(meshes)
        self.vert = np.zeros(self.k + 1, dtype=np.int64)
        self.cell = np.zeros(self.k + 1, dtype=np.int64)
        self.dof = np.zeros(self.k + 1, dtype=np.int64)
        for i, m in enumerate(meshes):
            self.vert[i + 1] = self.vert[i] + m.nv
            self.cell[i + 1] = self.cell[i] + m.nc
            self.dof[i + 1] = self.dof[i] + m.ndof
        self.facet = None  # exterior facet ranges, filled from the built mesh

    def instance_of_cell(self, cells):
        return np.searchsorted(self.cell, cells, side="right") - 1


def concat_meshes(meshes):
    """One HostMesh holding all instances, and the offset table."""
    t = InstanceTable(meshes)
    xyz = np.concatenate([m.xyz for m in meshes])
    cells = np.concatenate([m.cells + np.int32(t.vert[i]) for i, m in enumerate(meshes)])
    fb = np.concatenate([m.facet_bnd for m in meshes])
    vdof = np.concatenate([m.vdof + np.int32(t.dof[i]) for i, m in enumerate(meshes)])
    return HostMesh(xyz, cells, fb, vdof, int(t.dof[-1])), t


class BatchedPopulation:
    """K instances in one device population.

    species_per_instance[i]: list of Species of instance i (ns per instance).
    seeds[i]: the seed instance i uses for loading and injection."""

    def __init__(self, host_meshes, species_per_instance, seeds, capacity):
        hm, table = concat_meshes(host_meshes)
        self.host = hm
        self.table = table
        self.mesh = Mesh(hm)
        ef = self.mesh.exterior_facet_arrays()
        inst = table.instance_of_cell(ef["cell"])
        table.facet = np.searchsorted(inst, np.arange(table.k + 1), side="left")
        self.ns = len(species_per_instance[0])
        self.species = [s for sp in species_per_instance for s in sp]
        self.seeds = [int(x) for x in seeds]
        self.pop = Population(self.mesh, [(s.q, s.m) for s in self.species], capacity=capacity)
        L = N.lib()
        for g in range(len(self.species)):
            i, s = divmod(g, self.ns)
            check(L.punc_gpu_species_instance(self.pop._h, g, s, self.seeds[i], int(table.cell[i]),
                                              int(table.cell[i + 1]), int(table.facet[i]),
                                              int(table.facet[i + 1])))

    def register_injectors(self):
        L = N.lib()
        for g, s in enumerate(self.species):
            check(L.punc_gpu_injector_set(self.pop._h, g, s.vdf.vth(), dptr(s.vdf.vd()), float(s.n),
                                          float(s.vdf.vdf_range)))

    def load(self):
        L = N.lib()
        for g, s in enumerate(self.species):
            pdf = s.pdf
            check(L.punc_gpu_pop_load(self.pop._h, g, int(s.num), 0, float(pdf.amp), float(pdf.kw),
                                      int(pdf.axis), s.vdf.vth(), dptr(s.vdf.vd())))

    def split_particles(self):
        """Particles per instance, with instance-local cells and species."""
        st = self.pop.get_particles()
        inst = st["species"] // self.ns
        out = []
        for i in range(self.table.k):
            m = inst == i
            out.append({"x": st["x"][m], "v": st["v"][m], "species": st["species"][m] - i * self.ns,
                        "cell": st["cell"][m] - self.table.cell[i]})
        return out

    def split_dofs(self, arr):
        return [arr[self.table.dof[i]:self.table.dof[i + 1]] for i in range(self.table.k)]

    def split_absorbed(self, ab):
        return [ab[i * self.ns:(i + 1) * self.ns] for i in range(self.table.k)]


class BiasSweep:
    """The probe-bias sweep workload: n_inst independent sphere-probe
    instances on [-10, 10]^3 with a sphere of radius 1 as the object, meshes
    of n^3 cubes x 6 tetrahedra with n drawn uniformly from {16, 20, 24, 28,
    32}, probe bias linspace(-5, 5) (in T_e/e) shaping each instance's frozen
    field E = 0.05 bias x / max(|x|, r)^3, density scale drawn uniformly from
    [0.5, 2] (scale * base_particles particles per instance, half electrons
    and half ions), Maxwellian velocities with drift (0, 0, 0.5 vth_e) and
    boundary injection, Boris push with B = (0, 0, 1), dt = 0.02.  Instance i
    loads and injects with seed seed_base + i."""

    def __init__(self, n_inst=64, seed=1, base_particles=2_000_000, seed_base=1000):
        from .punc import Maxwellian, Mesh, ParticleAmountType, Species, UniformPosition, box_mesh
        rng = np.random.default_rng(seed)
        self.n_inst = n_inst
        self.ns = rng.choice([16, 20, 24, 28, 32], size=n_inst)
        self.bias = np.linspace(-5.0, 5.0, n_inst)
        self.scale = rng.uniform(0.5, 2.0, size=n_inst)
        self.L, self.r, self.dt = 10.0, 1.0, 0.02
        self.vd = [0.0, 0.0, 0.5]
        self.B = np.array([0.0, 0.0, 1.0])
        self.seeds = [seed_base + i for i in range(n_inst)]
        self.host_meshes, self.meshes, self.species, self.fields = [], [], [], []
        for i in range(n_inst):
            hm = box_mesh(int(self.ns[i]), [-self.L] * 3, [self.L] * 3, sphere_r=self.r)
            mesh = Mesh(hm)
            dens = self.scale[i] * base_particles / 2 / mesh.volume()
            sp = [Species(q, m, dens, 1.0, ParticleAmountType.phys_per_sim, mesh, UniformPosition(mesh),
                          Maxwellian(v, self.vd))
                  for (q, m), v in zip([(-1.0, 1.0), (1.0, 1836.0)], [1.0, 1.0 / np.sqrt(1836.0)])]
            xyz = np.zeros((hm.ndof, 3))
            xyz[hm.vdof] = hm.xyz
            r2 = np.maximum((xyz ** 2).sum(1), self.r * self.r)
            self.fields.append(np.ascontiguousarray(0.05 * self.bias[i] * xyz / r2[:, None] ** 1.5))
            self.host_meshes.append(hm)
            self.meshes.append(mesh)
            self.species.append(sp)

    def particles(self, i):
        return sum(int(s.num) for s in self.species[i])

    def total_particles(self):
        return sum(self.particles(i) for i in range(self.n_inst))

    def plain_population(self, i, sort=True):
        """Instance i as a population of its own: loaded, (sorted,) injectors
        registered, field set, deposited once."""
        from .punc import Population, _register_injectors, element_volume
        L = N.lib()
        pop = Population(self.meshes[i], [(-1.0, 1.0), (1.0, 1836.0)],
                         capacity=int(self.particles(i) * 1.1) + 100000)
        for s, S in enumerate(self.species[i]):
            check(L.punc_gpu_pop_load(pop._h, s, int(S.num), self.seeds[i], 0.0, 0.0, 0, S.vdf.vth(),
                                      dptr(S.vdf.vd())))
        if sort:
            pop.sort()
        _register_injectors(pop, self.species[i])
        check(L.punc_gpu_set_field(pop._h, dptr(self.fields[i])))
        pop.dv_inv = np.ascontiguousarray(element_volume(self.meshes[i]))
        check(L.punc_gpu_distribute(pop._h, dptr(pop.dv_inv), None))
        return pop

    def batched_population(self, sort=True):
        """All instances in one population: loaded, (sorted,) injectors
        registered, field set, deposited once."""
        from .punc import element_volume
        L = N.lib()
        bp = BatchedPopulation(self.host_meshes, self.species, self.seeds,
                               capacity=int(self.total_particles() * 1.1) + 1_000_000)
        bp.load()
        if sort:
            bp.pop.sort()
        bp.register_injectors()
        check(L.punc_gpu_set_field(bp.pop._h, dptr(np.concatenate(self.fields))))
        bp.dv_inv = np.ascontiguousarray(element_volume(bp.mesh))
        check(L.punc_gpu_distribute(bp.pop._h, dptr(bp.dv_inv), None))
        return bp

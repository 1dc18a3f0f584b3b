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

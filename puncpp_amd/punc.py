"""Python mirror of the PUNC++ particle-path operator API over libpunc_gpu.so.

Names, argument meaning and error behaviour follow the reference's C++ API
(punc/include/punc/{population,pusher,distributor,injector,distributions,
mesh,object,diagnostics}.h); DOLFIN functions become numpy arrays of CG1 dofs
(E: ndof x 3, rho: ndof).  Every operator runs on the GPU through the C ABI;
there is no CPU fallback.
"""
import enum
import math

import numpy as np

from . import _native as N
from ._native import check, dptr, i32ptr, i64ptr, vec3


# ---------------------------------------------------------------------------
# Meshes
# ---------------------------------------------------------------------------
class HostMesh:
    """Raw tetrahedral mesh arrays: xyz (nv,3), cells (nc,4) ascending vertex
    ids, facet_bnd (nc,4) boundary id of local facet i (opposite vertex i),
    vdof (nv,) vertex -> CG1 dof, ndof."""

    def __init__(self, xyz, cells, facet_bnd, vdof=None, ndof=None):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
        self.facet_bnd = np.ascontiguousarray(facet_bnd, dtype=np.int32).reshape(-1, 4)
        nv = self.xyz.shape[0]
        if vdof is None:
            vdof = np.arange(nv, dtype=np.int32)
            ndof = nv
        self.vdof = np.ascontiguousarray(vdof, dtype=np.int32)
        self.ndof = int(ndof)

    @property
    def nv(self):
        return self.xyz.shape[0]

    @property
    def nc(self):
        return self.cells.shape[0]


def box_mesh(n, lo, hi, sphere_r=0.0, periodic=False):
    """n^3 cubes x 6 Kuhn tetrahedra on [lo,hi]; cells with centroid within
    sphere_r of the origin removed (object boundary 2; exterior is 1)."""
    L = N.lib()
    h = N.vp()
    lo3, hi3 = vec3(lo), vec3(hi)
    check(L.punc_gpu_meshgen_box(int(n), dptr(lo3), dptr(hi3), float(sphere_r), int(bool(periodic)),
                                 N.C.byref(h)))
    try:
        nv, nc, nd = N.C.c_int64(), N.C.c_int64(), N.C.c_int64()
        check(L.punc_gpu_hostmesh_sizes(h, N.C.byref(nv), N.C.byref(nc), N.C.byref(nd)))
        xyz = np.empty((nv.value, 3))
        cells = np.empty((nc.value, 4), dtype=np.int32)
        fb = np.empty((nc.value, 4), dtype=np.int32)
        vdof = np.empty(nv.value, dtype=np.int32)
        check(L.punc_gpu_hostmesh_copy(h, dptr(xyz), i32ptr(cells), i32ptr(fb), i32ptr(vdof)))
    finally:
        L.punc_gpu_hostmesh_free(h)
    return HostMesh(xyz, cells, fb, vdof, nd.value)


class ExteriorFacet:
    """Exterior facet (reference mesh.h:38-44) plus its owning cell."""

    def __init__(self, area, vertices, normal, basis, cell):
        self.area = area
        self.vertices = vertices
        self.normal = normal
        self.basis = basis
        self.cell = cell


class Mesh:
    """Device-resident mesh and localizer (reference Mesh, mesh.h:49-97, plus
    the localizer part of Population<3>, population.h:310-389)."""

    dim = 3
    ext_bnd_id = 1

    def __init__(self, host):
        self.host = host
        L = N.lib()
        h = N.vp()
        check(L.punc_gpu_mesh_create(host.nv, dptr(host.xyz), host.nc, i32ptr(host.cells),
                                     i32ptr(host.facet_bnd), i32ptr(host.vdof), host.ndof,
                                     N.C.byref(h)))
        self._h = h
        nv, nc, nd, ne = (N.C.c_int64() for _ in range(4))
        vol = N.C.c_double()
        check(L.punc_gpu_mesh_info(h, N.C.byref(nv), N.C.byref(nc), N.C.byref(nd), N.C.byref(ne),
                                   N.C.byref(vol)))
        self.num_vertices, self.num_cells, self.ndof = nv.value, nc.value, nd.value
        self.num_exterior = ne.value
        self._volume = vol.value
        bnd = set(np.unique(host.facet_bnd[self._boundary_mask()]).tolist())
        self.num_objects = len([b for b in bnd if b >= 2])
        self._facets = None

    def _boundary_mask(self):
        adj = self.localizer()["adj"]
        return adj < 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().punc_gpu_mesh_destroy(h)
            self._h = None

    def volume(self):
        return self._volume

    def domain_size(self):
        # Mesh::domain_size (mesh.cpp:124-145): per axis the largest coordinate >= 0
        return [max(0.0, float(self.host.xyz[:, d].max())) for d in range(3)]

    def localizer(self):
        nc = self.num_cells
        out = {"adj": np.empty((nc, 4), dtype=np.int64), "planes": np.empty((nc, 16)),
               "bary": np.empty((nc, 12)), "vcoords": np.empty((nc, 12)),
               "dofs": np.empty((nc, 4), dtype=np.int32)}
        check(N.lib().punc_gpu_mesh_localizer(self._h, i64ptr(out["adj"]), dptr(out["planes"]),
                                              dptr(out["bary"]), dptr(out["vcoords"]),
                                              i32ptr(out["dofs"])))
        return out

    def exterior_facet_arrays(self):
        ne = self.num_exterior
        a = {"area": np.empty(ne), "vertices": np.empty((ne, 9)), "normal": np.empty((ne, 3)),
             "basis": np.empty((ne, 9)), "cell": np.empty(ne, dtype=np.int32)}
        check(N.lib().punc_gpu_exterior_facets(self._h, dptr(a["area"]), dptr(a["vertices"]),
                                               dptr(a["normal"]), dptr(a["basis"]), i32ptr(a["cell"])))
        return a

    @property
    def exterior_facets(self):
        if self._facets is None:
            a = self.exterior_facet_arrays()
            self._facets = [ExteriorFacet(a["area"][j], a["vertices"][j], a["normal"][j],
                                          a["basis"][j], int(a["cell"][j]))
                            for j in range(self.num_exterior)]
        return self._facets

    def locate(self, xs):
        xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, 3)
        out = np.empty(xs.shape[0], dtype=np.int64)
        check(N.lib().punc_gpu_mesh_locate(self._h, xs.shape[0], dptr(xs), i64ptr(out)))
        return out


def element_volume(mesh, voronoi=True):
    """Reference distributor.cpp:35-69: (D+1)/patch volume per CG1 dof."""
    out = np.empty(mesh.ndof)
    check(N.lib().punc_gpu_element_volume(mesh._h, int(bool(voronoi)), dptr(out)))
    return out


# ---------------------------------------------------------------------------
# Distributions and species
# ---------------------------------------------------------------------------
class UniformPosition:
    """Reference distributions.h:186-221."""

    amp = 0.0
    kw = 0.0
    axis = 0

    def __init__(self, mesh=None):
        self.mesh = mesh

    def max(self):
        return 1.0


class CosinePosition(UniformPosition):
    """inside(x) * (1 + amp*cos(kw*x[axis])): a density-perturbed uniform
    position pdf (the Langmuir-wave pattern of demo/langmuir/langmuir.cpp:17-54)."""

    def __init__(self, mesh=None, amp=0.01, kw=1.0, axis=0):
        super().__init__(mesh)
        self.amp = float(amp)
        self.kw = float(kw)
        self.axis = int(axis)

    def max(self):
        return 1.0 + self.amp


class _Vdf:
    """Velocity distribution (reference Pdf, distributions.h:41-180): thermal
    speed, drift, the half width of the sampling box in units of vth, and the
    create_flux tables (num_particles, pdf_max per exterior facet) once
    create_flux has run.  Normalisation, Pdf::max() and the create_flux flags
    come from the library (punc_gpu_vdf_info)."""

    kind = 0

    def __init__(self, vth, vd, vdf_range, k=0.0, alpha=0.0):
        self._vth = float(vth)
        self._vd = vec3(vd)
        self.vdf_range = float(vdf_range)
        self.k = float(k)
        self.alpha = float(alpha)
        self.dim = 3
        self.num_particles = None
        self.pdf_max = None
        fac, mx = N.C.c_double(), N.C.c_double()
        fn, fm = N.C.c_int32(), N.C.c_int32()
        check(N.lib().punc_gpu_vdf_info(self.kind, self._vth, dptr(self._vd), self.k, self.alpha, self.vdf_range,
                                        N.C.byref(fac), N.C.byref(mx), N.C.byref(fn), N.C.byref(fm)))
        self.factor = fac.value
        self._max = mx.value
        self.has_flux_number = bool(fn.value)
        self.has_flux_max = bool(fm.value)
        self.has_icdf = self.kind == 0

    def vth(self):
        return self._vth

    def vd(self):
        return self._vd.copy()

    def max(self):
        return self._max

    def domain(self):
        """Pdf::update_domain (distributions.h:139-146): [lo x3, hi x3]."""
        return np.concatenate([self._vd - self.vdf_range * self._vth, self._vd + self.vdf_range * self._vth])

    def _debye_factor(self):
        return 1.0

    def debye(self, m, q, n, eps0):
        return math.sqrt(eps0 * m / (n * q * q * self._debye_factor())) * self._vth


class Maxwellian(_Vdf):
    """Reference distributions.h:233-249, distributions.cpp:23-77."""

    kind = 0

    def __init__(self, vth, vd, vdf_range=5.0):
        super().__init__(vth, vd, vdf_range)

    def debye(self, m, q, n, eps0):
        return math.sqrt(eps0 * m / (n * q * q)) * self._vth  # distributions.cpp:74-77


class Kappa(_Vdf):
    """Reference distributions.h:265-297, distributions.cpp:79-134 (kappa > 3/2).
    has_flux_number follows the reference's integer sum of the drift."""

    kind = 1

    def __init__(self, vth, vd, k, vdf_range=7.0):
        super().__init__(vth, vd, vdf_range, k=k)

    def _debye_factor(self):
        return (self.k - 0.5) / (self.k - 1.5)  # distributions.cpp:130-134


class Cairns(_Vdf):
    """Reference distributions.h:299-333, distributions.cpp:136-191."""

    kind = 2

    def __init__(self, vth, vd, alpha, vdf_range=7.0):
        super().__init__(vth, vd, vdf_range, alpha=alpha)

    def _debye_factor(self):
        return (1. + 3. * self.alpha) / (1. + 15 * self.alpha)  # distributions.cpp:187-191


class KappaCairns(_Vdf):
    """Reference distributions.h:335-360, distributions.cpp:193-243."""

    kind = 3

    def __init__(self, vth, vd, k, alpha, vdf_range=15.0):
        super().__init__(vth, vd, vdf_range, k=k, alpha=alpha)

    def _debye_factor(self):
        k, a = self.k, self.alpha  # distributions.cpp:238-243
        return ((k - 0.5) / (k - 1.5)) * ((1. + 3. * a * ((k - 1.5) / (k - 0.5))) / (1. + 15 * a * ((k - 1.5) / (k - 2.5))))


class ParticleAmountType(enum.Enum):
    per_cell = 0
    in_total = 1
    per_volume = 2
    phys_per_sim = 3


class Species:
    """Reference population.h:99-125, population.cpp:41-70 (the amount-type
    fall-through converting the amount into a statistical weight)."""

    def __init__(self, charge, mass, density, amount, type, mesh, pdf, vdf, eps0=1.0):
        t = ParticleAmountType(type)
        if t == ParticleAmountType.per_cell:
            amount *= mesh.num_cells
        if t in (ParticleAmountType.per_cell, ParticleAmountType.in_total):
            amount /= mesh.volume()
        if t in (ParticleAmountType.per_cell, ParticleAmountType.in_total,
                 ParticleAmountType.per_volume):
            amount = density / amount
        self.weight = amount
        self.q = charge * amount
        self.m = mass * amount
        self.n = density / amount
        self.num = int(self.n * mesh.volume())
        self.pdf = pdf
        self.vdf = vdf
        self.debye = vdf.debye(self.m, self.q, self.n, eps0)  # population.cpp:68


def create_flux(species, mesh, seed=0, n_iter=500000):
    """Reference injector.cpp:104-171: per species and exterior facet the
    expected number of particles per unit density and time and the flux-pdf
    maximum.  Closed forms where the distribution has them; otherwise n_iter
    Monte Carlo samples per facet on the device, from the Philox stream keyed
    by `seed` (the species' position in `species` names its substream)."""
    for i, s in enumerate(species):
        ne = mesh.num_exterior
        num, pmax = np.empty(ne), np.empty(ne)
        v = s.vdf
        check(N.lib().punc_gpu_vdf_flux(mesh._h, v.kind, v.vth(), dptr(v.vd()), v.k, v.alpha, v.vdf_range,
                                        int(seed), i, int(n_iter), dptr(num), dptr(pmax)))
        v.num_particles = num
        v.pdf_max = pmax


class Object:
    """Reference object.h:80-130 (charge/current bookkeeping only)."""

    def __init__(self, bnd_id):
        self.bnd_id = int(bnd_id)
        self.charge = 0.0
        self.current = 0.0
        self.potential = 0.0

    def set_potential(self, p):
        self.potential = p

    def get_potential(self):
        return self.potential


# ---------------------------------------------------------------------------
# Population
# ---------------------------------------------------------------------------
class Population:
    """Device-resident population (reference Population<3>,
    population.h:236-308).  species: list of Species (or of (q, m) pairs)."""

    g_dim = 3

    def __init__(self, mesh, species, capacity=None, stream=None):
        self.mesh = mesh
        qs, ms = [], []
        for s in species:
            q, m = (s.q, s.m) if hasattr(s, "q") else s
            qs.append(float(q))
            ms.append(float(m))
        self.q = np.array(qs)
        self.m = np.array(ms)
        self.nspecies = len(qs)
        if capacity is None:
            capacity = 1 << 20
        self.capacity = int(capacity)
        h = N.vp()
        check(N.lib().punc_gpu_pop_create(mesh._h, self.nspecies, dptr(self.q), dptr(self.m),
                                          self.capacity, stream, N.C.byref(h)))
        self._h = h
        self._inj_set = set()
        self.dv_inv_uploaded = False
        self._dens_len = None  # length of the last density computed on the device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().punc_gpu_pop_destroy(h)
            self._h = None

    def _sync_species(self):
        n = N.C.c_int32()
        check(N.lib().punc_gpu_pop_species(self._h, N.C.byref(n), None, None))
        q, m = np.empty(n.value), np.empty(n.value)
        check(N.lib().punc_gpu_pop_species(self._h, N.C.byref(n), dptr(q), dptr(m)))
        self.q, self.m, self.nspecies = q, m, n.value

    def species_index(self, q, m):
        for s in range(self.nspecies):
            if self.q[s] == q and self.m[s] == m:
                return s
        raise ValueError(f"no species with q={q}, m={m} in this population")

    def add_particles(self, xs, vs, q, m):
        """Reference population.h:411-435; points outside the mesh are skipped.
        Returns the number added."""
        s = self.species_index(q, m)
        xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, 3)
        vs = np.ascontiguousarray(vs, dtype=np.float64).reshape(-1, 3)
        if xs.shape != vs.shape:
            raise ValueError("xs and vs differ in shape")
        k = N.C.c_int64()
        check(N.lib().punc_gpu_pop_add(self._h, s, xs.shape[0], dptr(xs), dptr(vs), N.C.byref(k)))
        return k.value

    def add_located(self, xs, vs, species, cells):
        xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, 3)
        vs = np.ascontiguousarray(vs, dtype=np.float64).reshape(-1, 3)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        cl = np.ascontiguousarray(cells, dtype=np.int64)
        check(N.lib().punc_gpu_pop_add_located(self._h, xs.shape[0], dptr(xs), dptr(vs), i32ptr(sp),
                                               i64ptr(cl)))

    def locate(self, x):
        return int(self.mesh.locate(np.asarray(x, dtype=np.float64).reshape(1, 3))[0])

    def size(self):
        n = N.C.c_int64()
        check(N.lib().punc_gpu_pop_size(self._h, N.C.byref(n)))
        return n.value

    def counts(self):
        out = np.empty(self.nspecies, dtype=np.int64)
        check(N.lib().punc_gpu_pop_counts(self._h, i64ptr(out)))
        return out

    def num_of_particles(self):
        return int(self.counts().sum())

    def num_of_positives(self):
        c = self.counts()
        return int(c[self.q > 0].sum())

    def num_of_negatives(self):
        c = self.counts()
        return int(c[self.q < 0].sum())

    def get_particles(self):
        n = self.size()
        xs, vs = np.empty((n, 3)), np.empty((n, 3))
        sp = np.empty(n, dtype=np.int32)
        cl = np.empty(n, dtype=np.int64)
        check(N.lib().punc_gpu_pop_get(self._h, dptr(xs), dptr(vs), i32ptr(sp), i64ptr(cl)))
        return {"x": xs, "v": vs, "species": sp, "cell": cl}

    def update(self, objects=(), dt=1.0):
        """Reference population.h:641-699: relocate, remove leavers, and per
        object current = sum q of absorbed particles; charge += current;
        current /= dt.  Returns absorbed counts (nspecies x 64)."""
        ab = np.zeros((self.nspecies, N_MAX_BND), dtype=np.int64)
        check(N.lib().punc_gpu_update(self._h, i64ptr(ab)))
        _objects_update(objects, ab, self.q, dt)
        return ab

    def sort(self):
        check(N.lib().punc_gpu_sort(self._h))

    def digest(self):
        """(digest, counts) per species: an order-independent 64-bit hash sum
        of the particles' bit patterns and cells (instance-local for bound
        species), and the particle counts."""
        d = np.zeros(self.nspecies, dtype=np.uint64)
        c = np.zeros(self.nspecies, dtype=np.int64)
        check(N.lib().punc_gpu_pop_digest(self._h, d.ctypes.data_as(N.u64_p), i64ptr(c)))
        return d, c

    def near_face_count(self):
        """Particles relocated within 1e-12 of a facet plane of their cell
        (the walk's proj, population.h:547-565) since the last call."""
        k = N.C.c_int64()
        check(N.lib().punc_gpu_near_face(self._h, N.C.byref(k)))
        return k.value

    def set_potential(self, phi):
        """phi (ndof) for the potential energy each fused step computes at
        its starting positions (None switches it off)."""
        if phi is None:
            check(N.lib().punc_gpu_set_potential(self._h, None))
            return
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        if phi.shape != (self.mesh.ndof,):
            raise ValueError(f"phi must hold {self.mesh.ndof} dof values")
        check(N.lib().punc_gpu_set_potential(self._h, dptr(phi)))

    def step_potential_energy(self):
        """Potential energy computed by the last fused step (NaN if off)."""
        v = N.C.c_double()
        check(N.lib().punc_gpu_step_potential_energy(self._h, N.C.byref(v)))
        return v.value

    def save_file(self, fname, binary=True):
        """Population::save_file (population.h:807-839): every particle as
        x[3], v[3], q, m - 64-byte binary records or "%.17g" text lines - in
        cell order (the population is re-sorted by cell first)."""
        check(N.lib().punc_gpu_pop_save_file(self._h, str(fname).encode(), int(bool(binary))))

    def load_file(self, fname, binary=True):
        """Population::load_file (population.h:841-880): appends the file's
        particles (those inside the mesh), registering a species for every
        new (q, m).  Returns the number added."""
        k = N.C.c_int64()
        check(N.lib().punc_gpu_pop_load_file(self._h, str(fname).encode(), int(bool(binary)), N.C.byref(k)))
        self._sync_species()
        return k.value

    def statistics(self, stats=None):
        """Population::statistics (population.h:746-805): [mean, std] of the
        speed of the q < 0 particles, then of the q > 0 ones.  Entries of a
        class without particles keep the values passed in stats."""
        out = np.zeros(4) if stats is None else np.array(stats, dtype=np.float64)
        check(N.lib().punc_gpu_statistics(self._h, dptr(out)))
        return out

    def set_launch(self, grid, block=256, deposit_mode=1):
        check(N.lib().punc_gpu_set_launch(self._h, int(grid), int(block), int(deposit_mode)))

    def get_launch(self):
        g, b, d = N.C.c_int32(), N.C.c_int32(), N.C.c_int32()
        check(N.lib().punc_gpu_get_launch(self._h, N.C.byref(g), N.C.byref(b), N.C.byref(d)))
        return g.value, b.value, d.value

    def synchronize(self):
        check(N.lib().punc_gpu_synchronize(self._h))


N_MAX_BND = 64


def _objects_update(objects, absorbed, q, dt):
    for o in objects:
        o.current = 0.0
    for o in objects:
        o.current = float(np.dot(absorbed[:, o.bnd_id].astype(np.float64), q))
        o.charge += o.current
        o.current /= dt


# ---------------------------------------------------------------------------
# Operators (reference pusher.h, distributor.h, injector.h, diagnostics.h)
# ---------------------------------------------------------------------------
def _set_field(pop, E):
    E = np.ascontiguousarray(E, dtype=np.float64).reshape(pop.mesh.ndof, 3)
    check(N.lib().punc_gpu_set_field(pop._h, dptr(E)))


def accel_cg1(pop, E, dt):
    """pusher.h:154-201.  E: CG1 dofs (ndof, 3).  Returns the mid-step KE."""
    _set_field(pop, E)
    ke = N.C.c_double()
    check(N.lib().punc_gpu_accel(pop._h, float(dt), N.C.byref(ke)))
    return ke.value


def boris_cg1(pop, E, B, dt):
    """pusher.h:282-372."""
    _set_field(pop, E)
    ke = N.C.c_double()
    check(N.lib().punc_gpu_boris(pop._h, dptr(vec3(B)), float(dt), N.C.byref(ke)))
    return ke.value


def move(pop, dt):
    """pusher.h:648-662."""
    check(N.lib().punc_gpu_move(pop._h, float(dt)))


def move_periodic(pop, dt, Ld):
    """pusher.h:666-681."""
    check(N.lib().punc_gpu_move_periodic(pop._h, float(dt), dptr(vec3(Ld))))


def distribute_cg1(pop, dv_inv, rho=None):
    """distributor.h:168-210.  Returns rho (ndof); fills `rho` when given."""
    out = rho if rho is not None else np.empty(pop.mesh.ndof)
    dv = np.ascontiguousarray(dv_inv, dtype=np.float64)
    check(N.lib().punc_gpu_distribute(pop._h, dptr(dv), dptr(out)))
    return out


# ---------------------------------------------------------------------------
# Field solve (PoissonSolver, poisson.cpp:260-376; ESolver, efield.cpp:41-75)
# ---------------------------------------------------------------------------
class FieldSolver:
    """CG1 Poisson solve eps0 grad(phi).grad(v) dx = rho v dx and the vector
    CG1 projection E = -grad(phi), on the device.  exterior: Dirichlet on the
    exterior boundary (id 1); n_objects: objects (ids 2..) held at given
    potentials; remove_null_space: no Dirichlet condition (periodic box), the
    potential has zero mean.  Tolerances as the reference's Krylov solvers
    (relative 1e-12, absolute 1e-14).  preconditioner: "amg" (smoothed-
    aggregation multigrid, the default where a hierarchy can be built; the
    reference's PETSc solver takes hypre's AMG) or "jacobi"."""

    def __init__(self, mesh, eps0=1.0, exterior=True, n_objects=0, remove_null_space=False, warm_start=False,
                 preconditioner=None):
        self.mesh = mesh
        self.n_objects = int(n_objects)
        h = N.vp()
        check(N.lib().punc_gpu_field_create(mesh._h, float(eps0), int(bool(exterior)), self.n_objects,
                                            int(bool(remove_null_space)), N.C.byref(h)))
        self._h = h
        self.iterations = (0, 0)
        self.set_warm_start(warm_start)
        if preconditioner is not None:
            self.set_preconditioner(preconditioner)

    _PRECONDITIONERS = {"jacobi": 0, "amg": 1}

    def set_preconditioner(self, kind):
        """Preconditioner of the Poisson solves, "amg" or "jacobi"; "amg"
        raises the library's error when no hierarchy could be built."""
        if kind not in self._PRECONDITIONERS:
            raise ValueError(f"preconditioner must be one of {sorted(self._PRECONDITIONERS)}")
        check(N.lib().punc_gpu_field_set_preconditioner(self._h, self._PRECONDITIONERS[kind]))

    @property
    def preconditioner(self):
        """(kind, multigrid levels, coarsest dense size)."""
        k, lv, nc = N.C.c_int(), N.C.c_int(), N.C.c_int64()
        check(N.lib().punc_gpu_field_get_preconditioner(self._h, N.C.byref(k), N.C.byref(lv), N.C.byref(nc)))
        return ("amg" if k.value == 1 else "jacobi", lv.value, nc.value)

    def set_warm_start(self, warm):
        """Start each solve from the previous solution (the reference starts
        from zero; the result agrees to the solver tolerance)."""
        check(N.lib().punc_gpu_field_set_warm_start(self._h, int(bool(warm))))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().punc_gpu_field_destroy(h)
            self._h = None

    def _bc(self, exterior_values, object_potentials):
        ev = None if exterior_values is None else np.ascontiguousarray(exterior_values, dtype=np.float64)
        if ev is not None and ev.shape != (self.mesh.ndof,):
            raise ValueError("exterior_values must hold one value per dof")
        op = None if object_potentials is None else np.ascontiguousarray(object_potentials, dtype=np.float64)
        if op is not None and op.shape != (self.n_objects,):
            raise ValueError(f"object_potentials must hold {self.n_objects} values")
        return ev, op

    def solve(self, rho, exterior_values=None, object_potentials=None):
        """PoissonSolver::solve: rho (ndof) -> phi (ndof)."""
        ev, op = self._bc(exterior_values, object_potentials)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        phi = np.empty(self.mesh.ndof)
        it = N.C.c_int32()
        check(N.lib().punc_gpu_poisson_solve(self._h, dptr(rho), dptr(ev) if ev is not None else None,
                                             dptr(op) if op is not None else None, dptr(phi), N.C.byref(it)))
        self.iterations = (it.value, 0)
        return phi

    def efield(self, phi):
        """ESolver::solve: phi (ndof) -> E (ndof x 3)."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        E = np.empty((self.mesh.ndof, 3))
        it = N.C.c_int32()
        check(N.lib().punc_gpu_efield_solve(self._h, dptr(phi), dptr(E), N.C.byref(it)))
        self.iterations = (0, it.value)
        return E

    def efield_adjoint(self, gE):
        """Reverse of efield: dJ/dphi (ndof) = B^T M^-1 dJ/dE for gE = dJ/dE
        (ndof x 3), M the vector-CG1 mass matrix and B the projection's
        gradient form (E = M^-1 B phi)."""
        gE = np.ascontiguousarray(gE, dtype=np.float64).reshape(self.mesh.ndof, 3)
        gphi = np.empty(self.mesh.ndof)
        it = N.C.c_int32()
        check(N.lib().punc_gpu_efield_adjoint(self._h, dptr(gE), dptr(gphi), N.C.byref(it)))
        return gphi

    def object_charges(self, phi):
        """Charge on each object of the potential phi (Charge.ufl)."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        q = np.empty(self.n_objects)
        check(N.lib().punc_gpu_object_charges(self._h, dptr(phi), dptr(q)))
        return q

    def set_circuit(self, dt, vsources=(), isources=()):
        """CircuitCM (object_CM.cpp:53-357): sources as (node_a, node_b,
        value) with node_a = -1 for ground.  Returns the inverse capacitance
        matrix."""
        def arrays(src):
            a = np.array([int(x[0]) for x in src], dtype=np.int32)
            b = np.array([int(x[1]) for x in src], dtype=np.int32)
            v = np.array([float(x[2]) for x in src], dtype=np.float64)
            return a, b, v
        va, vb, vv = arrays(vsources)
        ia, ib, iv = arrays(isources)
        inv_c = np.empty((self.n_objects, self.n_objects))
        check(N.lib().punc_gpu_circuit_set(self._h, float(dt), len(va), i32ptr(va), i32ptr(vb), dptr(vv), len(ia),
                                           i32ptr(ia), i32ptr(ib), dptr(iv), dptr(inv_c)))
        return inv_c

    def solve_circuit(self, rho, objects):
        """PoissonSolver::solve_circuit: updates each object's charge and
        potential (objects: n_objects Object instances); returns phi."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        q = np.array([o.charge for o in objects], dtype=np.float64)
        v = np.empty(self.n_objects)
        phi = np.empty(self.mesh.ndof)
        it = N.C.c_int32()
        check(N.lib().punc_gpu_poisson_solve_circuit(self._h, dptr(rho), dptr(q), dptr(v), dptr(phi),
                                                     N.C.byref(it)))
        for o, qi, vi in zip(objects, q, v):
            o.charge = float(qi)
            o.set_potential(float(vi))
        self.iterations = (it.value, 0)
        return phi

    def solve_population_circuit(self, pop, objects, copy_out=True):
        """solve_circuit with the population's deposit, E left in the
        population's field; returns (phi, E) when copy_out."""
        q = np.array([o.charge for o in objects], dtype=np.float64)
        v = np.empty(self.n_objects)
        phi = np.empty(self.mesh.ndof) if copy_out else None
        E = np.empty((self.mesh.ndof, 3)) if copy_out else None
        its = (N.C.c_int32 * 2)()
        check(N.lib().punc_gpu_field_solve_pop_circuit(self._h, pop._h, dptr(q), dptr(v),
                                                       dptr(phi) if copy_out else None,
                                                       dptr(E) if copy_out else None, its))
        for o, qi, vi in zip(objects, q, v):
            o.charge = float(qi)
            o.set_potential(float(vi))
        self.iterations = (its[0], its[1])
        return (phi, E) if copy_out else None

    def solve_population(self, pop, exterior_values=None, object_potentials=None, copy_out=True):
        """The population's deposited rho -> phi -> E, left in the
        population's field for its next push.  Returns (phi, E) when
        copy_out."""
        ev, op = self._bc(exterior_values, object_potentials)
        phi = np.empty(self.mesh.ndof) if copy_out else None
        E = np.empty((self.mesh.ndof, 3)) if copy_out else None
        its = (N.C.c_int32 * 2)()
        check(N.lib().punc_gpu_field_solve_pop(self._h, pop._h, dptr(ev) if ev is not None else None,
                                               dptr(op) if op is not None else None,
                                               dptr(phi) if copy_out else None, dptr(E) if copy_out else None,
                                               its))
        self.iterations = (its[0], its[1])
        return (phi, E) if copy_out else None


# ---------------------------------------------------------------------------
# Run state and history files (diagnostics.h:39-117, diagnostics.cpp:35-126)
# ---------------------------------------------------------------------------
def _c_strtol(s, i):
    """C strtol(s + i, &end, 10): (value, end index); no number -> (0, i)."""
    j = i
    while j < len(s) and s[j] in " \t\n\v\f\r":
        j += 1
    k = j
    if k < len(s) and s[k] in "+-":
        k += 1
    d = k
    while d < len(s) and s[d].isdigit():
        d += 1
    if d == k:
        return 0, i
    return int(s[j:d]), d


def _c_strtod(s, i):
    """C strtod(s + i, &end) for decimal numbers, inf and nan: (value, end)."""
    import re
    m = re.compile(r"[ \t\n\v\f\r]*([+-]?(?:inf(?:inity)?|nan|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?))",
                   re.IGNORECASE).match(s, i)
    if not m:
        return 0.0, i
    return float(m.group(1)), m.end()


def _c_nan(x):
    # glibc prints a NaN with its sign ("-nan" for the x86 default NaN)
    return "-nan" if math.copysign(1.0, x) < 0 else "nan"


def _fmt_default(x):
    """A double through an ostream with default flags (precision 6, %g)."""
    return _c_nan(x) if math.isnan(x) else "%g" % x


def _fmt_sci16(x):
    """A double through an ostream set to std::scientific, precision 16."""
    return _c_nan(x) if math.isnan(x) else "%.16e" % x


class State:
    """Run state for restarts (reference State, diagnostics.h:39-71,
    diagnostics.cpp:35-68): one line "n+1<TAB>t<TAB>" followed, per object,
    by "charge<TAB>current<NEWLINE>".  load reads the first line only, so
    objects after the first read 0 (as the reference does)."""

    def __init__(self, fname):
        self.fname = str(fname)

    def load(self, objects=()):
        with open(self.fname) as f:
            line = f.readline()
        line = line[:-1] if line.endswith("\n") else line
        n, i = _c_strtol(line, 0)
        t, i = _c_strtod(line, i)
        for o in objects:
            o.charge, i = _c_strtod(line, i)
            o.current, i = _c_strtod(line, i)
        return n, t

    def save(self, n, t, objects=()):
        with open(self.fname, "w") as f:
            f.write(f"{int(n) + 1}\t{_fmt_default(t)}\t")
            for o in objects:
                f.write(f"{_fmt_default(o.charge)}\t")
                f.write(f"{_fmt_default(o.current)}\n")


class History:
    """Time-series file (reference History, diagnostics.h:73-170,
    diagnostics.cpp:70-126): a four-line header (unless continuing), then per
    save one line: n, then t, ne, ni, KE, PE, per object V, I, Q, and with
    stats the speed statistics, each as "%.16e"."""

    def __init__(self, fname, objects=(), dim=3, stats=False, continue_simulation=False, hex_output=False):
        self.dim = dim
        self.stats = stats
        self.hex_output = hex_output  # accepted; the reference writes decimal either way
        self._f = open(str(fname), "a" if continue_simulation else "w")
        if not continue_simulation:
            f = self._f
            nobj = len(objects)
            f.write("#:xaxis\tt\n")
            f.write("#:name\tn\tt\tne\tni\tKE\tPE")
            for i in range(nobj):
                f.write(f"\tV[{i}]\tI[{i}]\tQ[{i}]")
            if stats:
                f.write("\tmean_e\tstdev_e\tmean_i\tstdev_i")
            f.write("\n")
            f.write("#:long\ttimestep\ttime\t\"number of electrons\"\t")
            f.write("\"number of ions\"\t\"kinetic energy\"\t")
            f.write("\"potential energy\"")
            for _ in range(nobj):
                f.write("\tvoltage\tcurrent\tcharge")
            if stats:
                f.write("\tmean velocity of electrons\tstandard deviation of electron velocities"
                        "\tmean velocity of ions\tstandard deviation of ion velocities")
            f.write("\n")
            f.write("#:units\t1\ts\tm**(-3)\tm**(-3)\tJ\tJ")
            for _ in range(nobj):
                f.write("\tV")
                f.write({1: "\tA/m**2", 2: "\tA/m", 3: "\tA"}.get(dim, ""))
                f.write("\tC")
            if stats:
                f.write("\tm/s" * 4)
            f.write("\n")
            f.flush()

    def save(self, n, t, num_e, num_i, KE, PE, objects=(), pop=None):
        e = _fmt_sci16
        parts = [f"{int(n)}\t" + "\t".join(e(v) for v in (t, num_e, num_i, KE, PE))]
        for o in objects:
            parts.append(f"\t{e(o.get_potential())}\t{e(o.current)}\t{e(o.charge)}")
        if self.stats:
            st = pop.statistics(np.zeros(4))
            parts.append("".join(f"\t{e(v)}" for v in st))
        self._f.write("".join(parts) + "\n")
        self._f.flush()

    def close(self):
        self._f.close()


def kinetic_energy(pop):
    """diagnostics.h:248-265."""
    ke = N.C.c_double()
    check(N.lib().punc_gpu_kinetic_energy(pop._h, N.C.byref(ke)))
    return ke.value


def particle_potential_energy_cg1(pop, phi):
    """diagnostics.h:379-412: sum over the particles of 0.5 q phi(x_p), phi a
    CG1 function given by its ndof dof values."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    if phi.shape != (pop.mesh.ndof,):
        raise ValueError(f"phi must hold {pop.mesh.ndof} dof values")
    pe = N.C.c_double()
    check(N.lib().punc_gpu_potential_energy(pop._h, dptr(phi), N.C.byref(pe)))
    return pe.value


def density_cg1(V, pop, species, dv_inv=None):
    """diagnostics.h:492-555: CG1 number densities (ne, ni), each ndof long.
    V (the function space) is not needed: the dofs are the mesh's.  Weights
    are species[0].weight and species[1].weight; q < 0 counts as electron."""
    nd = pop.mesh.ndof
    ne, ni = np.empty(nd), np.empty(nd)
    dv = None
    if dv_inv is not None:
        dv = np.ascontiguousarray(dv_inv, dtype=np.float64)
        if dv.shape != (nd,):
            raise ValueError(f"dv_inv must hold {nd} values")
    check(N.lib().punc_gpu_density_cg1(pop._h, dptr(dv) if dv is not None else None,
                                       float(species[0].weight), float(species[1].weight),
                                       dptr(ne), dptr(ni)))
    pop._dens_len = nd
    return ne, ni


def density_dg0(Q, pop, species):
    """diagnostics.h:432-465: DG0 number densities (ne, ni), one value per cell
    (indexed by cell id).  q > 0 counts as ion."""
    nc = pop.mesh.num_cells
    ne, ni = np.empty(nc), np.empty(nc)
    check(N.lib().punc_gpu_density_dg0(pop._h, float(species[0].weight), float(species[1].weight),
                                       dptr(ne), dptr(ni)))
    pop._dens_len = nc
    return ne, ni


def rho_ema(pop, dt, tau, reset=False):
    """ema(rho, rho_ema, dt, tau) (diagnostics.h:573-588) over the population's
    last deposited charge density, the average kept on the device; returns it."""
    g = np.empty(pop.mesh.ndof)
    check(N.lib().punc_gpu_rho_ema(pop._h, float(dt), float(tau), int(bool(reset)), dptr(g)))
    return g


def density_ema(pop, dt, tau, reset=False):
    """ema(ne, ne_ema, dt, tau), ema(ni, ni_ema, dt, tau) over the last density
    computed (density_cg1 or density_dg0); returns (ne_ema, ni_ema)."""
    n = pop._dens_len
    if n is None:
        raise RuntimeError("density_ema: no density computed yet")
    ne, ni = np.empty(n), np.empty(n)
    check(N.lib().punc_gpu_density_ema(pop._h, float(dt), float(tau), int(bool(reset)), dptr(ne), dptr(ni)))
    return ne, ni


def _register_injectors(pop, species):
    for s in species:
        idx = pop.species_index(s.q, s.m)
        if idx in pop._inj_set:
            continue
        v = s.vdf
        tables = v.num_particles is not None and v.pdf_max is not None
        num = np.ascontiguousarray(v.num_particles, dtype=np.float64) if tables else None
        pmax = np.ascontiguousarray(v.pdf_max, dtype=np.float64) if tables else None
        check(N.lib().punc_gpu_injector_set_vdf(pop._h, idx, v.kind, v.vth(), dptr(v.vd()), v.k, v.alpha,
                                                float(s.n), float(v.vdf_range),
                                                dptr(num) if tables else None, dptr(pmax) if tables else None))
        pop._inj_set.add(idx)


def inject_particles(pop, species, facets=None, dt=1.0, seed=0, step=0):
    """injector.h:157-232 with a counter-based stream keyed by (seed, step).
    Returns the number of particles added."""
    _register_injectors(pop, species)
    k = N.C.c_int64()
    check(N.lib().punc_gpu_inject(pop._h, float(dt), int(seed), int(step), N.C.byref(k)))
    return k.value


def load_particles(pop, species, seed=0):
    """injector.h:244-278: species.num particles per species."""
    for s in species:
        idx = pop.species_index(s.q, s.m)
        pdf = s.pdf
        v = s.vdf
        check(N.lib().punc_gpu_pop_load_vdf(pop._h, idx, int(s.num), int(seed), float(pdf.amp), float(pdf.kw),
                                            int(pdf.axis), v.kind, v.vth(), dptr(v.vd()), v.k, v.alpha,
                                            float(v.vdf_range)))

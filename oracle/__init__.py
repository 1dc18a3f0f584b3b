"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/liboracle.so, the serial CPU restatement of the
PUNC++ particle path (oracle/oracle.cpp).  Only tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg may use it, and only as the checker.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

d_p = C.POINTER(C.c_double)
i32_p = C.POINTER(C.c_int32)
i64_p = C.POINTER(C.c_int64)
u64_p = C.POINTER(C.c_uint64)
vp = C.c_void_p

_SIGS = {
    "oracle_philox": (None, [u64_p, u64_p, u64_p]),
    "oracle_unif": (C.c_double, [C.c_uint64]),
    "oracle_erfc_inv": (C.c_double, [C.c_double]),
    "oracle_create": (vp, [C.c_int64, d_p, C.c_int64, i32_p, i32_p, i32_p, C.c_int64]),
    "oracle_destroy": (None, [vp]),
    "oracle_set_species": (None, [vp, C.c_int, d_p, d_p]),
    "oracle_localizer": (None, [vp, i64_p, d_p, d_p, d_p]),
    "oracle_num_exterior": (C.c_int64, [vp]),
    "oracle_exterior_facets": (None, [vp, d_p, d_p, d_p, d_p, i64_p]),
    "oracle_element_volume": (None, [vp, C.c_int, d_p]),
    "oracle_locate": (C.c_int64, [vp, d_p]),
    "oracle_locate_brute": (C.c_int64, [vp, d_p]),
    "oracle_relocate": (C.c_int64, [vp, d_p, C.c_int64]),
    "oracle_add_particles": (C.c_int64, [vp, C.c_int, C.c_int64, d_p, d_p]),
    "oracle_add_located": (None, [vp, C.c_int64, d_p, d_p, i32_p, i64_p]),
    "oracle_size": (C.c_int64, [vp]),
    "oracle_num_positives": (C.c_int64, [vp]),
    "oracle_num_negatives": (C.c_int64, [vp]),
    "oracle_get": (None, [vp, d_p, d_p, i32_p, i64_p]),
    "oracle_accel": (C.c_double, [vp, d_p, C.c_double]),
    "oracle_boris": (C.c_double, [vp, d_p, d_p, C.c_double]),
    "oracle_move": (None, [vp, C.c_double]),
    "oracle_move_periodic": (None, [vp, C.c_double, d_p]),
    "oracle_update": (C.c_int64, [vp, i64_p, d_p]),
    "oracle_distribute": (None, [vp, d_p, d_p]),
    "oracle_kinetic_energy": (C.c_double, [vp]),
    "oracle_potential_energy": (C.c_double, [vp, d_p]),
    "oracle_density_cg1": (None, [vp, d_p, C.c_double, C.c_double, d_p, d_p]),
    "oracle_density_dg0": (None, [vp, C.c_double, C.c_double, d_p, d_p]),
    "oracle_statistics": (None, [vp, d_p]),
    "oracle_flux": (None, [C.c_double, d_p, d_p, C.c_double, d_p, d_p]),
    "oracle_set_injector": (None, [vp, C.c_int, C.c_double, d_p, C.c_double, C.c_double]),
    "oracle_save_file": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "oracle_load_file": (C.c_int64, [vp, C.c_char_p, C.c_int]),
    "oracle_set_injector_vdf": (None, [vp, C.c_int, C.c_int, C.c_double, d_p, C.c_double, C.c_double,
                                       C.c_double, C.c_double, d_p, d_p]),
    "oracle_vdf_flux": (None, [vp, C.c_int, C.c_double, d_p, C.c_double, C.c_double, C.c_double, C.c_uint64,
                               C.c_int, C.c_int64, d_p, d_p]),
    "oracle_vdf_value": (C.c_double, [C.c_int, C.c_double, d_p, C.c_double, C.c_double, d_p]),
    "oracle_vdf_max": (C.c_double, [C.c_int, C.c_double, d_p, C.c_double, C.c_double, C.POINTER(C.c_int)]),
    "oracle_load_vdf": (C.c_int64, [vp, C.c_int, C.c_int64, C.c_uint64, C.c_double, C.c_double, C.c_int, C.c_int,
                                    C.c_double, d_p, C.c_double, C.c_double, C.c_double]),
    "oracle_inject": (C.c_int64, [vp, C.c_double, C.c_uint64, C.c_uint64]),
    "oracle_load": (C.c_int64, [vp, C.c_int, C.c_int64, C.c_uint64, C.c_double, C.c_double, C.c_int,
                                C.c_double, d_p]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `python oracle/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(d_p)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint64)
    k = np.ascontiguousarray(key, dtype=np.uint64)
    out = np.empty(4, dtype=np.uint64)
    lib().oracle_philox(c.ctypes.data_as(u64_p), k.ctypes.data_as(u64_p), out.ctypes.data_as(u64_p))
    return out


def unif(x):
    return lib().oracle_unif(int(x))


def erfc_inv(y):
    return lib().oracle_erfc_inv(float(y))


def vdf_value(kind, vth, vd, v, k=0.0, alpha=0.0):
    """Pdf::operator()(v) of a Maxwellian (0), Kappa (1), Cairns (2) or KappaCairns (3)."""
    vd = np.ascontiguousarray(vd, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().oracle_vdf_value(int(kind), float(vth), vd.ctypes.data_as(d_p), float(k), float(alpha),
                                  v.ctypes.data_as(d_p))


def vdf_max(kind, vth, vd, k=0.0, alpha=0.0):
    """(Pdf::max(), has_flux_number, has_flux_max)."""
    vd = np.ascontiguousarray(vd, dtype=np.float64)
    flags = (C.c_int * 2)()
    m = lib().oracle_vdf_max(int(kind), float(vth), vd.ctypes.data_as(d_p), float(k), float(alpha), flags)
    return m, bool(flags[0]), bool(flags[1])


def flux(vth, vd, normal, area):
    num, pmax = C.c_double(), C.c_double()
    vd = np.ascontiguousarray(vd, dtype=np.float64)
    n = np.ascontiguousarray(normal, dtype=np.float64)
    lib().oracle_flux(float(vth), vd.ctypes.data_as(d_p), n.ctypes.data_as(d_p), float(area),
                      C.byref(num), C.byref(pmax))
    return num.value, pmax.value


class OraclePop:
    """CPU restatement of Population<3> and the particle operators."""

    def __init__(self, host_mesh, q, m):
        L = lib()
        hm = host_mesh
        self._keep = (hm.xyz, hm.cells, hm.facet_bnd, hm.vdof)
        self.ndof = hm.ndof
        self.nc = hm.nc
        self._h = L.oracle_create(hm.nv, hm.xyz.ctypes.data_as(d_p), hm.nc, hm.cells.ctypes.data_as(i32_p),
                                  hm.facet_bnd.ctypes.data_as(i32_p), hm.vdof.ctypes.data_as(i32_p), hm.ndof)
        self.q = np.array(q, dtype=np.float64)
        self.m = np.array(m, dtype=np.float64)
        self.nspecies = len(self.q)
        L.oracle_set_species(self._h, self.nspecies, self.q.ctypes.data_as(d_p), self.m.ctypes.data_as(d_p))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_destroy(h)
            self._h = None

    def localizer(self):
        nc = self.nc
        adj = np.empty((nc, 4), dtype=np.int64)
        planes, bary, vc = np.empty((nc, 16)), np.empty((nc, 12)), np.empty((nc, 12))
        lib().oracle_localizer(self._h, adj.ctypes.data_as(i64_p), planes.ctypes.data_as(d_p),
                               bary.ctypes.data_as(d_p), vc.ctypes.data_as(d_p))
        return {"adj": adj, "planes": planes, "bary": bary, "vcoords": vc}

    def exterior_facets(self):
        ne = lib().oracle_num_exterior(self._h)
        a = {"area": np.empty(ne), "vertices": np.empty((ne, 9)), "normal": np.empty((ne, 3)),
             "basis": np.empty((ne, 9)), "cell": np.empty(ne, dtype=np.int64)}
        lib().oracle_exterior_facets(self._h, a["area"].ctypes.data_as(d_p), a["vertices"].ctypes.data_as(d_p),
                                     a["normal"].ctypes.data_as(d_p), a["basis"].ctypes.data_as(d_p),
                                     a["cell"].ctypes.data_as(i64_p))
        return a

    def element_volume(self, voronoi=True):
        out = np.empty(self.ndof)
        lib().oracle_element_volume(self._h, int(bool(voronoi)), out.ctypes.data_as(d_p))
        return out

    def locate(self, x, brute=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        f = lib().oracle_locate_brute if brute else lib().oracle_locate
        return int(f(self._h, x.ctypes.data_as(d_p)))

    def relocate(self, x, cell):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return int(lib().oracle_relocate(self._h, x.ctypes.data_as(d_p), int(cell)))

    def add_particles(self, species, xs, vs):
        xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, 3)
        vs = np.ascontiguousarray(vs, dtype=np.float64).reshape(-1, 3)
        return lib().oracle_add_particles(self._h, int(species), xs.shape[0], xs.ctypes.data_as(d_p),
                                          vs.ctypes.data_as(d_p))

    def add_located(self, xs, vs, species, cells):
        xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, 3)
        vs = np.ascontiguousarray(vs, dtype=np.float64).reshape(-1, 3)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        cl = np.ascontiguousarray(cells, dtype=np.int64)
        lib().oracle_add_located(self._h, xs.shape[0], xs.ctypes.data_as(d_p), vs.ctypes.data_as(d_p),
                                 sp.ctypes.data_as(i32_p), cl.ctypes.data_as(i64_p))

    def size(self):
        return int(lib().oracle_size(self._h))

    def num_of_positives(self):
        return int(lib().oracle_num_positives(self._h))

    def num_of_negatives(self):
        return int(lib().oracle_num_negatives(self._h))

    def get_particles(self):
        n = self.size()
        xs, vs = np.empty((n, 3)), np.empty((n, 3))
        sp = np.empty(n, dtype=np.int32)
        cl = np.empty(n, dtype=np.int64)
        lib().oracle_get(self._h, xs.ctypes.data_as(d_p), vs.ctypes.data_as(d_p), sp.ctypes.data_as(i32_p),
                         cl.ctypes.data_as(i64_p))
        return {"x": xs, "v": vs, "species": sp, "cell": cl}

    def accel(self, E, dt):
        E = np.ascontiguousarray(E, dtype=np.float64)
        return lib().oracle_accel(self._h, E.ctypes.data_as(d_p), float(dt))

    def boris(self, E, B, dt):
        E = np.ascontiguousarray(E, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        return lib().oracle_boris(self._h, E.ctypes.data_as(d_p), B.ctypes.data_as(d_p), float(dt))

    def move(self, dt):
        lib().oracle_move(self._h, float(dt))

    def move_periodic(self, dt, Ld):
        Ld = np.ascontiguousarray(Ld, dtype=np.float64)
        lib().oracle_move_periodic(self._h, float(dt), Ld.ctypes.data_as(d_p))

    def update(self):
        ab = np.zeros((self.nspecies, 64), dtype=np.int64)
        ch = np.zeros(64)
        lost = lib().oracle_update(self._h, ab.ctypes.data_as(i64_p), ch.ctypes.data_as(d_p))
        return ab, ch, int(lost)

    def distribute(self, dv_inv):
        dv = np.ascontiguousarray(dv_inv, dtype=np.float64)
        rho = np.empty(self.ndof)
        lib().oracle_distribute(self._h, dv.ctypes.data_as(d_p), rho.ctypes.data_as(d_p))
        return rho

    def kinetic_energy(self):
        return lib().oracle_kinetic_energy(self._h)

    def potential_energy(self, phi):
        return lib().oracle_potential_energy(self._h, _d(phi))

    def density_cg1(self, dv_inv, weight_e=1.0, weight_i=1.0):
        nd = self.ndof
        ne, ni = np.zeros(nd), np.zeros(nd)
        lib().oracle_density_cg1(self._h, _d(dv_inv), float(weight_e), float(weight_i), _d(ne), _d(ni))
        return ne, ni

    def density_dg0(self, weight_e=1.0, weight_i=1.0):
        nc = self.nc
        ne, ni = np.zeros(nc), np.zeros(nc)
        lib().oracle_density_dg0(self._h, float(weight_e), float(weight_i), _d(ne), _d(ni))
        return ne, ni

    def statistics(self, stats=None):
        st = np.full(4, np.nan) if stats is None else np.array(stats, dtype=np.float64)
        lib().oracle_statistics(self._h, _d(st))
        return st

    def set_injector(self, species, vth, vd, n, vrange=5.0):
        vd = np.ascontiguousarray(vd, dtype=np.float64)
        lib().oracle_set_injector(self._h, int(species), float(vth), vd.ctypes.data_as(d_p), float(n),
                                  float(vrange))

    def set_injector_vdf(self, species, kind, vth, vd, k, alpha, n, vrange, num, pmax):
        vd = np.ascontiguousarray(vd, dtype=np.float64)
        num = np.ascontiguousarray(num, dtype=np.float64)
        pmax = np.ascontiguousarray(pmax, dtype=np.float64)
        lib().oracle_set_injector_vdf(self._h, int(species), int(kind), float(vth), vd.ctypes.data_as(d_p),
                                      float(k), float(alpha), float(n), float(vrange), num.ctypes.data_as(d_p),
                                      pmax.ctypes.data_as(d_p))

    def vdf_flux(self, kind, vth, vd, k=0.0, alpha=0.0, vrange=5.0, seed=0, rng_sp=0, n_iter=500000):
        """create_flux tables (num, pdf_max) per exterior facet."""
        vd = np.ascontiguousarray(vd, dtype=np.float64)
        nf = lib().oracle_num_exterior(self._h)
        num, pmax = np.empty(nf), np.empty(nf)
        lib().oracle_vdf_flux(self._h, int(kind), float(vth), vd.ctypes.data_as(d_p), float(k), float(alpha),
                              float(vrange), int(seed), int(rng_sp), int(n_iter), _d(num), _d(pmax))
        return num, pmax

    def load_vdf(self, species, N, seed, amp, kw, axis, kind, vth, vd, k=0.0, alpha=0.0, vrange=5.0):
        vd = np.ascontiguousarray(vd, dtype=np.float64)
        return int(lib().oracle_load_vdf(self._h, int(species), int(N), int(seed), float(amp), float(kw), int(axis),
                                         int(kind), float(vth), vd.ctypes.data_as(d_p), float(k), float(alpha),
                                         float(vrange)))

    def save_file(self, fname, binary=True):
        assert lib().oracle_save_file(self._h, str(fname).encode(), int(bool(binary))) == 0

    def load_file(self, fname, binary=True):
        n = int(lib().oracle_load_file(self._h, str(fname).encode(), int(bool(binary))))
        assert n >= 0, f"cannot open {fname}"
        return n

    def inject(self, dt, seed, step):
        return int(lib().oracle_inject(self._h, float(dt), int(seed), int(step)))

    def load(self, species, N, seed, amp, kw, axis, vth, vd):
        vd = np.ascontiguousarray(vd, dtype=np.float64)
        return int(lib().oracle_load(self._h, int(species), int(N), int(seed), float(amp), float(kw),
                                     int(axis), float(vth), vd.ctypes.data_as(d_p)))

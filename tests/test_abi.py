"""C-ABI checks that need no GPU: the library loads and exports every entry
point include/punc_gpu.h declares; its host-side setup (mesh generation,
localizer, patch volumes, exterior facets, location, flux tables) is
bit-identical to the oracle; errors come back as status codes with messages."""
import math
import os
import re

import numpy as np
import pytest

import oracle as O
import puncpp_amd as P
from puncpp_amd import _native as N
from helpers import MI, ME, QE, QI

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "punc_gpu.h")).read()
    return sorted(set(re.findall(r"\b(punc_gpu_\w+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    import ctypes
    lib = ctypes.CDLL(N.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the ctypes binding covers them all
    assert set(syms) <= set(N._SIGS), set(syms) - set(N._SIGS)


def test_abi_version():
    assert N.lib().punc_gpu_abi_version() == 1


@pytest.mark.parametrize("n,r", [(4, 0.0), (6, 1.3), (8, 2.1)])
def test_meshgen_sizes(n, r):
    hm = P.box_mesh(n, [-3.0] * 3, [3.0] * 3, sphere_r=r)
    c = hm.xyz[hm.cells].mean(axis=1)
    assert np.all(np.linalg.norm(c, axis=1) >= r)
    removed = 6 * n ** 3 - hm.nc
    if r == 0:
        assert hm.nc == 6 * n ** 3 and hm.nv == (n + 1) ** 3
    else:
        assert removed > 0
    assert np.all(np.diff(hm.cells, axis=1) > 0)  # ascending vertex ids
    used = np.zeros(hm.nv, bool)
    used[hm.cells.ravel()] = True
    assert used.all()
    # Kuhn cells have volume h^3/6
    h = 6.0 / n
    v = hm.xyz[hm.cells]
    vol = np.abs(np.linalg.det(v[:, 1:] - v[:, :1])) / 6
    assert np.allclose(vol, h ** 3 / 6)


def test_meshgen_periodic_dofs():
    hm = P.box_mesh(4, [0, 0, 0], [2 * np.pi] * 3, periodic=True)
    assert hm.ndof == 64 and hm.nv == 125
    assert np.bincount(hm.vdof).min() >= 1
    # opposite faces share dofs
    i0 = np.where(np.isclose(hm.xyz[:, 0], 0) & np.isclose(hm.xyz[:, 1], 1.5707963267948966 * 1)
                  & np.isclose(hm.xyz[:, 2], 0))[0]
    i1 = np.where(np.isclose(hm.xyz[:, 0], 2 * np.pi) & np.isclose(hm.xyz[:, 1], np.pi / 2)
                  & np.isclose(hm.xyz[:, 2], 2 * np.pi))[0]
    assert hm.vdof[i0[0]] == hm.vdof[i1[0]]


@pytest.mark.parametrize("kind", ["sphere", "periodic", "box"])
def test_localizer_bit_identical_to_oracle(kind):
    if kind == "sphere":
        hm = P.box_mesh(7, [-3.0] * 3, [3.0] * 3, sphere_r=1.1)
    elif kind == "periodic":
        hm = P.box_mesh(5, [0.0] * 3, [2 * np.pi] * 3, periodic=True)
    else:
        hm = P.box_mesh(3, [-1.0, -2.0, -0.5], [2.0, 1.0, 3.5])
    m = P.Mesh(hm)
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    a, b = m.localizer(), o.localizer()
    for k in ("adj", "planes", "bary", "vcoords"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["dofs"], hm.vdof[hm.cells])
    assert np.array_equal(P.element_volume(m), o.element_volume())
    assert np.array_equal(P.element_volume(m, voronoi=False), o.element_volume(voronoi=False))
    ef, oe = m.exterior_facet_arrays(), o.exterior_facets()
    for k in ("area", "vertices", "normal", "basis", "cell"):
        assert np.array_equal(ef[k], oe[k]), k


def test_exterior_facets_cover_the_box():
    n, L = 5, 2.0
    hm = P.box_mesh(n, [-L] * 3, [L] * 3, sphere_r=0.9)
    m = P.Mesh(hm)
    ef = m.exterior_facet_arrays()
    assert m.num_exterior == 6 * 2 * n * n
    assert ef["area"].sum() == pytest.approx(6 * (2 * L) ** 2, rel=1e-13)
    # inward normals: a step along the normal from the centroid goes inside the box
    cen = ef["vertices"].reshape(-1, 3, 3).mean(axis=1)
    assert np.all(np.abs(cen + 1e-3 * ef["normal"]) < L)
    assert np.allclose(np.linalg.norm(ef["normal"], axis=1), 1.0)
    assert m.num_objects == 1


def test_locate_matches_oracle_brute_force():
    hm = P.box_mesh(5, [-2.0] * 3, [2.0] * 3, sphere_r=0.8)
    m = P.Mesh(hm)
    o = O.OraclePop(hm, [QE], [ME])
    rng = np.random.default_rng(8)
    xs = rng.uniform(-2.3, 2.3, size=(4000, 3))
    got = m.locate(xs)
    ref = np.array([o.locate(x, brute=True) for x in xs])
    assert np.array_equal(got, ref)
    assert (got < 0).any() and (got >= 0).any()


def test_locate_non_finite_is_outside():
    """NaN or infinite coordinates are outside every cell in both the
    library's point location and the oracle's (a NaN passes every plane test,
    and its bucket index would be out of range)."""
    hm = P.box_mesh(4, [-2.0] * 3, [2.0] * 3, sphere_r=0.8)
    m = P.Mesh(hm)
    o = O.OraclePop(hm, [QE], [ME])
    bad = np.array([[np.nan, 0.1, 0.1], [0.1, np.inf, 0.1], [0.1, 0.1, -np.inf], [np.nan] * 3])
    assert (m.locate(bad) == -1).all()
    for x in bad:
        assert o.locate(x) == -1 and o.locate(x, brute=True) == -1
    assert o.add_particles(0, bad, np.zeros_like(bad)) == 0


def test_facet_flux_form_matches_mesh_form():
    """create_flux over a facet list (the reference's form) gives the mesh
    form's closed-form tables bit for bit."""
    hm = P.box_mesh(3, [-1.0] * 3, [1.0] * 3, sphere_r=0.5)
    m = P.Mesh(hm)
    ef = m.exterior_facet_arrays()
    nf = m.num_exterior
    for vth, vd in [(1.0, [0, 0, 0.5]), (0.3, [0.2, -0.1, 0.0])]:
        vd = np.array(vd, dtype=np.float64)
        a_num, a_max, b_num, b_max = (np.empty(nf) for _ in range(4))
        N.check(N.lib().punc_gpu_vdf_flux(m._h, 0, vth, N.dptr(vd), 0.0, 0.0, 5.0, 0, 0, 500000,
                                          N.dptr(a_num), N.dptr(a_max)))
        nrm = np.ascontiguousarray(ef["normal"])
        area = np.ascontiguousarray(ef["area"])
        N.check(N.lib().punc_gpu_vdf_flux_facets(nf, N.dptr(nrm), N.dptr(area), 0, vth, N.dptr(vd), 0.0, 0.0,
                                                 5.0, 0, 0, 500000, N.dptr(b_num), N.dptr(b_max)))
        assert np.array_equal(a_num, b_num) and np.array_equal(a_max, b_max)


def test_maxwellian_flux_matches_oracle():
    hm = P.box_mesh(3, [-1.0] * 3, [1.0] * 3)
    m = P.Mesh(hm)
    for vth, vd in [(1.0, [0, 0, 0.5]), (0.023, [0.1, 0, 0.5]), (2.0, [0, 0, 0])]:
        sp = P.Species(-1.0, 1.0, 1.0, 1.0, P.ParticleAmountType.phys_per_sim, m,
                       P.UniformPosition(m), P.Maxwellian(vth, vd))
        P.create_flux([sp], m)
        ef = m.exterior_facet_arrays()
        for j in range(0, m.num_exterior, 5):
            num, pmax = O.flux(vth, vd, ef["normal"][j], ef["area"][j])
            assert sp.vdf.num_particles[j] == num
            assert sp.vdf.pdf_max[j] == pmax


def test_species_amount_types():
    hm = P.box_mesh(2, [0.0] * 3, [2.0] * 3)
    m = P.Mesh(hm)
    vdf = P.Maxwellian(1.0, [0, 0, 0])
    # population.cpp:46-67: per_cell -> in_total -> per_volume -> weight
    s = P.Species(-1.0, 1.0, 100.0, 4, P.ParticleAmountType.per_cell, m, P.UniformPosition(m), vdf)
    per_volume = 4 * m.num_cells / m.volume()
    assert s.weight == pytest.approx(100.0 / per_volume)
    assert s.q == pytest.approx(-1.0 * s.weight) and s.m == pytest.approx(s.weight)
    assert s.num == int(round(4 * m.num_cells))
    s2 = P.Species(1.0, 1836.0, 7.0, 1.0, P.ParticleAmountType.phys_per_sim, m, P.UniformPosition(m), vdf)
    assert s2.q == 1.0 and s2.m == 1836.0 and s2.n == 7.0


def test_errors_are_reported_not_raised_across_abi():
    L = N.lib()
    h = N.vp()
    lo = np.zeros(3)
    hi = np.ones(3)
    assert L.punc_gpu_meshgen_box(0, N.dptr(lo), N.dptr(hi), 0.0, 0, N.C.byref(h)) != 0
    assert b"n must be" in L.punc_gpu_last_error()
    # non-ascending cell
    xyz = np.eye(4, 3).copy()
    bad = np.array([[0, 2, 1, 3]], dtype=np.int32)
    fb = np.ones((1, 4), dtype=np.int32)
    assert L.punc_gpu_mesh_create(4, N.dptr(xyz), 1, N.i32ptr(bad), N.i32ptr(fb), None, 0,
                                  N.C.byref(h)) != 0
    assert b"ascending" in L.punc_gpu_last_error()
    # boundary facet without an id
    good = np.array([[0, 1, 2, 3]], dtype=np.int32)
    zero = np.zeros((1, 4), dtype=np.int32)
    assert L.punc_gpu_mesh_create(4, N.dptr(xyz), 1, N.i32ptr(good), N.i32ptr(zero), None, 0,
                                  N.C.byref(h)) != 0
    assert b"boundary" in L.punc_gpu_last_error()
    with pytest.raises(N.PuncGpuError):
        P.HostMesh(xyz, bad, fb) and P.Mesh(P.HostMesh(xyz, bad, fb))


def test_population_without_gpu_fails_loudly():
    from conftest import has_gpu
    if has_gpu():
        pytest.skip("a GPU is visible")
    hm = P.box_mesh(2, [0.0] * 3, [1.0] * 3)
    m = P.Mesh(hm)
    with pytest.raises(N.PuncGpuError):
        P.Population(m, [(QE, ME)], capacity=10)

"""GPU parity of the particle diagnostics against the oracle: potential
energy, CG1 and DG0 number densities, speed statistics and the exponential
moving averages, on identical populations (electrons, ions and a neutral
species, which the CG1 density counts with the ions and the DG0 density with
the electrons).

Tolerances: DG0 densities and the averages are bit-exact (integer counts,
weights summed one particle at a time, the same volume formula, IEEE
division); potential energy and CG1 densities are sums in another order,
within 1e-11 of the sum of magnitudes; the statistics merge per-thread
running moments (Chan et al.) instead of one serial pass, within 1e-11
relative."""
import math

import numpy as np
import pytest

import oracle as O
import puncpp_amd as P
from helpers import MI, ME, QE, QI, sphere_mesh

pytestmark = pytest.mark.gpu

SPECIES = [(QE, ME), (QI, MI), (0.0, 1.0)]


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


class _W:
    def __init__(self, w):
        self.weight = w


def _pair(hm, n=20000, seed=4):
    o = O.OraclePop(hm, [q for q, _ in SPECIES], [m for _, m in SPECIES])
    for s, vth in enumerate((1.0, 0.05, 0.3)):
        assert o.load(s, n, seed, 0.0, 0.0, 0, vth, [0.0, 0.0, 0.5]) == 0
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, SPECIES, capacity=4 * n + 1000)
    st = o.get_particles()
    pop.add_located(st["x"], st["v"], st["species"], st["cell"])
    return o, mesh, pop, st


@pytest.fixture(scope="module")
def pair():
    return _pair(sphere_mesh(8))


def test_potential_energy(gpu, pair):
    o, mesh, pop, st = pair
    hm = mesh.host
    rng = np.random.default_rng(1)
    q = np.array([s[0] for s in SPECIES])[st["species"]]
    for k, phi in enumerate([hm.xyz @ np.array([0.3, -1.1, 0.6]) + 0.2, rng.normal(size=hm.ndof)]):
        _print_case("potential_energy", phi=k)
        pe = P.particle_potential_energy_cg1(pop, phi)
        pe_o = o.potential_energy(phi)
        scale = 0.5 * np.sum(np.abs(q)) * np.abs(phi).max()
        assert abs(pe - pe_o) <= 1e-11 * scale, (pe, pe_o)
    with pytest.raises(ValueError):
        P.particle_potential_energy_cg1(pop, np.zeros(hm.ndof + 1))


def test_density_cg1(gpu, pair):
    o, mesh, pop, st = pair
    dv = P.element_volume(mesh)
    we, wi = 2.5, 0.7
    _print_case("density_cg1", weights=(we, wi))
    ne, ni = P.density_cg1(None, pop, [_W(we), _W(wi)], dv)
    ne_o, ni_o = o.density_cg1(dv, we, wi)
    # per dof at most (count of particles around it) terms of magnitude <= w dv_inv
    cnt = np.bincount(st["cell"], minlength=mesh.num_cells).max()
    for got, ref, w in ((ne, ne_o, we), (ni, ni_o, wi)):
        tol = 1e-11 * w * dv * 24 * cnt
        assert np.all(np.abs(got - ref) <= tol)
    # with dv_inv omitted the uploaded one is kept
    ne2, _ = P.density_cg1(None, pop, [_W(we), _W(wi)])
    assert np.all(np.abs(ne2 - ne_o) <= 1e-11 * we * dv * 24 * cnt)


def test_density_dg0_bit_exact(gpu, pair):
    o, mesh, pop, _ = pair
    for we, wi in ((1.0, 1.0), (0.1, 3.7)):
        _print_case("density_dg0", weights=(we, wi))
        ne, ni = P.density_dg0(None, pop, [_W(we), _W(wi)])
        ne_o, ni_o = o.density_dg0(we, wi)
        assert np.array_equal(ne, ne_o) and np.array_equal(ni, ni_o)


def test_statistics(gpu, pair):
    o, mesh, pop, _ = pair
    _print_case("statistics")
    got = pop.statistics()
    ref = o.statistics(np.zeros(4))
    np.testing.assert_allclose(got, ref, rtol=1e-11)


def test_statistics_edge_cases(gpu):
    hm = sphere_mesh(4)
    mesh = P.Mesh(hm)
    _print_case("statistics_edge", particles=1)
    pop = P.Population(mesh, SPECIES, capacity=1000)
    # no particles: every entry kept
    assert np.array_equal(pop.statistics([-7.0] * 4), [-7.0] * 4)
    # one electron: mean = speed, 0/0 standard deviation (the reference's
    # formula); the ion class keeps its entries; neutrals count in neither
    x = np.array([[2.5, 2.5, 2.5], [2.6, 2.5, 2.5]])
    v = np.array([[3.0, 4.0, 0.0], [1.0, 1.0, 1.0]])
    pop.add_particles(x[:1], v[:1], QE, ME)
    pop.add_particles(x[1:], v[1:], 0.0, 1.0)
    s = pop.statistics([-7.0] * 4)
    assert s[0] == 5.0 and np.isnan(s[1]) and s[2] == -7.0 and s[3] == -7.0


def test_diagnostics_after_sort(gpu, pair):
    # a pending cell order is applied before the diagnostics read the particles
    o, mesh, pop, _ = pair
    _print_case("diagnostics_after_sort")
    ne0, ni0 = P.density_dg0(None, pop, [_W(1.0), _W(1.0)])
    s0 = pop.statistics()
    pop.sort()
    ne1, ni1 = P.density_dg0(None, pop, [_W(1.0), _W(1.0)])
    assert np.array_equal(ne0, ne1) and np.array_equal(ni0, ni1)
    np.testing.assert_allclose(pop.statistics(), s0, rtol=1e-11)
    assert np.all(np.diff(pop.get_particles()["cell"]) >= 0)


def test_ema(gpu, pair):
    o, mesh, pop, _ = pair
    _print_case("ema")
    dv = P.element_volume(mesh)
    rho = P.distribute_cg1(pop, dv)
    dt, tau = 0.05, 0.4
    w = 1.0 - math.exp(-dt / tau)  # the C library's exp, as the host code calls
    g = P.rho_ema(pop, dt, tau, reset=True)
    ref = w * rho + (1.0 - w) * np.zeros_like(rho)
    assert np.array_equal(g, ref)
    g2 = P.rho_ema(pop, dt, tau)
    assert np.array_equal(g2, w * rho + (1.0 - w) * ref)
    # densities: CG1, then a DG0 density (another length) needs a reset
    ne, ni = P.density_cg1(None, pop, [_W(1.0), _W(1.0)], dv)
    a, b = P.density_ema(pop, dt, tau, reset=True)
    assert np.array_equal(a, w * ne) and np.array_equal(b, w * ni)
    a2, b2 = P.density_ema(pop, dt, tau)
    assert np.array_equal(a2, w * ne + (1.0 - w) * a) and np.array_equal(b2, w * ni + (1.0 - w) * b)
    P.density_dg0(None, pop, [_W(1.0), _W(1.0)])
    with pytest.raises(P._native.PuncGpuError):
        P.density_ema(pop, dt, tau)
    ne0, _ = P.density_dg0(None, pop, [_W(1.0), _W(1.0)])
    a3, _ = P.density_ema(pop, dt, tau, reset=True)
    assert np.array_equal(a3, w * ne0)


def test_large_density_properties(gpu):
    # 64^3 x 6 cells, 16M particles: no oracle at this size; the moments
    # that hold whatever the size (ions of charge 2 so that sum q != 0)
    hm = P.box_mesh(64, [-20.0] * 3, [20.0] * 3, sphere_r=2.0)
    mesh = P.Mesh(hm)
    nps = 8 * 1024 * 1024
    _print_case("large_density", particles=2 * nps)
    pop = P.Population(mesh, [(QE, ME), (2.0, MI)], capacity=2 * nps + 1000)
    L = P._native.lib()
    for s, vth in enumerate((1.0, 0.05)):
        P._native.check(L.punc_gpu_pop_load(pop._h, s, nps, 0, 0.0, 0.0, 0, vth,
                                            P._native.dptr(np.array([0.0, 0.0, 0.5]))))
    pop.sort()
    dv = P.element_volume(mesh)
    ne, ni = P.density_cg1(None, pop, [_W(1.0), _W(2.0)], dv)
    assert np.sum(ne / dv) == pytest.approx(nps, rel=1e-12)
    assert np.sum(ni / dv) == pytest.approx(2.0 * nps, rel=1e-12)
    ne0, ni0 = P.density_dg0(None, pop, [_W(1.0), _W(2.0)])
    vc = hm.xyz[hm.cells]
    vol = np.abs(np.linalg.det(vc[:, 1:] - vc[:, :1])) / 6.0
    assert np.sum(ne0 * vol) == pytest.approx(nps, rel=1e-12)
    assert np.sum(ni0 * vol) == pytest.approx(2.0 * nps, rel=1e-12)
    pe = P.particle_potential_energy_cg1(pop, np.ones(hm.ndof))
    assert pe == pytest.approx(0.5 * (QE + 2.0) * nps, rel=1e-12)
    s = pop.statistics()
    # drifting Maxwellian speeds, a = |drift| / vth:
    # E|v| = vth (sqrt(2/pi) exp(-a^2/2) + (a + 1/a) erf(a/sqrt(2)))
    assert s[0] == pytest.approx(1.66144, rel=1e-3) and s[2] == pytest.approx(0.505, rel=1e-3)


@pytest.mark.parametrize("graph", [False, True])
def test_fused_step_potential_energy(gpu, graph):
    # the potential energy computed inside the fused step equals the oracle's
    # particle_potential_energy_cg1 at the positions the step starts from,
    # and the step's particles are unchanged by computing it
    from helpers import assert_same_sets, loaded_pair, synthetic_field
    from puncpp_amd import _native as N
    _print_case("fused_step_pe", graph=graph)
    hm = sphere_mesh(n=6, L=3.0, r=1.0)
    o, mesh, pop = loaded_pair(hm, (4000, 4000))
    E = synthetic_field(hm, scale=0.2)
    phi = hm.xyz @ np.array([0.2, -0.5, 0.3]) + np.random.default_rng(6).normal(size=hm.ndof) * 0.1
    dv = P.element_volume(mesh)
    B = np.array([0.0, 0.0, 1.0])
    L = N.lib()
    N.check(L.punc_gpu_set_field(pop._h, N.dptr(E)))
    N.check(L.punc_gpu_distribute(pop._h, N.dptr(dv), None))
    pop.set_potential(phi)
    dt = 0.05
    q = np.array([QE, QI])
    for n in range(3):
        pe_o = o.potential_energy(phi)
        if graph:
            if n == 0:
                N.check(L.punc_gpu_step_graph(pop._h, 2, dt, N.dptr(B), dt, None, 0, 0, 0))
            N.check(L.punc_gpu_graph_launch(pop._h, 1))
        else:
            N.check(L.punc_gpu_step(pop._h, 2, dt, N.dptr(B), dt, None, 0, 0, n))
        pe = pop.step_potential_energy()
        st = o.get_particles()
        scale = 0.5 * np.sum(np.abs(q[st["species"]])) * np.abs(phi).max()
        assert abs(pe - pe_o) <= 1e-11 * scale, (pe, pe_o)
        o.boris(E, B, dt)
        o.move(dt)
        o.update()
        assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"pe step {n}")
    pop.set_potential(None)
    N.check(L.punc_gpu_step(pop._h, 2, dt, N.dptr(B), dt, None, 0, 0, 9))
    assert np.isnan(pop.step_potential_energy())

"""GPU parity of the non-Maxwellian velocity distributions (Kappa, Cairns,
KappaCairns) against the oracle: create_flux tables (closed forms and the
device Monte Carlo), boundary injection and initial loading.

Tolerances: closed-form tables and Monte Carlo maxima equal to 1e-14
relative (the same formulas; pow and exp may differ in the last bit between
the device and the C library); Monte Carlo flux numbers within 1e-12
relative (block-reduction order); injected and loaded particle sets
bit-identical (the samplers accept uniform draws, the pdf only decides)."""
import numpy as np
import pytest

import oracle as O
import puncpp_amd as P
from puncpp_amd import _native as N
from helpers import MI, ME, QE, QI, assert_same_sets, loaded_pair, sphere_mesh

pytestmark = pytest.mark.gpu

# (kind, vth, vd, kappa, alpha, range)
CASES = {
    "kappa": (1, 1.0, [0.0, 0.0, 0.5], 4.0, 0.0, 7.0),       # integer drift sum 0: closed-form flux number
    "kappa_drift": (1, 1.0, [0.0, 0.0, 1.5], 3.0, 0.0, 7.0),  # Monte Carlo flux number
    "cairns": (2, 1.0, [0.0, 0.0, 0.5], 0.0, 0.4, 7.0),      # Monte Carlo maximum
    "kappa_cairns": (3, 1.0, [0.2, 0.0, 0.5], 4.0, 0.3, 15.0),  # both by Monte Carlo
}


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _gpu_flux(mesh, case, seed, rng_sp, n_iter):
    kind, vth, vd, k, a, rr = case
    ne = mesh.num_exterior
    num, pmax = np.empty(ne), np.empty(ne)
    N.check(N.lib().punc_gpu_vdf_flux(mesh._h, kind, vth, N.dptr(np.array(vd, dtype=np.float64)), k, a, rr,
                                      seed, rng_sp, n_iter, N.dptr(num), N.dptr(pmax)))
    return num, pmax


@pytest.mark.parametrize("name", sorted(CASES))
def test_vdf_flux_tables(gpu, name):
    case = CASES[name]
    _print_case("vdf_flux", vdf=name)
    hm = sphere_mesh(n=4, L=2.0, r=0.6)
    mesh = P.Mesh(hm)
    o = O.OraclePop(hm, [QE], [ME])
    kind, vth, vd, k, a, rr = case
    num, pmax = _gpu_flux(mesh, case, 5, 1, 30000)
    num_o, pmax_o = o.vdf_flux(kind, vth, vd, k, a, rr, seed=5, rng_sp=1, n_iter=30000)
    _, fn, fm = O.vdf_max(kind, vth, vd, k, a)
    np.testing.assert_allclose(num, num_o, rtol=1e-14 if fn else 1e-12)
    np.testing.assert_allclose(pmax, pmax_o, rtol=1e-14)
    assert np.all(num > 0) and np.all(pmax > 0)


def _vdf_species(mesh, name, q, m, density):
    kind, vth, vd, k, a, rr = CASES[name]
    cls = {1: lambda: P.Kappa(vth, vd, k, rr), 2: lambda: P.Cairns(vth, vd, a, rr),
           3: lambda: P.KappaCairns(vth, vd, k, a, rr)}[kind]
    return P.Species(q, m, density, 1.0, P.ParticleAmountType.phys_per_sim, mesh, P.UniformPosition(mesh), cls())


@pytest.mark.parametrize("names", [("kappa", "cairns"), ("kappa_drift", "kappa_cairns")])
def test_vdf_inject(gpu, names):
    _print_case("vdf_inject", vdfs=names)
    hm = sphere_mesh(n=6, L=3.0, r=1.0)
    o, mesh, pop = loaded_pair(hm, (10, 10))
    species = [_vdf_species(mesh, names[0], QE, ME, 40.0), _vdf_species(mesh, names[1], QI, MI, 40.0)]
    P.create_flux(species, mesh, seed=2, n_iter=20000)
    for s, sp in enumerate(species):
        v = sp.vdf
        o.set_injector_vdf(s, v.kind, v.vth(), v.vd(), v.k, v.alpha, sp.n, v.vdf_range, v.num_particles,
                           v.pdf_max)
    dt = 0.1
    total = 0
    for step in range(3):
        k_g = P.inject_particles(pop, species, None, dt, seed=7, step=step)
        k_o = o.inject(dt, 7, step)
        assert k_g == k_o, (k_g, k_o)
        total += k_g
        assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"vdf inject {step}")
    assert total > 0


@pytest.mark.parametrize("name", sorted(CASES))
def test_vdf_load(gpu, name):
    _print_case("vdf_load", vdf=name)
    kind, vth, vd, k, a, rr = CASES[name]
    hm = sphere_mesh()
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=20000)
    for s in range(2):
        assert o.load_vdf(s, 2500, 11, 0.0, 0.0, 0, kind, vth, vd, k, a, rr) == 0
        N.check(N.lib().punc_gpu_pop_load_vdf(pop._h, s, 2500, 11, 0.0, 0.0, 0, kind, vth,
                                              N.dptr(np.array(vd, dtype=np.float64)), k, a, rr))
    assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"vdf load {name}")


def test_vdf_flux_large_mesh(gpu):
    # the large workload's boundary (64^3 cubes: 49152 exterior facets), the
    # reference's 500000 samples per facet, both KappaCairns tables by Monte
    # Carlo: facets sharing a normal estimate the same integral
    hm = P.box_mesh(64, [-20.0] * 3, [20.0] * 3, sphere_r=2.0)
    mesh = P.Mesh(hm)
    _print_case("vdf_flux_large", facets=mesh.num_exterior)
    num, pmax = _gpu_flux(mesh, CASES["kappa_cairns"], 0, 0, 500000)
    f = mesh.exterior_facet_arrays()
    dens = num / f["area"]
    for axis in range(3):
        for sgn in (1.0, -1.0):
            sel = np.isclose(f["normal"][:, axis], sgn)
            assert sel.sum() == 2 * 64 * 64
            d = dens[sel]
            assert d.std() / d.mean() < 0.05
            assert pmax[sel].std() / pmax[sel].mean() < 0.05

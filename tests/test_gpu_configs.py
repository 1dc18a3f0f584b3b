"""The workload configurations at their own sizes.

* Periodic box (16^3 cubes x 6 tets, 32 particles per cell and species, the
  1 + 0.01 cos(x) density, accel + move_periodic, half push at step 0): the
  fused step against the oracle's reference-ordered sequence at full size,
  with a re-sort in between - particle sets bit-identical, KE and rho within
  the tolerances.
* Sphere probe (32^3 cubes x 6 tets minus the r = 1 sphere, Boris with
  B = (0, 0, 1), drifting Maxwellian flux injection on both species): the
  same comparison with absorption and injection, at 300k particles.
* Large box (64^3 cubes x 6 tets minus the r = 2 sphere, 64M particles, 40
  graph-replayed steps with a re-sort every 20): no oracle at this size; the
  properties that hold whatever the size - the
  count balance n' = n - absorbed + injected every step, charge conservation
  of the deposit, every particle inside its cell (checked with the oracle's
  point location on a sample), non-decreasing cells right after a re-sort.
"""
import numpy as np
import pytest

import oracle as O
import puncpp_amd as P
from puncpp_amd import _native as N
from helpers import MI, ME, QE, QI, assert_same_sets, rho_scale, synthetic_field

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _results(pop):
    ke = N.C.c_double()
    ab = np.zeros((2, 64), dtype=np.int64)
    kin = N.C.c_int64()
    N.check(N.lib().punc_gpu_step_results(pop._h, N.C.byref(ke), N.i64ptr(ab), N.C.byref(kin)))
    return ke.value, ab, kin.value


def _rho(pop, ndof):
    rho = np.empty(ndof)
    N.check(N.lib().punc_gpu_get_rho(pop._h, N.dptr(rho)))
    return rho


def test_periodic_config_full_size(gpu):
    _print_case("periodic_config", particles=2 * 32 * 24576)
    hm = P.box_mesh(16, [0.0] * 3, [2 * np.pi] * 3, periodic=True)
    assert hm.cells.shape[0] == 24576
    nps = 32 * 24576
    vth = (1.0, 1.0 / np.sqrt(1836.0))
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    for s in range(2):
        assert o.load(s, nps, 0, 0.01, 1.0, 0, vth[s], [0.0, 0.0, 0.0]) == 0
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=2 * nps + 1000)
    st = o.get_particles()
    pop.add_located(st["x"], st["v"], st["species"], st["cell"])
    E = synthetic_field(hm, scale=0.5)
    dv = P.element_volume(mesh)
    Ld = np.array([2 * np.pi] * 3)
    L = N.lib()
    N.check(L.punc_gpu_set_field(pop._h, N.dptr(E)))
    N.check(L.punc_gpu_distribute(pop._h, N.dptr(dv), None))
    dt = 0.05
    for n in range(4):
        if n == 2:
            pop.sort()  # the next fused step reads through the new order
        push_dt = (1.0 - 0.5 * (n == 0)) * dt
        N.check(L.punc_gpu_step(pop._h, 1, push_dt, None, dt, N.dptr(Ld), 0, 0, n))
        ke, ab, _ = _results(pop)
        ke_o = o.accel(E, push_dt)
        o.move_periodic(dt, Ld)
        ab_o, _, lost = o.update()
        assert lost == 0 and ab.sum() == 0 and ab_o.sum() == 0
        assert abs(ke - ke_o) <= 1e-11 * abs(ke_o)
        assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"periodic step {n}")
        rho_o = o.distribute(dv)
        assert np.all(np.abs(_rho(pop, hm.ndof) - rho_o) <= 1e-11 * rho_scale(o, hm, dv))


def test_probe_config_with_injection(gpu):
    _print_case("probe_config", particles=300000)
    hm = P.box_mesh(32, [-10.0] * 3, [10.0] * 3, sphere_r=1.0)
    nps = 150000
    vth = (1.0, 1.0 / np.sqrt(1836.0))
    vd = [0.0, 0.0, 0.5]
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    for s in range(2):
        assert o.load(s, nps, 3, 0.0, 0.0, 0, vth[s], vd) == 0
    mesh = P.Mesh(hm)
    density = nps / mesh.volume()
    species = [P.Species(q, m, density, 1.0, P.ParticleAmountType.phys_per_sim, mesh,
                         P.UniformPosition(mesh), P.Maxwellian(vth[s], vd))
               for s, (q, m) in enumerate([(QE, ME), (QI, MI)])]
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=2 * nps + 200000)
    st = o.get_particles()
    pop.add_located(st["x"], st["v"], st["species"], st["cell"])
    P.punc._register_injectors(pop, species)
    for s in range(2):
        o.set_injector(s, vth[s], vd, species[s].n, 5.0)
    E = synthetic_field(hm, scale=0.05)
    B = np.array([0.0, 0.0, 1.0])
    dv = P.element_volume(mesh)
    L = N.lib()
    N.check(L.punc_gpu_set_field(pop._h, N.dptr(E)))
    N.check(L.punc_gpu_distribute(pop._h, N.dptr(dv), None))
    dt = 0.02
    total_abs = 0
    for n in range(5):
        if n == 3:
            pop.sort()
        push_dt = (1.0 - 0.5 * (n == 0)) * dt
        n_before = pop.size()
        N.check(L.punc_gpu_step(pop._h, 2, push_dt, N.dptr(B), dt, None, 1, 7, n))
        ke, ab, kin = _results(pop)
        ke_o = o.boris(E, B, push_dt)
        o.move(dt)
        ab_o, _, lost = o.update()
        k_o = o.inject(dt, 7, n)
        assert lost == 0
        np.testing.assert_array_equal(ab, ab_o)
        assert kin == k_o and kin > 0
        assert pop.size() == n_before - ab.sum() + kin
        assert abs(ke - ke_o) <= 1e-11 * abs(ke_o)
        assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"probe step {n}")
        rho_o = o.distribute(dv)
        assert np.all(np.abs(_rho(pop, hm.ndof) - rho_o) <= 1e-11 * rho_scale(o, hm, dv))
        total_abs += ab.sum()
    assert total_abs > 0  # the exterior boundary and the probe both absorb


def test_large_config_properties(gpu):
    _print_case("large_config", particles=64 * 1024 * 1024)
    hm = P.box_mesh(64, [-20.0] * 3, [20.0] * 3, sphere_r=2.0)
    mesh = P.Mesh(hm)
    nps = 32 * 1024 * 1024
    vth = [1.0, 1.0 / np.sqrt(1836.0)]
    vd = [0.0, 0.0, 0.5]
    qm = [(QE, ME), (QI, MI)]
    density = nps / mesh.volume()
    species = [P.Species(q, m, density, 1.0, P.ParticleAmountType.phys_per_sim, mesh,
                         P.UniformPosition(mesh), P.Maxwellian(vth[s], vd))
               for s, (q, m) in enumerate(qm)]
    pop = P.Population(mesh, qm, capacity=2 * nps + 2_000_000)
    L = N.lib()
    for s in range(2):
        N.check(L.punc_gpu_pop_load(pop._h, s, nps, 0, 0.0, 0.0, 0, vth[s], N.dptr(np.array(vd))))
    pop.sort()
    P.punc._register_injectors(pop, species)
    E = synthetic_field(hm, scale=0.02)
    N.check(L.punc_gpu_set_field(pop._h, N.dptr(E)))
    dv = P.element_volume(mesh)
    N.check(L.punc_gpu_distribute(pop._h, N.dptr(dv), None))
    B = np.array([0.0, 0.0, 1.0])
    dt = 0.02
    N.check(L.punc_gpu_step_graph(pop._h, 2, dt, N.dptr(B), dt, None, 1, 0, 0))
    for n in range(40):  # the configuration's re-sort interval: every 20 steps
        if n % 20 == 19:
            pop.sort()
        n_before = pop.size()
        N.check(L.punc_gpu_graph_launch(pop._h, 1))
        _, ab, kin = _results(pop)
        assert pop.size() == n_before - ab.sum() + kin, f"count balance at step {n}"
        assert kin > 0 and ab.sum() > 0
    # charge conservation of the deposit: sum_j rho_j / dv_inv_j = sum_p q_p
    c = pop.counts()
    tot = np.sum(_rho(pop, hm.ndof) / dv)
    assert abs(tot - (c[0] * QE + c[1] * QI)) <= 1e-9 * (c[0] + c[1])
    # a re-sort leaves the cells non-decreasing; every particle is inside its
    # cell (the oracle's own point location on a sample; points within 1e-12
    # of a face may legitimately belong to the neighbour)
    pop.sort()
    got = pop.get_particles()
    assert np.all(np.diff(got["cell"]) >= 0)
    assert np.all(np.isfinite(got["x"])) and np.all(np.isfinite(got["v"]))
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    loc = o.localizer()
    bary = loc["bary"].reshape(-1, 3, 4)
    sample = np.random.default_rng(5).choice(got["x"].shape[0], 20000, replace=False)
    mism = 0
    for k in sample:
        x, c0 = got["x"][k], got["cell"][k]
        if o.locate(x) != c0:
            A = bary[c0]
            lam = A[:, 0] + A[:, 1:] @ x
            lam = np.append(lam, 1.0 - lam.sum())
            assert lam.min() > -1e-12, f"particle {k} outside its cell {c0}: {lam}"
            mism += 1
    assert mism < 5

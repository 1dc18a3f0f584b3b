"""Longer runs of the workload configurations with the field solved on the
device every step, as the reference's time loop does (interaction.cpp:
342-423: distribute, solve, push, move, update, inject).

* Periodic box at full size (16^3 cubes x 6 tets, 32 particles per cell and
  species, density 1 + 0.01 cos(x), accel + move_periodic, dt = 0.05): 100
  self-consistent steps, the Poisson problem (periodic, zero-mean potential)
  and the E projection solved on the device from the device deposit every
  step.  The oracle follows the first 20 steps with the device's E as its
  input, so the particle path is compared bit for bit on the self-consistent
  trajectory; all 100 steps keep N, the deposit's total charge and finite
  states, and every particle ends inside its cell.
* Sphere probe at its own size (32^3 cubes x 6 tets minus the r = 1 sphere,
  2M particles per species, Boris with B = (0, 0, 1), drifting Maxwellian
  flux injection): 3 steps against the oracle, absorbed and injected counts
  exact, particle sets bit-identical, rho within 1e-11 of its scale.
* Sphere probe with the floating object coupled in through the capacitance-
  matrix circuit (CircuitCM, no sources: the object keeps the charge it
  collects): 30 steps at 1M particles, the object's charge the running sum of
  the charge absorbed through its boundary, the circuit's potential and field
  from the device solve each step, the oracle following the first 3 steps
  with the device's E.
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


def _inside_sample(o, hm, got, k=20000, seed=5):
    """Every sampled particle inside its cell: the oracle's own location, or
    (within 1e-12 of a face) non-negative barycentric coordinates."""
    bary = o.localizer()["bary"].reshape(-1, 3, 4)
    idx = np.random.default_rng(seed).choice(got["x"].shape[0], min(k, got["x"].shape[0]), replace=False)
    mism = 0
    for j in idx:
        x, c0 = got["x"][j], got["cell"][j]
        if o.locate(x) != c0:
            A = bary[c0]
            lam = A[:, 0] + A[:, 1:] @ x
            lam = np.append(lam, 1.0 - lam.sum())
            assert lam.min() > -1e-12, f"particle {j} outside its cell {c0}: {lam}"
            mism += 1
    assert mism < 5


def test_periodic_config_100_steps_self_consistent(gpu):
    nps = 32 * 24576
    _print_case("periodic_100", particles=2 * nps)
    hm = P.box_mesh(16, [0.0] * 3, [2 * np.pi] * 3, periodic=True)
    vth = (1.0, 1.0 / np.sqrt(1836.0))
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    for s in range(2):
        assert o.load(s, nps, 0, 0.01, 1.0, 0, vth[s], [0.0, 0.0, 0.0]) == 0
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=2 * nps + 1000)
    st = o.get_particles()
    pop.add_located(st["x"], st["v"], st["species"], st["cell"])
    dv = P.element_volume(mesh)
    field = P.FieldSolver(mesh, eps0=1.0, exterior=False, remove_null_space=True)
    Ld = np.array([2 * np.pi] * 3)
    L = N.lib()
    dt = 0.05
    N.check(L.punc_gpu_distribute(pop._h, N.dptr(dv), None))
    follow, check_sets = 20, (0, 4, 9, 19)
    for n in range(100):
        if n % 20 == 19:
            pop.sort()  # the next fused step reads through the new order
        res = field.solve_population(pop, copy_out=n < follow)
        E = res[1] if res is not None else None
        assert max(field.iterations) < 1000
        push_dt = (1.0 - 0.5 * (n == 0)) * dt
        N.check(L.punc_gpu_step(pop._h, 1, push_dt, None, dt, N.dptr(Ld), 0, 0, n))
        ke, ab, _ = _results(pop)
        assert ab.sum() == 0 and pop.size() == 2 * nps and np.isfinite(ke)
        if n < follow:
            ke_o = o.accel(E, push_dt)
            o.move_periodic(dt, Ld)
            ab_o, _, lost = o.update()
            assert lost == 0 and ab_o.sum() == 0
            assert abs(ke - ke_o) <= 1e-11 * abs(ke_o)
            if n in check_sets:
                assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"periodic step {n}")
                rho_o = o.distribute(dv)
                assert np.all(np.abs(_rho(pop, hm.ndof) - rho_o) <= 1e-11 * rho_scale(o, hm, dv))
        if n % 10 == 9:
            # the deposit holds the particles' charge: sum_j rho_j / dv_inv_j = sum_p q_p = 0
            assert abs(np.sum(_rho(pop, hm.ndof) / dv)) <= 1e-9 * 2 * nps
    got = pop.get_particles()
    assert np.all(np.isfinite(got["x"])) and np.all(np.isfinite(got["v"]))
    assert np.all((got["x"] >= 0.0) & (got["x"] <= 2 * np.pi))  # x - L floor(x / L) may round to L
    assert np.bincount(got["species"], minlength=2).tolist() == [nps, nps]
    _inside_sample(o, hm, got)


def test_probe_config_4M_particles(gpu):
    nps = 2 * 1024 * 1024
    _print_case("probe_4M", particles=2 * nps)
    hm = P.box_mesh(32, [-10.0] * 3, [10.0] * 3, sphere_r=1.0)
    vth = (1.0, 1.0 / np.sqrt(1836.0))
    vd = [0.0, 0.0, 0.5]
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    for s in range(2):
        assert o.load(s, nps, 4, 0.0, 0.0, 0, vth[s], vd) == 0
    mesh = P.Mesh(hm)
    density = nps / mesh.volume()
    species = [P.Species(q, m, density, 1.0, P.ParticleAmountType.phys_per_sim, mesh,
                         P.UniformPosition(mesh), P.Maxwellian(vth[s], vd))
               for s, (q, m) in enumerate([(QE, ME), (QI, MI)])]
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=2 * nps + 400000)
    st = o.get_particles()
    pop.add_located(st["x"], st["v"], st["species"], st["cell"])
    pop.sort()
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
    for n in range(3):
        push_dt = (1.0 - 0.5 * (n == 0)) * dt
        n_before = pop.size()
        N.check(L.punc_gpu_step(pop._h, 2, push_dt, N.dptr(B), dt, None, 1, 11, n))
        ke, ab, kin = _results(pop)
        ke_o = o.boris(E, B, push_dt)
        o.move(dt)
        ab_o, _, lost = o.update()
        k_o = o.inject(dt, 11, n)
        assert lost == 0
        np.testing.assert_array_equal(ab, ab_o)
        assert kin == k_o and kin > 0 and ab[:, 2].sum() > 0  # injection, and the probe absorbs
        assert pop.size() == n_before - ab.sum() + kin
        assert abs(ke - ke_o) <= 1e-11 * abs(ke_o)
        assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"probe 4M step {n}")
        rho_o = o.distribute(dv)
        assert np.all(np.abs(_rho(pop, hm.ndof) - rho_o) <= 1e-11 * rho_scale(o, hm, dv))


def test_probe_config_floating_object_circuit(gpu):
    nps = 512 * 1024
    _print_case("probe_circuit", particles=2 * nps)
    hm = P.box_mesh(32, [-10.0] * 3, [10.0] * 3, sphere_r=1.0)
    vth = (1.0, 1.0 / np.sqrt(1836.0))
    vd = [0.0, 0.0, 0.5]
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    for s in range(2):
        assert o.load(s, nps, 6, 0.0, 0.0, 0, vth[s], vd) == 0
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
    dv = P.element_volume(mesh)
    dt = 0.02
    field = P.FieldSolver(mesh, eps0=1.0, exterior=True, n_objects=1)
    inv_c = field.set_circuit(dt)  # a floating object: no sources
    assert inv_c.shape == (1, 1) and inv_c[0, 0] > 0
    obj = [P.Object(2)]
    B = np.array([0.0, 0.0, 1.0])
    L = N.lib()
    N.check(L.punc_gpu_distribute(pop._h, N.dptr(dv), None))
    collected = 0.0
    follow = 3
    for n in range(30):
        res = field.solve_population_circuit(pop, obj, copy_out=n < follow)
        E = res[1] if res is not None else None
        assert np.isfinite(obj[0].get_potential())
        push_dt = (1.0 - 0.5 * (n == 0)) * dt
        n_before = pop.size()
        N.check(L.punc_gpu_step(pop._h, 2, push_dt, N.dptr(B), dt, None, 1, 13, n))
        ke, ab, kin = _results(pop)
        assert pop.size() == n_before - ab.sum() + kin
        # Population::update's object bookkeeping (population.h:672-683)
        q_in = float(ab[0, 2] * QE + ab[1, 2] * QI)
        P.punc._objects_update(obj, ab, np.array([QE, QI]), dt)
        collected += q_in
        assert obj[0].charge == pytest.approx(collected, abs=1e-9)
        assert obj[0].current == pytest.approx(q_in / dt, abs=1e-9)
        if n < follow:
            ke_o = o.boris(E, B, push_dt)
            o.move(dt)
            ab_o, ch_o, lost = o.update()
            k_o = o.inject(dt, 13, n)
            assert lost == 0 and kin == k_o
            np.testing.assert_array_equal(ab, ab_o)
            assert ch_o[2] == pytest.approx(q_in, abs=1e-9)
            assert abs(ke - ke_o) <= 1e-11 * abs(ke_o)
            assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"circuit step {n}")
    c = pop.counts()
    assert abs(np.sum(_rho(pop, hm.ndof) / dv) - (c[0] * QE + c[1] * QI)) <= 1e-9 * (c[0] + c[1])
    assert collected != 0.0  # the probe collected charge

"""GPU parity: every particle operator through the C ABI against the oracle on
the same seeded inputs.  Cells, counts and absorbed counts must be bit-exact;
positions and velocities within 1e-12 relative (expected: bit-identical, the
library is built without FMA contraction); KE within 1e-11 relative; rho and
object charge within 1e-11 of a magnitude scale (atomic summation order)."""
import numpy as np
import pytest

import puncpp_amd as P
from puncpp_amd import _native as N
from helpers import (MI, ME, QE, QI, assert_same_sets, loaded_pair, near_face, periodic_mesh,
                     rho_scale, sphere_mesh, synthetic_field)

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def test_initial_state_roundtrip(gpu):
    _print_case("roundtrip")
    hm = sphere_mesh()
    o, mesh, pop = loaded_pair(hm, (1500, 1500))
    assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what="roundtrip")
    c = pop.counts()
    assert c.tolist() == [1500, 1500]
    assert pop.num_of_negatives() == o.num_of_negatives()
    assert pop.num_of_positives() == o.num_of_positives()


@pytest.mark.parametrize("half", [False, True])
def test_accel_cg1(gpu, half):
    _print_case("accel", half=half)
    hm = sphere_mesh()
    o, mesh, pop = loaded_pair(hm, (3000, 3000))
    E = synthetic_field(hm)
    dt = 0.05 * (0.5 if half else 1.0)
    ke_g = P.accel_cg1(pop, E, dt)
    ke_o = o.accel(E, dt)
    a, b = pop.get_particles(), o.get_particles()
    # accel does not reorder: arrays correspond element-wise
    np.testing.assert_array_equal(a["v"], b["v"])
    np.testing.assert_array_equal(a["x"], b["x"])
    assert abs(ke_g - ke_o) <= 1e-11 * abs(ke_o)


def test_boris_cg1(gpu):
    _print_case("boris")
    hm = sphere_mesh()
    o, mesh, pop = loaded_pair(hm, (3000, 3000))
    E = synthetic_field(hm)
    B = [0.0, 0.0, 1.0]
    ke_g = P.boris_cg1(pop, E, B, 0.02)
    ke_o = o.boris(E, B, 0.02)
    a, b = pop.get_particles(), o.get_particles()
    np.testing.assert_array_equal(a["v"], b["v"])
    assert abs(ke_g - ke_o) <= 1e-11 * abs(ke_o)
    # oblique B exercises every component of t and s
    B2 = [0.3, -0.7, 1.1]
    ke_g = P.boris_cg1(pop, E, B2, 0.03)
    ke_o = o.boris(E, B2, 0.03)
    np.testing.assert_array_equal(pop.get_particles()["v"], o.get_particles()["v"])
    assert abs(ke_g - ke_o) <= 1e-11 * abs(ke_o)


def test_kinetic_energy(gpu):
    _print_case("kinetic_energy")
    hm = sphere_mesh()
    o, mesh, pop = loaded_pair(hm, (2000, 2500))
    ke_g, ke_o = P.kinetic_energy(pop), o.kinetic_energy()
    assert abs(ke_g - ke_o) <= 1e-11 * abs(ke_o)


@pytest.mark.parametrize("steps", [1, 4])
def test_move_update(gpu, steps):
    _print_case("move_update", steps=steps)
    hm = sphere_mesh()
    o, mesh, pop = loaded_pair(hm, (4000, 4000), vth=(1.0, 0.3))
    objects = [P.Object(2)]
    total = np.zeros((2, 64), dtype=np.int64)
    for k in range(steps):
        P.move(pop, 0.2)
        o.move(0.2)
        ab_g = pop.update(objects, 0.2)
        ab_o, ch_o, lost = o.update()
        assert lost == 0
        np.testing.assert_array_equal(ab_g, ab_o)
        total += ab_o
        a, b = pop.get_particles(), o.get_particles()
        assert_same_sets(a, b, rel=0.0, what=f"step {k}")
    assert total[:, 1].sum() > 0, "test should see exterior leavers"
    assert total[:, 2].sum() > 0, "test should see object absorption"
    print(f"near-face particles (<1e-12): {near_face(o, o.get_particles())}", flush=True)
    # object charge bookkeeping (object.h:118-121, population.h:695-698)
    assert objects[0].charge == pytest.approx(float(total[0, 2] * QE + total[1, 2] * QI), rel=1e-12)


def test_move_periodic_update(gpu):
    _print_case("move_periodic")
    hm = periodic_mesh(6)
    o, mesh, pop = loaded_pair(hm, (3000, 3000), vth=(1.0, 0.5), vd=(0.0, 0.0, 0.0),
                               amp=0.01, kw=1.0)
    Ld = [2 * np.pi] * 3
    for k in range(3):
        P.move_periodic(pop, 0.3, Ld)
        o.move_periodic(0.3, Ld)
        ab_g = pop.update([], 0.3)
        ab_o, _, lost = o.update()
        assert lost == 0
        assert ab_g.sum() == 0 and ab_o.sum() == 0
        assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"periodic step {k}")


def test_distribute_cg1(gpu):
    _print_case("distribute")
    hm = sphere_mesh()
    o, mesh, pop = loaded_pair(hm, (5000, 5000))
    dv = P.element_volume(mesh)
    rho_g = P.distribute_cg1(pop, dv)
    rho_o = o.distribute(dv)
    scale = rho_scale(o, hm, dv)
    err = np.abs(rho_g - rho_o)
    assert np.all(err <= 1e-11 * np.maximum(scale, 1e-300)), err.max()
    # charge conservation: sum rho_j / dv_inv_j = sum q
    tot = np.sum(rho_g / dv)
    assert tot == pytest.approx(5000 * QE + 5000 * QI, abs=1e-9)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_distribute_unsorted_modes(gpu, mode):
    """Particles in random order: equal cells recur in non-adjacent lanes."""
    _print_case("distribute_unsorted", mode=mode)
    hm = sphere_mesh(n=4, L=2.0, r=0.7)
    import oracle as O
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    o.load(0, 6000, 4, 0.0, 0.0, 0, 1.0, [0, 0, 0])
    o.load(1, 6000, 4, 0.0, 0.0, 0, 1.0, [0, 0, 0])
    st = o.get_particles()
    perm = np.random.default_rng(0).permutation(st["x"].shape[0])
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=20000)
    pop.add_located(st["x"][perm], st["v"][perm], st["species"][perm], st["cell"][perm])
    pop.set_launch(pop.get_launch()[0], 256, mode)
    dv = P.element_volume(mesh)
    rho_g, rho_o = P.distribute_cg1(pop, dv), o.distribute(dv)
    scale = rho_scale(o, hm, dv)
    assert np.all(np.abs(rho_g - rho_o) <= 1e-11 * scale)


def test_distribute_periodic(gpu):
    _print_case("distribute_periodic")
    hm = periodic_mesh(5)
    o, mesh, pop = loaded_pair(hm, (4000, 3000), vd=(0.0, 0.0, 0.0))
    dv = P.element_volume(mesh)
    rho_g, rho_o = P.distribute_cg1(pop, dv), o.distribute(dv)
    scale = rho_scale(o, hm, dv)
    assert np.all(np.abs(rho_g - rho_o) <= 1e-11 * scale)


def test_inject(gpu):
    _print_case("inject")
    hm = sphere_mesh(n=6, L=3.0, r=1.0)
    o, mesh, pop = loaded_pair(hm, (10, 10))
    vth = [1.0, 1.0 / np.sqrt(1836.0)]
    vd = [0.0, 0.0, 0.5]
    n_dens = [40.0, 40.0]
    species = []
    for s, (q, m) in enumerate([(QE, ME), (QI, MI)]):
        sp = P.Species(q, m, n_dens[s], 1.0, P.ParticleAmountType.phys_per_sim, mesh,
                       P.UniformPosition(mesh), P.Maxwellian(vth[s], vd))
        species.append(sp)
        o.set_injector(s, vth[s], vd, sp.n, 5.0)
    dt = 0.1
    for step in range(3):
        k_g = P.inject_particles(pop, species, None, dt, seed=7, step=step)
        k_o = o.inject(dt, 7, step)
        assert k_g == k_o, (k_g, k_o)
        assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"inject {step}")
    assert k_o > 0


def test_device_load_matches_oracle(gpu):
    _print_case("device_load")
    hm = sphere_mesh()
    import oracle as O
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=20000)
    vth = [1.0, 0.05]
    for s in range(2):
        assert o.load(s, 3000, 11, 0.01, 1.0, 0, vth[s], [0.0, 0.0, 0.5]) == 0
        N.check(N.lib().punc_gpu_pop_load(
            pop._h, s, 3000, 11, 0.01, 1.0, 0, vth[s], N.dptr(np.array([0.0, 0.0, 0.5]))))
    a, b = pop.get_particles(), o.get_particles()
    # positions are exact; velocities go through erfc_inv (1-2 ulp between implementations)
    ca, cb = a, b
    order_a = np.lexsort((ca["x"][:, 2], ca["x"][:, 1], ca["x"][:, 0], ca["species"]))
    order_b = np.lexsort((cb["x"][:, 2], cb["x"][:, 1], cb["x"][:, 0], cb["species"]))
    np.testing.assert_array_equal(ca["x"][order_a], cb["x"][order_b])
    np.testing.assert_array_equal(ca["cell"][order_a], cb["cell"][order_b])
    va, vb = ca["v"][order_a], cb["v"][order_b]
    assert np.max(np.abs(va - vb) / np.maximum(np.abs(vb), 1e-300)) <= 1e-12


def test_sort_preserves_set(gpu):
    _print_case("sort")
    hm = sphere_mesh()
    o, mesh, pop = loaded_pair(hm, (3000, 3000))
    before = pop.get_particles()
    pop.sort()
    after = pop.get_particles()
    assert_same_sets(after, before, rel=0.0, what="sort")
    assert np.all(np.diff(after["cell"]) >= 0)


@pytest.mark.parametrize("pusher,mode,grid", [(1, 2, 0), (2, 2, 0), (2, 3, 0), (1, 3, 7), (2, 1, 3), (2, 4, 0),
                                               (1, 4, 7)])
def test_fused_step_matches_sequence(gpu, pusher, mode, grid):
    """punc_gpu_step (push+move+update+inject+deposit) against the oracle's
    reference-ordered sequence (interaction.cpp:384-422, then distribute), for
    deposit modes 1-4; grid 0 keeps the default grid, a small grid gives each
    block a long chunk (several prefetched groups and a long walk list)."""
    _print_case("fused_step", pusher=pusher, mode=mode, grid=grid)
    hm = sphere_mesh(n=6, L=3.0, r=1.0)
    o, mesh, pop = loaded_pair(hm, (4000, 4000))
    pop.set_launch(grid or pop.get_launch()[0], 256, mode)
    E = synthetic_field(hm, scale=0.2)
    B = np.array([0.0, 0.0, 1.0])
    dv = P.element_volume(mesh)
    vth = [1.0, 1.0 / np.sqrt(1836.0)]
    vd = [0.0, 0.0, 0.5]
    species = [P.Species(q, m, 30.0, 1.0, P.ParticleAmountType.phys_per_sim, mesh,
                         P.UniformPosition(mesh), P.Maxwellian(vth[s], vd))
               for s, (q, m) in enumerate([(QE, ME), (QI, MI)])]
    P.punc._register_injectors(pop, species)
    for s in range(2):
        o.set_injector(s, vth[s], vd, species[s].n, 5.0)
    L = N
    L.check(L.lib().punc_gpu_set_field(pop._h, L.dptr(E)))
    L.check(L.lib().punc_gpu_distribute(pop._h, L.dptr(dv), None))
    dt = 0.05
    for n in range(4):
        push_dt = (1.0 - 0.5 * (n == 0)) * dt
        L.check(L.lib().punc_gpu_step(pop._h, pusher, push_dt, L.dptr(B), dt, None, 1, 5, n))
        ke = L.C.c_double()
        ab = np.zeros((2, 64), dtype=np.int64)
        kin = L.C.c_int64()
        L.check(L.lib().punc_gpu_step_results(pop._h, L.C.byref(ke), L.i64ptr(ab), L.C.byref(kin)))
        ke_o = o.accel(E, push_dt) if pusher == 1 else o.boris(E, B, push_dt)
        o.move(dt)
        ab_o, _, lost = o.update()
        k_o = o.inject(dt, 5, n)
        assert lost == 0
        assert abs(ke.value - ke_o) <= 1e-11 * abs(ke_o)
        np.testing.assert_array_equal(ab, ab_o)
        assert kin.value == k_o
        assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"fused step {n}")
        rho_g = np.empty(hm.ndof)
        L.check(L.lib().punc_gpu_get_rho(pop._h, L.dptr(rho_g)))
        rho_o = o.distribute(dv)
        scale = rho_scale(o, hm, dv)
        assert np.all(np.abs(rho_g - rho_o) <= 1e-11 * scale)


@pytest.mark.parametrize("mode", [2, 3, 4])
def test_fused_step_table_overflow(gpu, mode):
    """One block processes the whole population (70 groups) on a mesh with
    more dofs than the LDS deposit table has slots: the table is flushed
    whenever it is 5/8 full at an 8-group boundary, and adds that find it
    full in between take the direct HBM path.  The block's walk list (a few
    thousand walkers) also outgrows its LDS part and continues in the
    block's global segment.  Results as the oracle's."""
    _print_case("fused_step_overflow", mode=mode)
    hm = sphere_mesh(n=14, L=4.0, r=1.0)
    assert hm.ndof > 2048
    o, mesh, pop = loaded_pair(hm, (9000, 9000))
    pop.set_launch(1, 256, mode)
    E = synthetic_field(hm, scale=0.2)
    B = np.array([0.0, 0.0, 1.0])
    dv = P.element_volume(mesh)
    L = N
    L.check(L.lib().punc_gpu_set_field(pop._h, L.dptr(E)))
    L.check(L.lib().punc_gpu_distribute(pop._h, L.dptr(dv), None))
    for n in range(2):
        L.check(L.lib().punc_gpu_step(pop._h, 2, 0.05, L.dptr(B), 0.05, None, 0, 5, n))
        ab = np.zeros((2, 64), dtype=np.int64)
        L.check(L.lib().punc_gpu_step_results(pop._h, None, L.i64ptr(ab), None))
        o.boris(E, B, 0.05)
        o.move(0.05)
        ab_o, _, lost = o.update()
        np.testing.assert_array_equal(ab, ab_o)
        assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"overflow step {n}")
        rho_g = np.empty(hm.ndof)
        L.check(L.lib().punc_gpu_get_rho(pop._h, L.dptr(rho_g)))
        rho_o = o.distribute(dv)
        assert np.all(np.abs(rho_g - rho_o) <= 1e-11 * rho_scale(o, hm, dv))


@pytest.mark.parametrize("fused", [False, True])
def test_mass_exit_removal(gpu, fused):
    """Most of a 400k-particle population leaves through the exterior in one
    step (thermal speed 30 per component, dt 0.2, box half-width 4): some
    350k holes, more than one pass of the removal's grid-stride loops (four
    blocks of 256 threads per compute unit), where holes and survivors meet
    through per-block tickets.  Survivors and absorbed counts as the
    oracle's, through update and through the fused step."""
    _print_case("mass_exit", fused=fused)
    hm = sphere_mesh()
    o, mesh, pop = loaded_pair(hm, (200_000, 200_000), vth=(30.0, 30.0))
    dt = 0.2
    if fused:
        E = np.zeros((hm.ndof, 3))
        dv = P.element_volume(mesh)
        N.check(N.lib().punc_gpu_set_field(pop._h, N.dptr(E)))
        N.check(N.lib().punc_gpu_distribute(pop._h, N.dptr(dv), None))
        B = np.array([0.0, 0.0, 1.0])
        N.check(N.lib().punc_gpu_step(pop._h, 1, dt, N.dptr(B), dt, None, 0, 5, 0))
        ab_g = np.zeros((2, 64), dtype=np.int64)
        N.check(N.lib().punc_gpu_step_results(pop._h, None, N.i64ptr(ab_g), None))
        o.accel(E, dt)
    else:
        P.move(pop, dt)
        ab_g = pop.update([P.Object(2)], dt)
    o.move(dt)
    ab_o, _, lost = o.update()
    assert lost == 0
    np.testing.assert_array_equal(ab_g, ab_o)
    left = int(ab_o.sum())
    assert left > 300_000, f"only {left} particles left"
    assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what="mass exit")


def test_graph_replay_matches_eager(gpu):
    """A captured step replayed one step at a time (with host work and
    synchronization between replays) gives what the eager step gives."""
    _print_case("graph_replay")
    hm = sphere_mesh(n=6, L=3.0, r=1.0)
    o, mesh, popA = loaded_pair(hm, (3000, 3000))
    st = o.get_particles()
    popB = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=100000)
    popB.add_located(st["x"], st["v"], st["species"], st["cell"])
    E = synthetic_field(hm, scale=0.2)
    B = np.array([0.0, 0.0, 1.0])
    dv = P.element_volume(mesh)
    vd = [0.0, 0.0, 0.5]
    species = [P.Species(q, m, 30.0, 1.0, P.ParticleAmountType.phys_per_sim, mesh,
                         P.UniformPosition(mesh), P.Maxwellian(v, vd))
               for (q, m), v in zip([(QE, ME), (QI, MI)], [1.0, 0.05])]
    L = N
    for pop in (popA, popB):
        P.punc._register_injectors(pop, species)
        L.check(L.lib().punc_gpu_set_field(pop._h, L.dptr(E)))
        L.check(L.lib().punc_gpu_distribute(pop._h, L.dptr(dv), None))
    dt = 0.05
    L.check(L.lib().punc_gpu_step_graph(popB._h, 2, dt, L.dptr(B), dt, None, 1, 9, 0))
    for n in range(4):
        L.check(L.lib().punc_gpu_step(popA._h, 2, dt, L.dptr(B), dt, None, 1, 9, n))
        L.check(L.lib().punc_gpu_graph_launch(popB._h, 1))
        L.check(L.lib().punc_gpu_synchronize(popB._h))
        ka, kb = L.C.c_double(), L.C.c_double()
        aba, abb = np.zeros((2, 64), np.int64), np.zeros((2, 64), np.int64)
        L.check(L.lib().punc_gpu_step_results(popA._h, L.C.byref(ka), L.i64ptr(aba), None))
        L.check(L.lib().punc_gpu_step_results(popB._h, L.C.byref(kb), L.i64ptr(abb), None))
        np.testing.assert_array_equal(aba, abb)
        assert ka.value == pytest.approx(kb.value, rel=1e-12)
        assert_same_sets(popB.get_particles(), popA.get_particles(), rel=0.0, what=f"graph step {n}")
        ra, rb = np.empty(hm.ndof), np.empty(hm.ndof)
        L.check(L.lib().punc_gpu_get_rho(popA._h, L.dptr(ra)))
        L.check(L.lib().punc_gpu_get_rho(popB._h, L.dptr(rb)))
        assert np.allclose(ra, rb, rtol=1e-11, atol=1e-11 * np.abs(ra).max())


@pytest.mark.parametrize("mode,graph", [(2, False), (3, False), (2, True), (4, True)])
def test_fused_step_after_sort(gpu, mode, graph):
    """A re-sort leaves its cell order pending and the next fused step reads
    the particles through it (writing the alternate buffers); the result is
    the oracle's sequence.  Particles are added in random order first, so the
    order is far from sorted.  graph: the steps are replays of a captured
    step with re-sorts between them."""
    _print_case("fused_step_after_sort", mode=mode, graph=graph)
    hm = sphere_mesh(n=6, L=3.0, r=1.0)
    import oracle as O
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    o.load(0, 5000, 11, 0.0, 0.0, 0, 1.0, [0.0, 0.0, 0.5])
    o.load(1, 5000, 11, 0.0, 0.0, 0, 0.05, [0.0, 0.0, 0.5])
    st = o.get_particles()
    perm = np.random.default_rng(4).permutation(st["x"].shape[0])
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=40000)
    pop.add_located(st["x"][perm], st["v"][perm], st["species"][perm], st["cell"][perm])
    pop.set_launch(5, 256, mode)
    E = synthetic_field(hm, scale=0.2)
    B = np.array([0.0, 0.0, 1.0])
    dv = P.element_volume(mesh)
    L = N
    L.check(L.lib().punc_gpu_set_field(pop._h, L.dptr(E)))
    L.check(L.lib().punc_gpu_distribute(pop._h, L.dptr(dv), None))
    dt = 0.05
    if graph:
        L.check(L.lib().punc_gpu_step_graph(pop._h, 2, dt, L.dptr(B), dt, None, 0, 3, 0))
    for n in range(3):
        L.check(L.lib().punc_gpu_sort(pop._h))
        if graph:
            L.check(L.lib().punc_gpu_graph_launch(pop._h, 1))
        else:
            L.check(L.lib().punc_gpu_step(pop._h, 2, dt, L.dptr(B), dt, None, 0, 3, n))
        ab = np.zeros((2, 64), dtype=np.int64)
        ke = L.C.c_double()
        L.check(L.lib().punc_gpu_step_results(pop._h, L.C.byref(ke), L.i64ptr(ab), None))
        ke_o = o.boris(E, B, dt)
        o.move(dt)
        ab_o, _, lost = o.update()
        np.testing.assert_array_equal(ab, ab_o)
        assert abs(ke.value - ke_o) <= 1e-11 * abs(ke_o)
        got = pop.get_particles()
        assert_same_sets(got, o.get_particles(), rel=0.0, what=f"sorted step {n}")
        rho_g = np.empty(hm.ndof)
        L.check(L.lib().punc_gpu_get_rho(pop._h, L.dptr(rho_g)))
        rho_o = o.distribute(dv)
        assert np.all(np.abs(rho_g - rho_o) <= 1e-11 * rho_scale(o, hm, dv))
    # a re-sort applied by a drop-in read (no fused step): cells in order
    L.check(L.lib().punc_gpu_sort(pop._h))
    got = pop.get_particles()
    assert np.all(np.diff(got["cell"]) >= 0)
    assert_same_sets(got, o.get_particles(), rel=0.0, what="sorted read")


@pytest.mark.parametrize("in_step", [False, True])
@pytest.mark.parametrize("graph", [False, True])
def test_steps_with_injection_and_resort_match_oracle(gpu, graph, in_step):
    """Eight fused steps back to back with no download in between (eager, or
    replays of the captured step), each removing the particles that leave and
    appending the injected ones, with a re-sort after the fifth that the next
    step reads through.  Per step absorbed and injected counts exact and the
    count balanced; at the end the particle set and rho equal the oracle's.
    in_step: the injection generator runs inside the step kernel's grid
    instead of on a second stream."""
    _print_case("steps_inject_resort", graph=graph, in_step=in_step)
    hm = sphere_mesh(n=6, L=3.0, r=1.0)
    o, mesh, pop = loaded_pair(hm, (4000, 4000))
    E = synthetic_field(hm, scale=0.2)
    B = np.array([0.0, 0.0, 1.0])
    dv = P.element_volume(mesh)
    vth = [1.0, 1.0 / np.sqrt(1836.0)]
    vd = [0.0, 0.0, 0.5]
    species = [P.Species(q, m, 30.0, 1.0, P.ParticleAmountType.phys_per_sim, mesh,
                         P.UniformPosition(mesh), P.Maxwellian(vth[s], vd))
               for s, (q, m) in enumerate([(QE, ME), (QI, MI)])]
    P.punc._register_injectors(pop, species)
    for s in range(2):
        o.set_injector(s, vth[s], vd, species[s].n, 5.0)
    L = N
    L.check(L.lib().punc_gpu_set_inject_in_step(pop._h, int(in_step)))
    got_in_step = L.C.c_int32()
    L.check(L.lib().punc_gpu_get_inject_in_step(pop._h, L.C.byref(got_in_step)))
    assert got_in_step.value == int(in_step)
    L.check(L.lib().punc_gpu_set_field(pop._h, L.dptr(E)))
    L.check(L.lib().punc_gpu_distribute(pop._h, L.dptr(dv), None))
    dt = 0.05
    if graph:
        L.check(L.lib().punc_gpu_step_graph(pop._h, 2, dt, L.dptr(B), dt, None, 1, 5, 0))
    total_abs = 0
    for n in range(8):
        if n == 5:
            pop.sort()  # the next step reads the new order
            if graph:
                L.check(L.lib().punc_gpu_step_graph(pop._h, 2, dt, L.dptr(B), dt, None, 1, 5, n))
        n_before = pop.size()
        if graph:
            L.check(L.lib().punc_gpu_graph_launch(pop._h, 1))
        else:
            L.check(L.lib().punc_gpu_step(pop._h, 2, dt, L.dptr(B), dt, None, 1, 5, n))
        ab = np.zeros((2, 64), dtype=np.int64)
        kin = L.C.c_int64()
        L.check(L.lib().punc_gpu_step_results(pop._h, None, L.i64ptr(ab), L.C.byref(kin)))
        o.boris(E, B, dt)
        o.move(dt)
        ab_o, _, lost = o.update()
        k_o = o.inject(dt, 5, n)
        assert lost == 0
        np.testing.assert_array_equal(ab, ab_o)
        assert kin.value == k_o
        assert pop.size() == n_before - ab.sum() + kin.value == o.size()
        total_abs += ab.sum()
    assert total_abs > 0
    assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what="after 8 steps with a re-sort")
    rho_g = np.empty(hm.ndof)
    L.check(L.lib().punc_gpu_get_rho(pop._h, L.dptr(rho_g)))
    assert np.all(np.abs(rho_g - o.distribute(dv)) <= 1e-11 * rho_scale(o, hm, dv))

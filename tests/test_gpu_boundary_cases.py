"""GPU parity at the edges of the relocation path: the first-positive-facet
order of the walk (population.h:547-565) on a hand-built mesh where it
decides the outcome, add_particles with points outside the mesh
(population.h:411-446: the reference would index out of bounds, both paths
skip and count), and the near-face count (particles relocated within 1e-12
of a facet plane of their cell) against the oracle's planes."""
import numpy as np
import pytest

import oracle as O
import puncpp_amd as P
from puncpp_amd import _native as N
from helpers import ME, MI, QE, QI, assert_same_sets, first_positive_meshes, near_face, sphere_mesh

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.parametrize("fused", [False, True])
def test_walk_first_positive_facet(gpu, fused):
    (m0, want0), (m1, want1), p = first_positive_meshes()
    for lab, (hm, want) in enumerate(((m0, want0), (m1, want1))):
        _print_case("first_positive", labeling=lab, fused=fused)
        mesh = P.Mesh(hm)
        pop = P.Population(mesh, [(QE, ME)], capacity=1024)
        o = O.OraclePop(hm, [QE], [ME])
        x0 = np.array([[0.25, 0.25, 0.25]])
        v0 = (p[None, :] - x0) / 0.5
        pop.add_located(x0, v0, [0], [0])
        o.add_located(x0, v0, [0], [0])
        o.move(0.5)
        ab_o, ch_o, _ = o.update()
        if fused:
            E = np.zeros((hm.ndof, 3))
            N.check(N.lib().punc_gpu_set_field(pop._h, N.dptr(E)))
            N.check(N.lib().punc_gpu_distribute(pop._h, N.dptr(np.ones(hm.ndof)), None))
            N.check(N.lib().punc_gpu_step(pop._h, 1, 0.0, None, 0.5, None, 0, 0, 0))
            ab = np.zeros((1, 64), dtype=np.int64)
            N.check(N.lib().punc_gpu_step_results(pop._h, None, N.i64ptr(ab), None))
        else:
            P.move(pop, 0.5)
            ab = pop.update()
        np.testing.assert_array_equal(ab, ab_o)
        if want < 0:
            assert pop.size() == 0 and ab[0, -want] == 1
        else:
            assert pop.size() == 1 and pop.get_particles()["cell"][0] == want
        assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what=f"labeling {lab}")


def test_add_particles_outside_mesh(gpu):
    """Points beyond the box and inside the object's hole are skipped by
    both paths; the kept sets and the counts agree."""
    _print_case("add_particles_outside")
    hm = sphere_mesh(n=8, L=4.0, r=1.5)
    rng = np.random.default_rng(17)
    xs = np.concatenate([rng.uniform(-5.0, 5.0, size=(6000, 3)),      # about half beyond the box
                         rng.uniform(-1.0, 1.0, size=(1500, 3)) * 0.8,  # inside the sphere hole
                         np.array([[4.0, 0.3, 0.2], [-4.0, -4.0, -4.0], [4.0 + 1e-9, 0.0, 0.0],
                                   [np.nan, 0.0, 0.0], [np.inf, 0.0, 0.0]])])
    vs = rng.normal(size=xs.shape)
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=20000)
    kept_o = o.add_particles(0, xs, vs)
    kept_g = pop.add_particles(xs, vs, QE, ME)
    assert kept_g == kept_o
    inside_box = np.all(np.abs(xs) <= 4.0, axis=1)
    assert 0 < kept_o < inside_box.sum()  # the hole and the outside were skipped
    assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what="add_particles")
    # a second batch for the other species, appended after the first
    xs2 = rng.uniform(-4.5, 4.5, size=(3000, 3))
    vs2 = rng.normal(size=xs2.shape) * 0.1
    assert pop.add_particles(xs2, vs2, QI, MI) == o.add_particles(1, xs2, vs2)
    assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what="add_particles, second batch")
    assert pop.counts().tolist() == [kept_o, o.size() - kept_o]


@pytest.mark.parametrize("fused", [False, True])
def test_near_face_count(gpu, fused):
    """Particles placed on the mesh's facet planes (grid planes x = k and the
    cubes' diagonal planes x = y) stay put; the count of those within 1e-12
    of a facet plane of their final cell equals the oracle's."""
    _print_case("near_face", fused=fused)
    hm = sphere_mesh(n=8, L=4.0, r=1.5)
    rng = np.random.default_rng(23)
    n = 3000
    xs = rng.uniform(-3.9, 3.9, size=(n, 3))
    on = rng.uniform(size=n) < 0.3
    k = rng.integers(-3, 4, size=n).astype(np.float64)
    xs[on, 0] = k[on]                       # grid planes
    diag = (~on) & (rng.uniform(size=n) < 0.2)
    xs[diag, 1] = xs[diag, 0]               # a diagonal plane inside each cube
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    o.add_particles(0, xs, np.zeros_like(xs))
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=20000)
    st = o.get_particles()
    pop.add_located(st["x"], st["v"], st["species"], st["cell"])
    assert pop.near_face_count() == 0
    if fused:
        E = np.zeros((hm.ndof, 3))
        N.check(N.lib().punc_gpu_set_field(pop._h, N.dptr(E)))
        N.check(N.lib().punc_gpu_distribute(pop._h, N.dptr(np.ones(hm.ndof)), None))
        N.check(N.lib().punc_gpu_step(pop._h, 1, 0.0, None, 0.1, None, 0, 0, 0))
        N.check(N.lib().punc_gpu_step_results(pop._h, None, None, None))
    else:
        P.move(pop, 0.1)
        pop.update()
    o.move(0.1)
    o.update()
    assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what="near-face particles")
    want = near_face(o, o.get_particles())
    assert want > 500  # the placement put many particles on facets
    assert pop.near_face_count() == want
    assert pop.near_face_count() == 0  # read resets


def test_graph_dropped_when_captured_state_changes(gpu):
    """A captured step is dropped (and graph_launch fails with a message)
    when what it captured is freed or changes: a new species, an injector
    change, an eager step with another dt that rebuilds the injection
    tables.  A fresh capture replays again, matching eager steps."""
    _print_case("graph_invalidation")
    hm = sphere_mesh(n=6, L=3.0, r=1.0)
    mesh = P.Mesh(hm)
    sp = [P.Species(QE, ME, 30.0, 1.0, P.ParticleAmountType.phys_per_sim, mesh, P.UniformPosition(mesh),
                    P.Maxwellian(1.0, [0.0, 0.0, 0.5])),
          P.Species(QI, MI, 30.0, 1.0, P.ParticleAmountType.phys_per_sim, mesh, P.UniformPosition(mesh),
                    P.Maxwellian(1.0 / np.sqrt(MI), [0.0, 0.0, 0.5]))]
    P.create_flux(sp, mesh)
    pop = P.Population(mesh, sp, capacity=200000)
    P.load_particles(pop, sp, seed=1)
    P.punc._register_injectors(pop, sp)
    E = np.zeros((hm.ndof, 3))
    N.check(N.lib().punc_gpu_set_field(pop._h, N.dptr(E)))
    N.check(N.lib().punc_gpu_distribute(pop._h, N.dptr(P.element_volume(mesh)), None))
    B = np.array([0.0, 0.0, 1.0])
    L = N.lib()

    def capture(dt=0.05):
        N.check(L.punc_gpu_step_graph(pop._h, 2, dt, N.dptr(B), dt, None, 1, 7, 100))

    def launch_fails():
        rc = L.punc_gpu_graph_launch(pop._h, 1)
        msg = L.punc_gpu_last_error().decode()
        return rc != 0 and "dropped" in msg

    capture()
    N.check(L.punc_gpu_graph_launch(pop._h, 2))
    # an eager step with another dt rebuilds the injection tables
    N.check(L.punc_gpu_step(pop._h, 2, 0.03, N.dptr(B), 0.03, None, 1, 7, 0))
    assert launch_fails()
    capture()
    N.check(L.punc_gpu_graph_launch(pop._h, 1))
    # an injector change
    N.check(L.punc_gpu_injector_set(pop._h, 0, 1.1, N.dptr(np.array([0.0, 0.0, 0.4])), 30.0, 5.0))
    assert launch_fails()
    capture()
    N.check(L.punc_gpu_graph_launch(pop._h, 1))
    # a new species
    idx = N.C.c_int()
    N.check(L.punc_gpu_pop_add_species(pop._h, -2.0, 3.0, N.C.byref(idx)))
    assert idx.value == 2
    assert launch_fails()
    capture()
    N.check(L.punc_gpu_graph_launch(pop._h, 3))
    N.check(L.punc_gpu_synchronize(pop._h))
    # the replays of the fresh capture still agree with eager steps from the same state
    st = pop.get_particles()
    g2 = P.Population(mesh, [(QE, ME), (QI, MI), (-2.0, 3.0)], capacity=200000)
    g2.add_located(st["x"], st["v"], st["species"], st["cell"])
    N.check(L.punc_gpu_set_field(g2._h, N.dptr(E)))
    N.check(L.punc_gpu_distribute(g2._h, N.dptr(P.element_volume(mesh)), None))
    N.check(L.punc_gpu_injector_set(g2._h, 0, 1.1, N.dptr(np.array([0.0, 0.0, 0.4])), 30.0, 5.0))
    P.punc._register_injectors(g2, sp[1:])
    N.check(L.punc_gpu_step_graph(pop._h, 2, 0.05, N.dptr(B), 0.05, None, 1, 7, 500))
    N.check(L.punc_gpu_graph_launch(pop._h, 2))
    for step in (500, 501):
        N.check(L.punc_gpu_step(g2._h, 2, 0.05, N.dptr(B), 0.05, None, 1, 7, step))
    N.check(L.punc_gpu_synchronize(pop._h))
    assert_same_sets(pop.get_particles(), g2.get_particles(), rel=0.0, what="replay after re-capture")


def test_pinned_host_transfers(gpu):
    # The field from page-locked memory (punc_gpu_host_alloc) pushes exactly
    # as from pageable memory, and the deposit reads back the same either way.
    _print_case("pinned_host_transfers")
    from helpers import loaded_pair, synthetic_field
    hm = sphere_mesh()
    o, mesh, pop = loaded_pair(hm, (2000, 2000))
    st = o.get_particles()
    pop2 = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=20000)
    pop2.add_located(st["x"], st["v"], st["species"], st["cell"])
    E = synthetic_field(hm, scale=0.2)
    pin = N.PinnedArray(E.shape)
    pin.array[:] = E
    L = N.lib()
    B = np.array([0.0, 0.0, 1.0])
    dv = P.element_volume(mesh)
    ke = [N.C.c_double(), N.C.c_double()]
    for k, (pp, src) in enumerate(((pop, E), (pop2, pin.array))):
        N.check(L.punc_gpu_set_field(pp._h, N.dptr(src)))
        N.check(L.punc_gpu_boris(pp._h, N.dptr(B), 0.05, N.C.byref(ke[k])))
        N.check(L.punc_gpu_distribute(pp._h, N.dptr(dv), None))
    assert ke[0].value == ke[1].value
    assert_same_sets(pop.get_particles(), pop2.get_particles(), rel=0.0, what="pinned field")
    rho = np.empty(hm.ndof)
    rho_pin = N.PinnedArray((hm.ndof,))
    N.check(L.punc_gpu_get_rho(pop2._h, N.dptr(rho)))
    N.check(L.punc_gpu_get_rho(pop2._h, N.dptr(rho_pin.array)))
    assert np.array_equal(rho, rho_pin.array) and np.any(rho != 0)

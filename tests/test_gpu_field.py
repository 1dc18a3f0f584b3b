"""GPU field solve against the oracle (oracle/field.py: the same forms and
boundary conditions, solved directly): Poisson with a grounded exterior and
an object at a potential, with per-dof exterior values, the periodic box with
the null space removed, the E projection, and the device-resident path from
a population's deposit to the field its next push reads.

Tolerances: both solve to relative 1e-12 (the GPU) or exactly (the oracle);
potentials agree within 1e-9 of max |phi|, fields within 1e-8 of max |E|."""
import numpy as np
import pytest

import oracle as O
from oracle.field import FieldOracle
import puncpp_amd as P
from helpers import MI, ME, QE, QI, loaded_pair, sphere_mesh

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _oracle(hm, eps0):
    o = O.OraclePop(hm, [1.0], [1.0])
    return FieldOracle(hm, eps0, o.localizer()["adj"])


def _close(a, b, rel):
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= rel * scale, np.abs(a - b).max() / scale


def test_poisson_object_and_exterior(gpu):
    _print_case("poisson_object")
    hm = sphere_mesh(n=10, L=4.0, r=1.5)
    mesh = P.Mesh(hm)
    rng = np.random.default_rng(2)
    rho = rng.normal(size=hm.ndof)
    ext = 0.1 * hm.xyz[:, 0]  # per-dof exterior values (dofs = vertices here)
    f = P.FieldSolver(mesh, eps0=1.7, exterior=True, n_objects=1)
    phi = f.solve(rho, exterior_values=ext, object_potentials=[-2.0])
    assert f.iterations[0] > 0
    ref = _oracle(hm, 1.7).poisson(rho, exterior=True, exterior_values=ext, object_potentials=[-2.0])
    _close(phi, ref, 1e-9)
    E = f.efield(phi)
    _close(E, _oracle(hm, 1.7).efield(ref), 1e-8)


def test_poisson_periodic_null_space(gpu):
    _print_case("poisson_periodic")
    hm = P.box_mesh(12, [0.0] * 3, [2 * np.pi] * 3, periodic=True)
    mesh = P.Mesh(hm)
    rng = np.random.default_rng(3)
    rho = rng.normal(size=hm.ndof) + 0.3  # a nonzero mean is projected out
    f = P.FieldSolver(mesh, eps0=1.0, exterior=False, remove_null_space=True)
    phi = f.solve(rho)
    assert abs(phi.mean()) < 1e-12 * np.abs(phi).max()
    ref = _oracle(hm, 1.0).poisson(rho, exterior=False, remove_null_space=True)
    _close(phi, ref, 1e-9)


def test_field_from_population(gpu):
    # deposit -> solve -> E in the population; the potential and field match
    # the oracle's, and a push with the device-resident field equals the
    # oracle's push with that field handed over from the host
    _print_case("field_from_population")
    hm = sphere_mesh(n=8, L=4.0, r=1.5)
    o, mesh, pop = loaded_pair(hm, (20000, 20000))
    dv = P.element_volume(mesh)
    rho = P.distribute_cg1(pop, dv)
    f = P.FieldSolver(mesh, eps0=1.0, exterior=True, n_objects=1)
    phi, E = f.solve_population(pop, object_potentials=[0.5])
    fo = _oracle(hm, 1.0)
    phi_o = fo.poisson(o.distribute(dv), exterior=True, object_potentials=[0.5])
    _close(phi, phi_o, 1e-9)
    E_o = fo.efield(phi_o)
    _close(E, E_o, 1e-8)
    # the push reads the solved field straight from the device
    from puncpp_amd import _native as N
    kd = N.C.c_double()
    N.check(N.lib().punc_gpu_accel(pop._h, 0.05, N.C.byref(kd)))
    ke_o = o.accel(np.ascontiguousarray(E), 0.05)  # the same field, from the host
    assert abs(kd.value - ke_o) <= 1e-11 * abs(ke_o)
    a, b = pop.get_particles(), o.get_particles()
    order_a = np.lexsort((a["x"][:, 2], a["x"][:, 1], a["x"][:, 0]))
    order_b = np.lexsort((b["x"][:, 2], b["x"][:, 1], b["x"][:, 0]))
    np.testing.assert_array_equal(a["v"][order_a], b["v"][order_b])


def test_field_large_mesh(gpu):
    # the large workload's mesh (1.57M cells, 275k dofs): grounded box,
    # sphere at 1 V, random charge; the residual of the eliminated system
    # is checked through the oracle's matrices
    hm = P.box_mesh(64, [-20.0] * 3, [20.0] * 3, sphere_r=2.0)
    mesh = P.Mesh(hm)
    _print_case("field_large", dofs=hm.ndof)
    rho = np.random.default_rng(4).normal(size=hm.ndof) * 1e-2
    f = P.FieldSolver(mesh, eps0=1.0, exterior=True, n_objects=1)
    phi = f.solve(rho, object_potentials=[1.0])
    print(f"poisson iterations {f.iterations[0]}", flush=True)
    fo = _oracle(hm, 1.0)
    d_ext, d_obj = fo.boundary_dofs(1), fo.boundary_dofs(2)
    assert np.all(phi[d_ext] == 0.0) and np.all(phi[d_obj] == 1.0)
    free = np.ones(hm.ndof, bool)
    free[d_ext] = free[d_obj] = False
    res = (fo.K @ phi - fo.M @ rho)[free]
    assert np.linalg.norm(res) <= 1e-10 * np.linalg.norm((fo.M @ rho)[free] - (fo.K @ np.where(free, 0, phi))[free])
    E = f.efield(phi)
    assert np.all(np.isfinite(E))


def test_warm_start_same_answer(gpu):
    _print_case("warm_start")
    hm = sphere_mesh(n=10, L=4.0, r=1.5)
    mesh = P.Mesh(hm)
    rng = np.random.default_rng(7)
    rho1 = rng.normal(size=hm.ndof)
    rho2 = rho1 + 0.01 * rng.normal(size=hm.ndof)
    cold = P.FieldSolver(mesh, exterior=True, n_objects=1)
    warm = P.FieldSolver(mesh, exterior=True, n_objects=1, warm_start=True)
    warm.solve(rho1, object_potentials=[0.3])
    warm.efield(warm.solve(rho1, object_potentials=[0.3]))
    phi_w = warm.solve(rho2, object_potentials=[0.3])
    it_w = warm.iterations[0]
    phi_c = cold.solve(rho2, object_potentials=[0.3])
    assert it_w < cold.iterations[0]
    _close(phi_w, phi_c, 1e-9)
    _close(warm.efield(phi_w), cold.efield(phi_c), 1e-8)


def _two_object_mesh(n=8):
    # the sphere's upper facets become a second object (boundary id 3)
    hm = sphere_mesh(n=n, L=4.0, r=1.5)
    fv = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))
    adj = O.OraclePop(hm, [1.0], [1.0]).localizer()["adj"]
    for c, f in zip(*np.nonzero(adj == -2)):
        if hm.xyz[hm.cells[c, list(fv[f])]].mean(axis=0)[2] > 0:
            hm.facet_bnd[c, f] = 3
    return hm


@pytest.mark.parametrize("case", ["floating", "fixed", "pair"])
def test_circuit_matches_oracle(gpu, case):
    from oracle.field import CircuitOracle
    _print_case("circuit", case=case)
    hm = _two_object_mesh() if case == "pair" else sphere_mesh(n=8, L=4.0, r=1.5)
    nobj = 2 if case == "pair" else 1
    vs = {"floating": [], "fixed": [(-1, 0, 3.0)], "pair": [(0, 1, 2.0)]}[case]
    isrc = [(-1, 0, 0.5)] if case == "pair" else []
    mesh = P.Mesh(hm)
    f = P.FieldSolver(mesh, eps0=1.0, exterior=True, n_objects=nobj)
    inv_c = f.set_circuit(0.1, vs, isrc)
    fo = _oracle(hm, 1.0)
    co = CircuitOracle(fo, nobj, 0.1, vs, isrc)
    np.testing.assert_allclose(inv_c, co.inv_c, rtol=1e-9)
    rho = np.random.default_rng(11).normal(size=hm.ndof)
    objs = [P.Object(2 + j) for j in range(nobj)]
    for j, o in enumerate(objs):
        o.charge = 1.0 + j
    q0 = [o.charge for o in objs]
    phi = f.solve_circuit(rho, objs)
    phi_o, pot_o, q_o = co.solve(rho, q0)
    np.testing.assert_allclose([o.get_potential() for o in objs], pot_o, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose([o.charge for o in objs], q_o, rtol=1e-9, atol=1e-12)
    _close(phi, phi_o, 1e-9)
    np.testing.assert_allclose(f.object_charges(phi), fo.object_charges(phi_o, nobj), rtol=1e-8, atol=1e-10)


def test_circuit_population_path(gpu):
    # the deposit of a population through the circuit, E left for the push
    _print_case("circuit_population")
    hm = sphere_mesh(n=8, L=4.0, r=1.5)
    o, mesh, pop = loaded_pair(hm, (5000, 5000))
    dv = P.element_volume(mesh)
    P.distribute_cg1(pop, dv)
    f = P.FieldSolver(mesh, eps0=1.0, exterior=True, n_objects=1)
    f.set_circuit(0.1)
    obj = [P.Object(2)]
    obj[0].charge = -3.0
    phi, E = f.solve_population_circuit(pop, obj)
    from oracle.field import CircuitOracle
    fo = _oracle(hm, 1.0)
    phi_o, pot_o, q_o = CircuitOracle(fo, 1, 0.1).solve(o.distribute(dv), [-3.0])
    assert obj[0].get_potential() == pytest.approx(pot_o[0], rel=1e-9)
    _close(phi, phi_o, 1e-9)
    _close(E, fo.efield(phi_o), 1e-8)


@pytest.mark.parametrize("case", ["object", "periodic"])
def test_multigrid_and_jacobi_agree(gpu, case):
    # the multigrid-preconditioned solve (the default) and the Jacobi one
    # reach the oracle's solution; the multigrid one in far fewer iterations
    _print_case("multigrid", case=case)
    if case == "object":
        hm = sphere_mesh(n=12, L=4.0, r=1.5)
        kw, okw = dict(exterior=True, n_objects=1), dict(exterior=True, object_potentials=[0.7])
        skw = dict(object_potentials=[0.7])
    else:
        hm = P.box_mesh(14, [0.0] * 3, [2 * np.pi] * 3, periodic=True)
        kw, okw, skw = dict(exterior=False, remove_null_space=True), dict(exterior=False, remove_null_space=True), {}
    mesh = P.Mesh(hm)
    rho = np.random.default_rng(21).normal(size=hm.ndof)
    f = P.FieldSolver(mesh, **kw)
    kind, levels, coarse = f.preconditioner
    assert kind == "amg" and levels >= 1 and 0 < coarse <= 2048
    phi_m = f.solve(rho, **skw)
    it_m = f.iterations[0]
    f.set_preconditioner("jacobi")
    assert f.preconditioner[0] == "jacobi"
    phi_j = f.solve(rho, **skw)
    it_j = f.iterations[0]
    print(f"iterations multigrid {it_m} jacobi {it_j}", flush=True)
    assert it_m < it_j / 2
    ref = _oracle(hm, 1.0).poisson(rho, **okw)
    _close(phi_m, ref, 1e-9)
    _close(phi_j, ref, 1e-9)
    f.set_preconditioner("amg")  # back again: its captured batch is reused
    _close(f.solve(rho, **skw), ref, 1e-9)
    assert f.iterations[0] == it_m


def test_multigrid_small_operator_stays_jacobi(gpu):
    # at most 256 dofs: no hierarchy, Jacobi, and asking for multigrid fails
    _print_case("multigrid_small")
    hm = P.box_mesh(4, [0.0] * 3, [1.0] * 3)
    f = P.FieldSolver(P.Mesh(hm), exterior=True)
    assert f.preconditioner == ("jacobi", 0, 0)
    from puncpp_amd._native import PuncGpuError
    with pytest.raises(PuncGpuError):
        f.set_preconditioner("amg")
    rho = np.random.default_rng(5).normal(size=hm.ndof)
    _close(f.solve(rho), _oracle(hm, 1.0).poisson(rho, exterior=True), 1e-9)

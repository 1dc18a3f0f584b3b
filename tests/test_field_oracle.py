"""Pins for the field-solve oracle (oracle/field.py): what the mathematics
fixes for CG1 finite elements on the generated tetrahedral meshes.
* A linear potential is reproduced exactly (Dirichlet data from it, rho = 0),
  and its L2-projected field is the constant -grad(phi): pins the stiffness,
  the boundary rows and the gradient/mass pair.
* Manufactured solutions converge at second order at the nodes: Dirichlet
  sin(x)sin(y)sin(z) on [0, pi]^3 with eps0 = 2 (pins eps0 and the mass
  matrix of the right-hand side), periodic cos(x) with the null space
  removed (zero mean).
* An object held at V0 in a grounded box: the boundary values, the discrete
  maximum principle (the Kuhn tetrahedra are non-obtuse) and the mesh's
  symmetries."""
import numpy as np
import pytest

import oracle as O
from oracle.field import FieldOracle
import puncpp_amd as P


def _field(hm, eps0=1.0):
    o = O.OraclePop(hm, [1.0], [1.0])
    return FieldOracle(hm, eps0, o.localizer()["adj"])


def test_linear_potential_exact():
    hm = P.box_mesh(4, [-1.0] * 3, [2.0] * 3)
    f = _field(hm, eps0=3.0)
    a, c = np.array([0.3, -1.2, 0.7]), 0.4
    lin = hm.xyz @ a + c
    phi = f.poisson(np.zeros(hm.ndof), exterior=True, exterior_values=lin)
    np.testing.assert_allclose(phi, lin, atol=1e-12)
    E = f.efield(phi)
    np.testing.assert_allclose(E, np.tile(-a, (hm.ndof, 1)), atol=1e-11)


def _sine_error(n, eps0):
    hm = P.box_mesh(n, [0.0] * 3, [np.pi] * 3)
    f = _field(hm, eps0)
    x = hm.xyz
    exact = np.sin(x[:, 0]) * np.sin(x[:, 1]) * np.sin(x[:, 2])
    phi = f.poisson(3.0 * eps0 * exact, exterior=True)
    return np.abs(phi - exact).max()


def test_manufactured_dirichlet_second_order():
    e8, e16 = _sine_error(8, 2.0), _sine_error(16, 2.0)
    assert e16 < 0.02
    assert 3.2 < e8 / e16 < 5.0


def _cos_error(n):
    hm = P.box_mesh(n, [0.0] * 3, [2 * np.pi] * 3, periodic=True)
    f = _field(hm)
    x = np.zeros((hm.ndof, 3))
    x[hm.vdof] = hm.xyz
    exact = np.cos(x[:, 0])
    phi = f.poisson(exact, exterior=False, remove_null_space=True)
    assert abs(phi.mean()) < 1e-13
    return np.abs(phi - exact).max()


def test_manufactured_periodic_second_order():
    e8, e16 = _cos_error(8), _cos_error(16)
    assert e16 < 0.03
    assert 3.2 < e8 / e16 < 5.0


def test_object_potential():
    hm = P.box_mesh(8, [-4.0] * 3, [4.0] * 3, sphere_r=1.5)
    f = _field(hm)
    phi = f.poisson(np.zeros(hm.ndof), exterior=True, object_potentials=[2.5])
    obj, ext = f.boundary_dofs(2), f.boundary_dofs(1)
    assert len(obj) > 0 and len(ext) > 0
    np.testing.assert_allclose(phi[obj], 2.5, atol=1e-13)
    np.testing.assert_allclose(phi[ext], 0.0, atol=1e-13)
    assert phi.min() >= -1e-12 and phi.max() <= 2.5 + 1e-12
    # the Kuhn triangulation (cube diagonals along (1, 1, 1)) and the problem
    # are invariant under p -> -p and under the cyclic axis permutation
    idx = {tuple(np.round(p, 9)): i for i, p in enumerate(hm.xyz)}
    for f_map in (lambda p: -p, lambda p: p[[1, 2, 0]]):
        j = np.array([idx[tuple(np.round(f_map(p), 9))] for p in hm.xyz])
        np.testing.assert_allclose(phi[j], phi, atol=1e-10)


# ---------------------------------------------------------------------------
# Object charges and the capacitance-matrix circuit
# ---------------------------------------------------------------------------
from oracle.field import CircuitOracle, charge_sharing_sets  # noqa: E402


def test_object_charge_of_linear_potential_vanishes():
    # a closed object surface: the flux of a constant field is zero
    hm = P.box_mesh(8, [-4.0] * 3, [4.0] * 3, sphere_r=1.5)
    f = _field(hm, eps0=2.0)
    lin = hm.xyz @ np.array([0.3, -1.2, 0.7]) + 5.0
    assert abs(f.object_charges(lin, 1)[0]) < 1e-13


def test_object_charge_planar_face_closed_form():
    # the exterior face x = 4 relabelled as an object: the charge of a linear
    # potential a.x is eps0 a_x times the face's area (outward normal +x)
    hm = P.box_mesh(4, [-4.0] * 3, [4.0] * 3)
    fv = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))
    adj = O.OraclePop(hm, [1.0], [1.0]).localizer()["adj"]
    for c, f in zip(*np.nonzero(adj == -1)):
        if np.allclose(hm.xyz[hm.cells[c, list(fv[f])]][:, 0], 4.0):
            hm.facet_bnd[c, f] = 2
    f = _field(hm, eps0=2.0)
    lin = hm.xyz @ np.array([0.7, 0.3, -0.2]) + 1.0
    assert f.object_charges(lin, 1)[0] == pytest.approx(2.0 * 0.7 * 64.0, rel=1e-13)


def test_charge_sharing_sets():
    # object.cpp:90-135: sources chain objects; a chain reaching ground is a
    # fixed-voltage set, not a charge-sharing group
    assert charge_sharing_sets([], 3) == [[0], [1], [2]]
    assert charge_sharing_sets([(0, 1, 2.0)], 3) == [[0, 1], [2]]
    assert charge_sharing_sets([(-1, 0, 1.0), (1, 2, 3.0)], 3) == [[1, 2]]


def _two_object_mesh(n=8):
    # the sphere's upper facets become a second object (boundary id 3)
    hm = P.box_mesh(n, [-4.0] * 3, [4.0] * 3, sphere_r=1.5)
    fv = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))
    o = O.OraclePop(hm, [1.0], [1.0])
    adj = o.localizer()["adj"]
    for c, f in zip(*np.nonzero(adj == -2)):
        if hm.xyz[hm.cells[c, list(fv[f])]].mean(axis=0)[2] > 0:
            hm.facet_bnd[c, f] = 3
    return hm


def test_circuit_closed_forms():
    # no space charge: a floating object at Q floats at Q/C and carries Q;
    # an object held by a source at 3 V carries C 3; two objects joined by a
    # 2 V source differ by 2 V and share the collected charge
    hm = P.box_mesh(8, [-4.0] * 3, [4.0] * 3, sphere_r=1.5)
    f = _field(hm)
    rho0 = np.zeros(hm.ndof)
    c = CircuitOracle(f, 1, 0.1)
    C = c.capacitance[0, 0]
    phi, pot, q = c.solve(rho0, [2.0])
    assert pot[0] == pytest.approx(2.0 / C, rel=1e-12) and q[0] == pytest.approx(2.0, rel=1e-12)
    assert f.object_charges(phi, 1)[0] == pytest.approx(2.0, rel=1e-9)
    cf = CircuitOracle(f, 1, 0.1, vsources=[(-1, 0, 3.0)])
    phi, pot, q = cf.solve(rho0, [0.0])
    assert pot[0] == pytest.approx(3.0, rel=1e-12) and q[0] == pytest.approx(3.0 * C, rel=1e-12)
    h2 = _two_object_mesh()
    f2 = _field(h2)
    c2 = CircuitOracle(f2, 2, 0.1, vsources=[(0, 1, 2.0)], isources=[(-1, 0, 0.5)])
    phi, pot, q = c2.solve(np.zeros(h2.ndof), [1.0, 0.5])
    assert pot[1] - pot[0] == pytest.approx(2.0, abs=1e-10)
    # the current source added 0.5 * dt to object 0 before the solve
    assert q.sum() == pytest.approx(1.0 + 0.5 + 0.05, rel=1e-12)
    np.testing.assert_allclose(f2.object_charges(phi, 2).sum(), q.sum(), rtol=1e-8)

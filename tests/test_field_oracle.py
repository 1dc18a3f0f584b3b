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

"""The relocation kernel keeps a particle in its cell without walking when
all four barycentric coordinates of its new position exceed the mesh margin
``inside_tol``.  That is only valid if every such point also passes the
reference's plane test (population.h:539-569: no facet with proj > 0).

These CPU checks evaluate both tests with the kernel's own operation order
(numpy float64, no fused multiply-add) at points placed just above the margin
next to every facet, and check that the margin is small enough to matter."""
import ctypes as C

import numpy as np
import pytest

from puncpp_amd import _native as N
from puncpp_amd import punc as P
from tests.helpers import periodic_mesh, sphere_mesh


def inside_tol(mesh):
    t = C.c_double()
    N.check(N.lib().punc_internal_inside_tol(mesh._h, C.byref(t)))
    return t.value


def bary4(A, x):
    """Kernel order: l_r = ((A0 + A1 x) + A2 y) + A3 z; l3 = ((1 - l0) - l1) - l2."""
    l = [((A[:, 4 * r] + A[:, 4 * r + 1] * x[:, 0]) + A[:, 4 * r + 2] * x[:, 1]) + A[:, 4 * r + 3] * x[:, 2]
         for r in range(3)]
    l.append(((1 - l[0]) - l[1]) - l[2])
    return np.stack(l, axis=1)


def proj(Pl, x):
    """plane_proj order: ((p0 + p1 x) + p2 y) + p3 z, per facet."""
    return np.stack([((Pl[:, 4 * f] + Pl[:, 4 * f + 1] * x[:, 0]) + Pl[:, 4 * f + 2] * x[:, 1])
                     + Pl[:, 4 * f + 3] * x[:, 2] for f in range(4)], axis=1)


@pytest.mark.parametrize("make", [lambda: sphere_mesh(10), lambda: periodic_mesh(7)], ids=["sphere", "periodic"])
def test_margin_implies_plane_test(make):
    mesh = P.Mesh(make())
    tol = inside_tol(mesh)
    assert 0 < tol < 1e-9, tol
    loc = mesh.localizer()
    nc = mesh.num_cells
    V = loc["vcoords"].reshape(nc, 4, 3)
    rng = np.random.default_rng(7)
    checked = 0
    for face in range(4):
        for scale in (1.0, 1.5, 4.0, 1e3):
            # barycentric point with coordinate `face` at scale*tol, the rest random
            b = rng.random((nc, 4)) + 0.05
            b[:, face] = 0.0
            b /= b.sum(axis=1, keepdims=True)
            b *= 1.0 - scale * tol
            b[:, face] = scale * tol
            x = np.einsum("cj,cjd->cd", b, V)
            lam = bary4(loc["bary"], x)
            sel = np.all(lam > tol, axis=1)
            pr = proj(loc["planes"][sel], x[sel])
            assert not np.any(pr > 0.0), (face, scale, int(np.sum(pr > 0)))
            checked += int(sel.sum())
    assert checked > nc  # the margin admits most of these points


def test_margin_is_needed():
    """Points a few ulps inside a facet can pass the barycentric test while
    the plane test sends them through the facet: some margin is required."""
    mesh = P.Mesh(sphere_mesh(10))
    loc = mesh.localizer()
    nc = mesh.num_cells
    V = loc["vcoords"].reshape(nc, 4, 3)
    rng = np.random.default_rng(3)
    bad = 0
    for face in range(4):
        for _ in range(4):
            b = rng.random((nc, 4)) + 0.05
            b[:, face] = 0.0
            b /= b.sum(axis=1, keepdims=True)
            b[:, face] = 1e-16
            x = np.einsum("cj,cjd->cd", b, V)
            lam = bary4(loc["bary"], x)
            sel = np.all(lam > 0.0, axis=1)
            bad += int(np.any(proj(loc["planes"][sel], x[sel]) > 0.0, axis=1).sum())
    assert bad > 0


@pytest.mark.parametrize("make", [lambda: sphere_mesh(10), lambda: periodic_mesh(7)], ids=["sphere", "periodic"])
def test_margin_decides_facets_outside(make):
    """The walk decides facet i from lambda_i alone when |lambda_i| exceeds
    m = inside_tol * max(1, sum |lambda_j|): lambda_i < -m must give a positive
    plane projection, lambda_i > m a negative one.  Checked at points beyond
    one facet by a few margins up to half a cell, where the sum exceeds 1."""
    mesh = P.Mesh(make())
    tol = inside_tol(mesh)
    loc = mesh.localizer()
    nc = mesh.num_cells
    V = loc["vcoords"].reshape(nc, 4, 3)
    rng = np.random.default_rng(11)
    decided_neg = decided_pos = 0
    for face in range(4):
        for t in (1.0 * tol, 1.5 * tol, 4.0 * tol, 1e3 * tol, 1e-3, 0.5):
            b = rng.random((nc, 4)) + 0.05
            b[:, face] = 0.0
            b /= b.sum(axis=1, keepdims=True)
            b *= 1.0 + t
            b[:, face] = -t
            x = np.einsum("cj,cjd->cd", b, V)
            lam = bary4(loc["bary"], x)
            m = tol * np.maximum(1.0, np.abs(lam).sum(axis=1))
            pr = proj(loc["planes"], x)
            neg = lam < -m[:, None]
            pos = lam > m[:, None]
            assert np.all(pr[neg] > 0.0), (face, t)
            assert np.all(pr[pos] < 0.0), (face, t)
            decided_neg += int(neg.sum())
            decided_pos += int(pos.sum())
    assert decided_neg > 2 * nc and decided_pos > 8 * nc

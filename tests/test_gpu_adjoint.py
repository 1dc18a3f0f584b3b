"""Adjoint kernels (reverse mode through K steps of Boris push + CG1 deposit)
against central finite differences of the GPU forward and of the oracle's CPU
path, for J = ||rho||^2 and a charge-near-the-object functional, with respect
to the initial velocities, the CG1 field dofs and (through E = G phi) the
vertex potentials.  FD evaluations recompute w on the host from the final
state in a fixed order, so + and - evaluations round alike."""
import math

import numpy as np
import pytest
import torch

import oracle as O
import puncpp_amd as P
from puncpp_amd.adjoint import PushDeposit, gradient_operator, push_deposit
from helpers import MI, ME, QE, QI

pytestmark = pytest.mark.gpu


def _setup(n_mesh=6, npart=(6000, 6000), K=20, every=5, periodic=True, seed=2):
    if periodic:
        hm = P.box_mesh(n_mesh, [0, 0, 0], [2 * np.pi] * 3, periodic=True)
        Ld = [2 * np.pi] * 3
    else:
        hm = P.box_mesh(n_mesh, [-4.0] * 3, [4.0] * 3, sphere_r=1.0)
        Ld = None
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    o.load(0, npart[0], seed, 0.0, 0.0, 0, 1.0, [0, 0, 0.2])
    o.load(1, npart[1], seed, 0.0, 0.0, 0, 0.3, [0, 0, 0.2])
    st = o.get_particles()
    mesh = P.Mesh(hm)
    rng = np.random.default_rng(seed)
    xyz = np.zeros((hm.ndof, 3))
    xyz[hm.vdof] = hm.xyz
    phi = 0.3 * np.sin(xyz[:, 0]) * np.cos(xyz[:, 1]) + 0.05 * rng.normal(size=hm.ndof)
    G = gradient_operator(mesh)
    E = (G @ phi).reshape(-1, 3)
    runner = PushDeposit(mesh, [QE, QI], [ME, MI], st["species"], st["x"], st["cell"], [0, 0, 1.0], 0.05, K,
                         every=every, Ld=Ld)
    return hm, mesh, st, E, phi, G, runner, Ld


def _w_host(hm, mesh, st, xK, cK):
    loc = mesh.localizer()
    alive = cK >= 0
    c = cK[alive]
    x = xK[alive]
    q = np.where(st["species"][alive] == 0, QE, QI)
    bary = loc["bary"].reshape(-1, 3, 4)[c]
    l3 = bary[:, :, 0] + np.einsum("nij,nj->ni", bary[:, :, 1:], x)
    lam = np.concatenate([l3, 1 - l3.sum(axis=1, keepdims=True)], axis=1)
    w = np.zeros(hm.ndof)
    np.add.at(w, loc["dofs"][c].ravel(), (q[:, None] * lam).ravel())
    return w


def _J2(w, dv):
    return math.fsum((dv * w) ** 2)


def test_forward_matches_oracle(gpu):
    print("[case] adjoint forward vs oracle", flush=True)
    hm, mesh, st, E, phi, G, runner, Ld = _setup(npart=(3000, 3000), K=10)
    w, xK, vK, cK = runner.forward(st["v"], E, final_state=True)
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    o.add_located(st["x"], st["v"], st["species"], st["cell"])
    for _ in range(10):
        o.boris(E, [0, 0, 1.0], 0.05)
        o.move_periodic(0.05, Ld)
        o.update()
    w_o = o.distribute(np.ones(hm.ndof))
    assert np.allclose(w, w_o, rtol=1e-11, atol=1e-11 * np.abs(w_o).max())
    assert np.allclose(_w_host(hm, mesh, st, xK, cK), w_o, rtol=1e-11, atol=1e-11 * np.abs(w_o).max())


@pytest.mark.parametrize("target", ["rho2", "near_surface"])
def test_adjoint_vs_central_fd(gpu, target):
    print(f"[case] adjoint vs FD {target}", flush=True)
    periodic = target == "rho2"
    hm, mesh, st, E, phi, G, runner, Ld = _setup(periodic=periodic, n_mesh=6 if periodic else 8)
    dv = P.element_volume(mesh)
    if target == "rho2":
        def J_t(w):
            return ((torch.from_numpy(dv) * w) ** 2).sum()

        def J_np(w):
            return _J2(w, dv)
    else:
        # differentiable surrogate of object-collected charge: charge carried
        # by the CG1 basis functions of the object-surface vertices
        adj = mesh.localizer()["adj"]
        cdofs = hm.vdof[hm.cells]
        mask = np.zeros(hm.ndof)
        for c, f in zip(*np.where(adj == -2)):
            for i in range(4):
                if i != f:  # facet f is opposite local vertex f
                    mask[cdofs[c, i]] = 1.0
        assert mask.sum() > 0

        def J_t(w):
            return (torch.from_numpy(mask) * w).sum()

        def J_np(w):
            return math.fsum(mask * w)

    v0 = torch.tensor(st["v"], requires_grad=True)
    Et = torch.tensor(E, requires_grad=True)
    J = J_t(push_deposit(v0, Et, runner))
    J.backward()
    gv = v0.grad.numpy()
    gE = Et.grad.numpy()
    gphi = G.T @ gE.ravel()

    def J_of(v, e):
        w, xK, vK, cK = runner.forward(v, e, final_state=True)
        return J_np(_w_host(hm, mesh, st, xK, cK)), cK

    rng = np.random.default_rng(2)
    h = 1e-6
    # Directions are unit vectors, so |grad| bounds |<grad, d>|; the error is
    # measured relative to that bound (a directional derivative can be
    # arbitrarily close to zero) and also reported plainly.
    gnorm = {0: np.linalg.norm(gv), 1: np.linalg.norm(gE), 2: np.linalg.norm(gphi)}
    worst = 0.0
    worst_plain = 0.0
    flipped = 0
    for k in range(32):
        kind = k % 3
        if kind == 0:
            d = rng.normal(size=st["v"].shape)
            d /= np.linalg.norm(d)
            Jp, cp = J_of(st["v"] + h * d, E)
            Jm, cm = J_of(st["v"] - h * d, E)
            ad = float(np.sum(gv * d))
        elif kind == 1:
            d = rng.normal(size=E.shape)
            d /= np.linalg.norm(d)
            Jp, cp = J_of(st["v"], E + h * d)
            Jm, cm = J_of(st["v"], E - h * d)
            ad = float(np.sum(gE * d))
        else:
            d = rng.normal(size=phi.shape)
            d /= np.linalg.norm(d)
            Jp, cp = J_of(st["v"], (G @ (phi + h * d)).reshape(-1, 3))
            Jm, cm = J_of(st["v"], (G @ (phi - h * d)).reshape(-1, 3))
            ad = float(np.dot(gphi, d))
        if not np.array_equal(cp >= 0, cm >= 0):
            flipped += 1  # a particle crossed the object boundary: J jumps
            continue
        fd = (Jp - Jm) / (2 * h)
        plain = abs(fd - ad) / max(abs(ad), 1e-300)
        rel = abs(fd - ad) / max(abs(ad), gnorm[kind])
        worst = max(worst, rel)
        worst_plain = max(worst_plain, plain)
        print(f"dir {k} kind {kind}: fd {fd:.12e} ad {ad:.12e} rel {rel:.2e} (plain {plain:.2e})", flush=True)
    print(f"worst relative error {worst:.3e} (plain {worst_plain:.3e}), "
          f"directions excluded for absorption flips: {flipped}")
    assert flipped <= 4
    assert worst <= 1e-6
    assert worst_plain <= 1e-4


def test_adjoint_vs_oracle_fd(gpu):
    """Central differences of the oracle (the serial CPU restatement of the
    reference path: boris_cg1, move_periodic, update, distribute_cg1) against
    the adjoint gradient of J = ||rho||^2, in directions of v0 and of E."""
    print("[case] adjoint vs oracle FD", flush=True)
    hm, mesh, st, E, phi, G, runner, Ld = _setup(npart=(2000, 2000), K=8, every=4)
    dv = P.element_volume(mesh)
    v0 = torch.tensor(st["v"], requires_grad=True)
    Et = torch.tensor(E, requires_grad=True)
    J = ((torch.from_numpy(dv) * push_deposit(v0, Et, runner)) ** 2).sum()
    J.backward()
    gv, gE = v0.grad.numpy(), Et.grad.numpy()

    def J_oracle(v, e):
        o = O.OraclePop(hm, [QE, QI], [ME, MI])
        o.add_located(st["x"], v, st["species"], st["cell"])
        for _ in range(8):
            o.boris(e, [0, 0, 1.0], 0.05)
            o.move_periodic(0.05, Ld)
            o.update()
        rho = o.distribute(dv)
        return math.fsum(rho ** 2)

    rng = np.random.default_rng(5)
    h = 1e-6
    worst = 0.0
    for k in range(8):
        if k % 2 == 0:
            d = rng.normal(size=st["v"].shape)
            d /= np.linalg.norm(d)
            fd = (J_oracle(st["v"] + h * d, E) - J_oracle(st["v"] - h * d, E)) / (2 * h)
            ad, gn = float(np.sum(gv * d)), np.linalg.norm(gv)
        else:
            d = rng.normal(size=E.shape)
            d /= np.linalg.norm(d)
            fd = (J_oracle(st["v"], E + h * d) - J_oracle(st["v"], E - h * d)) / (2 * h)
            ad, gn = float(np.sum(gE * d)), np.linalg.norm(gE)
        rel = abs(fd - ad) / max(abs(ad), gn)
        worst = max(worst, rel)
        print(f"dir {k}: oracle fd {fd:.12e} ad {ad:.12e} rel {rel:.2e}", flush=True)
    print(f"worst relative error vs oracle FD {worst:.3e}")
    assert worst <= 1e-6

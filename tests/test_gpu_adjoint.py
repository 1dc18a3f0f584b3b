"""Adjoint kernels (reverse mode through K steps of Boris push + CG1 deposit)
against central finite differences of the GPU forward and of the oracle's CPU
path.  Objectives: J1 = ||rho||^2 = sum_j (dv_inv_j w_j)^2 of the final
deposit, and J2 = Q_2, the time average over the K steps of the charge the
object (boundary 2) collects (the collected charge's continuous form:
PushDeposit.collected).  Inputs: the initial velocities and the vertex
potentials phi, chained through the device's vector-CG1 E projection
(dJ/dphi = B^T M^-1 dJ/dE).

Errors are relative to the directional derivative itself.  Directions are
d = normalize(r/|r| + g1/|g1| + g2/|g2|) with r random (seed 2), so that
<g, d> is a sizeable part of |g| for both objectives and the random part
still probes every component.  FD evaluations recompute w on the host from the
final state in a fixed order and Q from per-particle weights in particle order,
so + and - evaluations round alike; the projection is linear, so
E(phi +- h d) = E(phi) +- h E(d) (one solve per direction: the CG tolerance
does not enter the difference)."""
import math

import numpy as np
import pytest
import torch

import oracle as O
import puncpp_amd as P
from puncpp_amd.adjoint import PushDeposit, efield_projection, push_deposit, push_deposit_collect
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
    if periodic:
        field = P.FieldSolver(mesh, exterior=False, remove_null_space=True)
        phi = 0.3 * np.sin(xyz[:, 0]) * np.cos(xyz[:, 1]) + 0.05 * rng.normal(size=hm.ndof)
    else:
        field = P.FieldSolver(mesh, exterior=True, n_objects=1)
        # a probe at -1.5 in a grounded box, plus noise
        phi = field.solve(np.zeros(hm.ndof), object_potentials=[-1.5]) + 0.02 * rng.normal(size=hm.ndof)
    E = field.efield(phi)
    runner = PushDeposit(mesh, [QE, QI], [ME, MI], st["species"], st["x"], st["cell"], [0, 0, 1.0], 0.05, K,
                         every=every, Ld=Ld)
    return hm, mesh, st, E, phi, field, runner, Ld


def w_host(hm, mesh, st, xK, cK):
    """w_j = sum_p q_p lambda_j(x_K) in a fixed order (np.add.at)."""
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


def gradients(runner, field, v0, phi, dv):
    """(dJ1/dv0, dJ1/dphi), (dJ2/dv0, dJ2/dphi) through torch autograd."""
    out = []
    for which in (1, 2):
        vt = torch.tensor(v0, requires_grad=True)
        pt = torch.tensor(phi, requires_grad=True)
        E = efield_projection(pt, field)
        w, Q = push_deposit_collect(vt, E, runner)
        J = ((torch.from_numpy(dv) * w) ** 2).sum() if which == 1 else Q[2]
        gv, gp = torch.autograd.grad(J, [vt, pt])
        out.append((gv.numpy(), gp.numpy()))
    return out


def fd_check(runner, field, hm, mesh, st, v0, phi, E, dv, grads, ndir, h=1e-6, seed=2, objectives=(1, 2),
             tol=1e-6, max_excluded=4, label=""):
    """Central differences of the GPU forward along mixed directions; returns
    the worst relative errors and the number of directions excluded because
    a particle's absorption (boundary or step) differs between the two sides."""
    rng = np.random.default_rng(seed)

    def J_of(v, e):
        w, xK, vK, cK = runner.forward(v, e, final_state=True)
        wt, b, stp, _ = runner.collected()
        J1 = math.fsum((dv * w_host(hm, mesh, st, xK, cK)) ** 2)
        J2 = math.fsum(runner.q[runner.species][b == 2] * wt[b == 2])
        return (J1, J2), (b.copy(), stp.copy())

    worst = {k: 0.0 for k in objectives}
    excluded = 0
    for k in range(ndir):
        kind = k % 2  # 0: initial velocities, 1: vertex potentials
        shape = v0.shape if kind == 0 else phi.shape
        r = rng.normal(size=shape)
        d = r / np.linalg.norm(r)
        for j in objectives:
            g = grads[j - 1][kind]
            d = d + g / np.linalg.norm(g)
        d /= np.linalg.norm(d)
        if kind == 0:
            (Jp, ap), (Jm, am) = J_of(v0 + h * d, E), J_of(v0 - h * d, E)
        else:
            Ed = field.efield(d)
            (Jp, ap), (Jm, am) = J_of(v0, E + h * Ed), J_of(v0, E - h * Ed)
        if not (np.array_equal(ap[0], am[0]) and np.array_equal(ap[1], am[1])):
            excluded += 1  # an absorption moved across a step or a boundary: J1 jumps, J2 has a kink
            continue
        for j in objectives:
            fd = (Jp[j - 1] - Jm[j - 1]) / (2 * h)
            ad = float(np.sum(grads[j - 1][kind] * d))
            rel = abs(fd - ad) / abs(ad)
            worst[j] = max(worst[j], rel)
            print(f"{label} dir {k} ({'v0' if kind == 0 else 'phi'}) J{j}: fd {fd:.12e} ad {ad:.12e} rel {rel:.2e}",
                  flush=True)
    print(f"{label} worst relative errors {worst}, directions excluded for absorption changes: {excluded}",
          flush=True)
    assert excluded <= max_excluded
    for j in objectives:
        assert worst[j] <= tol, (j, worst[j])
    return worst, excluded


def test_forward_matches_oracle(gpu):
    print("[case] adjoint forward vs oracle", flush=True)
    hm, mesh, st, E, phi, field, runner, Ld = _setup(npart=(3000, 3000), K=10)
    w, xK, vK, cK = runner.forward(st["v"], E, final_state=True)
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    o.add_located(st["x"], st["v"], st["species"], st["cell"])
    for _ in range(10):
        o.boris(E, [0, 0, 1.0], 0.05)
        o.move_periodic(0.05, Ld)
        o.update()
    w_o = o.distribute(np.ones(hm.ndof))
    assert np.allclose(w, w_o, rtol=1e-11, atol=1e-11 * np.abs(w_o).max())
    assert np.allclose(w_host(hm, mesh, st, xK, cK), w_o, rtol=1e-11, atol=1e-11 * np.abs(w_o).max())


def test_collected_weights_match_oracle_absorption(gpu):
    """The particles the forward records as collected by the object are the
    ones the oracle's update absorbs there, each at its step; weights lie in
    (0, 1] and follow the step: (K - n - s)/K with s in [0, 1)."""
    print("[case] adjoint collected record", flush=True)
    hm, mesh, st, E, phi, field, runner, Ld = _setup(n_mesh=8, npart=(6000, 6000), K=12, every=4, periodic=False)
    runner.forward(st["v"], E)
    wt, b, stp, irregular = runner.collected()
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    o.add_located(st["x"], st["v"], st["species"], st["cell"])
    absorbed = np.zeros((2, 64), dtype=np.int64)
    per_step = []
    for _ in range(12):
        o.boris(E, [0, 0, 1.0], 0.05)
        o.move(0.05)
        ab, _, _ = o.update()
        absorbed += ab
        per_step.append(ab)
    for bnd in (1, 2):
        for s in (0, 1):
            assert np.sum((b == bnd) & (st["species"] == s)) == absorbed[s, bnd]
    for n, ab in enumerate(per_step):
        assert np.sum((b == 2) & (stp == n)) == ab[:, 2].sum()
    m = b > 0
    assert np.all((wt[m] > 0) & (wt[m] <= 1)) and np.all(wt[~m] == 0)
    assert np.all(wt[m] <= (12 - stp[m]) / 12) and np.all(wt[m] > (11 - stp[m]) / 12)
    assert absorbed[:, 2].sum() > 20
    assert irregular <= max(1, absorbed.sum() // 100), irregular


@pytest.mark.parametrize("periodic", [True, False])
def test_adjoint_vs_central_fd(gpu, periodic):
    print(f"[case] adjoint vs FD periodic={periodic}", flush=True)
    hm, mesh, st, E, phi, field, runner, Ld = _setup(periodic=periodic, n_mesh=6 if periodic else 8)
    dv = P.element_volume(mesh)
    grads = gradients(runner, field, st["v"], phi, dv)
    objectives = (1,) if periodic else (1, 2)  # no object on the periodic box
    fd_check(runner, field, hm, mesh, st, st["v"], phi, E, dv, grads, 32, objectives=objectives,
             label=f"periodic={periodic}")


def test_push_deposit_matches_collect_form(gpu):
    """The deposit-only autograd form and the deposit + collected-charge form
    give the same J1 gradient."""
    print("[case] adjoint forms agree", flush=True)
    hm, mesh, st, E, phi, field, runner, Ld = _setup(npart=(2000, 2000), K=6, every=3)
    dv = P.element_volume(mesh)
    v0 = torch.tensor(st["v"], requires_grad=True)
    J = ((torch.from_numpy(dv) * push_deposit(v0, torch.tensor(E), runner)) ** 2).sum()
    J.backward()
    (g1v, _), _ = gradients(runner, field, st["v"], phi, dv)
    assert np.allclose(v0.grad.numpy(), g1v, rtol=1e-9, atol=1e-9 * np.abs(g1v).max())


def test_adjoint_vs_oracle_fd(gpu):
    """Central differences of the oracle (the serial CPU restatement of the
    reference path: boris_cg1, move_periodic, update, distribute_cg1) against
    the adjoint gradient of J1 = ||rho||^2, in directions of v0 and phi."""
    print("[case] adjoint vs oracle FD", flush=True)
    hm, mesh, st, E, phi, field, runner, Ld = _setup(npart=(2000, 2000), K=8, every=4)
    dv = P.element_volume(mesh)
    grads = gradients(runner, field, st["v"], phi, dv)

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
        kind = k % 2
        g = grads[0][kind]
        r = rng.normal(size=g.shape)
        d = r / np.linalg.norm(r) + g / np.linalg.norm(g)
        d /= np.linalg.norm(d)
        if kind == 0:
            fd = (J_oracle(st["v"] + h * d, E) - J_oracle(st["v"] - h * d, E)) / (2 * h)
        else:
            Ed = field.efield(d)
            fd = (J_oracle(st["v"], E + h * Ed) - J_oracle(st["v"], E - h * Ed)) / (2 * h)
        ad = float(np.sum(g * d))
        rel = abs(fd - ad) / abs(ad)
        worst = max(worst, rel)
        print(f"dir {k}: oracle fd {fd:.12e} ad {ad:.12e} rel {rel:.2e}", flush=True)
    print(f"worst relative error vs oracle FD {worst:.3e}")
    assert worst <= 1e-6

"""The adjoint at the size of the adjoint workload: reverse mode through K = 50
steps of Boris push + CG1 deposit with checkpoints every 10 steps, 1M
particles on the sphere-probe mesh ([-10, 10]^3 of 32^3 cubes x 6
tetrahedra, the r = 1 sphere removed: 35,936 vertex potentials; the
configuration's "config-1 mesh" has no object, so the probe mesh carries the
object-collected charge), vertex potentials phi from a probe held at -2 in a
grounded box, E their device projection.  Gradients of J1 = ||rho||^2 and of
the time-averaged collected charge on the probe with respect to the initial
velocities and the vertex potentials, checked by central differences
(h = 1e-6, 32 directions, seed 2) of the GPU forward, errors relative to the
directional derivative; and J1 against central differences of the oracle's
CPU path at 100k particles on the same mesh."""
import math
import time

import numpy as np
import pytest

import oracle as O
import puncpp_amd as P
from puncpp_amd import _native as N
from puncpp_amd.adjoint import PushDeposit
from helpers import MI, ME, QE, QI
from test_gpu_adjoint import fd_check, gradients

pytestmark = pytest.mark.gpu

K, EVERY, DT = 50, 10, 0.02
B = [0.0, 0.0, 1.0]


def _workload(nper, seed=2):
    hm = P.box_mesh(32, [-10.0] * 3, [10.0] * 3, sphere_r=1.0)
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=2 * nper + 1000)
    L = N.lib()
    vd = np.array([0.0, 0.0, 0.5])
    for s, vth in enumerate((1.0, 1.0 / np.sqrt(MI))):
        N.check(L.punc_gpu_pop_load(pop._h, s, nper, seed, 0.0, 0.0, 0, vth, N.dptr(vd)))
    st = pop.get_particles()
    field = P.FieldSolver(mesh, exterior=True, n_objects=1)
    phi = field.solve(np.zeros(hm.ndof), object_potentials=[-2.0])
    E = field.efield(phi)
    return hm, mesh, st, field, phi, E


def test_adjoint_1M_particles_K50(gpu):
    t0 = time.time()
    print("[case] adjoint 1M particles K=50", flush=True)
    hm, mesh, st, field, phi, E = _workload(500_000)
    assert hm.nv == 33 ** 3 - 1 and st["x"].shape[0] == 1_000_000
    runner = PushDeposit(mesh, [QE, QI], [ME, MI], st["species"], st["x"], st["cell"], B, DT, K, every=EVERY)
    dv = P.element_volume(mesh)
    grads = gradients(runner, field, st["v"], phi, dv)
    _, b, _, irregular = runner.collected()
    n_obj = int((b == 2).sum())
    print(f"[adjoint] gradients in {time.time() - t0:.1f}s; {n_obj} particles collected by the probe, "
          f"{int((b == 1).sum())} left through the exterior, irregular exits {irregular}", flush=True)
    assert n_obj > 100
    fd_check(runner, field, hm, mesh, st, st["v"], phi, E, dv, grads, 32, label="1M K=50")
    print(f"[adjoint] done in {time.time() - t0:.1f}s", flush=True)


def test_adjoint_100k_vs_oracle_fd(gpu):
    """J1 against central differences of the oracle's boris_cg1 / move /
    update / distribute_cg1 at 100k particles, K = 50."""
    print("[case] adjoint 100k vs oracle FD", flush=True)
    hm, mesh, st, field, phi, E = _workload(50_000, seed=3)
    runner = PushDeposit(mesh, [QE, QI], [ME, MI], st["species"], st["x"], st["cell"], B, DT, K, every=EVERY)
    dv = P.element_volume(mesh)
    (g1v, g1p), _ = gradients(runner, field, st["v"], phi, dv)

    def J_oracle(v, e):
        o = O.OraclePop(hm, [QE, QI], [ME, MI])
        o.add_located(st["x"], v, st["species"], st["cell"])
        for _ in range(K):
            o.boris(e, B, DT)
            o.move(DT)
            o.update()
        rho = o.distribute(dv)
        return math.fsum(rho ** 2), o.size()

    rng = np.random.default_rng(2)
    h = 1e-6
    worst, used = 0.0, 0
    for k in range(6):
        kind = k % 2
        g = g1v if kind == 0 else g1p
        r = rng.normal(size=g.shape)
        d = r / np.linalg.norm(r) + g / np.linalg.norm(g)
        d /= np.linalg.norm(d)
        if kind == 0:
            (Jp, np_), (Jm, nm) = J_oracle(st["v"] + h * d, E), J_oracle(st["v"] - h * d, E)
        else:
            Ed = field.efield(d)
            (Jp, np_), (Jm, nm) = J_oracle(st["v"], E + h * Ed), J_oracle(st["v"], E - h * Ed)
        if np_ != nm:
            print(f"dir {k}: an absorption differs between the two sides, excluded", flush=True)
            continue
        fd = (Jp - Jm) / (2 * h)
        ad = float(np.sum(g * d))
        rel = abs(fd - ad) / abs(ad)
        worst = max(worst, rel)
        used += 1
        print(f"dir {k} ({'v0' if kind == 0 else 'phi'}): oracle fd {fd:.12e} ad {ad:.12e} rel {rel:.2e}",
              flush=True)
    print(f"worst relative error vs oracle FD {worst:.3e} over {used} directions", flush=True)
    assert used >= 4 and worst <= 1e-6

"""The batched bias sweep at its own size: 64 sphere-probe instances (meshes
n^3 x 6 tetrahedra, n drawn from {16, 20, 24, 28, 32}; bias linspace(-5, 5);
density scale in [0.5, 2], about 164M particles) advanced by one launch per
phase, against every instance advanced as a population of its own (the
reference's one-instance-per-process loop, interaction.cpp:342-423), and four
of the n = 16 instances against the oracle directly.

Per instance and step: absorbed counts exact; after steps 1, 5 and 10 the
particle sets equal bit for bit (order-independent device digests, the
digest itself checked against a numpy digest of downloaded particles);
count balance N' = N - absorbed + injected; rho within 1e-11 of the
per-dof magnitude scale (ne + ni: sum |q| lambda dv_inv) and charge
conserved.  The first five batched steps are eager calls, the last five
replays of a captured step graph."""
import threading
import time

import numpy as np
import pytest

import oracle as O
import puncpp_amd as P
from puncpp_amd import _native as N
from puncpp_amd.batch import BiasSweep
from helpers import ME, MI, QE, QI, digest_np

pytestmark = pytest.mark.gpu

STEPS = 10
CHECK_AT = (0, 4, 9)


def _log(t0, msg):
    print(f"[sweep {time.time() - t0:6.1f}s] {msg}", flush=True)


def _oracle_instance(w, i, pop):
    """The oracle holding instance i's initial particles (the device's
    load_particles matches the oracle's to 1e-12 in v, not bit for bit:
    test_device_load_matches_oracle), with the instance's injectors."""
    o = O.OraclePop(w.host_meshes[i], [QE, QI], [ME, MI])
    st = pop.get_particles()
    o.add_located(st["x"], st["v"], st["species"], st["cell"])
    for s, S in enumerate(w.species[i]):
        o.set_injector(s, S.vdf.vth(), S.vdf.vd(), S.n, 5.0)
    return o


def _oracle_steps(w, i, o, out):
    ab_tot = np.zeros((2, 64), dtype=np.int64)
    for n in range(STEPS):
        o.boris(w.fields[i], w.B, w.dt)
        o.move(w.dt)
        ab, _, lost = o.update()
        assert lost == 0
        ab_tot += ab
        o.inject(w.dt, w.seeds[i], n)
    out[i] = ab_tot


def test_bias_sweep_64_instances(gpu):
    t0 = time.time()
    print("[case] bias_sweep_64", flush=True)
    L = N.lib()
    w = BiasSweep(64, seed=1, base_particles=2_000_000)
    _log(t0, f"workload built: {w.total_particles()} particles")
    plain = [w.plain_population(i) for i in range(w.n_inst)]
    bp = w.batched_population()
    _log(t0, "populations loaded")
    n16 = sorted([i for i in range(w.n_inst) if w.ns[i] == 16], key=w.particles)[:4]
    assert len(n16) == 4
    oracles = {i: _oracle_instance(w, i, plain[i]) for i in n16}
    ab_oracle = {}
    threads = [threading.Thread(target=_oracle_steps, args=(w, i, oracles[i], ab_oracle)) for i in n16]
    for t in threads:
        t.start()
    _log(t0, f"oracle instances {n16} loaded and running")

    # the digest itself against numpy over downloaded particles (one small instance)
    k = int(np.argmin([w.particles(i) for i in range(w.n_inst)]))
    d, c = plain[k].digest()
    dn, cn = digest_np(plain[k].get_particles(), 2)
    assert np.array_equal(d, dn) and np.array_equal(c, cn)

    B = w.B
    ab_gpu = [np.zeros((2, 64), dtype=np.int64) for _ in range(w.n_inst)]
    counts_prev = np.array([p.size() for p in plain])
    for n in range(STEPS):
        for i, pop in enumerate(plain):
            N.check(L.punc_gpu_step(pop._h, 2, w.dt, N.dptr(B), w.dt, None, 1, w.seeds[i], n))
        if n < 5:
            N.check(L.punc_gpu_step(bp.pop._h, 2, w.dt, N.dptr(B), w.dt, None, 1, 0, n))
        else:
            if n == 5:
                N.check(L.punc_gpu_step_graph(bp.pop._h, 2, w.dt, N.dptr(B), w.dt, None, 1, 0, 5))
            N.check(L.punc_gpu_graph_launch(bp.pop._h, 1))
        ab_b = np.zeros((2 * w.n_inst, 64), dtype=np.int64)
        N.check(L.punc_gpu_step_results(bp.pop._h, None, N.i64ptr(ab_b), None))
        counts = np.empty(w.n_inst, dtype=np.int64)
        for i, pop in enumerate(plain):
            ab = np.zeros((2, 64), dtype=np.int64)
            ninj = N.C.c_int64()
            N.check(L.punc_gpu_step_results(pop._h, None, N.i64ptr(ab), N.C.byref(ninj)))
            np.testing.assert_array_equal(ab_b[2 * i:2 * i + 2], ab, err_msg=f"step {n} instance {i}")
            ab_gpu[i] += ab
            counts[i] = pop.size()
            assert counts[i] == counts_prev[i] - ab.sum() + ninj.value, f"count balance, step {n} instance {i}"
        counts_prev = counts
        if n in CHECK_AT:
            db, cb = bp.pop.digest()
            rho_b = np.empty(bp.mesh.ndof)
            N.check(L.punc_gpu_get_rho(bp.pop._h, N.dptr(rho_b)))
            rho_split = bp.split_dofs(rho_b)
            for i, pop in enumerate(plain):
                dp, cp = pop.digest()
                assert np.array_equal(cb[2 * i:2 * i + 2], cp), f"counts, step {n} instance {i}"
                assert np.array_equal(db[2 * i:2 * i + 2], dp), f"particle sets, step {n} instance {i}"
                rho_p = np.empty(w.meshes[i].ndof)
                N.check(L.punc_gpu_get_rho(pop._h, N.dptr(rho_p)))
                ne, ni = P.density_cg1(None, pop, w.species[i], pop.dv_inv)
                scale = ne + ni  # sum |q| lambda_j dv_inv_j (weights 1, |q| = 1)
                assert np.all(np.abs(rho_split[i] - rho_p) <= 1e-11 * scale + 1e-300), f"rho, step {n} instance {i}"
                qsum = float(cp[1] - cp[0])
                assert abs(np.sum(rho_p / pop.dv_inv) - qsum) <= 1e-9 * cp.sum(), f"charge, step {n} instance {i}"
            _log(t0, f"step {n}: 64 instances identical to their plain runs")
    for t in threads:
        t.join()
    _log(t0, "oracle instances done")
    for i in n16:
        np.testing.assert_array_equal(ab_oracle[i], ab_gpu[i], err_msg=f"absorbed vs oracle, instance {i}")
        do, co = digest_np(oracles[i].get_particles(), 2)
        dp, cp = plain[i].digest()
        assert np.array_equal(co, cp) and np.array_equal(do, dp), f"particle set vs oracle, instance {i}"
        db, cb = bp.pop.digest()
        assert np.array_equal(db[2 * i:2 * i + 2], do)
    assert sum(int(a[:, 2].sum()) for a in ab_gpu) > 0, "the probes should absorb particles"
    _log(t0, "done")

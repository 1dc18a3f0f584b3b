#!/usr/bin/env python3
"""Throughput of the particle half of the PUNC PIC step on one MI355X.

Workload (default): the large single-instance configuration - box [-20,20]^3
of 64^3 cubes x 6 tetrahedra with the r=2 sphere object removed (1,572,864
tets before removal), 64M particles (32M electrons q=-1 m=1 vth=1, 32M ions
q=+1 m=1836 vth=1/sqrt(1836), both drifting (0,0,0.5)), Boris push with
B=(0,0,1), dt=0.02, Maxwellian flux injection at the exterior boundary, CG1
deposit each step, cell re-sort every 20 steps (also timed at 12).  The
field is a frozen synthetic CG1 field resident on the GPU (the host field
solve is outside this path).  One step = push + move + relocation walk with
absorption + removal + injection + deposit, all on the device.

Prints one JSON line (metric, value in particle-steps/s, roofline of the fused
particle kernel, the same for the sphere-probe configuration (4M particles),
CPU baseline of the oracle on a bounded sample, per-phase times of the
drop-in calls and of the oracle under the reference's Timer task names).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s & %HBM peak, 1 MI355X; batched sweep vs plain loop"
HBM_PEAK = 8.0e12

CONFIGS = {
    # name: (n, L, r, particles per species, B, description)
    "large": (64, 20.0, 2.0, 32 * 1024 * 1024,
              "box [-20,20]^3, 64^3 cubes x 6 tets, r=2 sphere removed, 64M particles, Boris B=(0,0,1), dt=0.02"),
    "probe": (32, 10.0, 1.0, 2 * 1024 * 1024,
              "box [-10,10]^3, 32^3 cubes x 6 tets, r=1 sphere removed, 4M particles, Boris B=(0,0,1), dt=0.02"),
    "small": (16, 10.0, 1.0, 256 * 1024,
              "box [-10,10]^3, 16^3 cubes x 6 tets, r=1 sphere removed, 512K particles (smoke size)"),
}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_step_main_traffic.json")


def measured_traffic(cfg, tune):
    """HBM bytes per launch of the fused kernel from the committed counter runs
    (FETCH_SIZE and WRITE_SIZE, each scaled by the copy-kernel factor), used
    only when the run has the workload and launch shape those runs measured."""
    if not tune or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        t = json.load(f)
    shape = f"{tune['grid']},{tune['block']},{tune['deposit_mode']}" + (",1" if tune.get("inject_in_step") else "")
    rec = t.get("runs", {}).get(f"{cfg}:{shape}")
    return rec["traffic_bytes"] if rec else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def synthetic_field(hm, seed=1):
    rng = np.random.default_rng(seed)
    xyz = np.zeros((hm.ndof, 3))
    xyz[hm.vdof] = hm.xyz
    r2 = np.maximum((xyz ** 2).sum(1), 1.0)
    # a weak probe-like radial field plus noise
    E = 0.05 * xyz / r2[:, None] ** 1.5 + 0.01 * rng.normal(size=(hm.ndof, 3))
    return np.ascontiguousarray(E)


def run_gpu(cfg, steps, warmup, sort_every, seed=0, autotune=True, launch=None, extras=True, alt_sort_every=0):
    import puncpp_amd as P
    from puncpp_amd import _native as N

    n, L, r, nps, desc = CONFIGS[cfg]
    t0 = time.time()
    hm = P.box_mesh(n, [-L] * 3, [L] * 3, sphere_r=r)
    mesh = P.Mesh(hm)
    log(f"[bench] mesh {cfg}: {mesh.num_cells} cells, {mesh.ndof} dofs, {mesh.num_exterior} exterior facets "
        f"({time.time() - t0:.1f}s)")
    vth = [1.0, 1.0 / np.sqrt(1836.0)]
    vd = [0.0, 0.0, 0.5]
    qm = [(-1.0, 1.0), (1.0, 1836.0)]
    density = nps / mesh.volume()
    species = [P.Species(q, m, density, 1.0, P.ParticleAmountType.phys_per_sim, mesh,
                         P.UniformPosition(mesh), P.Maxwellian(vth[s], vd))
               for s, (q, m) in enumerate(qm)]
    cap = int(2 * nps * 1.05) + 1_000_000
    pop = P.Population(mesh, qm, capacity=cap)
    for s in range(2):
        N.check(N.lib().punc_gpu_pop_load(pop._h, s, nps, seed, 0.0, 0.0, 0, vth[s], N.dptr(np.array(vd))))
    pop.sort()
    P.punc._register_injectors(pop, species)
    E = synthetic_field(hm)
    N.check(N.lib().punc_gpu_set_field(pop._h, N.dptr(E)))
    dv = P.element_volume(mesh)
    N.check(N.lib().punc_gpu_distribute(pop._h, N.dptr(dv), None))
    log(f"[bench] loaded {pop.size()} particles ({time.time() - t0:.1f}s)")
    B = np.array([0.0, 0.0, 1.0])
    dt = 0.02
    lib = N.lib()

    # Launch shape from the autotuner (runs on a snapshot; the state is kept).
    tune = None
    if launch is not None:
        N.check(lib.punc_gpu_set_launch(pop._h, *launch[:3]))
        N.check(lib.punc_gpu_set_inject_in_step(pop._h, launch[3] if len(launch) > 3 else 0))
        tune = {"grid": launch[0], "block": launch[1], "deposit_mode": launch[2], "step_ms": None}
    elif autotune:
        g, b, md, ms = N.C.c_int32(), N.C.c_int32(), N.C.c_int32(), N.C.c_double()
        N.check(lib.punc_gpu_autotune(pop._h, 2, dt, N.dptr(B), dt, None, 1, seed, 3, N.C.byref(g),
                                      N.C.byref(b), N.C.byref(md), N.C.byref(ms)))
        tune = {"grid": g.value, "block": b.value, "deposit_mode": md.value, "step_ms": ms.value}
    if tune is not None:
        # where the Maxwellian injection generator runs: inside the step
        # kernel's grid, or on a second stream beside it
        ins = N.C.c_int32()
        N.check(lib.punc_gpu_get_inject_in_step(pop._h, N.C.byref(ins)))
        tune["inject_in_step"] = bool(ins.value)
        log(f"[bench] launch: {tune}")
    # Step 0 is the half push (interaction.cpp:386,388); later steps replay a
    # captured graph of the full step; the cell re-sort runs between replays.
    N.check(lib.punc_gpu_step(pop._h, 2, 0.5 * dt, N.dptr(B), dt, None, 1, seed, 0))
    N.check(lib.punc_gpu_step_graph(pop._h, 2, dt, N.dptr(B), dt, None, 1, seed, 1))
    k_done = [1]

    def run(nsteps, every):
        for _ in range(nsteps):
            N.check(lib.punc_gpu_graph_launch(pop._h, 1))
            k_done[0] += 1
            if every and k_done[0] % every == 0:
                N.check(lib.punc_gpu_sort(pop._h))

    run(max(warmup - 1, 0), sort_every)
    N.check(lib.punc_gpu_synchronize(pop._h))
    nf = N.C.c_int64()
    N.check(lib.punc_gpu_near_face(pop._h, N.C.byref(nf)))  # resets the count
    n_start = pop.size()
    log(f"[bench] warmup done, {n_start} particles")
    t1 = time.perf_counter()
    run(steps, sort_every)
    N.check(lib.punc_gpu_synchronize(pop._h))
    t2 = time.perf_counter()
    n_end = pop.size()
    # particles each timed step left within 1e-12 of a facet plane of their
    # cell (counted by the relocation, reported only)
    N.check(lib.punc_gpu_near_face(pop._h, N.C.byref(nf)))
    near_face = {"particles": int(nf.value), "steps": steps, "per_particle_step": nf.value / max(1.0, 0.5 * (n_start + n_end) * steps)}
    alt = None
    if alt_sort_every and alt_sort_every != sort_every:
        # the same steps with another re-sort interval, from a fresh sort
        N.check(lib.punc_gpu_sort(pop._h))
        k_done[0] += -k_done[0] % alt_sort_every
        na = pop.size()
        ta = time.perf_counter()
        run(steps, alt_sort_every)
        N.check(lib.punc_gpu_synchronize(pop._h))
        tb = time.perf_counter()
        nb = pop.size()
        alt = {"sort_every": alt_sort_every, "value": 0.5 * (na + nb) * steps / (tb - ta),
               "ms_per_step": 1e3 * (tb - ta) / steps}
        log(f"[bench] re-sort every {alt_sort_every}: {alt['value']:.4g} particle-steps/s, "
            f"{alt['ms_per_step']:.3f} ms per step")
    # Kernel timing pass: one full re-sort cycle of eager steps with HIP events
    # around the fused particle kernel on the population's stream, so that the
    # average covers freshly sorted and aged particle orders alike.
    N.check(lib.punc_gpu_sort(pop._h))
    N.check(lib.punc_gpu_timing(pop._h, 1))
    for k in range(max(sort_every, 1)):
        N.check(lib.punc_gpu_step(pop._h, 2, dt, N.dptr(B), dt, None, 1, seed, k_done[0] + k))
    ms_tot = N.C.c_double()
    nl = N.C.c_int64()
    N.check(lib.punc_gpu_timing_read(pop._h, N.C.byref(ms_tot), N.C.byref(nl)))
    N.check(lib.punc_gpu_timing(pop._h, 0))
    ke = N.C.c_double()
    ab = np.zeros((2, 64), dtype=np.int64)
    kin = N.C.c_int64()
    N.check(lib.punc_gpu_step_results(pop._h, N.C.byref(ke), N.i64ptr(ab), N.C.byref(kin)))
    # Host hand-over per step of the drop-in time loop (not part of `value`):
    # field dofs in (plus the per-cell field copy), charge dofs out.
    rho = np.empty(mesh.ndof)
    t_in = time.perf_counter()
    for _ in range(5):
        N.check(lib.punc_gpu_set_field(pop._h, N.dptr(E)))
    t_mid = time.perf_counter()
    for _ in range(5):
        N.check(lib.punc_gpu_get_rho(pop._h, N.dptr(rho)))
    t_out = time.perf_counter()
    # the same from page-locked host memory (the drop-in layer's Function values)
    pin_E, pin_rho = N.PinnedArray(E.shape), N.PinnedArray((mesh.ndof,))
    pin_E.array[:] = E
    N.check(lib.punc_gpu_set_field(pop._h, N.dptr(pin_E.array)))
    t_pin = time.perf_counter()
    for _ in range(5):
        N.check(lib.punc_gpu_set_field(pop._h, N.dptr(pin_E.array)))
    t_pmid = time.perf_counter()
    for _ in range(5):
        N.check(lib.punc_gpu_get_rho(pop._h, N.dptr(pin_rho.array)))
    t_pout = time.perf_counter()
    assert np.array_equal(pin_rho.array, rho)
    xfer = {"field_in_ms": 1e3 * (t_mid - t_in) / 5, "rho_out_ms": 1e3 * (t_out - t_mid) / 5,
            "field_in_ms_pinned": 1e3 * (t_pmid - t_pin) / 5, "rho_out_ms_pinned": 1e3 * (t_pout - t_pmid) / 5,
            "field_bytes": 24 * mesh.ndof, "rho_bytes": 8 * mesh.ndof,
            "note": "field_in includes the per-cell field copy kernel; *_pinned: from/to punc_gpu_host_alloc memory"}
    del pin_E, pin_rho
    phases = run_dropin_phases(pop, mesh, dv, dt, B, seed, k_done[0] + max(sort_every, 1)) if extras else None
    diag = run_diagnostics(pop, mesh, dv) if extras else None
    flux_mc = run_flux_mc(mesh) if extras else None
    field = run_field_solve(pop, mesh) if extras else None
    elapsed = t2 - t1
    n_avg = 0.5 * (n_start + n_end)
    value = n_avg * steps / elapsed
    kern_ms = ms_tot.value / max(nl.value, 1)
    b_launch = 104.0 * n_avg + 128.0 * mesh.num_cells + 32.0 * mesh.ndof
    achieved = b_launch / (kern_ms * 1e-3)
    log(f"[bench] {steps} steps in {elapsed:.3f}s, {value:.4g} particle-steps/s, fused kernel {kern_ms:.3f} ms "
        f"({achieved / 1e9:.0f} GB/s), injected last step {kin.value}, absorbed {ab.sum()}")
    return {
        "value": value, "ms_per_step": 1e3 * elapsed / steps, "kern_ms": kern_ms, "b_launch": b_launch,
        "near_face": near_face,
        "achieved": achieved, "n_avg": n_avg, "desc": desc, "cells": mesh.num_cells, "ndof": mesh.ndof,
        "hm": hm, "nps": nps, "vth": vth, "vd": vd, "density": density, "tune": tune, "xfer": xfer,
        "diag": diag, "flux_mc": flux_mc, "field": field, "phases": phases, "alt": alt, "cfg": cfg,
    }


# create_flux workload for the non-Maxwellian distributions: KappaCairns
# (kind 3, vth 1, drift (0.2, 0, 0.5), kappa 4, alpha 0.3, range 15) has no
# closed form for either table, so both come from the Monte Carlo.
KC_FLUX = (3, 1.0, [0.2, 0.0, 0.5], 4.0, 0.3, 15.0)


def P_box_mesh(n):
    import puncpp_amd as P
    return P.box_mesh(n, [0.0] * 3, [1.0] * 3)


def run_flux_mc(mesh, n_iter=500000):
    """Device create_flux Monte Carlo over every exterior facet of the mesh,
    the reference's n_iter samples per facet and table (2 tables)."""
    from puncpp_amd import _native as N
    lib = N.lib()
    ne = mesh.num_exterior
    num, pmax = np.empty(ne), np.empty(ne)
    kind, vth, vd, k, a, rr = KC_FLUX

    def go(n):
        N.check(lib.punc_gpu_vdf_flux(mesh._h, kind, vth, N.dptr(np.array(vd)), k, a, rr, 0, 0, n, N.dptr(num),
                                      N.dptr(pmax)))

    go(1000)
    t = time.perf_counter()
    go(n_iter)
    ms = 1e3 * (time.perf_counter() - t)
    samples = 2.0 * ne * n_iter
    log(f"[bench] flux Monte Carlo: {ne} facets x 2 x {n_iter} samples in {ms:.1f} ms "
        f"({samples / ms / 1e6:.3g} G samples/s)")
    return {"facets": ne, "samples": samples, "ms": ms, "samples_per_s": samples / (ms * 1e-3),
            "distribution": "KappaCairns kappa=4 alpha=0.3 drift (0.2,0,0.5) range 15"}


def run_field_solve(pop, mesh, reps=3):
    """Device field solve from the population's deposit: Poisson (grounded
    exterior, the sphere at 0 V) then the E projection, the result left in
    the population's field.  Zero initial guess every call, as the
    reference's Krylov solver starts."""
    import puncpp_amd as P
    t = time.perf_counter()
    f = P.FieldSolver(mesh, eps0=1.0, exterior=True, n_objects=1)
    setup_s = time.perf_counter() - t
    pre = f.preconditioner
    res = {}
    for kind in ("amg", "jacobi"):  # the default (multigrid), then Jacobi for comparison
        f.set_preconditioner(kind)
        f.solve_population(pop, object_potentials=[0.0], copy_out=False)
        t = time.perf_counter()
        for _ in range(reps):
            f.solve_population(pop, object_potentials=[0.0], copy_out=False)
        res[kind] = (1e3 * (time.perf_counter() - t) / reps, f.iterations)
        log(f"[bench] field solve ({kind}) {res[kind][0]:.2f} ms (Poisson {res[kind][1][0]} + E {res[kind][1][1]} "
            f"CG iterations, {mesh.ndof} dofs)")
    f.set_preconditioner("amg")
    ms, its = res["amg"]
    # warm start from the previous solution after one step of the particles:
    # the time loop's case (the reference starts from zero)
    f.set_warm_start(True)
    from puncpp_amd import _native as N
    lib = N.lib()
    B = np.array([0.0, 0.0, 1.0])
    wms = []
    for k in range(reps):
        N.check(lib.punc_gpu_step(pop._h, 2, 0.02, N.dptr(B), 0.02, None, 1, 0, 10000 + k))
        N.check(lib.punc_gpu_synchronize(pop._h))
        t = time.perf_counter()
        f.solve_population(pop, object_potentials=[0.0], copy_out=False)
        wms.append(1e3 * (time.perf_counter() - t))
    log(f"[bench] field solve, warm start after a step: {np.mean(wms):.2f} ms (Poisson {f.iterations[0]} + "
        f"E {f.iterations[1]} CG iterations)")
    # the fused step's potential energy (on since the field solve set phi):
    # the particle kernel's average with and without it, over 6 steps each
    kern = {}
    for on in (True, False):
        if not on:
            N.check(lib.punc_gpu_set_potential(pop._h, None))
        N.check(lib.punc_gpu_sort(pop._h))
        N.check(lib.punc_gpu_timing(pop._h, 1))
        for k in range(6):
            N.check(lib.punc_gpu_step(pop._h, 2, 0.02, N.dptr(B), 0.02, None, 1, 0, 20000 + k))
        ms_tot, nl = N.C.c_double(), N.C.c_int64()
        N.check(lib.punc_gpu_timing_read(pop._h, N.C.byref(ms_tot), N.C.byref(nl)))
        N.check(lib.punc_gpu_timing(pop._h, 0))
        kern["pe_on" if on else "pe_off"] = ms_tot.value / max(nl.value, 1)
    log(f"[bench] fused step kernel with the potential energy {kern['pe_on']:.3f} ms, without {kern['pe_off']:.3f} ms")
    return {"ms": ms, "dofs": mesh.ndof, "poisson_iterations": its[0], "efield_iterations": its[1],
            "warm_start_ms": float(np.mean(wms)), "warm_poisson_iterations": f.iterations[0],
            "warm_efield_iterations": f.iterations[1], "rtol": 1e-12,
            "preconditioner": f"smoothed-aggregation AMG ({pre[1]} levels + dense {pre[2]}) for Poisson, "
                              "Jacobi for the E projection",
            "setup_s": setup_s, "jacobi_ms": res["jacobi"][0], "jacobi_poisson_iterations": res["jacobi"][1][0],
            "step_kernel_ms_with_potential_energy": kern["pe_on"], "step_kernel_ms_without": kern["pe_off"]}


# The reference loop's Timer tasks on the particle path (interaction.cpp:
# 324-334), in loop order.
PHASES = ("distributor", "counting particles", "accelerator", "move", "update", "injector")


def run_dropin_phases(pop, mesh, dv, dt, B, seed, step0, steps=3):
    """Wall ms per call of the drop-in calls of the reference's time loop
    (each returns its host result, so each ends with a synchronize), keyed
    by the reference's Timer task names; the fused, graph-replayed step that
    `value` times does the same work in one launch sequence."""
    from puncpp_amd import _native as N
    lib = N.lib()
    h = pop._h
    rho = np.empty(mesh.ndof)
    counts = np.zeros(2, dtype=np.int64)
    ke = N.C.c_double()
    ab = np.zeros((2, 64), dtype=np.int64)
    ninj = N.C.c_int64()
    tot = dict.fromkeys(PHASES, 0.0)
    for k in range(steps + 1):  # the first pass is untimed
        calls = {
            "distributor": lambda: lib.punc_gpu_distribute(h, N.dptr(dv), N.dptr(rho)),
            "counting particles": lambda: lib.punc_gpu_pop_counts(h, N.i64ptr(counts)),
            "accelerator": lambda: lib.punc_gpu_boris(h, N.dptr(B), dt, N.C.byref(ke)),
            "move": lambda: lib.punc_gpu_move(h, dt),
            "update": lambda: lib.punc_gpu_update(h, N.i64ptr(ab)),
            "injector": lambda: lib.punc_gpu_inject(h, dt, seed, step0 + k, N.C.byref(ninj)),
        }
        for name in PHASES:
            t = time.perf_counter()
            N.check(calls[name]())
            if k:
                tot[name] += 1e3 * (time.perf_counter() - t)
    out = {name: tot[name] / steps for name in PHASES}
    log("[bench] drop-in phases (ms per call): " + ", ".join(f"{k} {v:.3f}" for k, v in out.items()))
    return out


def run_diagnostics(pop, mesh, dv, reps=5):
    """Wall time per call of the particle diagnostics over the whole
    population (each call ends with a stream synchronize; results stay on the
    device except the scalars).  Algorithmic HBM bytes per call: the particle
    fields read (x, y, z, cs: 28 B; statistics: vx, vy, vz, cs), the cell
    records (128 B per cell) and the dof/cell arrays read or written."""
    from puncpp_amd import _native as N
    lib = N.lib()
    n = pop.size()
    nc, nd = mesh.num_cells, mesh.ndof
    phi = np.ones(nd)
    out = N.C.c_double()
    stats = np.zeros(4)

    def timed(f):
        f()
        t = time.perf_counter()
        for _ in range(reps):
            f()
        return 1e3 * (time.perf_counter() - t) / reps

    res = {
        "potential_energy": (timed(lambda: N.check(lib.punc_gpu_potential_energy(pop._h, N.dptr(phi),
                                                                                 N.C.byref(out)))),
                             28.0 * n + 128.0 * nc + 8.0 * nd),
        "density_cg1": (timed(lambda: N.check(lib.punc_gpu_density_cg1(pop._h, N.dptr(dv), 1.0, 1.0, None,
                                                                       None))),
                        28.0 * n + 128.0 * nc + 2 * 8.0 * nd * 3),
        "density_dg0": (timed(lambda: N.check(lib.punc_gpu_density_dg0(pop._h, 1.0, 1.0, None, None))),
                        4.0 * n + 2 * 8.0 * nc * 3 + 8.0 * nc),
        "statistics": (timed(lambda: N.check(lib.punc_gpu_statistics(pop._h, N.dptr(stats)))), 28.0 * n),
    }
    out_d = {"particles": n}
    for k, (ms, b) in res.items():
        out_d[k] = {"ms_per_call": ms, "bytes": b, "GBps": b / (ms * 1e-3) / 1e9}
    log(f"[bench] diagnostics: " + ", ".join(f"{k} {v['ms_per_call']:.3f} ms ({v['GBps']:.0f} GB/s)"
                                             for k, v in out_d.items() if k != "particles"))
    return out_d


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo, as lscpu prints it)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def run_cpu_baseline(res, sample_per_species=1_000_000, steps=3, seed=0):
    """Oracle (serial, 1 core) on a bounded sample: the same mesh and species,
    1M particles per species, `steps` full particle steps."""
    import oracle as O

    hm = res["hm"]
    t0 = time.time()
    o = O.OraclePop(hm, [-1.0, 1.0], [1.0, 1836.0])
    for s in range(2):
        o.load(s, sample_per_species, seed, 0.0, 0.0, 0, res["vth"][s], res["vd"])
        o.set_injector(s, res["vth"][s], res["vd"], res["density"], 5.0)
    E = synthetic_field(hm)
    dv = o.element_volume()
    B = np.array([0.0, 0.0, 1.0])
    dt = 0.02
    log(f"[bench] oracle set up ({time.time() - t0:.1f}s)")
    t1 = time.perf_counter()
    ps = 0
    ph = dict.fromkeys(PHASES, 0.0)

    def timed(name, f):
        t = time.perf_counter()
        f()
        ph[name] += time.perf_counter() - t

    for k in range(steps):
        ps += o.size()
        timed("distributor", lambda: o.distribute(dv))
        timed("counting particles", lambda: (o.num_of_negatives(), o.num_of_positives(), o.size()))
        timed("accelerator", lambda: o.boris(E, B, dt))
        timed("move", lambda: o.move(dt))
        timed("update", lambda: o.update())
        timed("injector", lambda: o.inject(dt, seed, k))
    t2 = time.perf_counter()
    v = ps / (t2 - t1)
    phases_ms = {name: 1e3 * ph[name] / steps for name in PHASES}
    log(f"[bench] oracle {ps} particle-steps in {t2 - t1:.2f}s -> {v:.4g}/s; per step (ms): " +
        ", ".join(f"{k} {x:.1f}" for k, x in phases_ms.items()))
    # field solve: the oracle's assembly + sparse direct solves (Poisson and
    # the E projection) on a 16^3-cube sphere mesh (direct factorisations of
    # the larger meshes take minutes)
    from oracle.field import FieldOracle
    import puncpp_amd as P
    hmf = P.box_mesh(16, [-10.0] * 3, [10.0] * 3, sphere_r=1.0)
    of = O.OraclePop(hmf, [1.0], [1.0])
    fo = FieldOracle(hmf, 1.0, of.localizer()["adj"])
    rho_f = np.random.default_rng(0).normal(size=hmf.ndof)
    t5 = time.perf_counter()
    fo.efield(fo.poisson(rho_f, exterior=True, object_potentials=[0.0]))
    field_s = time.perf_counter() - t5
    log(f"[bench] oracle field solve {field_s:.2f} s ({hmf.ndof} dofs)")
    # create_flux Monte Carlo (KappaCairns: both tables sampled) on a sample
    # of facets: a 2^3-cube box has 48 exterior facets
    ob = O.OraclePop(P_box_mesh(2), [-1.0], [1.0])
    nf = 48
    t3 = time.perf_counter()
    ob.vdf_flux(*KC_FLUX[:6], seed=0, rng_sp=0, n_iter=100000)
    t4 = time.perf_counter()
    flux_rate = nf * 2 * 100000 / (t4 - t3)
    log(f"[bench] oracle flux Monte Carlo {flux_rate:.4g} samples/s")
    return {"value": v, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(), "phases_ms_per_step": phases_ms,
            "phases_particles": ps / steps,
            "flux_mc_samples_per_s": flux_rate,
            "field_solve_s_16cube_mesh": field_s, "field_solve_dofs": hmf.ndof,
            "sample": f"oracle (serial C++ restatement, g++ -O2) on the same mesh with "
                      f"{2 * sample_per_species} particles, {steps} steps of boris+move+update+inject+distribute; "
                      f"host {cpu_model()}, {os.cpu_count()} cores visible, 1 used"}


def run_batched(n_inst=64, steps=20, warmup=3, base_particles=2_000_000, seed=1):
    """Bias sweep: n_inst independent sphere-probe instances, meshes n^3 x 6
    tets with n ~ U{16,20,24,28,32}, bias linspace(-5, 5) shaping each
    instance's (frozen, synthetic) probe field, density scale ~ U[0.5, 2].
    Boris dt=0.02 with injection.

    Three ways to advance all instances by one step:
      plain loop  - per instance, the drop-in call sequence of the reference's
                    time loop: boris (returns KE), move, update (returns the
                    absorbed counts), inject, distribute (returns rho);
      fused loop  - per instance, one fused single-instance step call;
      batched     - all instances in one population (instance-offset tables),
                    one launch per phase, graph-replayed, autotuned.
    The plain populations run `warmup + steps` drop-in steps and then `steps`
    fused steps; the batched population runs the same total, and the
    per-instance particle counts are compared at the end."""
    import puncpp_amd as P
    from puncpp_amd import _native as N
    from puncpp_amd.batch import BatchedPopulation

    rng = np.random.default_rng(seed)
    ns = rng.choice([16, 20, 24, 28, 32], size=n_inst)
    bias = np.linspace(-5.0, 5.0, n_inst)
    scale = rng.uniform(0.5, 2.0, size=n_inst)
    L, r, dt = 10.0, 1.0, 0.02
    vd = [0.0, 0.0, 0.5]
    B = np.array([0.0, 0.0, 1.0])
    lib = N.lib()
    t0 = time.time()
    hms, meshes, species, fields = [], [], [], []
    for i in range(n_inst):
        hm = P.box_mesh(int(ns[i]), [-L] * 3, [L] * 3, sphere_r=r)
        mesh = P.Mesh(hm)
        dens = scale[i] * base_particles / 2 / mesh.volume()
        sp = [P.Species(q, m, dens, 1.0, P.ParticleAmountType.phys_per_sim, mesh, P.UniformPosition(mesh),
                        P.Maxwellian(v, vd))
              for (q, m), v in zip([(-1.0, 1.0), (1.0, 1836.0)], [1.0, 1.0 / np.sqrt(1836.0)])]
        xyz = np.zeros((hm.ndof, 3))
        xyz[hm.vdof] = hm.xyz
        r2 = np.maximum((xyz ** 2).sum(1), r * r)
        E = np.ascontiguousarray(0.05 * bias[i] * xyz / r2[:, None] ** 1.5)
        hms.append(hm)
        meshes.append(mesh)
        species.append(sp)
        fields.append(E)
    total = sum(int(s.num) for sp in species for s in sp)
    log(f"[bench] batched: {n_inst} instances, {total} particles, meshes built ({time.time() - t0:.1f}s)")
    plain, dvs = [], []
    for i in range(n_inst):
        pop = P.Population(meshes[i], [(-1.0, 1.0), (1.0, 1836.0)],
                           capacity=int(sum(s.num for s in species[i]) * 1.1) + 100000)
        for s, S in enumerate(species[i]):
            N.check(lib.punc_gpu_pop_load(pop._h, s, int(S.num), 1000 + i, 0.0, 0.0, 0, S.vdf.vth(),
                                          N.dptr(S.vdf.vd())))
        pop.sort()
        P.punc._register_injectors(pop, species[i])
        N.check(lib.punc_gpu_set_field(pop._h, N.dptr(fields[i])))
        dvs.append(np.ascontiguousarray(P.element_volume(meshes[i])))
        N.check(lib.punc_gpu_distribute(pop._h, N.dptr(dvs[i]), None))
        plain.append(pop)
    bp = BatchedPopulation(hms, species, [1000 + i for i in range(n_inst)], capacity=int(total * 1.1) + 1_000_000)
    bp.load()
    bp.pop.sort()
    bp.register_injectors()
    N.check(lib.punc_gpu_set_field(bp.pop._h, N.dptr(np.concatenate(fields))))
    N.check(lib.punc_gpu_distribute(bp.pop._h, N.dptr(P.element_volume(bp.mesh)), None))
    g, b, md, ms = N.C.c_int32(), N.C.c_int32(), N.C.c_int32(), N.C.c_double()
    N.check(lib.punc_gpu_autotune(bp.pop._h, 2, dt, N.dptr(B), dt, None, 1, 0, 3, N.C.byref(g), N.C.byref(b),
                                  N.C.byref(md), N.C.byref(ms)))
    tune = {"grid": g.value, "block": b.value, "deposit_mode": md.value, "step_ms": ms.value}
    ins = N.C.c_int32()
    N.check(lib.punc_gpu_get_inject_in_step(bp.pop._h, N.C.byref(ins)))
    tune["inject_in_step"] = bool(ins.value)
    log(f"[bench] batched: populations loaded, batched autotune {tune} ({time.time() - t0:.1f}s)")
    ke = N.C.c_double()
    absorbed = np.zeros((2, 64), dtype=np.int64)
    ninj = N.C.c_int64()
    rhos = [np.empty(m.ndof) for m in meshes]

    def dropin_steps(k0, k):
        for n in range(k0, k0 + k):
            for i, pop in enumerate(plain):
                h = pop._h
                N.check(lib.punc_gpu_boris(h, N.dptr(B), dt, N.C.byref(ke)))
                N.check(lib.punc_gpu_move(h, dt))
                N.check(lib.punc_gpu_update(h, N.i64ptr(absorbed)))
                N.check(lib.punc_gpu_inject(h, dt, 1000 + i, n, N.C.byref(ninj)))
                N.check(lib.punc_gpu_distribute(h, N.dptr(dvs[i]), N.dptr(rhos[i])))

    def fused_steps(k0, k):
        for n in range(k0, k0 + k):
            for i, pop in enumerate(plain):
                N.check(lib.punc_gpu_step(pop._h, 2, dt, N.dptr(B), dt, None, 1, 1000 + i, n))
        for pop in plain:
            N.check(lib.punc_gpu_synchronize(pop._h))

    def count(pops):
        return sum(p.size() for p in pops)

    dropin_steps(0, warmup)
    n0 = count(plain)
    t1 = time.perf_counter()
    dropin_steps(warmup, steps)
    t2 = time.perf_counter()
    n1 = count(plain)
    dropin_rate = 0.5 * (n0 + n1) * steps / (t2 - t1)
    t3 = time.perf_counter()
    fused_steps(warmup + steps, steps)
    t4 = time.perf_counter()
    n2 = count(plain)
    fused_rate = 0.5 * (n1 + n2) * steps / (t4 - t3)
    N.check(lib.punc_gpu_step_graph(bp.pop._h, 2, dt, N.dptr(B), dt, None, 1, 0, 0))
    N.check(lib.punc_gpu_graph_launch(bp.pop._h, warmup))
    N.check(lib.punc_gpu_synchronize(bp.pop._h))
    m0 = bp.pop.size()
    t5 = time.perf_counter()
    N.check(lib.punc_gpu_graph_launch(bp.pop._h, steps))
    N.check(lib.punc_gpu_synchronize(bp.pop._h))
    t6 = time.perf_counter()
    m1 = bp.pop.size()
    batched_rate = 0.5 * (m0 + m1) * steps / (t6 - t5)
    N.check(lib.punc_gpu_graph_launch(bp.pop._h, steps))
    N.check(lib.punc_gpu_synchronize(bp.pop._h))
    # identical evolution: per instance and species, the order-independent
    # digest of the particles (bit patterns and instance-local cells) and the
    # count, batched against the instance's own population after the same steps
    dig_b, cnt_b = bp.pop.digest()
    dig_p = np.concatenate([p.digest()[0] for p in plain])
    cnt_p = np.concatenate([p.digest()[1] for p in plain])
    same = bool(np.array_equal(cnt_b, cnt_p))
    same_dig = bool(np.array_equal(dig_b, dig_p))
    log(f"[bench] batched {batched_rate:.4g} vs drop-in loop {dropin_rate:.4g} ({batched_rate / dropin_rate:.2f}x) "
        f"vs fused loop {fused_rate:.4g} ({batched_rate / fused_rate:.2f}x) particle-steps/s; "
        f"per-instance counts identical: {same}, particle digests identical: {same_dig}")
    return {"instances": n_inst, "particles": int(0.5 * (m0 + m1)), "steps": steps,
            "value": batched_rate, "unit": "particle-steps/s",
            "plain_loop_value": dropin_rate, "speedup": batched_rate / dropin_rate,
            "plain_loop": "per instance: boris, move, update, inject, distribute (drop-in calls, host results)",
            "fused_loop_value": fused_rate, "speedup_vs_fused_loop": batched_rate / fused_rate,
            "per_instance_counts_identical": same, "per_instance_digests_identical": same_dig,
            "launch": tune,
            "ms_per_step_batched": 1e3 * (t6 - t5) / steps, "ms_per_step_plain": 1e3 * (t2 - t1) / steps,
            "ms_per_step_fused_loop": 1e3 * (t4 - t3) / steps}


def roofline(res):
    """Roofline of the fused particle kernel: algorithmic bytes per launch
    (104 B per particle, 128 B per cell record, 32 B per dof) over its
    average launch time from HIP events on the population's stream."""
    return {"bound": "hbm", "achieved": res["achieved"] / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": res["achieved"] / HBM_PEAK,
            "traffic": measured_traffic(res["cfg"], res["tune"]),
            "kernel": "k_step_main<boris,move> (push, move, relocation, deposit)",
            "bytes_per_launch": res["b_launch"], "avg_launch_ms": res["kern_ms"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="large", choices=sorted(CONFIGS))
    ap.add_argument("--sort-every", type=int, default=20)
    ap.add_argument("--alt-sort-every", type=int, default=12,
                    help="also time the same steps with this re-sort interval (0: skip)")
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the sphere-probe configuration's line (4M particles)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the diagnostics, flux Monte Carlo and field-solve timings (profiling runs)")
    ap.add_argument("--launch", default=None,
                    help="GRID,BLOCK,MODE[,INSTEP]: fixed launch shape instead of the autotuner (profiling "
                         "runs); INSTEP 1: the injection generator runs inside the step kernel's grid")
    ap.add_argument("--batched-instances", type=int, default=64)
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the streaming-copy reference kernel (1 GiB, 20 copies)")
    args = ap.parse_args()
    if args.copy_probe:
        from puncpp_amd import _native as N
        ms = N.C.c_double()
        nbytes = 1 << 30
        N.check(N.lib().punc_gpu_copy_probe(nbytes, 20, N.C.byref(ms)))
        print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms.value,
                          "GBps_read_plus_write": 2 * nbytes / ms.value / 1e6}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              f"sharding of this path is a design note, not a build target.", file=sys.stderr)
        sys.exit(2)
    launch = tuple(int(v) for v in args.launch.split(",")) if args.launch else None
    res = run_gpu(args.config, args.steps, args.warmup, args.sort_every, autotune=not args.no_autotune,
                  launch=launch, extras=not args.no_extras, alt_sort_every=args.alt_sort_every)
    probe = None
    if not args.no_probe and args.config != "probe":
        pr = run_gpu("probe", args.steps, args.warmup, args.sort_every, autotune=not args.no_autotune,
                      extras=False)
        probe = {"workload": f"probe: {pr['desc']}", "value": pr["value"], "unit": "particle-steps/s",
                 "ms_per_step": pr["ms_per_step"], "particles": int(pr["n_avg"]), "cells": pr["cells"],
                 "dofs": pr["ndof"], "sort_every": args.sort_every, "launch": pr["tune"], "near_face": pr["near_face"],
                 "roofline": roofline(pr)}
    batched = None if args.no_batched else run_batched(args.batched_instances, args.batched_steps)
    cpu = None if args.no_cpu_baseline else run_cpu_baseline(res)
    out = {
        "metric": METRIC,
        "value": res["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {res['desc']}", "cells": res["cells"], "dofs": res["ndof"],
                   "particles": int(res["n_avg"]), "sort_every": args.sort_every, "launch": res["tune"],
                   "field": "frozen synthetic CG1 field on device"},
        "roofline": roofline(res),
        "near_face": res["near_face"],
        "resort_interval_alt": res["alt"],
        "probe": probe,
        "cpu_baseline": cpu,
        "host_transfer": res["xfer"],
        "dropin_phases_ms": res["phases"],
        "diagnostics": res["diag"],
        "flux_monte_carlo": res["flux_mc"],
        "field_solve": res["field"],
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

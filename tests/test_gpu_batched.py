"""Batched instances (one concatenated mesh, one launch per phase) against the
plain single-instance path on the same GPU: every instance must evolve
exactly as it does on its own (particles bit-identical as sets, absorbed
counts exact, rho within the deposit tolerance)."""
import numpy as np
import pytest

import puncpp_amd as P
from puncpp_amd import _native as N
from puncpp_amd.batch import BatchedPopulation
from helpers import MI, ME, QE, QI, assert_same_sets

pytestmark = pytest.mark.gpu


def _instances():
    rng = np.random.default_rng(1)
    out = []
    for i, n in enumerate([4, 6, 5]):
        hm = P.box_mesh(n, [-3.0] * 3, [3.0] * 3, sphere_r=0.8 + 0.1 * i)
        dens = float(rng.uniform(20, 60))
        out.append((hm, dens, 100 + i))
    return out


def _species(mesh, dens):
    vd = [0.0, 0.0, 0.5]
    return [P.Species(QE, ME, dens, 1.0, P.ParticleAmountType.phys_per_sim, mesh, P.UniformPosition(mesh),
                      P.Maxwellian(1.0, vd)),
            P.Species(QI, MI, dens, 1.0, P.ParticleAmountType.phys_per_sim, mesh, P.UniformPosition(mesh),
                      P.Maxwellian(1.0 / np.sqrt(1836.0), vd))]


def _field(hm, i):
    rng = np.random.default_rng(50 + i)
    return np.ascontiguousarray(0.05 * rng.normal(size=(hm.ndof, 3)))


def test_batched_equals_plain(gpu):
    print("[case] batched_equals_plain", flush=True)
    inst = _instances()
    L = N.lib()
    B = np.array([0.0, 0.0, 1.0])
    dt = 0.05
    steps = 4
    # plain: one population per instance
    plain = []
    for i, (hm, dens, seed) in enumerate(inst):
        mesh = P.Mesh(hm)
        sp = _species(mesh, dens)
        pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=200000)
        for s, S in enumerate(sp):
            N.check(L.punc_gpu_pop_load(pop._h, s, int(S.num), seed, 0.0, 0.0, 0, S.vdf.vth(),
                                        N.dptr(S.vdf.vd())))
        P.punc._register_injectors(pop, sp)
        N.check(L.punc_gpu_set_field(pop._h, N.dptr(_field(hm, i))))
        N.check(L.punc_gpu_distribute(pop._h, N.dptr(P.element_volume(mesh)), None))
        plain.append((mesh, sp, pop))
    # batched
    meshes = [hm for hm, _, _ in inst]
    species = [_species(plain[i][0], inst[i][1]) for i in range(len(inst))]
    bp = BatchedPopulation(meshes, species, [seed for _, _, seed in inst], capacity=600000)
    bp.load()
    bp.register_injectors()
    E = np.concatenate([_field(hm, i) for i, hm in enumerate(meshes)])
    N.check(L.punc_gpu_set_field(bp.pop._h, N.dptr(E)))
    N.check(L.punc_gpu_distribute(bp.pop._h, N.dptr(P.element_volume(bp.mesh)), None))
    # identical initial states
    split = bp.split_particles()
    for i, (mesh, sp, pop) in enumerate(plain):
        assert_same_sets(split[i], pop.get_particles(), rel=0.0, what=f"load instance {i}")
    for n in range(steps):
        push_dt = (1.0 - 0.5 * (n == 0)) * dt
        for i, (mesh, sp, pop) in enumerate(plain):
            N.check(L.punc_gpu_step(pop._h, 2, push_dt, N.dptr(B), dt, None, 1, inst[i][2], n))
        N.check(L.punc_gpu_step(bp.pop._h, 2, push_dt, N.dptr(B), dt, None, 1, 0, n))
        ab = np.zeros((bp.pop.nspecies, 64), dtype=np.int64)
        N.check(L.punc_gpu_step_results(bp.pop._h, None, N.i64ptr(ab), None))
        rho_b = np.empty(bp.mesh.ndof)
        N.check(L.punc_gpu_get_rho(bp.pop._h, N.dptr(rho_b)))
        split = bp.split_particles()
        rho_split = bp.split_dofs(rho_b)
        ab_split = bp.split_absorbed(ab)
        for i, (mesh, sp, pop) in enumerate(plain):
            abp = np.zeros((2, 64), dtype=np.int64)
            N.check(L.punc_gpu_step_results(pop._h, None, N.i64ptr(abp), None))
            np.testing.assert_array_equal(ab_split[i], abp)
            assert_same_sets(split[i], pop.get_particles(), rel=0.0, what=f"step {n} instance {i}")
            rho_p = np.empty(mesh.ndof)
            N.check(L.punc_gpu_get_rho(pop._h, N.dptr(rho_p)))
            scale = np.abs(rho_p).max()
            assert np.allclose(rho_split[i], rho_p, rtol=1e-11, atol=1e-11 * scale)

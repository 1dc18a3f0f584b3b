"""GPU parity of the population files (Population::save_file / load_file)
against the oracle: the same particles give the same 64-byte records and the
same "%.17g" text lines (as multisets; both write cell by cell, the order
within a cell may differ); files the oracle wrote load into the GPU
population as the oracle's own particles, bit for bit; a round trip at 8M
particles crosses several device chunks."""
import os
import time

import numpy as np
import pytest

import oracle as O
import puncpp_amd as P
from puncpp_amd import _native as N
from helpers import MI, ME, QE, QI, assert_same_sets, loaded_pair, sphere_mesh

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _sorted_rows(a):
    return a[np.lexsort(a.T[::-1])]


@pytest.fixture(scope="module")
def pair():
    return loaded_pair(sphere_mesh(), (3000, 2000))


def test_save_binary_matches_oracle(gpu, pair, tmp_path):
    _print_case("save_binary")
    o, mesh, pop = pair
    pop.save_file(tmp_path / "g.bin", binary=True)
    o.save_file(tmp_path / "o.bin", binary=True)
    g = np.fromfile(tmp_path / "g.bin", dtype=np.float64).reshape(-1, 8)
    r = np.fromfile(tmp_path / "o.bin", dtype=np.float64).reshape(-1, 8)
    assert g.shape == r.shape == (5000, 8)
    np.testing.assert_array_equal(_sorted_rows(g), _sorted_rows(r))
    cells = mesh.locate(g[:, :3])
    assert np.all(np.diff(cells) >= 0)  # cell order
    # saving re-sorted the population; its particle set is unchanged
    assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what="after save")


def test_save_text_matches_oracle(gpu, pair, tmp_path):
    _print_case("save_text")
    o, mesh, pop = pair
    pop.save_file(tmp_path / "g.txt", binary=False)
    o.save_file(tmp_path / "o.txt", binary=False)
    g = open(tmp_path / "g.txt").read().split("\n")
    r = open(tmp_path / "o.txt").read().split("\n")
    assert g[-1] == "" and len(g) == len(r) == 5001
    assert sorted(g) == sorted(r)


@pytest.mark.parametrize("binary", [True, False])
def test_load_oracle_file(gpu, pair, tmp_path, binary):
    _print_case("load_file", binary=binary)
    o, mesh, _ = pair
    f = tmp_path / ("o.bin" if binary else "o.txt")
    o.save_file(f, binary=binary)
    pop = P.Population(mesh, [], capacity=20000)  # species come from the file
    n = pop.load_file(f, binary=binary)
    assert n == 5000
    assert sorted(zip(pop.q, pop.m)) == sorted([(QE, ME), (QI, MI)])
    o2 = O.OraclePop(mesh.host, [], [])
    assert o2.load_file(f, binary=binary) == 5000
    assert_same_sets(pop.get_particles(), o2.get_particles(), rel=0.0, what="load")


def test_text_load_semantics(gpu, tmp_path):
    _print_case("text_semantics")
    hm = P.box_mesh(2, [0.0] * 3, [1.0] * 3)
    f = tmp_path / "hand.txt"
    f.write_text("0.1 0.2 0.3 1 2 3 -1 1 99 98\n0.4 0.5 0.6\n\n0.7\t0.7\t0.7\t4\t5\t6\t2.5\t7\n5 5 5 0 0 0 -1 1")
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(-1.0, 1.0)], capacity=100)
    assert pop.load_file(f, binary=False) == 4
    o = O.OraclePop(hm, [-1.0], [1.0])
    assert o.load_file(f, binary=False) == 4
    assert_same_sets(pop.get_particles(), o.get_particles(), rel=0.0, what="hand text")


def test_round_trip_large(gpu, tmp_path):
    # 8M particles: two full device chunks of 4M records and one partial
    hm = P.box_mesh(32, [-10.0] * 3, [10.0] * 3, sphere_r=1.0)
    mesh = P.Mesh(hm)
    nps = 4_500_000
    _print_case("round_trip_large", particles=2 * nps)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=2 * nps + 10)
    L = N.lib()
    for s, vth in enumerate((1.0, 0.05)):
        N.check(L.punc_gpu_pop_load(pop._h, s, nps, 3, 0.0, 0.0, 0, vth, N.dptr(np.array([0.0, 0.0, 0.5]))))
    f = tmp_path / "big.bin"
    t0 = time.perf_counter()
    pop.save_file(f, binary=True)
    t1 = time.perf_counter()
    pop2 = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=2 * nps + 10)
    assert pop2.load_file(f, binary=True) == 2 * nps
    t2 = time.perf_counter()
    print(f"save {t1 - t0:.2f} s, load {t2 - t1:.2f} s for {os.path.getsize(f) / 1e9:.2f} GB", flush=True)
    a, b = pop.get_particles(), pop2.get_particles()
    # the load keeps the file's (cell) order: identical arrays
    for k in ("x", "v", "species", "cell"):
        np.testing.assert_array_equal(a[k], b[k])
    assert np.all(np.diff(b["cell"]) >= 0)


def test_failed_load_leaves_population_unchanged(gpu, tmp_path):
    # An error part way through leaves the particles and species as they were,
    # and the population keeps working: a later load that fits succeeds.
    _print_case("failed_load_rollback")
    hm = P.box_mesh(2, [0.0] * 3, [1.0] * 3)
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(-1.0, 1.0)], capacity=64)
    rng = np.random.default_rng(5)
    x0 = rng.uniform(0.05, 0.95, (10, 3))
    assert pop.add_particles(x0, rng.normal(size=(10, 3)), -1.0, 1.0) == 10
    before = pop.get_particles()

    def recs(n, q, m):
        r = np.zeros((n, 8))
        r[:, :3] = rng.uniform(0.05, 0.95, (n, 3))
        r[:, 3:6] = rng.normal(size=(n, 3))
        r[:, 6], r[:, 7] = q, m
        return r

    # 1) capacity: 100 records into 54 free slots, with a new species among them
    f = tmp_path / "big.bin"
    np.concatenate([recs(50, -1.0, 1.0), recs(50, 2.0, 3.0)]).tofile(f)
    with pytest.raises(N.PuncGpuError, match="capacity"):
        pop.load_file(f, binary=True)
    # 2) species limit: more distinct (q, m) pairs than the packed field holds
    f2 = tmp_path / "many.bin"
    np.concatenate([recs(1, 1.0 + k, 1.0) for k in range(200)]).tofile(f2)
    with pytest.raises(N.PuncGpuError, match="species limit"):
        pop.load_file(f2, binary=True)
    assert pop.size() == 10 and list(zip(pop.q, pop.m)) == [(-1.0, 1.0)]
    assert_same_sets(pop.get_particles(), before, rel=0.0, what="after failed loads")
    f3 = tmp_path / "fits.bin"
    extra = np.concatenate([recs(20, -1.0, 1.0), recs(20, 2.0, 3.0)])
    extra.tofile(f3)
    assert pop.load_file(f3, binary=True) == 40
    assert pop.size() == 50 and sorted(zip(pop.q, pop.m)) == [(-1.0, 1.0), (2.0, 3.0)]

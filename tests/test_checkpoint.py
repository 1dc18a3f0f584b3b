"""Checkpoint and run-record formats on the CPU: the oracle's
Population::save_file / load_file restatement (binary records and text
lines), and the State / History writers of the Python mirror and of the C++
drop-in headers, against known-answer text built from the C formatting rules
("%.17g", "%.16e", the default ostream "%g")."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
import puncpp_amd as P
from helpers import MI, ME, QE, QI

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def pop_pair():
    hm = P.box_mesh(4, [-2.0] * 3, [2.0] * 3, sphere_r=0.8)
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    o.load(0, 700, 1, 0.0, 0.0, 0, 1.0, [0.0, 0.0, 0.5])
    o.load(1, 500, 1, 0.0, 0.0, 0, 0.05, [0.0, 0.0, 0.5])
    return hm, o


def _records(st):
    q = np.where(st["species"] == 0, QE, QI)
    m = np.where(st["species"] == 0, ME, MI)
    return np.column_stack([st["x"], st["v"], q, m])


def _sorted_rows(a):
    return a[np.lexsort(a.T[::-1])]


def test_binary_file_layout(pop_pair, tmp_path):
    hm, o = pop_pair
    f = tmp_path / "pop.bin"
    o.save_file(f, binary=True)
    st = o.get_particles()
    assert os.path.getsize(f) == 64 * len(st["x"])
    rec = np.fromfile(f, dtype=np.float64).reshape(-1, 8)
    np.testing.assert_array_equal(_sorted_rows(rec), _sorted_rows(_records(st)))
    # written cell by cell
    cells = [o.locate(r[:3]) for r in rec]
    assert np.all(np.diff(cells) >= 0)


def test_text_file_layout(pop_pair, tmp_path):
    hm, o = pop_pair
    fb, ft = tmp_path / "pop.bin", tmp_path / "pop.txt"
    o.save_file(fb, binary=True)
    o.save_file(ft, binary=False)
    rec = np.fromfile(fb, dtype=np.float64).reshape(-1, 8)
    lines = open(ft).read().split("\n")
    assert lines[-1] == "" and len(lines) - 1 == rec.shape[0]
    pat = re.compile(r"^(?:[^\t]+\t){6}[^\t]+\t [^\t]+\t$")
    for ln, r in zip(lines[:-1], rec):
        assert pat.match(ln), ln
        # "%.17g" of each value, and it reads back exactly
        f = ln.replace("\t ", "\t").split("\t")[:8]
        assert f == ["%.17g" % v for v in r]
        assert [float(v) for v in f] == list(r)


@pytest.mark.parametrize("binary", [True, False])
def test_round_trip(pop_pair, tmp_path, binary):
    hm, o = pop_pair
    f = tmp_path / ("pop.bin" if binary else "pop.txt")
    o.save_file(f, binary=binary)
    o2 = O.OraclePop(hm, [QE, QI], [ME, MI])
    n = o2.load_file(f, binary=binary)
    st, st2 = o.get_particles(), o2.get_particles()
    assert n == len(st["x"])
    np.testing.assert_array_equal(_sorted_rows(_records(st2)), _sorted_rows(_records(st)))
    np.testing.assert_array_equal(np.sort(st2["cell"]), np.sort(st["cell"]))


def test_text_load_semantics(tmp_path):
    # the reference's reader: numbers fill x, v, q, m in turn; missing fields
    # keep the previous line's values (zeros at first); numbers past the
    # eighth are ignored; an empty line repeats the previous particle; points
    # outside the mesh are skipped; a new (q, m) becomes a new species
    hm = P.box_mesh(2, [0.0] * 3, [1.0] * 3)
    f = tmp_path / "hand.txt"
    f.write_text("0.1 0.2 0.3 1 2 3 -1 1 99 98\n"
                 "0.4 0.5 0.6\n"
                 "\n"
                 "0.7\t0.7\t0.7\t4\t5\t6\t2.5\t7\n"
                 "5 5 5 0 0 0 -1 1\n")
    o = O.OraclePop(hm, [-1.0], [1.0])
    assert o.load_file(f, binary=False) == 4
    st = o.get_particles()
    got = _sorted_rows(np.column_stack([st["x"], st["v"], st["species"]]))
    want = _sorted_rows(np.array([[0.1, 0.2, 0.3, 1, 2, 3, 0], [0.4, 0.5, 0.6, 1, 2, 3, 0],
                                  [0.4, 0.5, 0.6, 1, 2, 3, 0], [0.7, 0.7, 0.7, 4, 5, 6, 1]]))
    np.testing.assert_array_equal(got, want)


class _Obj:
    def __init__(self, charge, current, potential):
        self.charge, self.current, self.potential = charge, current, potential

    def get_potential(self):
        return self.potential


def test_state_known_answer(tmp_path):
    f = tmp_path / "state.dat"
    P.State(f).save(41, 0.1, [_Obj(1.5, -2e-7, 0.0)])
    assert f.read_text() == "42\t0.1\t1.5\t-2e-07\n"
    o = _Obj(0.0, 0.0, 0.0)
    assert P.State(f).load([o]) == (42, 0.1)
    assert (o.charge, o.current) == (1.5, -2e-7)
    # two objects: one line per object, and the loader reads the first line
    # only, so the second object reads zeros (as the reference's loader does)
    P.State(f).save(9, 123456.789, [_Obj(1.0, 2.0, 0), _Obj(3.0, 4.0, 0)])
    assert f.read_text() == "10\t123457\t1\t2\n3\t4\n"
    a, b = _Obj(7, 7, 0), _Obj(7, 7, 0)
    assert P.State(f).load([a, b]) == (10, 123457.0)
    assert (a.charge, a.current, b.charge, b.current) == (1.0, 2.0, 0.0, 0.0)


HEADER_1OBJ_STATS = (
    "#:xaxis\tt\n"
    "#:name\tn\tt\tne\tni\tKE\tPE\tV[0]\tI[0]\tQ[0]\tmean_e\tstdev_e\tmean_i\tstdev_i\n"
    "#:long\ttimestep\ttime\t\"number of electrons\"\t\"number of ions\"\t\"kinetic energy\"\t"
    "\"potential energy\"\tvoltage\tcurrent\tcharge\tmean velocity of electrons\t"
    "standard deviation of electron velocities\tmean velocity of ions\tstandard deviation of ion velocities\n"
    "#:units\t1\ts\tm**(-3)\tm**(-3)\tJ\tJ\tV\tA\tC\tm/s\tm/s\tm/s\tm/s\n")


class _StatsPop:
    def statistics(self, stats):
        return np.array([1.25, 0.5, 0.125, float("nan")])


def test_history_known_answer(tmp_path):
    f = tmp_path / "history.dat"
    h = P.History(f, [_Obj(0, 0, 0)], 3, stats=True)
    h.save(3, 0.1, 1000.0, 999.0, 2.5e-3, -1.0 / 3.0, [_Obj(-4.0, 1e-9, 0.25)], _StatsPop())
    h.close()
    line = ("3\t1.0000000000000001e-01\t1.0000000000000000e+03\t9.9900000000000000e+02\t"
            "2.5000000000000001e-03\t-3.3333333333333331e-01\t2.5000000000000000e-01\t"
            "1.0000000000000001e-09\t-4.0000000000000000e+00\t1.2500000000000000e+00\t"
            "5.0000000000000000e-01\t1.2500000000000000e-01\tnan\n")
    assert f.read_text() == HEADER_1OBJ_STATS + line
    # continuing appends lines without a header
    h = P.History(f, [], 3, continue_simulation=True)
    h.save(4, 0.2, 1.0, 2.0, 3.0, 4.0, [], None)
    h.close()
    assert f.read_text().endswith("4\t2.0000000000000001e-01\t1.0000000000000000e+00\t2.0000000000000000e+00\t"
                                  "3.0000000000000000e+00\t4.0000000000000000e+00\n")


CPP = r"""
#include <punc.h>
#include <cmath>
using namespace punc;
struct StatsPop { void statistics(double* s) { s[0] = 1.25; s[1] = 0.5; s[2] = 0.125; s[3] = NAN; } };
int main(int argc, char** argv) {
    std::string dir = argv[1];
    auto o = std::make_shared<Object>(2);
    o->charge = -4.0; o->current = 1e-9; o->set_potential(0.25);
    ObjectVector objs{o};
    {
        History h(dir + "/history.dat", objs, 3, true);
        StatsPop p;
        h.save(3, 0.1, 1000.0, 999.0, 2.5e-3, -1.0 / 3.0, objs, p);
    }
    auto a = std::make_shared<Object>(2), b = std::make_shared<Object>(3);
    a->charge = 1.0; a->current = 2.0; b->charge = 3.0; b->current = 4.0;
    State(dir + "/state.dat").save(9, 123456.789, ObjectVector{a, b});
    std::size_t n; double t;
    State(dir + "/state.dat").load(n, t, ObjectVector{a, b});
    printf("%zu %.17g %g %g %g %g\n", n, t, a->charge, a->current, b->charge, b->current);
    return 0;
}
"""


def test_cpp_state_history_match(tmp_path):
    src = tmp_path / "sh.cpp"
    src.write_text(CPP)
    exe = tmp_path / "sh"
    lib = os.path.join(ROOT, "puncpp_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), str(src), "-L" + lib,
                        "-lpunc_gpu", "-Wl,-rpath," + lib, "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["10", "123457", "1", "2", "0", "0"]
    h = (tmp_path / "history.dat").read_text()
    assert h.startswith(HEADER_1OBJ_STATS)
    assert h.split("\n")[4].split("\t")[:4] == ["3", "1.0000000000000001e-01", "1.0000000000000000e+03",
                                                 "9.9900000000000000e+02"]
    assert h.split("\n")[4].split("\t")[-1] in ("nan", "-nan")
    assert (tmp_path / "state.dat").read_text() == "10\t123457\t1\t2\n3\t4\n"

"""The multigrid setup of the field solve (puncpp_amd/csrc/amg.cpp) on the
host, without a GPU: tests/cpp/amg_check.cpp builds hierarchies of 7-point
Laplacians (Dirichlet, periodic, too small for one) and checks R = P^T, the
Galerkin operators, symmetry, Dirichlet rows outside the aggregates, P 1 = 1
for the singular operator, the coarsest size and the iteration counts of
CG with the V-cycle (<= 30 to 1e-12) against Jacobi."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_amg_hierarchy(tmp_path):
    exe = str(tmp_path / "amg_check")
    src = [os.path.join(ROOT, "tests", "cpp", "amg_check.cpp"), os.path.join(ROOT, "puncpp_amd", "csrc", "amg.cpp")]
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "puncpp_amd", "csrc")] + src +
                       ["-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "amg_check: ok" in r.stdout

"""The drop-in C++ API (include/punc.h) compiles against the C ABI with a
plain g++ and runs the reference's time-loop order on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "dropin_loop.cpp")
LIBDIR = os.path.join(ROOT, "puncpp_amd")


def _compile(out):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-I" + os.path.join(ROOT, "include"), SRC, "-L" + LIBDIR, "-lpunc_gpu", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_dropin_api_compiles(tmp_path):
    _compile(str(tmp_path / "dropin_loop"))


@pytest.mark.gpu
def test_dropin_loop_runs(gpu, tmp_path):
    exe = _compile(str(tmp_path / "dropin_loop"))
    env = dict(os.environ, LD_LIBRARY_PATH=LIBDIR + ":" + os.environ.get("LD_LIBRARY_PATH", ""))
    print("[case] dropin_loop", flush=True)
    r = subprocess.run([exe, "4"], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and r.stdout.strip().endswith("OK")

"""Host C++ under AddressSanitizer and UndefinedBehaviorSanitizer (g++,
host code only, no GPU): the mesh set-up of the product library
(puncpp_amd/csrc/mesh_host.cpp) and the oracle's particle path
(oracle/oracle.cpp) through tests/cpp/host_sanitize.cpp, and the multigrid
set-up (puncpp_amd/csrc/amg.cpp) through tests/cpp/amg_check.cpp.  Any
sanitizer finding aborts the program (-fno-sanitize-recover=all), so a
clean exit with the driver's "ok" line is the pass."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "puncpp_amd", "csrc")
FLAGS = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
         "-fno-sanitize-recover=all", "-I" + CSRC]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _build_run(sources, exe, args, token):
    r = subprocess.run(["g++"] + FLAGS + sources + ["-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=600, env=ENV)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0 and token in r.stdout, r.stderr[-4000:]


def test_mesh_setup_and_oracle_sanitized(tmp_path):
    _build_run([os.path.join(ROOT, "tests", "cpp", "host_sanitize.cpp"), os.path.join(CSRC, "mesh_host.cpp"),
                os.path.join(ROOT, "oracle", "oracle.cpp")], str(tmp_path / "host_sanitize"), [str(tmp_path)],
               "host_sanitize: ok")


def test_multigrid_setup_sanitized(tmp_path):
    _build_run([os.path.join(ROOT, "tests", "cpp", "amg_check.cpp"), os.path.join(CSRC, "amg.cpp")],
               str(tmp_path / "amg_check"), [], "amg_check: ok")

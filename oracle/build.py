"""Build recipe for the oracle (test infrastructure only).

Compiles oracle/oracle.cpp into oracle/liboracle.so with g++ -O2
-ffp-contract=off (no FMA contraction, like the reference's x86-64 build).
The reference itself cannot be compiled here (it needs DOLFIN, Boost, PETSc
and FFC-generated headers), so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    src = os.path.join(HERE, "oracle.cpp")
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
           "-Wall", "-Wno-unused-variable", src, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)

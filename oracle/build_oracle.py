"""Build recipe for the oracle's C restatement: oracle/liboracle.so.

ORACLE — test infrastructure only.  The reference is MATLAB source and has
nothing to compile, so there is no ``oracle/_ref`` build; this compiles only
the project's own C restatement (``sid_oracle.c``).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose: bool = False) -> str:
    src = os.path.join(HERE, "sid_oracle.c")
    out = os.path.join(HERE, "liboracle.so")
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c11",
           "-o", out + ".tmp", src, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

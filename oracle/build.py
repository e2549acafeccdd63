"""Build the CPU oracle (test infrastructure only).

``python -m oracle.build`` compiles ``oracle/gfmc_oracle.c`` with gcc into
``oracle/liboracle.so`` (OpenMP on, FMA contraction off, no fast-math) so
that its FP64 arithmetic is operation-for-operation the plain C evaluation
of each expression.  The reference (PyGFMC) is pure Python/JAX with no
compiled sources, so there is no ``oracle/_ref`` build.  ``build_sanitized``
writes an AddressSanitizer / UndefinedBehaviorSanitizer build of the same
source for the tests.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SRC = os.path.join(HERE, "gfmc_oracle.c")


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", "-Wall", "-Wno-unknown-pragmas", SRC, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"gcc failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


SAN_LIB = os.path.join(HERE, "liboracle_san.so")


def build_sanitized(verbose: bool = False) -> str:
    """The same source with AddressSanitizer and UndefinedBehaviorSanitizer
    (any finding aborts), for running the oracle's tests under the
    sanitizers: load it with PYGFMC_ORACLE_LIB and the ASan runtime preloaded
    (``gcc -print-file-name=libasan.so``)."""
    tmp = SAN_LIB + ".tmp"
    cmd = ["gcc", "-O1", "-g", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-Wall", "-Wno-unknown-pragmas", SRC, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"gcc failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, SAN_LIB)
    return SAN_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

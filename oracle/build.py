"""Build recipe of the oracle's C restatement (test infrastructure only).

    python oracle/build.py      # -> oracle/_build/ess_port

gcc -O2 (the reference's own build passes no -O flag) with contraction of multiply-adds off,
so its arithmetic rounds as the Python oracle's does; portable x86-64 code (no -march), since
the binary also runs on the GPU box's host for the CPU baseline.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "ess_port.c")
OUT = os.path.join(HERE, "_build", "ess_port")


def build_port(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-Wall", "-Wextra", "-pthread", SRC, "-lm",
           "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


def port_path(build_if_missing=True):
    """The built restatement (built here first when missing and gcc is present)."""
    if not os.path.exists(OUT) and build_if_missing:
        build_port()
    return OUT


if __name__ == "__main__":
    print(build_port(verbose="-v" in sys.argv))

"""Build the ncpamodes HIP library (gfx950) in-tree.

    python ncpaprop-release_amd/build.py            # libncpamodes.so next to this file

Every HIP source is compiled with hipcc for gfx950 only; packed-FP32 instruction
selection is disabled (-packed-fp32-ops): only scalar FP32/FP64 VALU forms are emitted.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libncpamodes.so")
SOURCES = ["csrc/ctx.cpp", "csrc/ess.hip", "csrc/wide.hip", "csrc/field2d.hip", "csrc/broadband.hip", "csrc/plan.cpp",
           "host/atmosphere.cpp", "host/profile_parse.cpp", "host/dispersion.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-pthread",
]


def sources():
    return [os.path.join(HERE, s) for s in SOURCES]


def build(verbose=False, jobs=8):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [HIPCC, "-x", "hip", "-c", src, "-o", obj, "-I", CSRC,
               "-I", os.path.join(ROOT, "include")] + FLAGS
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            sys.stderr.write(text)
            raise RuntimeError("hipcc failed: " + " ".join(cmd))
        if verbose and text.strip():
            sys.stderr.write(text)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=gfx950", "-fPIC", "-pthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_cli(objdir, verbose)
    return LIB


def build_cli(objdir, verbose=False):
    """The Modess / WMod / ModBB command lines (host C++ over the C ABI), linked against the library."""
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    common = ["cli/params.cpp", "cli/ess_run.cpp"]
    progs = {"Modess": common + ["cli/modess_main.cpp"],
             "WMod": common + ["cli/wmod_main.cpp"],
             "ModBB": common + ["cli/bb_run.cpp", "cli/modbb_main.cpp"]}
    objs = {}
    procs = []
    for src in sorted({s for v in progs.values() for s in v}):
        obj = os.path.join(objdir, "cli_" + os.path.basename(src) + ".o")
        objs[src] = obj
        cmd = ["g++", "-O2", "-std=c++17", "-c", os.path.join(HERE, src), "-o", obj, "-I", CSRC,
               "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError("g++ failed: " + " ".join(cmd))
    for name, srcs in progs.items():
        exe = os.path.join(bindir, name)
        cmd = ["g++", "-o", exe + ".tmp"] + [objs[s] for s in srcs] + \
              ["-L", HERE, "-lncpamodes", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

"""Build the HIP library (libfemhip.so) in-tree for gfx950 with hipcc.

Every source compiles to its own object in parallel (the stepping kernel is
split into one translation unit per exchange mode), then one link step makes
the shared library.  Objects go to a temporary directory; only the library
is written into the package.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(REPO, "include")
OUT = os.path.join(HERE, "libfemhip.so")

SOURCES = ["api.hip", "scan.hip", "ctx.hip", "batch.hip", "step.hip", "step_plain.hip", "step_coherent.hip",
           "step_fenced.hip", "stream.hip", "plain.hip", "adjoint.hip", "operator.hip", "gl_compat.hip", "noalloc.hip", "surface.cpp"]

# -packed-fp32-ops: keep v_pk_*_f32 out of every code object.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def compile_command(src: str, obj: str):
    return [_hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", os.path.join(CSRC, src), "-o", obj]


def link_command(objs, out: str):
    return [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out, *objs]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)


def build(out: str = OUT, verbose: bool = False, jobs: int | None = None) -> str:
    jobs = jobs or min(len(SOURCES), max(1, min(16, os.cpu_count() or 1)))
    with tempfile.TemporaryDirectory(prefix="femhip_build_") as tmp:
        objs = [os.path.join(tmp, os.path.splitext(s)[0] + ".o") for s in SOURCES]
        with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
            futs = [ex.submit(_run, compile_command(s, o), verbose) for s, o in zip(SOURCES, objs)]
            for f in futs:
                f.result()
        _run(link_command(objs, out + ".tmp"), verbose)
    os.replace(out + ".tmp", out)
    return out


DRIVER_SRC = os.path.join(REPO, "tests", "cpp", "solver_driver.cpp")
DRIVER = os.path.join(REPO, "tests", "cpp", "solver_driver")


def driver_command(out: str = DRIVER, lib_dir: str = HERE):
    """g++ line of the headless Solver-adapter driver (links libfemhip.so,
    found at run time relative to the binary)."""
    rel = os.path.relpath(lib_dir, os.path.dirname(os.path.abspath(out)))
    return ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{INCLUDE}", DRIVER_SRC, f"-L{lib_dir}", "-lfemhip",
            f"-Wl,-rpath,$ORIGIN/{rel}", "-Wl,-rpath-link,/opt/rocm/lib", "-o", out]


def build_driver(out: str = DRIVER, verbose: bool = False) -> str:
    _run(driver_command(out), verbose)
    return out


if __name__ == "__main__":
    print(build(verbose=True))

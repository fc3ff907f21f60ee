"""Builds the gfx950 HIP library ``libqgn.so`` in-tree (and the Fortran host).

``python -m qgn_fortran_amd.build`` or ``build_all()`` from Python.  The
library is compiled with hipcc for gfx950 only; packed-FP32 instructions are
disabled (``-packed-fp32-ops``) for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libqgn.so")
HOST = os.path.join(PKG, "host")
HOST_BIN = os.path.join(HOST, "qgn_driver")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

DEVICE_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build_lib(verbose=False) -> str:
    """Compile csrc/*.hip and csrc/*.cpp into libqgn.so (gfx950)."""
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for name in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, name)
        if name.endswith(".hip"):
            obj = os.path.join(objdir, name + ".o")
            cmds.append([HIPCC, *DEVICE_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", src, "-o", obj])
            objs.append(obj)
        elif name.endswith(".cpp"):
            obj = os.path.join(objdir, name + ".o")
            cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-I", INCLUDE, "-c", src, "-o", obj])
            objs.append(obj)
    # the translation units compile in parallel
    procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True) for cmd in cmds]
    failed = None
    for cmd, pr in zip(cmds, procs):
        out, _ = pr.communicate()
        if pr.returncode != 0 and failed is None:
            sys.stderr.write(out)
            failed = cmd
    if failed:
        raise RuntimeError("build step failed: " + " ".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_host(verbose=False) -> str | None:
    """Compile the Fortran host (iso_c_binding module + driver) against libqgn.so."""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found; Fortran host not built")
        return None
    srcs = [os.path.join(HOST, f) for f in ("qgn_c.f90", "qgn_host.f90", "qgn_driver.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    objdir = os.path.join(HOST, "_build")
    os.makedirs(objdir, exist_ok=True)
    _run([FLANG, "-O2", "-J", objdir, "-I", HOST, *srcs, "-L", PKG, "-lqgn",
          "-Wl,-rpath,$ORIGIN/..", "-o", HOST_BIN], cwd=objdir)
    if verbose:
        print("built", HOST_BIN)
    # the DFTI module and the test program that drives it (tests/fortran)
    check_src = os.path.join(ROOT, "tests", "fortran", "dfti_check.f90")
    if os.path.exists(check_src):
        cdir = os.path.join(ROOT, "tests", "fortran", "_build")
        os.makedirs(cdir, exist_ok=True)
        _run([FLANG, "-O2", "-J", cdir, os.path.join(HOST, "qgn_dfti.f90"), check_src, "-L", PKG, "-lqgn",
              "-Wl,-rpath,$ORIGIN/../../../qgn_fortran_amd", "-o", os.path.join(cdir, "dfti_check")], cwd=cdir)
        if verbose:
            print("built", os.path.join(cdir, "dfti_check"))
    return HOST_BIN


def build_oracle(verbose=False) -> str | None:
    """Compile the oracle's C++/OpenMP restatement (oracle/qgn_cpu.cpp, test
    infrastructure: the CPU check and baseline, never loaded by the product)
    into oracle/libqgn_cpu.so.  It loads MKL at run time."""
    src = os.path.join(ROOT, "oracle", "qgn_cpu.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "oracle", "libqgn_cpu.so")
    _run(["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", src, "-o", out + ".tmp",
          "-ldl"])
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", out)
    return out


def build_all(verbose=False):
    build_lib(verbose)
    build_host(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)

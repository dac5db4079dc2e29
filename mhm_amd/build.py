"""Builds the native libraries of mhm_amd in-tree.

  mhm_amd/libmhmhip.so  HIP kernels + C-ABI (include/mhmhip.h), gfx950
  mhm_amd/libmhmsyn.so  host-only synthetic-domain tools (priority-flood D8)
  tests/native/libmhmcheck.so  host + device evaluation of the chain's log/exp/pow
                               (test-only, built from tests/native/mathcheck.hip)
  mhm_amd/fortran/mhmhip_driver  Fortran host on the iso_c_binding module

Run ``python -m mhm_amd.build`` (or ``__graft_entry__.build()``).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(PKG, "_build")

HIP_LIB = os.path.join(PKG, "libmhmhip.so")
SYN_LIB = os.path.join(PKG, "libmhmsyn.so")

HIP_SOURCES = ["kernels.hip", "api.hip", "network.cpp"]
ARCH = os.environ.get("MHMHIP_ARCH", "gfx950")


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def hip_flags(defines=()):
    return [f"-D{d}" for d in defines] + [
        f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
        # every + - * / rounds on its own, as in the reference (no FMA contraction)
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", CSRC,
    ]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build failed: {os.path.basename(cmd[-1])}")
    return r.stdout


def build_hip(verbose=False, out=None, defines=(), obj_dir=None):
    """Compile and link the HIP library (out: other path, defines: -D macros)."""
    obj = obj_dir or OBJ
    os.makedirs(obj, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        o = os.path.join(obj, src + ".o")
        objs.append(o)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        jobs.append([hipcc] + hip_flags(defines) + ["-c"] + lang + [os.path.join(CSRC, src), "-o", o])
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for log in ex.map(_run, jobs):
            if verbose and log.strip():
                print(log)
    target = out or HIP_LIB
    tmp = target + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, target)
    return target


TEST_NATIVE = os.path.join(ROOT, "tests", "native")
CHECK_LIB = os.path.join(TEST_NATIVE, "libmhmcheck.so")


def build_check():
    """Host + device evaluation of the chain's log/exp/pow (accuracy tests)."""
    obj = os.path.join(OBJ, "mathcheck.hip.o")
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    _run([hipcc] + hip_flags() + ["-c", "-x", "hip", os.path.join(TEST_NATIVE, "mathcheck.hip"),
                                  "-o", obj])
    tmp = CHECK_LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, obj])
    os.replace(tmp, CHECK_LIB)
    return CHECK_LIB


def build_syn():
    tmp = SYN_LIB + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-o", tmp,
          os.path.join(CSRC, "synth.cpp")])
    os.replace(tmp, SYN_LIB)
    return SYN_LIB


FORTRAN_DIR = os.path.join(PKG, "fortran")
FORTRAN_DRIVER = os.path.join(FORTRAN_DIR, "mhmhip_driver")


def build_fortran():
    """iso_c_binding module + the Fortran driver, linked to libmhmhip.so."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        return None
    obj = os.path.join(OBJ, "fortran")
    os.makedirs(obj, exist_ok=True)
    mods = ["mo_mhmhip.f90", "mhmhip_driver.f90"]
    objs = []
    for m in mods:
        o = os.path.join(obj, m + ".o")
        _run([fc, "-O2", "-J", obj, "-c", os.path.join(FORTRAN_DIR, m), "-o", o])
        objs.append(o)
    tmp = FORTRAN_DRIVER + ".tmp"
    _run([fc] + objs + ["-L" + PKG, "-lmhmhip", "-Wl,-rpath,$ORIGIN/..", "-o", tmp])
    os.replace(tmp, FORTRAN_DRIVER)
    return FORTRAN_DRIVER


def build_all(verbose=False):
    build_syn()
    build_hip(verbose)
    build_check()
    build_fortran()


if __name__ == "__main__":
    build_all(verbose=True)
    print("built", HIP_LIB, SYN_LIB)

"""Builds the native libraries in-tree.

* ``hirensi_amd/libhirensi.so``: HIP kernels + C ABI, gfx950, hipcc.
* ``oracle/liboracle.so``: CPU oracle (test infrastructure), g++.
* ``oracle/_ref/libhr_ref.so``: the reference's element layer, built with
  ROCm flang from its sources, only where ``/root/reference`` exists.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhirensi.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HR_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    # packed FP32 forms are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip(extra=None):
    srcs = [os.path.join(CSRC, f) for f in ("hr_kernels.hip", "hr_stage3.hip", "hr_reg.hip", "hr_api.hip",
                                            "hr_revolve.cpp", "hr_noalloc.hip", "hr_nn.hip")]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, *(extra or []), "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


FLANG = "/opt/rocm/lib/llvm/bin/flang"
FDRIVER = os.path.join(HERE, "fortran", "hr_driver")


def build_fortran():
    """The Fortran driver of the C ABI (hirensi_amd/fortran/hr_driver), linked
    against libhirensi.so next to it (rpath $ORIGIN/..), where ROCm flang exists."""
    if not os.path.exists(FLANG):
        print("ROCm flang absent: Fortran driver not built", flush=True)
        return False
    import tempfile
    fdir = os.path.join(HERE, "fortran")
    with tempfile.TemporaryDirectory() as tmp:
        _run([FLANG, "-O2", "-c", os.path.join(fdir, "hirensi_mod.F90"), "-o", "m.o"], cwd=tmp)
        _run([FLANG, "-O2", "-c", os.path.join(fdir, "hr_driver.F90"), "-o", "d.o"], cwd=tmp)
        rt = os.path.join(os.path.dirname(os.path.dirname(FLANG)), "lib")
        tmpexe = FDRIVER + ".tmp"
        _run([FLANG, "d.o", "m.o", "-o", tmpexe, f"-L{HERE}", "-lhirensi", "-Wl,-rpath,$ORIGIN/..",
              f"-Wl,-rpath,{rt}"], cwd=tmp)
        os.replace(tmpexe, FDRIVER)
    return True


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_reference():
    ref = "/root/reference/step_4_3D_Forward_and_Inverse"
    flang = "/opt/rocm/lib/llvm/bin/flang"
    if not (os.path.isdir(ref) and os.path.exists(flang)):
        print("reference sources or flang absent: oracle/_ref not built", flush=True)
        return False
    _run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "ref.mk")])
    return True


def build_all():
    build_hip()
    build_fortran()
    build_oracle()
    build_reference()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "oracle": build_oracle, "ref": build_reference,
         "fortran": build_fortran}[w]()

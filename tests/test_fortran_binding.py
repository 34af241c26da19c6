"""The ISO_C_BINDING module compiles with ROCm flang, and a Fortran program
linked against libhirensi.so calls the C ABI (host-only entry point, no GPU)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"

PROG = """
program t
  use hirensi
  implicit none
  integer :: i
  do i = 0, 8
    write(*, '(ES24.16)') hr_time_function(0.0125d0 * i, 2_c_int, 15.0d0)
  end do
end program t
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="ROCm flang absent")
def test_fortran_module_calls_the_c_abi(tmp_path):
    from hirensi_amd import native
    from oracle import oracle as O
    mod = os.path.join(ROOT, "hirensi_amd", "fortran", "hirensi_mod.F90")
    (tmp_path / "t.f90").write_text(PROG)
    libdir = os.path.dirname(native.LIB_PATH)
    subprocess.run([FLANG, "-c", mod, "-o", "m.o"], cwd=tmp_path, check=True)
    subprocess.run([FLANG, "t.f90", "m.o", "-o", "t", f"-L{libdir}", "-lhirensi", f"-Wl,-rpath,{libdir}"],
                   cwd=tmp_path, check=True)
    out = subprocess.run(["./t"], cwd=tmp_path, check=True, capture_output=True, text=True,
                         stdin=subprocess.DEVNULL).stdout.split()
    vals = np.array([float(v) for v in out])
    ref = np.array([O.time_function(0.0125 * i, 2, 15.0) for i in range(9)])
    assert np.allclose(vals, ref, rtol=1e-15, atol=1e-300)

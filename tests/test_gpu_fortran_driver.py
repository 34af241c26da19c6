"""A Fortran program drives the C ABI in the reference driver's order
(hr_ctx_create, hr_set_pml, hr_set_material, hr_forward, hr_adjoint_gradient;
Program_Main.F90:318-413) through the ISO_C_BINDING module, on the GPU, and
matches the Python mirror bit for bit on the same inputs (runs on the GPU;
the program is built by build() with ROCm flang)."""
import os
import subprocess

import numpy as np
import pytest

from hirensi_amd import solver as S
from hirensi_amd.build import FDRIVER
from tests.helpers import small_case

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(FDRIVER), reason="Fortran driver not built (ROCm flang absent)")
def test_fortran_driver_matches_python(tmp_path):
    mesh, model, rec_eq, src_eq, src_val = small_case(nex=8, ney=7, nez=6, npml=2, seed=13)
    nstep, dt, fr = 150, 1e-4, 60.0
    ftab = S.time_table(nstep, dt, 2, fr)
    rng = np.random.default_rng(4)
    dmeas = rng.standard_normal((nstep + 1, len(rec_eq))) * 1e-8
    ctx = S.Context(mesh, model)
    tr = S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, src_val, rec_eq).copy()
    gl, gm, mis = S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, rec_eq, dmeas)
    ctx.close()
    head = np.array([mesh.nex, mesh.ney, mesh.nez, *mesh.npml, 1 if mesh.top_stress_bc == "reference" else 0,
                     nstep, len(src_eq), len(rec_eq), 0], dtype=np.int32)
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(head.tobytes())
        f.write(np.array([*mesh.h, dt], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(np.asarray(mesh.pml_param()).T, dtype=np.float64).tobytes())   # (6,4) column-major
        for a in (model.lam, model.mu, model.rho, ftab):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(src_eq, dtype=np.int64).tobytes())
        f.write(np.ascontiguousarray(src_val, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(rec_eq, dtype=np.int64).tobytes())
        f.write(np.ascontiguousarray(dmeas, dtype=np.float64).tobytes())
    print("fortran driver case: running hr_driver", flush=True)
    r = subprocess.run([FDRIVER, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=300, stdin=subprocess.DEVNULL)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    ntr = (nstep + 1) * len(rec_eq)
    assert out.size == ntr + 2 * mesh.nnode + 1
    assert np.abs(tr).max() > 0 and np.abs(gm).max() > 0
    assert np.array_equal(out[:ntr].reshape(nstep + 1, len(rec_eq)), tr)
    assert np.array_equal(out[ntr:ntr + mesh.nnode], gl)
    assert np.array_equal(out[ntr + mesh.nnode:ntr + 2 * mesh.nnode], gm)
    assert out[-1] == mis

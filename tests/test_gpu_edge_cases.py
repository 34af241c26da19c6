"""Edge cases of the C ABI on the GPU: empty source and receiver sets, a shot
without sources, the smallest meshes, and argument errors that the host side
turns into error codes before any launch (HrError in the Python mirror)."""
import numpy as np
import pytest

from hirensi_amd import solver as S
from hirensi_amd.mesh import BoxMesh, receiver_dofs, surface_grid_nodes
from oracle import oracle as O
from tests.helpers import random_model, rel_l2, small_case

pytestmark = pytest.mark.gpu


def test_no_sources_gives_zero_traces_and_no_receivers_runs():
    mesh, model, rec_eq, src_eq, src_val = small_case(seed=6)
    print("edge case: empty source / receiver sets", flush=True)
    ctx = S.Context(mesh, model)
    nstep, dt = 40, 1e-4
    ftab = S.time_table(nstep, dt, 2, 60.0)
    tr = S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, np.zeros(0, np.int64), np.zeros(0), rec_eq)
    assert tr.shape == (nstep + 1, len(rec_eq)) and np.all(tr == 0.0)
    tr0 = S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, src_val, np.zeros(0, np.int64))
    assert tr0.shape[1] == 0
    gl, gm, j = S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, np.zeros(0, np.int64), np.zeros((nstep + 1, 0)))
    assert np.all(gl == 0.0) and np.all(gm == 0.0) and j == 0.0


def test_shot_without_sources_and_smallest_mesh():
    mesh = BoxMesh(3, 3, 2, npml=(1, 1, 1, 1, 1))
    model = random_model(mesh, seed=2)
    print(f"edge case: smallest mesh neq={mesh.neq}", flush=True)
    src_eq = np.array([mesh.dof(mesh.surface_node(0.0, 0.0), 2)], dtype=np.int64)
    rec_eq = receiver_dofs(mesh, surface_grid_nodes(mesh, 2, 2, -0.5, 0.5, -0.5, 0.5))
    nstep, dt, fr = 60, 1e-4, 60.0
    ctx = S.Context(mesh, model)
    ftab = np.stack([S.time_table(nstep, dt, 2, fr)] * 2)
    tr = S.explicit_rk4_petsc_pml_3d_shots(ctx, nstep, dt, ftab, np.array([1], np.int32), src_eq, np.ones(1), rec_eq)
    assert np.all(tr[0] == 0.0)                                  # shot 0 has no source
    B = np.zeros(mesh.neq)
    B[src_eq] = 1.0
    U = O.System(mesh, model).forward(B, nstep, dt, lcase=2, fr=fr)
    pos = {int(e): r for r, e in enumerate(mesh.store_eq)}
    ref = U[:, [pos[int(e)] for e in rec_eq]]
    assert np.abs(ref).max() > 0 and rel_l2(tr[1], ref) < 1e-10


def test_argument_errors_are_reported():
    mesh, model, rec_eq, src_eq, src_val = small_case(seed=1)
    print("edge case: argument errors", flush=True)
    ctx = S.Context(mesh, model)
    nstep, dt = 10, 1e-4
    ftab = S.time_table(nstep, dt, 2, 60.0)
    stress_eq = int(mesh.ID[mesh.ID[:, 3] >= 0][0, 3])         # a stress equation
    with pytest.raises(S.HrError):
        S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, np.array([mesh.neq], np.int64), src_val, rec_eq)
    with pytest.raises(S.HrError):
        S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, src_val, np.array([stress_eq], np.int64))
    with pytest.raises(S.HrError):
        S.explicit_rk4_petsc_pml_3d(ctx, nstep, -dt, ftab, src_eq, src_val, rec_eq)
    S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, src_val, rec_eq, history=S.HR_HISTORY_NONE)
    with pytest.raises(S.HrError):                              # no history kept
        S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, rec_eq, np.zeros((nstep + 1, len(rec_eq))))
    with pytest.raises(S.HrError):
        S.regularization(ctx, model.lam, model.mu, "TV", wreg=np.array([50.0, 1.0, 0.0]))
    with pytest.raises(S.HrError):
        ctx._check(ctx.L.hr_set_tiling(ctx.h, 2, 4, 3))        # tiles are 4 x 3 elements
    # the context still works after the errors
    tr = S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, src_val, rec_eq)
    assert np.isfinite(tr).all()

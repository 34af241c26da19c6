"""hr_misfit_noalloc (the misfit with its partial sums in a fixed device array,
for callers that cannot allocate) against hr_misfit on the same forward: equal
bit for bit (same kernels, same summation order), on two calls in a row, the
second with a smaller input; sizes beyond the fixed scratch are refused."""
import numpy as np
import pytest

from hirensi_amd import solver as S
from hirensi_amd.mesh import BoxMesh, receiver_dofs, surface_grid_nodes
from tests.helpers import random_model

pytestmark = pytest.mark.gpu


def _lib():
    from hirensi_amd import native
    return native.load()


def _case(nshot, nstep, seed):
    import torch
    dev = torch.device("cuda:0")
    mesh = BoxMesh(6, 6, 5, npml=(2,) * 5)
    model, true = random_model(mesh, seed=seed), random_model(mesh, seed=seed + 1)
    rec_eq = receiver_dofs(mesh, surface_grid_nodes(mesh, 3, 2, -1.0, 1.0, -1.0, 1.0))
    rec_eq = np.concatenate([rec_eq, rec_eq[:1]])                # a shared equation too
    src_eq = np.array([mesh.dof(mesh.surface_node(0.0, 0.0), 2)] * nshot, dtype=np.int64)
    T = lambda a, dt_=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt_, device=dev)
    dt = 1e-4
    ftab = T(np.stack([S.time_table(nstep, dt, 2, 40.0 + 10.0 * q) for q in range(nshot)]))
    ctx = S.Context(mesh, device_ptrs=True)
    shots = T(np.arange(nshot), torch.int32)
    args = (ftab, shots, T(src_eq, torch.int64), T(np.linspace(1.0, 2.0, nshot)), T(rec_eq, torch.int64))
    obs = torch.zeros((nshot, nstep + 1, len(rec_eq)), dtype=torch.float64, device=dev)
    tr = torch.zeros_like(obs)
    ctx.set_material_ptrs(T(true.lam), T(true.mu), T(true.rho))
    S.explicit_rk4_petsc_pml_3d_shots(ctx, nstep, dt, *args, history=S.HR_HISTORY_NONE, traces=obs, nshot=nshot)
    ctx.set_material_ptrs(T(model.lam), T(model.mu), T(model.rho))
    S.explicit_rk4_petsc_pml_3d_shots(ctx, nstep, dt, *args, history=S.HR_HISTORY_NONE, traces=tr, nshot=nshot)
    return ctx, obs, len(rec_eq), dt


def _both(ctx, L, nstep, dt, nrec, obs):
    import torch
    a = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    b = torch.full((1,), -1.0, dtype=torch.float64, device="cuda:0")
    ctx._check(L.hr_misfit(ctx.h, nstep, dt, nrec, obs.data_ptr(), a.data_ptr()))
    ctx._check(L.hr_misfit_noalloc(ctx.h, nstep, dt, nrec, obs.data_ptr(), b.data_ptr()))
    torch.cuda.synchronize()
    return float(a.item()), float(b.item())


def test_noalloc_misfit_matches_misfit_twice():
    L = _lib()
    print("no-alloc misfit: 3 shots x 300 steps, then 2 shots x 90 steps", flush=True)
    ctx, obs, nrec, dt = _case(3, 300, seed=81)
    a, b = _both(ctx, L, 300, dt, nrec, obs)
    assert a > 0 and a == b
    ctx.close()
    ctx, obs, nrec, dt = _case(2, 90, seed=83)
    a2, b2 = _both(ctx, L, 90, dt, nrec, obs)
    assert a2 > 0 and a2 == b2 and a2 != a
    ctx.close()


def test_noalloc_misfit_refuses_what_the_scratch_cannot_hold():
    L = _lib()
    # 17 shots x (61680 + 1) steps + 17 > 2^20 partial sums; the check precedes any launch
    nshot, nstep = 17, 61680
    print(f"no-alloc misfit size check: {nshot} shots x {nstep} steps on the smallest mesh", flush=True)
    import torch
    dev = torch.device("cuda:0")
    mesh = BoxMesh(3, 3, 2, npml=(1,) * 5)
    model = random_model(mesh, seed=2)
    T = lambda a, dt_=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt_, device=dev)
    ctx = S.Context(mesh, device_ptrs=True)
    ctx.set_material_ptrs(T(model.lam), T(model.mu), T(model.rho))
    rec_eq = T(np.array([mesh.dof(mesh.surface_node(0.0, 0.0), 2)]), torch.int64)
    ftab = torch.zeros((nshot, nstep, 3), dtype=torch.float64, device=dev)
    tr = torch.zeros((nshot, nstep + 1, 1), dtype=torch.float64, device=dev)
    S.explicit_rk4_petsc_pml_3d_shots(ctx, nstep, 1e-4, ftab, T(np.zeros(0), torch.int32), T(np.zeros(0), torch.int64),
                                      T(np.zeros(0)), rec_eq, history=S.HR_HISTORY_NONE, traces=tr, nshot=nshot)
    out = torch.zeros(1, dtype=torch.float64, device=dev)
    rc = L.hr_misfit_noalloc(ctx.h, nstep, 1e-4, 1, tr.data_ptr(), out.data_ptr())
    assert rc != 0 and b"fixed scratch" in L.hr_last_error(ctx.h)
    # a host-pointer context is refused as well
    hctx = S.Context(mesh, model)
    S.explicit_rk4_petsc_pml_3d(hctx, 5, 1e-4, S.time_table(5, 1e-4), np.zeros(0, np.int64), np.zeros(0),
                                np.array([mesh.dof(mesh.surface_node(0.0, 0.0), 2)]), history=S.HR_HISTORY_NONE)
    m = np.zeros(1)
    rc = L.hr_misfit_noalloc(hctx.h, 5, 1e-4, 1, np.zeros(6).ctypes.data, m.ctypes.data)
    assert rc != 0 and b"HR_DEVICE_PTRS" in L.hr_last_error(hctx.h)
    ctx.close()
    hctx.close()

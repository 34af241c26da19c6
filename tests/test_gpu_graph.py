"""Captured time loops (hr_set_graph): the forward and adjoint time loops run
as replays of a hipGraph of K RK4 steps whose per-step values (source row,
trace and history rows, adjoint misfit rows) come from a device-side step
index.  The same kernels run in the same order on the same streams, so traces,
gradients and misfits must equal the eager launches bit for bit.  Each graph is
replayed many times inside a solve, solves with graphs alternate with eager
solves on other materials (and the program's own synchronisations), and every
one is checked against its eager twin."""
import numpy as np
import pytest

from hirensi_amd import solver as S
from hirensi_amd.mesh import BoxMesh, surface_grid_nodes, receiver_dofs
from tests.helpers import random_model

pytestmark = pytest.mark.gpu


def _case():
    mesh = BoxMesh(10, 9, 7, npml=(2, 2, 2, 2, 2))
    rec_eq = receiver_dofs(mesh, surface_grid_nodes(mesh, 3, 3, -2.5, 2.5, -2.0, 2.0))
    shots = [(mesh.surface_node(-1.0, 0.0), 2, 45.0), (mesh.surface_node(1.5, -1.0), 0, 60.0)]
    return mesh, rec_eq, shots


def _solve(ctx, mesh, model, rec_eq, shots, nstep, dt, dobs, history=S.HR_HISTORY_STORE):
    nshot = len(shots)
    ftab = np.stack([S.time_table(nstep, dt, 2, fr) for _, _, fr in shots])
    src_eq = np.array([mesh.dof(nd, c) for nd, c, _ in shots], dtype=np.int64)
    ctx.set_material(model)
    tr = S.explicit_rk4_petsc_pml_3d_shots(ctx, nstep, dt, ftab, np.arange(nshot, dtype=np.int32), src_eq,
                                           np.ones(nshot), rec_eq, history=history, segment=16).copy()
    if dobs is None:
        return tr, None
    gl, gm, j = S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, rec_eq, dobs)
    return tr, (gl.copy(), gm.copy(), j)


@pytest.mark.parametrize("K,nstep", [(4, 37), (2, 20), (16, 61)])
def test_graph_time_loops_equal_eager_launches(K, nstep):
    mesh, rec_eq, shots = _case()
    dt = 1.0e-4
    print(f"graph case K={K} nstep={nstep} nshot={len(shots)} neq={mesh.neq}", flush=True)
    models = [random_model(mesh, seed=s) for s in (31, 32, 33)]
    eager = S.Context(mesh, models[0])
    dobs, _ = _solve(eager, mesh, models[2], rec_eq, shots, nstep, dt, None, history=S.HR_HISTORY_NONE)
    ref = [_solve(eager, mesh, m, rec_eq, shots, nstep, dt, dobs) for m in models[:2]]
    graph = S.Context(mesh, models[0])
    graph.set_graph(K)
    n0 = graph.graph_launches()
    for rnd in range(2):
        for i, m in enumerate(models[:2]):
            tr, (gl, gm, j) = _solve(graph, mesh, m, rec_eq, shots, nstep, dt, dobs)
            tr_e, (gl_e, gm_e, j_e) = ref[i]
            assert np.abs(tr).max() > 0 and np.abs(gm).max() > 0
            assert np.array_equal(tr, tr_e), (rnd, i)
            assert np.array_equal(gl, gl_e) and np.array_equal(gm, gm_e), (rnd, i)
            assert j == j_e
            # an eager solve in between (other material, no history) on the same context
            graph.set_graph(0)
            _solve(graph, mesh, models[2], rec_eq, shots, nstep, dt, None, history=S.HR_HISTORY_NONE)
            graph.set_graph(K)
    # forward: nstep // K replays; adjoint (steps nstep-1 .. 1): (nstep - 1) // K replays
    per_solve = nstep // K + (nstep - 1) // K
    assert graph.graph_launches() - n0 == 4 * per_solve
    eager.close()
    graph.close()


def test_graph_mode_leaves_other_loops_eager_and_rejects_odd_counts():
    mesh, rec_eq, shots = _case()
    nstep, dt = 40, 1.0e-4
    print("graph mode: checkpointed history stays eager; argument checks", flush=True)
    m0, m1 = random_model(mesh, seed=41), random_model(mesh, seed=42)
    eager = S.Context(mesh, m0)
    dobs, _ = _solve(eager, mesh, m1, rec_eq, shots, nstep, dt, None, history=S.HR_HISTORY_NONE)
    tr_e, (gl_e, gm_e, _) = _solve(eager, mesh, m0, rec_eq, shots, nstep, dt, dobs, history=S.HR_HISTORY_CHECKPOINT)
    graph = S.Context(mesh, m0)
    graph.set_graph(4)
    tr, (gl, gm, _) = _solve(graph, mesh, m0, rec_eq, shots, nstep, dt, dobs, history=S.HR_HISTORY_CHECKPOINT)
    assert np.array_equal(tr, tr_e) and np.array_equal(gl, gl_e) and np.array_equal(gm, gm_e)
    assert graph.graph_launches() == 0
    for bad in (3, -2, 5000):
        with pytest.raises(S.HrError):
            graph.set_graph(bad)
    eager.close()
    graph.close()

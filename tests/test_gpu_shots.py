"""Batched shots (hr_forward_shots) against the same shots solved one at a
time through the single-shot path: traces bit-identical (the same kernels
and summation order per shot), gradients and misfits summed over shots to
round-off (the batched path sums shots inside each step)."""
import numpy as np
import pytest

from hirensi_amd import solver as S
from hirensi_amd.mesh import BoxMesh, surface_grid_nodes, receiver_dofs
from tests.helpers import random_model, rel_l2

pytestmark = pytest.mark.gpu


def _case():
    mesh = BoxMesh(10, 9, 7, npml=(2, 2, 2, 2, 2))
    model = random_model(mesh, seed=21)
    true = random_model(mesh, seed=22)
    rec_eq = receiver_dofs(mesh, surface_grid_nodes(mesh, 3, 3, -2.5, 2.5, -2.0, 2.0))
    shots = [(mesh.surface_node(-1.0, 0.0), 2, 45.0), (mesh.surface_node(1.5, -1.0), 0, 60.0),
             (mesh.surface_node(0.0, 1.5), 1, 75.0)]
    return mesh, model, true, rec_eq, shots


@pytest.mark.parametrize("history,segment", [(S.HR_HISTORY_STORE, 64), (S.HR_HISTORY_CHECKPOINT, 23)])
def test_batched_shots_match_single_shots(history, segment):
    mesh, model, true, rec_eq, shots = _case()
    nstep, dt = 180, 1.0e-4
    nshot = len(shots)
    ftab = np.stack([S.time_table(nstep, dt, 2, fr) for _, _, fr in shots])
    src_eq = np.array([mesh.dof(nd, c) for nd, c, _ in shots], dtype=np.int64)
    src_shot = np.arange(nshot, dtype=np.int32)
    one = np.ones(nshot)
    print(f"shots case history={history} nshot={nshot} neq={mesh.neq}", flush=True)
    # observations from the true model, one shot at a time
    ct = S.Context(mesh, true)
    dobs = np.stack([S.explicit_rk4_petsc_pml_3d(ct, nstep, dt, ftab[q], src_eq[q:q + 1], one[:1], rec_eq,
                                                 history=S.HR_HISTORY_NONE).copy() for q in range(nshot)])
    # single shots
    c1 = S.Context(mesh, model)
    tr1, gl1, gm1, j1 = [], np.zeros(mesh.nnode), np.zeros(mesh.nnode), 0.0
    for q in range(nshot):
        tr1.append(S.explicit_rk4_petsc_pml_3d(c1, nstep, dt, ftab[q], src_eq[q:q + 1], one[:1], rec_eq,
                                               history=history, segment=segment).copy())
        gl, gm, j = S.explicit_rk4_petsc_pml_3d_adjoint_otd(c1, nstep, dt, rec_eq, dobs[q])
        gl1 += gl
        gm1 += gm
        j1 += j
    # batched
    cb = S.Context(mesh, model)
    trb = S.explicit_rk4_petsc_pml_3d_shots(cb, nstep, dt, ftab, src_shot, src_eq, one, rec_eq, history=history,
                                            segment=segment)
    for q in range(nshot):
        assert np.array_equal(trb[q], tr1[q]), q
    glb, gmb, jb = S.explicit_rk4_petsc_pml_3d_adjoint_otd(cb, nstep, dt, rec_eq, dobs)
    assert np.abs(gm1).max() > 0
    assert rel_l2(glb, gl1) < 1e-13 and rel_l2(gmb, gm1) < 1e-13
    assert abs(jb - j1) <= 1e-13 * abs(j1)
    assert abs(S.compute_misfit(cb, nstep, dt, dobs) - j1) <= 1e-13 * abs(j1)

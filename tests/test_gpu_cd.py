"""Central-difference variant (hr_forward_cd) against the oracle's
EXPLICIT_Central_Difference_3 restatement on the same seeded inputs (runs on
the GPU).  Bar: receiver traces <= 1e-10 relative L2 (FP64; the HIP path
applies M^-1 after the operator sum instead of pre-scaling the matrices)."""
import numpy as np
import pytest

from hirensi_amd import solver as S
from oracle import oracle as O
from tests.helpers import rel_l2, small_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,nstep,top", [(3, 300, "reference"), (5, 200, "none")])
def test_cd_traces_match_oracle(seed, nstep, top):
    mesh, model, rec_eq, src_eq, src_val = small_case(seed=seed, top=top)
    dt, fr = 1.0e-4, 60.0
    print(f"central-difference case seed={seed} nstep={nstep} top={top} neq={mesh.neq}", flush=True)
    ctx = S.Context(mesh, model)
    tr = S.explicit_central_difference_3(ctx, nstep, dt, src_eq, src_val, rec_eq, lcase=2, fr=fr)
    sysm = O.System(mesh, model)
    B = np.zeros(mesh.neq)
    np.add.at(B, src_eq, src_val)
    U = sysm.cd_forward(B, nstep, dt, lcase=2, fr=fr)
    pos = {int(e): r for r, e in enumerate(mesh.store_eq)}
    ref = U[:, [pos[int(e)] for e in rec_eq]]
    assert np.abs(ref).max() > 0
    assert rel_l2(tr, ref) < 1e-10
    # an RK4 solve on the same context still runs afterwards (state buffers restored)
    tr4 = S.explicit_rk4_petsc_pml_3d(ctx, 50, dt, S.time_table(50, dt, 2, fr), src_eq, src_val, rec_eq)
    ref4 = sysm.forward(B, 50, dt, lcase=2, fr=fr)[:, [pos[int(e)] for e in rec_eq]]
    assert rel_l2(tr4, ref4) < 1e-10

"""HIP path against the CPU oracle on the same seeded inputs (runs on the GPU).

Bars: operator applies <= 1e-12 relative (max error over max |A||x|), receiver
traces and lambda/mu gradients <= 1e-10 relative L2 (the path computes in
FP64; differences are summation order only).
"""
import numpy as np
import pytest

from hirensi_amd import solver as S
from hirensi_amd.mesh import BoxMesh
from oracle import oracle as O
from tests.helpers import small_case, rel_l2

pytestmark = pytest.mark.gpu

OPS = {S.HR_K: 0, S.HR_C: 1, S.HR_G: 2}
# stage-kernel configurations: (version, tile y-extent); 4 = split stage path
# (nodal PML pass and pull beside the tiled regular-domain kernel), 3 = fused tile classes
KCFG = [(3, 3), (4, 3)]


def _abs_bound(A, x, y):
    return np.abs(A).dot(np.abs(x)) + np.abs(y)


@pytest.mark.parametrize("shape,npml,top", [((8, 8, 6), 2, "reference"), ((7, 9, 5), (1, 2, 2, 1, 2), "none"),
                                            ((13, 10, 11), 3, "reference")])
@pytest.mark.parametrize("transpose", [0, 1])
@pytest.mark.parametrize("kcfg", KCFG[:1])   # single-operator applies always use the fused kernel
def test_apply_matches_oracle_csr(shape, npml, top, transpose, kcfg):
    npml = (npml,) * 5 if isinstance(npml, int) else npml
    mesh = BoxMesh(*shape, npml=npml, top_stress_bc=top)
    from tests.helpers import random_model
    model = random_model(mesh, seed=3)
    print(f"apply case shape={shape} npml={npml} top={top} transpose={transpose} kernel={kcfg}", flush=True)
    ctx = S.Context(mesh, model, version=kcfg[0], ty=kcfg[1])
    sysm = O.System(mesh, model)
    rng = np.random.default_rng(11)
    for which, w in OPS.items():
        A = sysm.csr(w)
        At = A.T.tocsr() if transpose else A
        x = rng.standard_normal(mesh.neq)
        y = rng.standard_normal(mesh.neq)
        z = S.ShellMat(ctx, which)._apply(transpose, x, y, None)
        zr = y + At.dot(x)
        err = np.max(np.abs(z - zr) / _abs_bound(At, x, y))
        assert err < 1e-13, (which, err)


def test_mass_matches_oracle():
    mesh, model, *_ = small_case()
    ctx = S.Context(mesh, model)
    m = ctx.diag_m()
    mr = O.System(mesh, model).diag_m()
    assert np.max(np.abs(m - mr) / mr) < 1e-13


def test_shell_scaling_chain():
    """MatDiagonalScale/MatScale recorded by the shell: the forward chain
    A <- -M^-1 A and the adjoint chain A <- diag(M) A diag(M^-1)."""
    mesh, model, *_ = small_case(seed=5)
    ctx = S.Context(mesh, model)
    sysm = O.System(mesh, model)
    M = sysm.diag_m()
    iM = 1.0 / M
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(mesh.neq), rng.standard_normal(mesh.neq)
    for which, w in OPS.items():
        sh = S.ShellMat(ctx, which)
        sh.diagonal_scale(iM, None)
        sh.scale(-1.0)
        A = sysm.csr(w)
        Af = -(A.multiply(iM[:, None])).tocsr()
        z = sh.mult_add(x, y)
        assert np.max(np.abs(z - (y + Af.dot(x))) / _abs_bound(Af, x, y)) < 1e-13
        if which != S.HR_K:
            sh.scale(-1.0)
            Af = -Af
        sh.diagonal_scale(M, iM)
        Aa = (Af.multiply(M[:, None]).multiply(iM[None, :])).tocsr()
        z = sh.mult_transpose_add(x, y)
        ref = y + Aa.T.dot(x)
        assert np.max(np.abs(z - ref) / _abs_bound(Aa.T.tocsr(), x, y)) < 1e-13
        sh.zero_entries()
        z = sh.mult_add(x, y)
        assert np.max(np.abs(z - (y + A.dot(x))) / _abs_bound(A, x, y)) < 1e-13


def test_dot_product_identity():
    mesh, model, *_ = small_case(nex=9, ney=7, nez=6, seed=7)
    ctx = S.Context(mesh, model)
    rng = np.random.default_rng(4)
    z0 = np.zeros(mesh.neq)
    for which in OPS:
        sh = S.ShellMat(ctx, which)
        x, y = rng.standard_normal(mesh.neq), rng.standard_normal(mesh.neq)
        lhs = np.dot(sh.mult_add(x, z0), y)
        rhs = np.dot(x, sh.mult_transpose_add(y, z0))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def _forward_both(mesh, model, rec_eq, src_eq, src_val, nstep, dt, fr=15.0, tz=0, kcfg=(0, 0)):
    ctx = S.Context(mesh, model, tz=tz, version=kcfg[0], ty=kcfg[1])
    ftab = S.time_table(nstep, dt, 2, fr)
    tr = S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, src_val, rec_eq)
    sysm = O.System(mesh, model)
    B = np.zeros(mesh.neq)
    np.add.at(B, src_eq, src_val)
    store = mesh.store_eq
    U = sysm.forward(B, nstep, dt, lcase=2, fr=fr)
    pos = {int(e): r for r, e in enumerate(store)}
    rows = np.array([pos[int(e)] for e in rec_eq])
    return ctx, sysm, tr, U, rows


@pytest.mark.parametrize("tz", [0, 1, 3])
@pytest.mark.parametrize("kcfg", KCFG)
def test_forward_traces_match_oracle(tz, kcfg):
    mesh, model, rec_eq, src_eq, src_val = small_case()
    nstep, dt = 300, 1.0e-4
    print(f"forward case tz={tz} kernel={kcfg} neq={mesh.neq} nstep={nstep}", flush=True)
    ctx, sysm, tr, U, rows = _forward_both(mesh, model, rec_eq, src_eq, src_val, nstep, dt, fr=60.0, tz=tz,
                                           kcfg=kcfg)
    ref = U[:, rows]
    assert np.abs(ref).max() > 0
    assert rel_l2(tr, ref) < 1e-10


def test_forward_config1_geometry():
    """Config-1 mesh (16x16x8, PML 4, homogeneous) at a reduced step count."""
    from hirensi_amd.mesh import config1, receiver_dofs
    c = config1()
    mesh, model = c["mesh"], c["model"]
    rec_eq = receiver_dofs(mesh, c["rec_nodes"])
    src_eq = np.array([mesh.dof(c["src_node"], 2)])
    nstep = 200
    print(f"config1 forward neq={mesh.neq} nstep={nstep}", flush=True)
    ctx, sysm, tr, U, rows = _forward_both(mesh, model, rec_eq, src_eq, np.array([1.0]), nstep, c["dt"], fr=60.0)
    assert rel_l2(tr, U[:, rows]) < 1e-10


def _observed(mesh, rec_eq, src_eq, src_val, nstep, dt, fr, seed):
    from tests.helpers import random_model
    true = random_model(mesh, seed=seed)
    sysm = O.System(mesh, true)
    B = np.zeros(mesh.neq)
    np.add.at(B, src_eq, src_val)
    U = sysm.forward(B, nstep, dt, lcase=2, fr=fr)
    pos = {int(e): r for r, e in enumerate(mesh.store_eq)}
    rows = np.array([pos[int(e)] for e in rec_eq])
    return U[:, rows].copy()


@pytest.mark.parametrize("history,kcfg", [(S.HR_HISTORY_STORE, (3, 3)), (S.HR_HISTORY_CHECKPOINT, (3, 3)),
                                          (S.HR_HISTORY_STORE, (4, 3)), (S.HR_HISTORY_CHECKPOINT, (4, 3))])
def test_adjoint_gradient_matches_oracle(history, kcfg):
    mesh, model, rec_eq, src_eq, src_val = small_case(seed=1)
    nstep, dt, fr = 240, 1.0e-4, 60.0
    print(f"adjoint case history={history} kernel={kcfg} neq={mesh.neq}", flush=True)
    dmeas = _observed(mesh, rec_eq, src_eq, src_val, nstep, dt, fr, seed=9)
    ctx = S.Context(mesh, model, version=kcfg[0], ty=kcfg[1])
    ftab = S.time_table(nstep, dt, 2, fr)
    S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, src_val, rec_eq, history=history, segment=37)
    gl, gm, mis = S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, rec_eq, dmeas)
    if history == S.HR_HISTORY_CHECKPOINT:
        assert ctx.recompute_ratio() > 1.5
    # oracle
    sysm = O.System(mesh, model)
    B = np.zeros(mesh.neq)
    np.add.at(B, src_eq, src_val)
    U = sysm.forward(B, nstep, dt, lcase=2, fr=fr)
    P = sysm.adjoint(U, nstep, dt, rec_eq, dmeas)
    glr, gmr = sysm.control(U, P, nstep, dt)
    pos = {int(e): r for r, e in enumerate(mesh.store_eq)}
    rows = np.array([pos[int(e)] for e in rec_eq])
    misr = O.misfit(U, nstep, dt, rows, dmeas)
    assert np.abs(glr).max() > 0 and np.abs(gmr).max() > 0
    assert rel_l2(gl, glr) < 1e-10
    assert rel_l2(gm, gmr) < 1e-10
    assert abs(mis - misr) <= 1e-12 * abs(misr)
    assert abs(S.compute_misfit(ctx, nstep, dt, dmeas) - misr) <= 1e-12 * abs(misr)


def test_checkpoint_gradient_equals_stored():
    mesh, model, rec_eq, src_eq, src_val = small_case(nex=9, ney=8, nez=7, seed=2)
    nstep, dt, fr = 150, 1.0e-4, 60.0
    dmeas = _observed(mesh, rec_eq, src_eq, src_val, nstep, dt, fr, seed=4)
    out = []
    for hist, seg in ((S.HR_HISTORY_STORE, 0), (S.HR_HISTORY_CHECKPOINT, 16), (S.HR_HISTORY_CHECKPOINT, 150)):
        ctx = S.Context(mesh, model)
        S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, S.time_table(nstep, dt, 2, fr), src_eq, src_val, rec_eq,
                                    history=hist, segment=max(seg, 1))
        out.append(S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, rec_eq, dmeas))
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1], out[0][1])


@pytest.mark.parametrize("shape,npml", [((5, 5, 4), (2, 2, 2, 2, 1)), ((9, 7, 6), (1, 3, 2, 2, 2))])
@pytest.mark.parametrize("kcfg", KCFG)
def test_forward_adjoint_uneven_boxes(shape, npml, kcfg):
    """A regular domain one element wide (rows of PML nodes on both sides of a
    one-node gap: the split path's x-line neighbours) and uneven PML, with
    receivers on the regular domain's surface nodes, its edges (interface nodes:
    adjoint sources in the tile kernel and in the nodal pull) included, and a
    second source in the PML."""
    from tests.helpers import random_model
    from hirensi_amd.mesh import receiver_dofs
    mesh = BoxMesh(*shape, npml=npml, top_stress_bc="reference")
    model = random_model(mesh, seed=6)
    lo, hi, top = mesh.rd_lo, mesh.rd_hi, mesh.NZ - 1
    rec_nodes = [mesh.node_index(i, j, top) for j in range(2 * lo[1], 2 * hi[1] + 1)
                 for i in range(2 * lo[0], 2 * hi[0] + 1, 2 if hi[0] - lo[0] > 2 else 1)]
    rec_eq = receiver_dofs(mesh, rec_nodes)
    src_eq = np.array([mesh.dof(mesh.surface_node(0.0, 0.0), 2), mesh.dof(mesh.node_index(1, 1, top), 0)],
                      dtype=np.int64)
    src_val = np.array([1.0, 0.5])
    nstep, dt, fr = 160, 1.0e-4, 60.0
    print(f"uneven box {shape} npml={npml} kernel={kcfg} neq={mesh.neq}", flush=True)
    dmeas = _observed(mesh, rec_eq, src_eq, src_val, nstep, dt, fr, seed=9)
    ctx = S.Context(mesh, model, version=kcfg[0], ty=kcfg[1])
    ftab = S.time_table(nstep, dt, 2, fr)
    tr = S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, src_val, rec_eq, history=S.HR_HISTORY_STORE)
    gl, gm, mis = S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, rec_eq, dmeas)
    sysm = O.System(mesh, model)
    B = np.zeros(mesh.neq)
    np.add.at(B, src_eq, src_val)
    U = sysm.forward(B, nstep, dt, lcase=2, fr=fr)
    P = sysm.adjoint(U, nstep, dt, rec_eq, dmeas)
    glr, gmr = sysm.control(U, P, nstep, dt)
    pos = {int(e): r for r, e in enumerate(mesh.store_eq)}
    rows = np.array([pos[int(e)] for e in rec_eq])
    assert np.abs(U[:, rows]).max() > 0 and np.abs(glr).max() > 0
    assert rel_l2(tr, U[:, rows]) < 1e-10
    assert rel_l2(gl, glr) < 1e-10
    assert rel_l2(gm, gmr) < 1e-10

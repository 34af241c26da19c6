"""Gradient post-processing on the GPU against the oracle (runs on the GPU).

hr_gradient_mass, hr_regularization (TV and TN, weighted and unweighted) and
hr_regularize_gradient against the oracle's Assem_R_Mg_3D / TV / TN
restatements and a numpy statement of Program_Main.F90:416-481 on top of them.
Bars (FP64, summation order differs): vectors <= 1e-12 relative L2, costs and
factors <= 1e-12 relative, 1/Mg <= 1e-14 relative, restriction exact.
"""
import numpy as np
import pytest

from hirensi_amd import solver as S
from hirensi_amd.mesh import BoxMesh
from oracle import oracle as O
from tests.helpers import random_model, rel_l2

pytestmark = pytest.mark.gpu

WREG = np.array([50.0, 1.0, 45.0])
MESHES = [((7, 6, 5), (2, 1, 1, 2, 2), (1.0, 1.25, 0.8)), ((12, 10, 9), (3, 3, 3, 3, 3), (1.0, 1.0, 1.0))]


def _case(k, seed=12):
    shape, npml, h = MESHES[k]
    mesh = BoxMesh(*shape, npml=npml, h=h)
    return mesh, random_model(mesh, seed=seed)


@pytest.mark.parametrize("k", [0, 1])
def test_gradient_mass_matches_oracle(k):
    mesh, model = _case(k)
    print(f"gradient mass case mesh={MESHES[k]}", flush=True)
    ctx = S.Context(mesh, model)
    img, res = S.gradient_mass(ctx)
    imr, rer = O.System(mesh, model).gradient_mass()
    assert np.array_equal(res, rer)
    on = rer > 0
    assert np.all(img[~on] == 0.0)
    assert np.max(np.abs(img[on] - imr[on]) / imr[on]) <= 1e-14


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("method", ["TV", "TN"])
@pytest.mark.parametrize("weighted", [False, True])
def test_regularization_matches_oracle(k, method, weighted):
    mesh, model = _case(k)
    wreg = WREG if weighted else None
    print(f"regularization case mesh={MESHES[k]} method={method} weighted={weighted}", flush=True)
    ctx = S.Context(mesh, model)
    rl, rm, cl, cm = S.regularization(ctx, model.lam, model.mu, method, wreg)
    rlr, rmr, clr, cmr = O.System(mesh, model).regularization(method, model.lam, model.mu, wreg)
    assert np.abs(rlr).max() > 0 and np.abs(rmr).max() > 0
    assert rel_l2(rl, rlr) <= 1e-12 and rel_l2(rm, rmr) <= 1e-12
    assert abs(cl - clr) <= 1e-12 * abs(clr) and abs(cm - cmr) <= 1e-12 * abs(cmr)


def _postprocess_oracle(sysm, it, gl, gm, lam, mu, method, factor, wreg):
    """Program_Main.F90:416-481 in numpy on the oracle's vectors."""
    rl, rm, cl, cm = sysm.regularization(method, lam, mu, wreg)
    img, _ = sysm.gradient_mass()
    fl = factor * np.linalg.norm(gl) / np.linalg.norm(rl)
    fm = factor * np.linalg.norm(gm) / np.linalg.norm(rm)
    rl, rm = fl * rl, fm * rm
    if it != 1:
        gl, gm = gl + rl, gm + rm
    gl, gm = gl * img, gm * img
    if method == "TN":
        cl, cm = 0.5 * rl @ lam, 0.5 * rm @ mu
    else:
        cl, cm = cl * fl, cm * fm
    if it == 1:
        cl = cm = 0.0
    return gl, gm, cl, cm, fl, fm


@pytest.mark.parametrize("it", [1, 4])
@pytest.mark.parametrize("method,weighted", [("TV", False), ("TN", False), ("TV", True)])
def test_regularize_gradient_matches_oracle(it, method, weighted):
    mesh, model = _case(1, seed=3)
    wreg = WREG if weighted else None
    print(f"regularize-gradient case it={it} method={method} weighted={weighted}", flush=True)
    rng = np.random.default_rng(21)
    box = mesh.rd_node.reshape(-1)
    gl0 = np.where(box, rng.standard_normal(mesh.nnode), 0.0) * 1e-3
    gm0 = np.where(box, rng.standard_normal(mesh.nnode), 0.0) * 2e-3
    ctx = S.Context(mesh, model)
    gl, gm = gl0.copy(), gm0.copy()
    cl, cm, fl, fm = S.regularize_gradient(ctx, it, gl, gm, model.lam, model.mu, method, 0.3, wreg)
    rgl, rgm, rcl, rcm, rfl, rfm = _postprocess_oracle(O.System(mesh, model), it, gl0, gm0, model.lam, model.mu,
                                                       method, 0.3, wreg)
    assert rel_l2(gl, rgl) <= 1e-12 and rel_l2(gm, rgm) <= 1e-12
    assert abs(fl - rfl) <= 1e-12 * rfl and abs(fm - rfm) <= 1e-12 * rfm
    if it == 1:
        assert cl == 0.0 and cm == 0.0
    else:
        assert abs(cl - rcl) <= 1e-12 * abs(rcl) and abs(cm - rcm) <= 1e-12 * abs(rcm)


def test_regularize_gradient_device_pointers():
    """The same call on a device-pointer context (torch tensors in HBM)."""
    import torch
    mesh, model = _case(0, seed=5)
    rng = np.random.default_rng(2)
    box = mesh.rd_node.reshape(-1)
    gl0 = np.where(box, rng.standard_normal(mesh.nnode), 0.0)
    gm0 = np.where(box, rng.standard_normal(mesh.nnode), 0.0)
    host = S.Context(mesh, model)
    gl, gm = gl0.copy(), gm0.copy()
    ref = S.regularize_gradient(host, 2, gl, gm, model.lam, model.mu, "TV", 0.3, WREG)
    dctx = S.Context(mesh, device_ptrs=True)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    dctx.set_material_ptrs(t(model.lam), t(model.mu), t(model.rho))
    dl, dm, dgl, dgm = t(model.lam), t(model.mu), t(gl0), t(gm0)
    out = S.regularize_gradient(dctx, 2, dgl, dgm, dl, dm, "TV", 0.3, WREG)
    torch.cuda.synchronize()
    assert np.array_equal(dgl.cpu().numpy(), gl) and np.array_equal(dgm.cpu().numpy(), gm)
    assert np.array_equal(np.array(out), np.array(ref))

"""The oracle's gradient post-processing (regularisation and gradient mass).

* ELEMENT_R_Mg_3D and ELEMENT_TV_TN_Reg_3D (Element_Matrices_3D.F90:1555-1718,
  1899-2075) bit for bit against the reference's routines: committed goldens
  (tests/golden/elements.npz) and, where oracle/_ref is built, live calls on
  random elements, weighted (Adjustments.F90:174-177) and unweighted;
* invariants that fix the routines independently of the reference: the
  Laplacian's constant null space and exact energy of a linear field, the GLL
  mass summing to the element volume, total variation of a constant field
  vanishing, its closed form for a linear field (with the reference's integer
  `numerator`, which truncates |grad m|^2 in the cost), and the TV vector being the
  derivative of the smoothed TV functional (central differences of a numpy
  quadrature of sqrt(|grad m|^2 + 0.01));
* the assembled system level (Assem_R_Mg_3D, Assem_TV_TN_Regularization,
  Assemble_PETSc_3D.F90:670-860): 1/Mg, the restriction mask and the TN/TV
  vectors against an independent scipy assembly of the element routines.
"""
import os

import numpy as np
import pytest

from hirensi_amd.mesh import BoxMesh, LOCAL_NODES
from oracle import oracle as O
from tests.helpers import random_model

GOLD = os.path.join(os.path.dirname(__file__), "golden", "elements.npz")
CASES = ["rd_cube", "rd_distorted", "rd_aniso", "pml_corner", "pml_edge", "pml_face", "pml_distorted"]
WREG = np.array([50.0, 1.0, 45.0])
EPS_TV = 0.01


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.fixture(scope="module")
def orc():
    return O.oracle_elem()


def _cube(center=(0.3, -0.2, -1.1), h=(1.0, 1.0, 1.0)):
    return np.asarray(center) + 0.5 * np.asarray(h) * LOCAL_NODES


@pytest.mark.parametrize("case", CASES)
def test_r_mg_and_tv_bitwise_vs_reference_goldens(gold, orc, case):
    xt, lam, mu = gold[f"{case}/xt"], gold[f"{case}/lam"], gold[f"{case}/mu"]
    for tag, w in (("", None), ("_w", WREG)):
        er, emg = orc.element_r_mg(xt, w)
        assert np.array_equal(er, gold[f"{case}/ER{tag}"]), tag
        assert np.array_equal(emg, gold[f"{case}/EMg{tag}"]), tag
        rm, rl, cost = orc.element_tv(xt, mu, lam, w)
        assert np.array_equal(rm, gold[f"{case}/TV_mu{tag}"]), tag
        assert np.array_equal(rl, gold[f"{case}/TV_lambda{tag}"]), tag
        assert np.array_equal(cost, gold[f"{case}/TV_cost{tag}"]), tag


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs /root/reference and flang)")
def test_r_mg_and_tv_bitwise_vs_reference_live(orc):
    r = O.ref_elem()
    rng = np.random.default_rng(31)
    mesh = BoxMesh(12, 10, 7, npml=(3, 2, 2, 3, 3), h=(1.0, 0.75, 1.25))
    for t in range(10):
        ex, ey, ez = rng.integers(0, mesh.nex), rng.integers(0, mesh.ney), rng.integers(0, mesh.nez)
        c = (mesh.xmin + (ex + 0.5) * mesh.h[0], mesh.ymin + (ey + 0.5) * mesh.h[1], mesh.zmin + (ez + 0.5) * mesh.h[2])
        xt = np.asarray(c) + 0.5 * np.asarray(mesh.h) * LOCAL_NODES
        if t % 2:
            xt = xt + 0.02 * rng.uniform(-1, 1, xt.shape)
        lam, mu = rng.uniform(50, 250, 27), rng.uniform(20, 120, 27)
        if t == 3:
            lam = np.full(27, 150.0)          # constant field: the TV vector is round-off only
        if t == 4:
            mu = 100.0 + 1e5 * rng.standard_normal(27)   # |grad mu|^2 beyond the int32 range
        w = None if t % 3 == 0 else rng.uniform([1, 1, 20], [60, 10, 60])
        a, b = orc.element_r_mg(xt, w), r.element_r_mg(xt, w)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), t
        a, b = orc.element_tv(xt, mu, lam, w), r.element_tv(xt, mu, lam, w)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), t


# ---------------------------------------------------------------- invariants
def _weight(z, w):
    return 1.0 if w is None else (1.0 - abs(z / w[2])) * w[0] + abs(z / w[2]) * w[1]


@pytest.mark.parametrize("wreg", [None, WREG])
def test_laplacian_null_space_symmetry_and_linear_energy(orc, wreg):
    h = (1.0, 0.5, 2.0)
    center = (0.3, -0.2, -3.1)
    xt = _cube(center, h)
    er, emg = orc.element_r_mg(xt, wreg)
    scale = np.abs(er).max()
    assert np.abs(er - er.T).max() <= 1e-13 * scale
    assert np.abs(er @ np.ones(27)).max() <= 1e-13 * scale
    V = h[0] * h[1] * h[2]
    # linear field a.x: f^T ER f = |a|^2 * integral of the weight; the weight is
    # linear in z on an element below z = 0, so its integral is V * w(z_centre)
    a = np.array([0.7, -1.3, 0.4])
    f = xt @ a
    assert abs(f @ er @ f - a @ a * V * _weight(center[2], wreg)) <= 1e-12 * (a @ a) * V * _weight(center[2], wreg)
    # GLL mass: the diagonal of the consistent mass at the quadrature nodes
    w1 = np.array([1 / 3, 4 / 3, 1 / 3])
    for n, (sx, sy, sz) in enumerate(LOCAL_NODES):
        assert abs(emg[n] - w1[sx + 1] * w1[sy + 1] * w1[sz + 1] * V / 8) <= 1e-15 * V
    assert abs(emg.sum() - V) <= 1e-14 * V


def test_tv_constant_and_linear_fields(orc):
    h = (1.0, 0.75, 1.5)
    center = (0.1, 0.4, -2.2)
    xt = _cube(center, h)
    V = h[0] * h[1] * h[2]
    rm, rl, cost = orc.element_tv(xt, np.full(27, 80.0), np.full(27, 150.0))
    assert np.abs(rm).max() < 1e-9 and np.abs(rl).max() < 1e-9 and np.abs(cost).max() < 1e-12
    a = np.array([1.3, -2.2, 0.5])
    b = np.array([-1.1, 0.4, 0.9])
    fm, fl = 80.0 + xt @ a, 150.0 + xt @ b
    er, _ = orc.element_r_mg(xt)
    for wreg in (None, WREG):
        rm, rl, cost = orc.element_tv(xt, fm, fl, wreg)
        wz = _weight(center[2], wreg)
        for g, r, c, f in ((b, rl, cost[0], fl), (a, rm, cost[1], fm)):
            den = np.sqrt(g @ g + EPS_TV)
            # the reference's numerator is an implicit INTEGER: |grad m|^2 truncated
            assert abs(c - np.trunc(g @ g) * V * wz / den) <= 1e-11 * abs(c)
            assert abs(r @ f - g @ g * V * wz / den) <= 1e-10 * abs(c)
            if wreg is None:
                assert np.abs(r - er @ f / den).max() <= 1e-14 * (np.abs(er) @ np.abs(f)).max() / den


def _tv_functional(xt, m, orc):
    """Smoothed total variation sum_q W_q |J| sqrt(|grad m(x_q)|^2 + eps) on an
    affine box element, from the shape-function derivatives."""
    h = xt.max(axis=0) - xt.min(axis=0)
    w1 = np.array([1 / 3, 4 / 3, 1 / 3])
    tot = 0.0
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                _, d = orc.shape_3_27(float(sx), float(sy), float(sz))
                grad = (d * (2.0 / h)).T @ m
                tot += w1[sx + 1] * w1[sy + 1] * w1[sz + 1] * np.prod(h) / 8 * np.sqrt(grad @ grad + EPS_TV)
    return tot


def test_tv_vector_is_derivative_of_smoothed_total_variation(orc):
    xt = _cube((0.5, -0.5, -1.5), (1.0, 1.0, 1.0))
    rng = np.random.default_rng(8)
    m = 100.0 + 3.0 * rng.standard_normal(27)
    rm, _, _ = orc.element_tv(xt, m, m)
    for _ in range(4):
        d = rng.standard_normal(27)
        hstep = 1e-4
        fd = (_tv_functional(xt, m + hstep * d, orc) - _tv_functional(xt, m - hstep * d, orc)) / (2 * hstep)
        assert abs(rm @ d - fd) <= 1e-7 * max(abs(fd), 1.0)


# ---------------------------------------------------------------- system level
def _assemble(mesh, orc, wreg):
    """Scipy assembly of ER and Mg over the regular-domain elements."""
    import scipy.sparse as sp
    xyz, inod = mesh.xyz(), mesh.inod()
    rows, cols, vals = [], [], []
    mg = np.zeros(mesh.nnode)
    for e in np.nonzero(mesh.mtel == 1)[0]:
        nodes = inod[e]
        er, emg = orc.element_r_mg(xyz[nodes], wreg)
        rows.append(np.repeat(nodes, 27))
        cols.append(np.tile(nodes, 27))
        vals.append(er.ravel())
        np.add.at(mg, nodes, emg)
    R = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(mesh.nnode, mesh.nnode))
    return R, mg


@pytest.mark.parametrize("wreg", [None, WREG])
def test_system_gradient_mass_and_regularization(orc, wreg):
    mesh = BoxMesh(7, 6, 5, npml=(2, 1, 1, 2, 2), h=(1.0, 1.25, 0.8))
    model = random_model(mesh, seed=12)
    sysm = O.System(mesh, model)
    R, mg = _assemble(mesh, orc, wreg)
    img, res = sysm.gradient_mass(wreg)
    touched = mesh.rd_node.reshape(-1)
    assert np.array_equal(res, touched.astype(float))
    assert np.allclose(img[touched], 1.0 / mg[touched], rtol=1e-15, atol=0)
    assert np.all(img[~touched] == 0.0)
    vol = np.prod(np.array(mesh.rd_hi) - np.array(mesh.rd_lo)) * np.prod(mesh.h)
    assert abs(mg.sum() - vol) <= 1e-12 * vol
    lam, mu = model.lam, model.mu
    rl, rm, cl, cm = sysm.regularization("TN", lam, mu, wreg)
    for r, m, c in ((rl, lam, cl), (rm, mu, cm)):
        ref = R @ m
        assert np.abs(r - ref).max() <= 1e-12 * np.abs(ref).max()
        assert abs(c - 0.5 * m @ ref) <= 1e-12 * abs(c)
    rl, rm, cl, cm = sysm.regularization("TV", lam, mu, wreg)
    xyz, inod = mesh.xyz(), mesh.inod()
    for r, m, c, which in ((rl, lam, cl, 1), (rm, mu, cm, 0)):
        ref = np.zeros(mesh.nnode)
        cref = 0.0
        for e in np.nonzero(mesh.mtel == 1)[0]:
            nodes = inod[e]
            out = orc.element_tv(xyz[nodes], mu[nodes], lam[nodes], wreg)
            np.add.at(ref, nodes, out[which])
            cref += out[2][1 - which]
        assert np.abs(r - ref).max() <= 1e-12 * np.abs(ref).max()
        assert abs(c - cref) <= 1e-12 * abs(cref)
        assert np.all(r[~touched] == 0.0)

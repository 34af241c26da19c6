"""The oracle's element layer against the reference's own element routines.

* goldens: tests/golden/elements.npz, written by tests/golden/make_goldens.py
  from the reference's Fortran sources compiled with flang (oracle/ref.mk);
* live: the same routines called directly where oracle/_ref is built;
* invariants that fix each routine independently of both (Kronecker and
  partition-of-unity properties, rigid-body null space, closed-form strain
  energy, the PML profile formula, Ricker peak and support).
"""
import os

import numpy as np
import pytest

from hirensi_amd.mesh import BoxMesh, LOCAL_NODES
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "elements.npz")
CASES = ["rd_cube", "rd_distorted", "rd_aniso", "pml_corner", "pml_edge", "pml_face", "pml_distorted"]


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.fixture(scope="module")
def orc():
    return O.oracle_elem()


@pytest.mark.parametrize("case", CASES)
def test_element_matrices_bitwise_vs_reference_goldens(gold, orc, case):
    pp = gold["pml_param"]
    xt, lam, mu = gold[f"{case}/xt"], gold[f"{case}/lam"], gold[f"{case}/mu"]
    assert np.array_equal(orc.element_k(xt, lam, mu), gold[f"{case}/K"])
    for f in (201, 202, 203, 204, 205):
        assert np.array_equal(np.diag(orc.element_m(xt, 2.0e-3, pp, f)), gold[f"{case}/M{f}"]), f
    for f in (301, 302, 303, 304):
        assert np.array_equal(np.diag(orc.element_en(xt, pp, f)), gold[f"{case}/EN{f}"]), f
    for f in range(101, 107):
        assert np.array_equal(orc.element_ea(xt, lam, mu, pp, f), gold[f"{case}/EA{f}"]), f
    gm, gl = orc.element_g(xt, gold[f"{case}/eu"], gold[f"{case}/ep"])
    assert np.array_equal(gm, gold[f"{case}/g_mu"])
    assert np.array_equal(gl, gold[f"{case}/g_lambda"])


def test_traction_pml_finder_shape_vs_goldens(gold, orc):
    assert np.array_equal(orc.element_fs(gold["fs/xt"], 5), gold["fs/face5"])
    assert np.array_equal(orc.element_fs(gold["fs/xt"], 6), gold["fs/face6"])
    pp = gold["pml_param"]
    for xg, pab in zip(gold["pml/xg"], gold["pml/pab"]):
        assert np.array_equal(orc.pml_finder_3d(xg, pp), pab)
    for p, fn, d in zip(gold["shape/pts"], gold["shape/fn"], gold["shape/dfxi"]):
        f2, d2 = orc.shape_3_27(*p)
        # GLL nodes are exact; off-node points may differ in the last bit
        # (operand grouping of the Fortran unary-minus products)
        assert np.allclose(f2, fn, rtol=0, atol=4e-16) and np.allclose(d2, d, rtol=0, atol=1e-15)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs /root/reference and flang)")
def test_element_matrices_bitwise_vs_reference_live(orc):
    r = O.ref_elem()
    rng = np.random.default_rng(77)
    mesh = BoxMesh(12, 10, 7, npml=(3, 2, 2, 3, 3), h=(1.0, 0.75, 1.25))
    pp = mesh.pml_param()
    hx, hy, hz = mesh.h
    for _ in range(12):
        ex, ey, ez = rng.integers(0, mesh.nex), rng.integers(0, mesh.ney), rng.integers(0, mesh.nez)
        c = (mesh.xmin + (ex + 0.5) * hx, mesh.ymin + (ey + 0.5) * hy, mesh.zmin + (ez + 0.5) * hz)
        xt = np.asarray(c) + 0.5 * np.asarray(mesh.h) * LOCAL_NODES
        lam, mu = rng.uniform(50, 250, 27), rng.uniform(20, 120, 27)
        assert np.array_equal(orc.element_k(xt, lam, mu), r.element_k(xt, lam, mu))
        for f in range(101, 107):
            assert np.array_equal(orc.element_ea(xt, lam, mu, pp, f), r.element_ea(xt, lam, mu, pp, f))
        for f in (201, 202, 203, 204, 205):
            assert np.array_equal(orc.element_m(xt, 2e-3, pp, f), r.element_m(xt, 2e-3, pp, f))
        for f in (301, 302, 303, 304):
            assert np.array_equal(orc.element_en(xt, pp, f), r.element_en(xt, pp, f))
        eu, ep = rng.standard_normal(81), rng.standard_normal(81)
        a, b = orc.element_g(xt, eu, ep), r.element_g(xt, eu, ep)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ---------------------------------------------------------------- invariants
def _cube(center=(0.3, -0.2, -1.1), h=(1.0, 1.0, 1.0)):
    return np.asarray(center) + 0.5 * np.asarray(h) * LOCAL_NODES


def test_shape_functions_kronecker_partition_and_derivative(orc):
    for a, (sx, sy, sz) in enumerate(LOCAL_NODES):
        fn, _ = orc.shape_3_27(float(sx), float(sy), float(sz))
        e = np.zeros(27)
        e[a] = 1.0
        assert np.array_equal(fn, e)
    rng = np.random.default_rng(1)
    for p in rng.uniform(-1, 1, (10, 3)):
        fn, d = orc.shape_3_27(*p)
        assert abs(fn.sum() - 1.0) < 1e-14
        assert np.abs(d.sum(axis=0)).max() < 1e-13
        # derivative against central differences of the shape functions
        for k in range(3):
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (orc.shape_3_27(*pp)[0] - orc.shape_3_27(*pm)[0]) / (2 * h)
            assert np.abs(fd - d[:, k]).max() < 1e-8


def test_stiffness_symmetry_rigid_modes_and_strain_energy(orc):
    h = (1.0, 0.5, 2.0)
    xt = _cube(h=h)
    rng = np.random.default_rng(3)
    lam, mu = rng.uniform(100, 200, 27), rng.uniform(50, 100, 27)
    K = orc.element_k(xt, lam, mu)
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    X = xt
    modes = []
    for c in range(3):
        m = np.zeros((3, 27))
        m[c] = 1.0
        modes.append(m.ravel())
    for (a, b) in ((0, 1), (0, 2), (1, 2)):
        m = np.zeros((3, 27))
        m[a] = -X[:, b]
        m[b] = X[:, a]
        modes.append(m.ravel())
    for m in modes:
        assert np.abs(K @ m).max() <= 1e-11 * np.abs(K).max() * np.abs(m).max()
    # homogeneous material, uniform strain: u^T K u = V (lam tr(E)^2 + 2 mu E:E)
    l0, m0 = 160.0, 80.0
    K0 = orc.element_k(xt, np.full(27, l0), np.full(27, m0))
    E = np.array([[1.0, 0.3, -0.2], [0.3, -0.5, 0.7], [-0.2, 0.7, 0.4]])
    u = (X @ E.T).T.ravel()
    V = h[0] * h[1] * h[2]
    energy = V * (l0 * np.trace(E) ** 2 + 2 * m0 * np.sum(E * E))
    assert abs(u @ K0 @ u - energy) <= 1e-12 * energy


def test_mass_and_pml_diagonal_blocks(orc):
    mesh = BoxMesh(16, 16, 8, npml=(4,) * 5)
    pp = mesh.pml_param()
    xt = _cube((-6.5, -7.5, -5.5))
    M1 = np.diag(orc.element_m(xt, 2e-3, pp, 201))
    assert abs(M1[:27].sum() - 2e-3) < 1e-15
    # PML masses are rho * {a,b,c,d}(x_q) * w_q |J| at each GLL node
    w = np.array([1 / 3, 4 / 3, 1 / 3])
    for f, idx in ((202, 6), (203, 7), (204, 8), (205, 9)):
        M = np.diag(orc.element_m(xt, 2e-3, pp, f))[:27]
        EN = np.diag(orc.element_en(xt, pp, 300 + (f - 201)))
        for a, (sx, sy, sz) in enumerate(LOCAL_NODES):
            pab = orc.pml_finder_3d(xt[a], pp)
            W = w[sx + 1] * w[sy + 1] * w[sz + 1] * 0.125
            assert abs(M[a] - 2e-3 * pab[idx] * W) <= 1e-15 * max(1.0, abs(M[a]))
            assert abs(EN[a] - pab[idx] * W) <= 1e-14 * max(1.0, abs(EN[a]))
            assert abs(EN[81 + a] - 2 * pab[idx] * W) <= 1e-14 * max(1.0, abs(EN[81 + a]))


def test_pml_profile_closed_form(orc):
    mesh = BoxMesh(16, 16, 8, npml=(4, 3, 2, 4, 3))
    pp = mesh.pml_param()
    rng = np.random.default_rng(9)
    for xg in rng.uniform([-8, -8, -8], [8, 8, 0], (60, 3)):
        pab = orc.pml_finder_3d(xg, pp)
        al, be = np.ones(3), np.zeros(3)
        for d, (fp, fm) in enumerate(((0, 1), (2, 3), (None, 5))):
            if fp is not None and xg[d] >= pp[fp, 3]:
                r = (xg[d] - pp[fp, 3]) / pp[fp, 2]
                al[d] += pp[fp, 0] * r * r
                be[d] += pp[fp, 1] * r * r
            elif xg[d] <= pp[fm, 3]:
                r = (pp[fm, 3] - xg[d]) / pp[fm, 2]
                al[d] += pp[fm, 0] * r * r
                be[d] += pp[fm, 1] * r * r
        assert np.allclose(pab[:6], np.r_[al, be], rtol=1e-14, atol=0)
        a = al.prod()
        b = al[0] * al[1] * be[2] + al[0] * al[2] * be[1] + al[1] * al[2] * be[0]
        c = al[0] * be[1] * be[2] + al[1] * be[0] * be[2] + al[2] * be[0] * be[1]
        d = be.prod()
        assert np.allclose(pab[6:10], [a, b, c, d], rtol=1e-13, atol=0)
        assert np.all(pab[10:] == 0.0)


def test_coupling_blocks_structure(orc):
    mesh = BoxMesh(16, 16, 8, npml=(4,) * 5)
    pp = mesh.pml_param()
    xt = _cube((-5.5, 6.5, -6.5))
    rng = np.random.default_rng(5)
    lam, mu = rng.uniform(100, 200, 27), rng.uniform(50, 100, 27)
    for f in (101, 102, 103):
        EA = orc.element_ea(xt, lam, mu, pp, f)
        # sum over the rows of one u component: derivative of the partition of unity
        s = EA.reshape(3, 27, 162).sum(axis=1)
        assert np.abs(s).max() <= 1e-12 * max(np.abs(EA).max(), 1e-300)
        # lower block with lambda = 0, mu = 1/2 has the unit weights of the upper block
        low = orc.element_ea(xt, np.zeros(27), np.full(27, 0.5), pp, f + 3)
        assert np.array_equal(low, EA)


def test_gradient_kernel_uniform_strain(orc):
    h = (1.0, 1.0, 1.0)
    xt = _cube(h=h)
    E = np.array([[0.4, 0.1, 0.0], [0.1, -0.2, 0.3], [0.0, 0.3, 0.5]])
    F = np.array([[-0.1, 0.2, 0.05], [0.2, 0.3, -0.4], [0.05, -0.4, 0.1]])
    u = (xt @ E.T).T.ravel()
    p = (xt @ F.T).T.ravel()
    gm, gl = orc.element_g(xt, u, p)
    assert abs(gl.sum() - np.trace(E) * np.trace(F)) < 1e-13
    assert abs(gm.sum() - 2 * np.sum(E * F)) < 1e-13


def test_ricker_time_function():
    fr = 15.0
    wr = 2 * np.pi * fr
    tpk = 3 * np.sqrt(6) / wr
    assert abs(O.time_function(tpk, 2, fr) + 1.0) < 1e-14
    assert abs(O.time_function(0.0, 2, fr)) < 1e-15
    assert O.time_function(2 * tpk + 1e-9, 2, fr) == 0.0
    for dt in (1e-3, 7e-3, 1.9e-2):
        assert abs(O.time_function(tpk + dt, 2, fr) - O.time_function(tpk - dt, 2, fr)) < 1e-13
    # Gaussian default (LCASE 3, L_Gaussian = 40)
    assert O.time_function(0.06, 3, 0.0) == 1.0

"""Host side of the learned gradient preconditioner (no GPU):

* the element-grid maps against brute-force loops over the lattice, and what
  the quadrature geometry fixes (a linear field's element mean is its centre
  value; constants are preserved both ways);
* the layer table: channel counts chain through encoder, skips and decoder,
  1.31 M parameters;
* the L-BFGS hook: an identity map reproduces the plain direction bit for bit,
  a scaled map scales only the mu part of H0, a map that declines leaves the
  plain direction; the inversion falls back to the plain direction (and then
  matches the plain run) when the line search rejects the learned step;
* LearnedH0's normalisation <P(q), q> = <q, q> and its descent safeguard.
"""
import numpy as np
import pytest

from hirensi_amd import fwi
from hirensi_amd import precond as PC
from hirensi_amd.mesh import BoxMesh
from tests.test_fwi_host import QuadProblem, _sequence


def _mesh():
    return BoxMesh(4, 6, 2, npml=(1, 1, 1, 2, 1))


def test_nodal_to_voxels_brute_force():
    m = _mesh()
    rng = np.random.default_rng(1)
    v = rng.standard_normal(m.nnode)
    got = PC.nodal_to_voxels(m, v)
    assert got.shape == (m.nez, m.ney, m.nex)
    for ez in range(m.nez):
        for ey in range(m.ney):
            for ex in range(m.nex):
                s = 0.0
                for k in range(2 * ez, 2 * ez + 3):
                    for j in range(2 * ey, 2 * ey + 3):
                        for i in range(2 * ex, 2 * ex + 3):
                            s += v[m.node_index(i, j, k)]
                assert abs(got[ez, ey, ex] - s / 27.0) < 1e-12


def test_voxels_to_nodal_brute_force():
    m = _mesh()
    rng = np.random.default_rng(2)
    vox = rng.standard_normal((m.nez, m.ney, m.nex))
    got = PC.voxels_to_nodal(m, vox)
    for k in range(m.NZ):
        for j in range(m.NY):
            for i in range(m.NX):
                vals = [vox[ez, ey, ex]
                        for ez in range(m.nez) if 2 * ez <= k <= 2 * ez + 2
                        for ey in range(m.ney) if 2 * ey <= j <= 2 * ey + 2
                        for ex in range(m.nex) if 2 * ex <= i <= 2 * ex + 2]
                assert abs(got[m.node_index(i, j, k)] - np.mean(vals)) < 1e-12


def test_maps_on_linear_and_constant_fields():
    m = _mesh()
    xyz = m.xyz()
    lin = 1.5 * xyz[:, 0] - 0.7 * xyz[:, 1] + 2.0 * xyz[:, 2] + 3.0
    vox = PC.nodal_to_voxels(m, lin)
    ex, ey, ez = np.meshgrid(np.arange(m.nex), np.arange(m.ney), np.arange(m.nez), indexing="ij")
    cx = m.xmin + (ex + 0.5) * m.h[0]
    cy = m.ymin + (ey + 0.5) * m.h[1]
    cz = m.zmin + (ez + 0.5) * m.h[2]
    centre = np.transpose(1.5 * cx - 0.7 * cy + 2.0 * cz + 3.0, (2, 1, 0))
    assert np.allclose(vox, centre, atol=1e-12)
    assert np.allclose(PC.voxels_to_nodal(m, np.full((m.nez, m.ney, m.nex), 4.25)), 4.25, atol=0)
    # the element-centre node takes its element's value
    vox = np.random.default_rng(3).standard_normal((m.nez, m.ney, m.nex))
    nod = PC.voxels_to_nodal(m, vox)
    for ez, ey, ex in [(0, 0, 0), (1, 5, 3), (1, 2, 1)]:
        assert nod[m.node_index(2 * ex + 1, 2 * ey + 1, 2 * ez + 1)] == vox[ez, ey, ex]


def test_rd_voxel_mask():
    m = _mesh()
    mask = PC.rd_voxel_mask(m)
    for ez in range(m.nez):
        for ey in range(m.ney):
            for ex in range(m.nex):
                rd = (m.rd_lo[0] <= ex < m.rd_hi[0] and m.rd_lo[1] <= ey < m.rd_hi[1] and ez >= m.rd_lo[2])
                assert mask[ez, ey, ex] == float(rd)


def test_layer_table():
    L = PC.LAYERS
    assert len(L) == 15
    # encoder and bottleneck: each level's first layer takes the previous level's width
    prev = 1
    for lev in range(PC.LEVELS):
        a, b = L[2 * lev], L[2 * lev + 1]
        assert a[1] == prev and a[2] == b[1] == b[2] and a[3] == b[3] == lev
        prev = b[2]
    # decoder: input = coarser output, output = the skip's width
    for lev in range(PC.LEVELS - 2, -1, -1):
        la = 2 * PC.LEVELS + 2 * (PC.LEVELS - 2 - lev)
        a, b = L[la], L[la + 1]
        assert a[1] == prev and a[2] == L[2 * lev + 1][2] and b[1] == b[2] == a[2] and a[3] == b[3] == lev
        prev = b[2]
    assert L[-1][1:] == (prev, 1, 0)
    assert PC.n_params() == 1310001 or 1.2e6 < PC.n_params() < 1.4e6
    assert PC.n_params() == sum(ci * co * 27 + co for _, ci, co, _ in L)


def test_random_case_layers_and_inclusions():
    m = BoxMesh(12, 12, 8, npml=(2,) * 5)
    rng = np.random.default_rng(5)
    for _ in range(4):
        true, start = PC.random_case(m, rng)
        vs_s = np.sqrt(start.mu / (true.rho.mean() * 0 + 1.0))
        assert np.all(start.mu > 0) and np.all(true.mu > 0)
        assert np.array_equal(true.rho, start.rho)
        d = (true.mu - start.mu) / start.mu
        assert 0.1 < np.abs(d).max() < 0.75       # at least one inclusion of 15-30 % in Vs
        # start: layered (mu constant on every horizontal lattice plane of the RD node box)
        mu = start.mu.reshape(m.NZ, m.NY, m.NX)
        assert np.all(mu.max(axis=(1, 2)) == mu.min(axis=(1, 2)))
        assert vs_s.size == m.nnode


def _lbfgs_pairs(rng, n, K, lb):
    seq = _sequence(rng, n, K)
    out = None
    for it in range(2, K + 2):
        (m0, g0), (m1, g1) = seq[it - 2], seq[it - 1]
        out = lb.direction(g1[:n], g1[n:], g1[:n] - g0[:n], g1[n:] - g0[n:], m1[:n] - m0[:n], m1[n:] - m0[n:],
                           np.ones(n))
    return out


def test_lbfgs_identity_map_is_plain_and_scaled_map_scales_mu_h0():
    n, K = 12, 5
    plain = _lbfgs_pairs(np.random.default_rng(7), n, K, fwi.CoupledLBFGS(4))
    ident = _lbfgs_pairs(np.random.default_rng(7), n, K, fwi.CoupledLBFGS(4, h0_map=lambda q: q.copy()))
    declined = _lbfgs_pairs(np.random.default_rng(7), n, K, fwi.CoupledLBFGS(4, h0_map=lambda q: None))
    for a, b, c in zip(plain, ident, declined):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    # with no stored pairs the result is -h0 P(q) on mu and -h0 q on lambda
    lb = fwi.CoupledLBFGS(4, h0_map=lambda q: 3.0 * q)
    lb.h0 = 0.5
    g_l, g_m = np.arange(1.0, 5.0), np.arange(2.0, 6.0)
    p_l, p_m = lb.two_loop(g_l, g_m)
    assert np.array_equal(p_l, -0.5 * g_l) and np.array_equal(p_m, -1.5 * g_m) and lb.mapped
    p_l, p_m = lb.two_loop(g_l, g_m, use_map=False)
    assert np.array_equal(p_m, -0.5 * g_m) and not lb.mapped


def _run(prob, h0_map, n_iter, params):
    inv = fwi.Inversion(prob, np.full(prob.n, 110.0), np.full(prob.n, 95.0), prob.restriction, params, h0_map=h0_map)
    return inv, [inv.iterate() for _ in range(n_iter)]


def _same_run(rec_p, inv_p, rec_s, inv_s):
    for a, b in zip(rec_p, rec_s):
        assert a["misfit"] == b["misfit"] and a["alpha"] == b["alpha"] and a["armijo_counter"] == b["armijo_counter"]
    assert np.array_equal(inv_p.mu, inv_s.mu) and np.array_equal(inv_p.lam, inv_s.lam)


def test_inversion_falls_back_when_the_first_trial_is_rejected():
    # steepest descent, mu only: P(g) = g + K z with z orthogonal to g (so <P(g), g> =
    # <g, g>); the capped trial step moves mu along z, where the misfit is quadratic
    # with positive curvature and has no first-order decrease: always rejected
    params = fwi.InversionParams(search_direction="SD", control_parameter="Mu", bound_lambda_lower=1.0,
                                 bound_lambda_upper=1e4, bound_mu_lower=1.0, bound_mu_upper=1e4)
    rng = np.random.default_rng(11)
    prob = QuadProblem(seed=3)

    def orth(q):
        z = rng.standard_normal(q.shape) * prob.restriction
        z -= (z @ q) / (q @ q) * q
        return q + 1e8 * np.linalg.norm(q) * z / np.linalg.norm(z)
    inv_s, rec_s = _run(prob, orth, 4, params)
    inv_p, rec_p = _run(QuadProblem(seed=3), None, 4, params)
    assert all(r["fallback"] and not r["learned_h0"] for r in rec_s)
    assert "fallback" not in rec_p[0]
    _same_run(rec_p, inv_p, rec_s, inv_s)


def test_inversion_falls_back_when_the_learned_direction_is_not_descent():
    # L-BFGS: P(q) = -1000 q makes the mu part of H negative definite, so the
    # direction fails the descent check and the plain L-BFGS direction is taken
    params = fwi.InversionParams(m_lbfgs=4, bound_lambda_lower=1.0, bound_lambda_upper=1e4, bound_mu_lower=1.0,
                                 bound_mu_upper=1e4)
    calls = []

    def neg(q):
        calls.append(1)
        return None if len(calls) == 1 else -1000.0 * q
    inv_s, rec_s = _run(QuadProblem(seed=5), neg, 5, params)
    inv_p, rec_p = _run(QuadProblem(seed=5), None, 5, params)
    assert not rec_s[0]["fallback"] and not rec_s[0]["learned_h0"]        # declined: plain
    assert all(r["fallback"] and not r["learned_h0"] for r in rec_s[1:])
    _same_run(rec_p, inv_p, rec_s, inv_s)


def test_inversion_keeps_an_accepted_learned_step():
    params = fwi.InversionParams(m_lbfgs=4, bias_lambda_search=False, bound_lambda_lower=1.0, bound_lambda_upper=1e4,
                                 bound_mu_lower=1.0, bound_mu_upper=1e4)
    prob = QuadProblem(seed=4)
    # the exact inverse of the mu block's diagonal: a good preconditioner
    def newton(q):
        r = np.zeros_like(q)
        r[:prob.n_rd] = q[:prob.n_rd] / prob.v
        return r * (float(q @ q) / float(r @ q))
    inv_l, rec_l = _run(prob, newton, 3, params)
    assert all(r["learned_h0"] and not r["fallback"] for r in rec_l)
    assert all(r["misfit"] < r["misfit_pre"] for r in rec_l)
    inv_p, rec_p = _run(QuadProblem(seed=4), None, 3, params)
    assert rec_l[0]["misfit"] != rec_p[0]["misfit"]        # the learned step was the one taken


class _IdentityNet:
    def infer(self, vol):
        return np.asarray(vol, np.float64) * (1.0 + 0.1 * np.sign(vol))


class _NegNet:
    def infer(self, vol):
        return -np.asarray(vol, np.float64)


def test_learned_h0_normalisation_and_safeguard():
    m = BoxMesh(8, 8, 8, npml=(2,) * 5)
    from hirensi_amd.solver import Context  # noqa: F401  (the restriction comes from the mesh here)
    restriction = m.rd_node.reshape(-1).astype(np.float64)
    rng = np.random.default_rng(9)
    q = rng.standard_normal(m.nnode) * restriction
    h = PC.LearnedH0(_IdentityNet(), m, restriction)
    r = h(q)
    assert r is not None and abs(float(r @ q) - float(q @ q)) < 1e-9 * float(q @ q)
    assert np.all(r[restriction == 0] == 0)
    assert PC.LearnedH0(_NegNet(), m, restriction)(q) is None
    assert h(np.zeros(m.nnode)) is None
    assert h.calls == 2 and h.used == 1

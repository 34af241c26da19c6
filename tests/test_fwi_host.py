"""The inversion driver (hirensi_amd/fwi.py) against the oracle's literal
restatement of Program_Main.F90:379-921 / LBFGS_Coupled.F90 (oracle/fwi_oracle.py),
and the oracle itself against what the mathematics fixes:

* the two-loop recursion equals the dense BFGS inverse-Hessian update of
  gamma*I (gamma = s.y / y.y of the newest pair) by the last m pairs, oldest
  first, on the stacked (lambda, mu) vector; with m smaller than the number of
  pairs only the newest m count;
* the newest secant equation H y = s holds;
* Armijo bookkeeping: step halving, the maximum update divided by 6 at the
  third cut (kept for later iterations), the step negated at the ninth cut,
  failure at the tenth.

The driver and the oracle run on the same synthetic problems (no GPU): a
coupled quadratic misfit with a PML-like extension map and a restriction mask.
"""
import numpy as np
import pytest

from hirensi_amd import fwi
from oracle import fwi_oracle as FO


def _dense_bfgs(pairs, g):
    """-H g with H from gamma*I updated by the pairs oldest first."""
    s_k, y_k = pairs[-1]
    H = (s_k @ y_k) / (y_k @ y_k) * np.eye(len(g))
    for s, y in pairs:
        rho = 1.0 / (y @ s)
        V = np.eye(len(g)) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return -H @ g


def _sequence(rng, n, K):
    """K+1 iterates (m, g) of a convex quadratic: g = A m - b."""
    A = rng.standard_normal((2 * n, 2 * n))
    A = A @ A.T + 2 * n * np.eye(2 * n)
    b = rng.standard_normal(2 * n)
    ms = [rng.standard_normal(2 * n) for _ in range(K + 1)]
    return [(m, A @ m - b) for m in ms]


@pytest.mark.parametrize("m,K", [(15, 5), (3, 7), (4, 4)])
def test_oracle_two_loop_equals_dense_bfgs(m, K):
    rng = np.random.default_rng(m * 10 + K)
    n = 6
    seq = _sequence(rng, n, K)
    lb = FO.LBFGSCoupled(n, m)
    ones = np.ones(n)
    pairs = []
    for it in range(1, K + 2):
        mm, g = seq[it - 1]
        if it == 1:
            p_l, p_m = lb.direction(it, 1, g[:n], g[n:], None, None, None, None, None, None, ones)
            assert np.array_equal(np.r_[p_l, p_m], -g)
        else:
            m0, g0 = seq[it - 2]
            p_l, p_m = lb.direction(it, 1, g[:n], g[n:], g0[:n], g0[n:], mm[:n], mm[n:], m0[:n], m0[n:], ones)
            pairs.append((mm - m0, g - g0))
            ref = _dense_bfgs(pairs[-m:], g)
            assert np.allclose(np.r_[p_l, p_m], ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
            # newest secant equation: the direction of y_k is -s_k
            s_k, y_k = pairs[-1]
            lb2 = FO.LBFGSCoupled(n, m)
            for j in range(1, it):
                mj, gj = seq[j - 1]
                if j == 1:
                    lb2.direction(j, 1, gj[:n], gj[n:], None, None, None, None, None, None, ones)
                else:
                    mi, gi = seq[j - 2]
                    lb2.direction(j, 1, gj[:n], gj[n:], gi[:n], gi[n:], mj[:n], mj[n:], mi[:n], mi[n:], ones)
            q_l, q_m = lb2.direction(it, 1, y_k[:n], y_k[n:], np.zeros(n), np.zeros(n), mm[:n], mm[n:],
                                     m0[:n], m0[n:], ones)
            # (the pair recorded here is (s_k, y_k) again: H y_k = s_k still holds)
            assert np.allclose(np.r_[q_l, q_m], -s_k, rtol=1e-9, atol=1e-11 * np.abs(s_k).max())


def test_product_lbfgs_matches_oracle_with_restriction():
    rng = np.random.default_rng(4)
    n, m, K = 40, 4, 9
    restriction = (rng.uniform(size=n) > 0.3).astype(float)
    seq = _sequence(rng, n, K)
    lo = FO.LBFGSCoupled(n, m)
    lp = fwi.CoupledLBFGS(m)
    for it in range(2, K + 2):
        (m0, g0), (m1, g1) = seq[it - 2], seq[it - 1]
        a = lo.direction(it, 1, g1[:n], g1[n:], g0[:n], g0[n:], m1[:n], m1[n:], m0[:n], m0[n:], restriction)
        b = lp.direction(g1[:n], g1[n:], g1[:n] - g0[:n], g1[n:] - g0[n:], m1[:n] - m0[:n], m1[n:] - m0[n:],
                         restriction)
        for x, y in zip(a, b):
            assert np.allclose(x, y, rtol=1e-12, atol=1e-13 * np.abs(x).max())


class QuadProblem:
    """Coupled quadratic misfit on nodal (lam, mu); nodes >= n_rd copy a mapped
    node (PML extension); the gradient is zero off the restriction."""

    def __init__(self, n=30, n_rd=24, seed=0, scale=1.0, never_decrease=False):
        rng = np.random.default_rng(seed)
        self.n, self.n_rd = n, n_rd
        self.src = np.r_[np.arange(n_rd), rng.integers(0, n_rd, n - n_rd)]
        self.restriction = (np.arange(n) < n_rd).astype(float)
        self.lt = rng.uniform(90, 140, n_rd)
        self.mt = rng.uniform(70, 120, n_rd)
        self.w = scale * rng.uniform(0.5, 2.0, n_rd)
        self.v = scale * rng.uniform(0.5, 2.0, n_rd)
        self.c = 0.2 * scale
        self.never = never_decrease
        self.calls = []

    def extend(self, x):
        return x[self.src]

    def forward_misfit(self, lam, mu):
        self.calls.append((lam.copy(), mu.copy()))
        if self.never:
            return 1.0e6 + len(self.calls)
        dl, dm = lam[:self.n_rd] - self.lt, mu[:self.n_rd] - self.mt
        return float(0.5 * (self.w * dl * dl).sum() + 0.5 * (self.v * dm * dm).sum() + self.c * (dl * dm).sum()
                     + 0.01 * ((dl ** 4).sum()))

    def gradient(self, it, lam, mu):
        dl, dm = lam[:self.n_rd] - self.lt, mu[:self.n_rd] - self.mt
        gl, gm = np.zeros(self.n), np.zeros(self.n)
        gl[:self.n_rd] = self.w * dl + self.c * dm + 0.04 * dl ** 3
        gm[:self.n_rd] = self.v * dm + self.c * dl
        return gl, gm, 0.0, 0.0


PARAM_CASES = [
    dict(),
    dict(search_direction="SD"),
    dict(search_direction="CG", n_cg_restart=4),
    dict(bias_lambda_search=False, m_lbfgs=3),
    dict(control_parameter="Mu"),
    dict(control_parameter="Lambda", bias_lambda_search=False),
    dict(n_iter_bias=4, fac_max_update_lambda=40.0, fac_max_update_mu=40.0),
]


@pytest.mark.parametrize("case", range(len(PARAM_CASES)))
def test_inversion_loop_matches_oracle(case):
    kw = PARAM_CASES[case]
    n_iter = 12
    po = QuadProblem(seed=case)
    pp = QuadProblem(seed=case)
    lam0 = po.extend(np.full(po.n, 100.0))
    mu0 = po.extend(np.full(po.n, 90.0))
    okw = dict(kw)
    if "bias_lambda_search" in okw:
        okw["bias_lambda_search"] = "Yes" if okw["bias_lambda_search"] else "No"
    lo, mo, ho = FO.invert(po, lam0, mu0, po.restriction, n_iter, FO.Params(**okw))
    lp, mp, hp = fwi.invert(pp, lam0, mu0, pp.restriction, n_iter, fwi.InversionParams(**kw))
    assert len(po.calls) == len(pp.calls)
    for (a, b), (c, d) in zip(po.calls, pp.calls):
        assert np.allclose(a, c, rtol=1e-12, atol=0) and np.allclose(b, d, rtol=1e-12, atol=0)
    for ro, rp in zip(ho, hp):
        assert ro["armijo_counter"] == rp["armijo_counter"]
        assert abs(ro["misfit"] - rp["misfit"]) <= 1e-12 * abs(ro["misfit"])
    # the misfit decreases monotonically and the bounds hold
    mis = [h["misfit"] for h in hp]
    assert all(b <= a for a, b in zip(mis, mis[1:]))
    assert lp.min() >= 50.0 and lp.max() <= 166.0 and mp.min() >= 50.0 and mp.max() <= 166.0
    if case == 0:
        assert sum(h["armijo_counter"] for h in hp) > 0       # the back-tracking path is exercised


def test_armijo_bookkeeping_and_failure():
    """A misfit that never decreases: 10 cuts of the first iteration, steps
    halving, negated at the ninth cut, then the loop stops."""
    for impl in ("oracle", "product"):
        pr = QuadProblem(seed=3, never_decrease=True)
        lam0, mu0 = np.full(pr.n, 100.0), np.full(pr.n, 90.0)
        g_l, g_m, _, _ = pr.gradient(1, lam0, mu0)
        with pytest.raises(FO.StopInversion if impl == "oracle" else fwi.ArmijoFailure):
            if impl == "oracle":
                FO.invert(pr, lam0, mu0, pr.restriction, 1)
            else:
                fwi.invert(pr, lam0, mu0, pr.restriction, 1)
        # calls: initial forward, first trial, 10 cuts
        assert len(pr.calls) == 12
        steps = [m - mu0 for _, m in pr.calls[1:]]
        a0 = 5.0 / np.abs(g_m).max()
        for k, s in enumerate(steps):
            expect = a0 * 0.5 ** k * (-1.0 if k >= 9 else 1.0)
            ref = np.clip(mu0 + expect * pr.extend(-g_m), 50.0, 166.0)
            assert np.allclose(s, ref - mu0, rtol=1e-12, atol=1e-12)


def test_max_update_reduction_persists():
    """Three or more cuts in one iteration divide the maximum update by 6 for
    every later iteration (Fac_Max_Update is not reset)."""
    pr = QuadProblem(seed=1)
    pr.lt = 100.0 + np.linspace(-3.0, 3.0, pr.n_rd)     # targets close to the start: a 60-unit step overshoots
    pr.mt = 90.0 + np.linspace(2.0, -2.0, pr.n_rd)
    inv = fwi.Inversion(pr, np.full(pr.n, 100.0), np.full(pr.n, 90.0), pr.restriction,
                        fwi.InversionParams(fac_max_update_lambda=60.0, fac_max_update_mu=60.0))
    rec = inv.iterate()
    assert rec["armijo_counter"] >= 3
    assert inv.fac_max == [10.0, 10.0]

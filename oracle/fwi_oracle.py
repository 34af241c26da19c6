"""Test infrastructure: a literal CPU restatement of HiReNSI's inversion loop
(step_4_3D_Forward_and_Inverse/Program_Main.F90:379-921) and its coupled
L-BFGS search direction (LBFGS_Coupled.F90:1-221).

Only tests/ load this module, as the checker of hirensi_amd/fwi.py.  It keeps
the reference's own storage (Ys/Ss columns indexed by N_store, Rhos, the
LBFGS_vector_sequence orderings), its control flow and its operation order,
on numpy vectors, so that a plausible slip in the product (wrong pair order,
missing restriction mask, wrong Armijo bookkeeping) shows up as a difference.

The physics is behind a ``problem`` object with two methods, the way the
reference includes Solve_Forward.F90 and the adjoint section:
  problem.forward_misfit(lam, mu) -> Cost_Misfit of a forward solve on (lam, mu)
  problem.gradient(it, lam, mu)   -> (g_lambda, g_mu, cost_reg_lambda, cost_reg_mu)
        the post-processed gradient of the last forward (Program_Main.F90:396-481)
  problem.extend(v)               -> v with PML nodes copied from their mapped
        regular-domain nodes (vscat_mat_extend, Program_Main.F90:633-636)
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Params:
    """Adjustments.F90:137-186 (inversion parameters)."""
    fac_max_update_lambda: float = 5.0
    fac_max_update_mu: float = 5.0
    bound_lambda_upper: float = 166.0
    bound_mu_upper: float = 166.0
    bound_lambda_lower: float = 50.0
    bound_mu_lower: float = 50.0
    control_parameter: str = "Lambda_Mu"
    bias_lambda_search: str = "Yes"
    n_iter_bias: int = 50
    m_lbfgs: int = 15
    search_direction: str = "LBFGS"
    n_cg_restart: int = 10


class StopInversion(Exception):
    """Armijo_Counter reached 10: the reference writes its output and STOPs."""


class LBFGSCoupled:
    """LBFGS_Coupled.F90: storage and two-loop recursion with one rho and one
    H0 for lambda and mu together."""

    def __init__(self, nj, m):
        self.M = m
        self.Ys_l = np.zeros((nj, m + 1))    # columns 1..M (Fortran indices)
        self.Ys_m = np.zeros((nj, m + 1))
        self.Ss_l = np.zeros((nj, m + 1))
        self.Ss_m = np.zeros((nj, m + 1))
        self.Rhos = np.zeros(m + 1)
        self.Alphas = np.zeros(m + 1)

    def direction(self, it, it_begin, g_l, g_m, g_l_pre, g_m_pre, lam, mu, lam_pre, mu_pre, restriction):
        if it == it_begin:
            return -g_l, -g_m
        K = it - it_begin
        y_l = g_l - g_l_pre
        s_l = lam - lam_pre
        rho_l = float(np.dot(y_l, s_l))
        y_m = g_m - g_m_pre
        s_m = mu - mu_pre
        rho_m = float(np.dot(y_m, s_m))
        rho_lm = 1.0 / (rho_l + rho_m)
        q_l = g_l.copy()
        q_m = g_m.copy()
        M = self.M
        if 0 < K <= M:
            n_use = K
            seq = list(range(K, 0, -1))
        else:
            n_use = M
            lend = K % M
            seq = list(range(lend, 0, -1))
            j = 0
            for _ in range(len(seq) + 1, M + 1):
                seq.append(M - j)
                j += 1
        n_store = K % M if K % M != 0 else M
        self.Rhos[n_store] = rho_lm
        self.Ys_l[:, n_store] = y_l
        self.Ys_m[:, n_store] = y_m
        self.Ss_l[:, n_store] = s_l * restriction
        self.Ss_m[:, n_store] = s_m * restriction
        for i_vec in range(n_use):
            i = seq[i_vec]
            self.Alphas[i] = (self.Rhos[i] * np.dot(q_l, self.Ss_l[:, i])
                              + self.Rhos[i] * np.dot(q_m, self.Ss_m[:, i]))
            q_l = q_l - self.Alphas[i] * self.Ys_l[:, i]
            q_m = q_m - self.Alphas[i] * self.Ys_m[:, i]
        h0_num = float(np.dot(y_l, s_l)) + float(np.dot(y_m, s_m))
        h0_den = float(np.dot(y_l, y_l)) + float(np.dot(y_m, y_m))
        r_l = (h0_num / h0_den) * q_l
        r_m = (h0_num / h0_den) * q_m
        for i_vec in range(n_use - 1, -1, -1):
            i = seq[i_vec]
            beta = (self.Rhos[i] * np.dot(r_l, self.Ys_l[:, i])
                    + self.Rhos[i] * np.dot(r_m, self.Ys_m[:, i]))
            r_l = r_l + self.Ss_l[:, i] * (self.Alphas[i] - beta)
            r_m = r_m + self.Ss_m[:, i] * (self.Alphas[i] - beta)
        return -r_l, -r_m


def _update(P, it, lam_pre, mu_pre, p_l, p_m, a_l, a_m, problem):
    """Program_Main.F90:590-660 (and :747-812): step, PML extension, bounds."""
    lam, mu = lam_pre.copy(), mu_pre.copy()
    if P.control_parameter == "Lambda":
        lam = lam_pre + a_l * p_l
    elif P.control_parameter == "Mu":
        mu = mu_pre + a_m * p_m
    else:
        if P.bias_lambda_search == "Yes" and it < P.n_iter_bias:
            pm2 = float(np.sqrt(np.dot(p_m, p_m)))
            pl2 = float(np.sqrt(np.dot(p_l, p_l)))
            W = 1.0 - float(it // P.n_iter_bias)          # Fortran integer division
            fb_mu = W * pl2 / pm2 * a_l
            fb_lam = (1.0 - W) * pl2 * a_l
            lam = lam_pre + (fb_mu * p_m + fb_lam * p_l)
            mu = mu_pre + a_m * p_m
        else:
            lam = lam_pre + a_l * p_l
            mu = mu_pre + a_m * p_m
    lam, mu = problem.extend(lam), problem.extend(mu)
    lam = np.where(lam < P.bound_lambda_lower, P.bound_lambda_lower,
                   np.where(lam > P.bound_lambda_upper, P.bound_lambda_upper, lam))
    mu = np.where(mu < P.bound_mu_lower, P.bound_mu_lower, np.where(mu > P.bound_mu_upper, P.bound_mu_upper, mu))
    return lam, mu


def _armijo(P, c, a_l, a_m, gs_l, gs_m):
    if P.control_parameter == "Lambda":
        return c * abs(a_l) * gs_l
    if P.control_parameter == "Mu":
        return c * abs(a_m) * gs_m
    return c * abs(a_m) * gs_m + c * abs(a_l) * gs_l


def invert(problem, lam, mu, restriction, n_iter, P=None, it_begin=1):
    """Iterations it_begin .. it_begin + n_iter - 1.  Returns the model and a
    record per iteration (misfit, regularisation costs, step lengths, Armijo
    counter)."""
    P = P or Params()
    fac_max_l, fac_max_m = P.fac_max_update_lambda, P.fac_max_update_mu
    nj = lam.shape[0]
    lb = LBFGSCoupled(nj, P.m_lbfgs)
    cost_misfit = problem.forward_misfit(lam, mu)
    g_l = np.zeros(nj)
    g_m = np.zeros(nj)
    p_l = np.zeros(nj)
    p_m = np.zeros(nj)
    lam_pre, mu_pre = lam.copy(), mu.copy()
    hist = []
    for it in range(it_begin, it_begin + n_iter):
        cost_misfit_pre = cost_misfit
        g_l_pre, g_m_pre, p_l_pre, p_m_pre = g_l.copy(), g_m.copy(), p_l.copy(), p_m.copy()
        g_l, g_m, c_rl, c_rm = problem.gradient(it, lam, mu)
        # VI- search direction
        if P.search_direction == "SD":
            p_l, p_m = -g_l, -g_m
        elif P.search_direction == "CG":
            if it == it_begin or it % P.n_cg_restart == 0:
                p_l, p_m = -g_l, -g_m
            else:
                b_m = np.sqrt(np.dot(g_m, g_m)) / np.sqrt(np.dot(g_m_pre, g_m_pre))
                b_l = np.sqrt(np.dot(g_l, g_l)) / np.sqrt(np.dot(g_l_pre, g_l_pre))
                p_m = -g_m + b_m * p_m_pre
                p_l = -g_l + b_l * p_l_pre
        else:
            p_l, p_m = lb.direction(it, it_begin, g_l, g_m, g_l_pre, g_m_pre, lam, mu, lam_pre, mu_pre,
                                    restriction)
        lam_pre, mu_pre = lam.copy(), mu.copy()
        # descent checks (Program_Main.F90:522-553)
        chk_l = chk_m = -1.0
        gs_l = float(np.dot(g_l, p_l))
        if gs_l > 0:
            chk_l = 1.0
            p_l = -g_l
            gs_l = float(np.dot(g_l, p_l))
        gs_m = float(np.dot(g_m, p_m))
        if gs_m > 0:
            chk_m = 1.0
            p_m = -g_m
            gs_m = float(np.dot(g_m, p_m))
        # step lengths (:556-588)
        pl_max = float(np.abs(p_l).max())
        pm_max = float(np.abs(p_m).max())
        a_l = fac_max_l / pl_max
        a_m = fac_max_m / pm_max
        a_l_sd, a_m_sd = a_l, a_m
        if P.search_direction == "LBFGS" and it != it_begin:
            a_l = a_m = 1.0
            if pl_max > fac_max_l:
                a_l = fac_max_l / pl_max
            if pm_max > fac_max_m:
                a_m = fac_max_m / pm_max
        if chk_l > 0 and P.search_direction == "LBFGS":
            a_l = a_l_sd
        if chk_m > 0 and P.search_direction == "LBFGS":
            a_m = a_m_sd
        lam, mu = _update(P, it, lam_pre, mu_pre, p_l, p_m, a_l, a_m, problem)
        cost_misfit = problem.forward_misfit(lam, mu)
        c, rho = 1.0e-4, 0.5
        armijo = _armijo(P, c, a_l, a_m, gs_l, gs_m)
        counter = 0
        while cost_misfit > cost_misfit_pre + armijo:
            counter += 1
            if counter == 3:
                fac_max_l /= 6.0
                fac_max_m /= 6.0
            a_l *= rho
            a_m *= rho
            if counter == 9:
                a_l, a_m = -a_l, -a_m
            lam, mu = _update(P, it, lam_pre, mu_pre, p_l, p_m, a_l, a_m, problem)
            cost_misfit = problem.forward_misfit(lam, mu)
            armijo = _armijo(P, c, a_l, a_m, gs_l, gs_m)
            if counter == 10:
                raise StopInversion(it)
        hist.append({"iter": it, "misfit": cost_misfit, "misfit_pre": cost_misfit_pre, "cost_reg": (c_rl, c_rm),
                     "alpha": (a_l, a_m), "armijo_counter": counter})
    return lam, mu, hist

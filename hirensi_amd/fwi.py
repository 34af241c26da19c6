"""Full-waveform inversion loop on top of the HIP forward/adjoint path.

Mirrors the inversion driver of HiReNSI (step_4_3D_Forward_and_Inverse/
Program_Main.F90:379-921) with its coupled L-BFGS search direction
(LBFGS_Coupled.F90:1-221):

* gradient: OtD adjoint + control pass (``hr_adjoint_gradient``), then the
  adaptive TV/TN regularisation and 1/M_g (``hr_regularize_gradient``);
* search direction: L-BFGS with one rho and one H0 for lambda and mu together
  (the memory keeps s restricted to the regular-domain nodes), or steepest
  descent, or the reference's norm-ratio conjugate gradient;
* a descent check per field (fall back to -g), step lengths capped by the
  maximum update, the lambda-search bias of the first iterations, PML material
  extension, bounds clamp;
* Armijo back-tracking on the misfit (c = 1e-4, rho = 0.5, the maximum update
  divided by 6 at the third cut, the step negated at the ninth, stop at the
  tenth).

Vectors are numpy arrays or torch tensors (host or device); every operation
below is one of +, -, *, dot, max |.|, clip and indexing, available on both.
The forward and gradient work is behind a *problem* object
(:class:`HipProblem` for the HIP path).
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import solver as S
from .mesh import Model


@dataclass
class InversionParams:
    """Inversion parameters of Adjustments.F90:137-186 (reference defaults)."""
    fac_max_update_lambda: float = 5.0
    fac_max_update_mu: float = 5.0
    bound_lambda_upper: float = 166.0
    bound_mu_upper: float = 166.0
    bound_lambda_lower: float = 50.0
    bound_mu_lower: float = 50.0
    control_parameter: str = "Lambda_Mu"        # Lambda | Mu | Lambda_Mu
    bias_lambda_search: bool = True
    n_iter_bias: int = 50
    m_lbfgs: int = 15
    search_direction: str = "LBFGS"             # SD | CG | LBFGS
    n_cg_restart: int = 10
    regularization: str = "TV"                  # TV | TN
    regularization_factor: float = 0.3
    armijo_c: float = 1.0e-4
    armijo_rho: float = 0.5


class ArmijoFailure(RuntimeError):
    """Ten back-tracking cuts without sufficient decrease (the reference stops)."""


def _dot(a, b):
    return float((a * b).sum())


def _copy(a):
    return a.copy() if isinstance(a, np.ndarray) else a.clone()


def _absmax(a):
    return float(abs(a).max())


class CoupledLBFGS:
    """Two-loop recursion over the last m pairs (newest first), one rho per pair
    and one H0 = (y.s)/(y.y) from the newest pair, with the dot products summed
    over lambda and mu.  The stored s is restricted to the regular-domain nodes;
    rho and H0 use the unrestricted s (LBFGS_Coupled.F90:13-215)."""

    def __init__(self, m, h0_map=None):
        self.pairs = deque(maxlen=m)     # (s_l, s_m, y_l, y_m, rho), newest at the right
        self.h0_map = h0_map             # q_mu -> P(q_mu) or None (learned H0; None: identity)
        self.h0 = None
        self.mapped = False

    def direction(self, g_l, g_m, d_g_l, d_g_m, d_l, d_m, restriction, use_map=True):
        """g: current gradient; d_g = g - g_pre; d = m - m_pre.  Returns p = -H g."""
        ys = _dot(d_g_l, d_l) + _dot(d_g_m, d_m)
        yy = _dot(d_g_l, d_g_l) + _dot(d_g_m, d_g_m)
        self.pairs.append((d_l * restriction, d_m * restriction, d_g_l, d_g_m, 1.0 / ys))
        self.h0 = ys / yy
        return self.two_loop(g_l, g_m, use_map)

    def two_loop(self, g_l, g_m, use_map=True):
        """p = -H g over the stored pairs with H0 = h0 I, or h0 P on the mu
        field when a learned map is set, used and returns a direction."""
        q_l, q_m = _copy(g_l), _copy(g_m)
        alphas = []
        for s_l, s_m, y_l, y_m, rho in reversed(self.pairs):
            a = rho * _dot(q_l, s_l) + rho * _dot(q_m, s_m)
            alphas.append(a)
            q_l = q_l - a * y_l
            q_m = q_m - a * y_m
        h0 = self.h0
        r_l, r_m = h0 * q_l, h0 * q_m
        self.mapped = False
        if use_map and self.h0_map is not None:
            pq = self.h0_map(q_m)
            if pq is not None:
                r_m = h0 * pq
                self.mapped = True
        for (s_l, s_m, y_l, y_m, rho), a in zip(self.pairs, reversed(alphas)):
            b = rho * _dot(r_l, y_l) + rho * _dot(r_m, y_m)
            r_l = r_l + s_l * (a - b)
            r_m = r_m + s_m * (a - b)
        return -r_l, -r_m


class Inversion:
    """State of one inversion run (Program_Main.F90:379-921)."""

    def __init__(self, problem, lam, mu, restriction, params: InversionParams | None = None, it_begin=1,
                 h0_map=None):
        self.problem = problem
        self.P = params or InversionParams()
        self.lam, self.mu = _copy(lam), _copy(mu)
        self.restriction = restriction
        self.it = it_begin
        self.it_begin = it_begin
        self.fac_max = [self.P.fac_max_update_lambda, self.P.fac_max_update_mu]
        # h0_map: learned stand-in for the identity in H0 (mu field), applied to
        # the L-BFGS and first (steepest-descent) directions; the plain direction
        # is the fallback when the line search rejects the first trial step
        self.h0_map = h0_map
        self.mapped = False
        self.lbfgs = CoupledLBFGS(self.P.m_lbfgs, h0_map)
        self.g = None
        self.p = None
        self.history = []
        self.misfit = problem.forward_misfit(self.lam, self.mu)     # Program_Main.F90:318-347

    # -------------------------------------------------------------- helpers
    def _step(self, lam0, mu0, p_l, p_m, a_l, a_m):
        P = self.P
        lam, mu = lam0, mu0
        if P.control_parameter in ("Lambda", "Lambda_Mu"):
            if P.control_parameter == "Lambda_Mu" and P.bias_lambda_search and self.it < P.n_iter_bias:
                # the lambda update follows the mu direction, scaled to |p_lambda|
                # (the weight W = 1 - (Iter / NIterBias) is 1 under integer division)
                lam = lam0 + (math.sqrt(_dot(p_l, p_l)) / math.sqrt(_dot(p_m, p_m)) * a_l) * p_m
            else:
                lam = lam0 + a_l * p_l
        if P.control_parameter in ("Mu", "Lambda_Mu"):
            mu = mu0 + a_m * p_m
        ext = self.problem.extend
        lam = ext(lam).clip(P.bound_lambda_lower, P.bound_lambda_upper)
        mu = ext(mu).clip(P.bound_mu_lower, P.bound_mu_upper)
        return lam, mu

    def _armijo(self, a_l, a_m, gs_l, gs_m):
        c = self.P.armijo_c
        cp = self.P.control_parameter
        return ((c * abs(a_m) * gs_m if cp != "Lambda" else 0.0)
                + (c * abs(a_l) * gs_l if cp != "Mu" else 0.0))

    def _direction(self, g_l, g_m, use_map=True):
        P = self.P
        self.mapped = False
        if P.search_direction == "SD" or self.it == self.it_begin:
            if use_map and self.h0_map is not None:
                pg = self.h0_map(g_m)
                if pg is not None:
                    self.mapped = True
                    return -g_l, -pg
            return -g_l, -g_m
        if P.search_direction == "CG":
            if self.it % P.n_cg_restart == 0:
                return -g_l, -g_m
            (gl0, gm0), (pl0, pm0) = self.g, self.p
            b_l = math.sqrt(_dot(g_l, g_l)) / math.sqrt(_dot(gl0, gl0))
            b_m = math.sqrt(_dot(g_m, g_m)) / math.sqrt(_dot(gm0, gm0))
            return -g_l + b_l * pl0, -g_m + b_m * pm0
        if not use_map:          # the fallback of a rejected learned step: the pairs are already stored
            return self.lbfgs.two_loop(g_l, g_m, use_map=False)
        gl0, gm0 = self.g
        p = self.lbfgs.direction(g_l, g_m, g_l - gl0, g_m - gm0, self.lam - self.lam_pre,
                                 self.mu - self.mu_pre, self.restriction)
        self.mapped = self.lbfgs.mapped
        return p

    # ------------------------------------------------------------ iteration
    def iterate(self):
        """One inversion iteration; returns its record."""
        misfit_pre = self.misfit
        g_l, g_m, c_rl, c_rm = self.problem.gradient(self.it, self.lam, self.mu)
        p_l, p_m = self._direction(g_l, g_m)
        mapped = self.mapped
        self.lam_pre, self.mu_pre = self.lam, self.mu
        res =self._line_search(g_l, g_m, p_l, p_m, misfit_pre, first_only=mapped)
        fallback = False
        if res is None:
            # the learned direction's first trial was rejected: plain direction
            fallback = True
            p_l, p_m = self._direction(g_l, g_m, use_map=False)
            res = self._line_search(g_l, g_m, p_l, p_m, misfit_pre)
        lam, mu, misfit, a, cuts, p_l, p_m = res
        self.lam, self.mu, self.misfit = lam, mu, misfit
        self.g, self.p = (g_l, g_m), (p_l, p_m)
        rec = {"iter": self.it, "misfit": misfit, "misfit_pre": misfit_pre, "cost_reg": (c_rl, c_rm),
               "alpha": (a[0], a[1]), "armijo_counter": cuts}
        if self.h0_map is not None:
            rec["learned_h0"] = mapped and not fallback
            rec["fallback"] = fallback
        self.history.append(rec)
        self.it += 1
        return rec

    def _line_search(self, g_l, g_m, p_l, p_m, misfit_pre, first_only=False):
        """Descent check, step lengths and Armijo back-tracking of one direction;
        first_only: return None if the direction fails the descent check or its
        first trial step is rejected."""
        P = self.P
        sd = [False, False]
        gs = [_dot(g_l, p_l), _dot(g_m, p_m)]
        if gs[0] > 0:
            sd[0], p_l = True, -g_l
            gs[0] = _dot(g_l, p_l)
        if gs[1] > 0:
            sd[1], p_m = True, -g_m
            gs[1] = _dot(g_m, p_m)
        if first_only and (sd[0] or sd[1]):
            return None          # a learned direction that is not a descent direction: fall back
        pmax = [_absmax(p_l), _absmax(p_m)]
        a_sd = [self.fac_max[0] / pmax[0], self.fac_max[1] / pmax[1]]
        a = list(a_sd)
        if P.search_direction == "LBFGS" and self.it != self.it_begin:
            a = [self.fac_max[k] / pmax[k] if pmax[k] > self.fac_max[k] else 1.0 for k in range(2)]
            a = [a_sd[k] if sd[k] else a[k] for k in range(2)]
        lam, mu = self._step(self.lam_pre, self.mu_pre, p_l, p_m, a[0], a[1])
        misfit = self.problem.forward_misfit(lam, mu)
        armijo = self._armijo(a[0], a[1], gs[0], gs[1])
        if first_only and misfit > misfit_pre + armijo:
            return None
        cuts = 0
        while misfit > misfit_pre + armijo:
            cuts += 1
            if cuts == 3:
                self.fac_max = [f / 6.0 for f in self.fac_max]
            a = [v * P.armijo_rho for v in a]
            if cuts == 9:
                a = [-v for v in a]
            lam, mu = self._step(self.lam_pre, self.mu_pre, p_l, p_m, a[0], a[1])
            misfit = self.problem.forward_misfit(lam, mu)
            armijo = self._armijo(a[0], a[1], gs[0], gs[1])
            if cuts == 10:
                raise ArmijoFailure(f"iteration {self.it}: no sufficient decrease after 10 cuts")
        return lam, mu, misfit, a, cuts, p_l, p_m


def invert(problem, lam, mu, restriction, n_iter, params: InversionParams | None = None, it_begin=1, h0_map=None):
    """Run n_iter iterations from (lam, mu); returns (lam, mu, records)."""
    inv = Inversion(problem, lam, mu, restriction, params, it_begin, h0_map=h0_map)
    for _ in range(n_iter):
        inv.iterate()
    return inv.lam, inv.mu, inv.history


@dataclass
class Survey:
    """Shots over one receiver set: ftab[nshot, nstep, 3]; source entry e
    belongs to shot src_shot[e]; dis_meas[nshot, nstep + 1, nrec]."""
    nstep: int
    dt: float
    ftab: np.ndarray
    src_shot: np.ndarray
    src_eq: np.ndarray
    src_val: np.ndarray
    rec_eq: np.ndarray
    dis_meas: np.ndarray | None = None


class HipProblem:
    """Forward misfit and post-processed gradient on the HIP path (host
    vectors): Solve_Forward + Compute_Misfit, then the adjoint section and the
    regularisation of Program_Main.F90:396-481."""

    def __init__(self, ctx: S.Context, survey: Survey, rho, params: InversionParams | None = None, wreg=None,
                 history=S.HR_HISTORY_STORE, segment=64, log=None):
        self.ctx, self.sv, self.rho = ctx, survey, rho
        self.P = params or InversionParams()
        self.wreg = wreg
        self.history, self.segment = history, segment
        self.src = ctx.mesh.material_map()
        self.nforward = 0
        self.log = log                   # log(message) after each forward and gradient, or None

    def _done(self, what, t0):
        if self.log is not None:
            self.log(f"{what} done ({time.perf_counter() - t0:.1f} s)")

    def extend(self, v):
        return v[self.src]

    def _forward(self, lam, mu, history):
        sv = self.sv
        t0 = time.perf_counter()
        self.ctx.set_material(Model(lam, mu, self.rho))
        self.nforward += 1
        tr = S.explicit_rk4_petsc_pml_3d_shots(self.ctx, sv.nstep, sv.dt, sv.ftab, sv.src_shot, sv.src_eq,
                                               sv.src_val, sv.rec_eq, history=history, segment=self.segment)
        self._done(f"forward {self.nforward}", t0)
        return tr

    def observe(self, lam, mu):
        """Synthetic observations of a model (traces of every shot)."""
        return self._forward(lam, mu, S.HR_HISTORY_NONE).copy()

    def forward_misfit(self, lam, mu):
        sv = self.sv
        self._forward(lam, mu, self.history)
        return S.compute_misfit(self.ctx, sv.nstep, sv.dt, sv.dis_meas)

    def gradient(self, it, lam, mu):
        sv = self.sv
        t0 = time.perf_counter()
        g_l, g_m, _ = S.explicit_rk4_petsc_pml_3d_adjoint_otd(self.ctx, sv.nstep, sv.dt, sv.rec_eq, sv.dis_meas,
                                                               with_misfit=False)
        self._done("adjoint gradient", t0)
        c_l, c_m, _, _ = S.regularize_gradient(self.ctx, it, g_l, g_m, lam, mu, self.P.regularization,
                                               self.P.regularization_factor, self.wreg)
        return g_l, g_m, c_l, c_m


class SequentialHipProblem(HipProblem):
    """The plain form of the same problem: one shot per solve, as the reference
    runs them (Solve_Forward.F90:117, one source).  The forward history is kept
    whole when it fits the HBM budget, else checkpointed (HR_HISTORY_REVOLVE).
    Misfit and gradient are summed over the shots."""

    def __init__(self, ctx: S.Context, survey: Survey, rho, params: InversionParams | None = None, wreg=None,
                 log=None):
        super().__init__(ctx, survey, rho, params, wreg, history=S.HR_HISTORY_REVOLVE, log=log)

    def _shot(self, q, history):
        sv = self.sv
        sel = sv.src_shot == q
        t0 = time.perf_counter()
        self.nforward += 1
        tr = S.explicit_rk4_petsc_pml_3d(self.ctx, sv.nstep, sv.dt, sv.ftab[q], sv.src_eq[sel], sv.src_val[sel],
                                         sv.rec_eq, history=history)
        self._done(f"forward {self.nforward} (shot {q})", t0)
        return tr

    def observe(self, lam, mu):
        self.ctx.set_material(Model(lam, mu, self.rho))
        return np.stack([self._shot(q, S.HR_HISTORY_NONE).copy() for q in range(self.sv.ftab.shape[0])])

    def forward_misfit(self, lam, mu):
        sv = self.sv
        self.ctx.set_material(Model(lam, mu, self.rho))
        total = 0.0
        for q in range(sv.ftab.shape[0]):
            self._shot(q, S.HR_HISTORY_NONE)
            total += S.compute_misfit(self.ctx, sv.nstep, sv.dt, sv.dis_meas[q])
        return total

    def gradient(self, it, lam, mu):
        sv = self.sv
        g_l = np.zeros(self.ctx.nnode)
        g_m = np.zeros(self.ctx.nnode)
        for q in range(sv.ftab.shape[0]):
            self._shot(q, self.history)
            t0 = time.perf_counter()
            gl, gm, _ = S.explicit_rk4_petsc_pml_3d_adjoint_otd(self.ctx, sv.nstep, sv.dt, sv.rec_eq, sv.dis_meas[q],
                                                                with_misfit=False)
            self._done(f"adjoint gradient (shot {q})", t0)
            g_l += gl
            g_m += gm
        c_l, c_m, _, _ = S.regularize_gradient(self.ctx, it, g_l, g_m, lam, mu, self.P.regularization,
                                               self.P.regularization_factor, self.wreg)
        return g_l, g_m, c_l, c_m


def survey_of(cfg, nstep=None):
    """Survey of a config dict (mesh, src_nodes, rec_nodes, fr or frs, dt, nstep):
    vertical unit point loads, Ricker time functions, shared receivers."""
    from .mesh import receiver_dofs
    mesh = cfg["mesh"]
    nstep = int(nstep or cfg["nstep"])
    nshot = len(cfg["src_nodes"])
    frs = cfg.get("frs", [cfg.get("fr", 15.0)] * nshot)
    ftab = np.stack([S.time_table(nstep, cfg["dt"], 2, float(fr)) for fr in frs])
    return Survey(nstep=nstep, dt=cfg["dt"], ftab=ftab, src_shot=np.arange(nshot, dtype=np.int32),
                  src_eq=np.array([mesh.dof(nd, 2) for nd in cfg["src_nodes"]], dtype=np.int64),
                  src_val=np.ones(nshot), rec_eq=receiver_dofs(mesh, cfg["rec_nodes"]))


def model_error(mesh, mu, mu_true):
    """Relative L2 error of mu over the regular-domain node box."""
    box = mesh.rd_node.reshape(-1)
    d = mu[box] - mu_true[box]
    return float(np.linalg.norm(d) / np.linalg.norm(mu_true[box]))

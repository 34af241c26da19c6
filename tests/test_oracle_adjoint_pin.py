"""Independent pins of the oracle's OtD adjoint, control pass and misfit.

* adjoint (or_rk4_adjoint, Solver_PETSc_Explicit.F90:1239-1364): a textbook
  classical RK4 with step -dt of the first-order system
  x1' = x2, x2' = x3, x3' = M^-1 (C^T x3 - K^T x2 + G^T x1 + r(t)),
  written on the assembled matrices (pinned against the dense restatement of
  the element routines), with the misfit source r(t_n) at the first stage,
  r(t_{n-1}) at the last and their mean in between, p stored before each step;
* control (or_control, Solver_PETSc_Explicit.F90:1474-1510 + Update_g): the
  derivative of the assembled stiffness with respect to one nodal lambda or mu
  is K assembled for that unit nodal field (K is linear in lambda and mu), so
  g(k) = -dt sum_{n=1}^{N-1} u_n^T (dK/dm_k) p_n, independent of ELEMENT_g_3D;
* misfit (or_misfit, Solver_PETSc_Inverse.F90:633-650): closed forms for
  residuals linear in the step index, including two receivers on one
  equation, which add before squaring (ADD_VALUES + VecDot);
* TIME_FUNCTION cases 4-7 and 10 (IO.F90:630-715): derivatives, continuity
  and integral properties that fix each formula.
"""
import math

import numpy as np
import pytest

from hirensi_amd.mesh import Model
from oracle import oracle as O
from tests.helpers import small_case, random_model, rel_l2


def _case(seed=2):
    mesh, model, rec_eq, src_eq, src_val = small_case(nex=6, ney=6, nez=5, npml=2, seed=seed)
    return mesh, model, rec_eq, src_eq, src_val


def _histories(mesh, model, rec_eq, src_eq, src_val, nstep, dt, fr, dup=True):
    s = O.System(mesh, model)
    B = np.zeros(mesh.neq)
    np.add.at(B, src_eq, src_val)
    U = s.forward(B, nstep, dt, lcase=2, fr=fr)
    rec = np.asarray(rec_eq, dtype=np.int64)
    if dup:
        # a second receiver on the first receiver's equation
        rec = np.concatenate([rec, rec[:1]])
    rng = np.random.default_rng(7)
    pos = {int(e): r for r, e in enumerate(mesh.store_eq)}
    rows = np.array([pos[int(e)] for e in rec])
    dmeas = U[:, rows] * (1.0 + 0.3 * rng.uniform(-1, 1, (nstep + 1, len(rec))))
    return s, U, rec, rows, dmeas


def _textbook_backward_rk4(s, mesh, U, rec, rows, dmeas, nstep, dt):
    K, C, G = (s.csr(w) for w in range(3))
    KT, CT, GT = K.T.tocsr(), C.T.tocsr(), G.T.tocsr()
    iM = 1.0 / s.diag_m()
    n = mesh.neq

    def r(step):
        v = np.zeros(n)
        np.add.at(v, rec, U[step, rows] - dmeas[step])
        return v

    def rhs(y, src):
        x1, x2, x3 = y
        return (x2, x3, iM * (CT @ x3 - KT @ x2 + GT @ x1 + src))

    y = (np.zeros(n), np.zeros(n), np.zeros(n))
    P = np.zeros((nstep + 1, len(mesh.store_eq)))
    h = -dt
    for step in range(nstep, 0, -1):
        P[step] = y[1][mesh.store_eq]
        ra, rb = r(step), r(step - 1)
        rm = 0.5 * (ra + rb)
        k1 = rhs(y, ra)
        k2 = rhs(tuple(a + 0.5 * h * b for a, b in zip(y, k1)), rm)
        k3 = rhs(tuple(a + 0.5 * h * b for a, b in zip(y, k2)), rm)
        k4 = rhs(tuple(a + h * b for a, b in zip(y, k3)), rb)
        y = tuple(a + h / 6.0 * (b1 + 2.0 * b2 + 2.0 * b3 + b4) for a, b1, b2, b3, b4 in zip(y, k1, k2, k3, k4))
    return P


def test_adjoint_matches_textbook_backward_rk4():
    mesh, model, rec_eq, src_eq, src_val = _case()
    nstep, dt, fr = 120, 1e-4, 60.0
    s, U, rec, rows, dmeas = _histories(mesh, model, rec_eq, src_eq, src_val, nstep, dt, fr)
    P = s.adjoint(U, nstep, dt, rec, dmeas)
    Pt = _textbook_backward_rk4(s, mesh, U, rec, rows, dmeas, nstep, dt)
    assert np.abs(Pt).max() > 0
    assert rel_l2(P, Pt) < 1e-12
    # a plausible slip, the two source times swapped, is far outside the bar
    Ps = s.adjoint(U, nstep, dt, rec, dmeas[::-1].copy())
    assert rel_l2(Ps, Pt) > 1e-3


def _unit_stiffness(mesh, rho, k, field):
    """dK/dm_k: the stiffness assembled for the unit nodal field e_k minus the
    one assembled for zero material (the PML mass-like blocks cancel)."""
    z = np.zeros(mesh.nnode)
    e = z.copy()
    e[k] = 1.0
    m1 = Model(e, z, rho) if field == "lam" else Model(z, e, rho)
    K1 = O.System(mesh, m1).csr(0)
    K0 = O.System(mesh, Model(z, z, rho)).csr(0)
    return (K1 - K0).tocsr()


def test_control_matches_stiffness_derivative():
    mesh, model, rec_eq, src_eq, src_val = _case(seed=5)
    nstep, dt, fr = 80, 1e-4, 60.0
    s, U, rec, rows, dmeas = _histories(mesh, model, rec_eq, src_eq, src_val, nstep, dt, fr, dup=False)
    P = s.adjoint(U, nstep, dt, rec, dmeas)
    gl, gm = s.control(U, P, nstep, dt)
    # full-length u_n and p_n: the stored regular-domain entries, zero elsewhere
    n = mesh.neq
    Uf = np.zeros((nstep + 1, n))
    Pf = np.zeros((nstep + 1, n))
    Uf[:, mesh.store_eq] = U
    Pf[:, mesh.store_eq] = P
    # nodes touched only by regular-domain elements (Update_g sums MTYPE == 1)
    rd_only = np.flatnonzero(~mesh.pml_node.reshape(-1))
    rng = np.random.default_rng(3)
    sample = rng.choice(rd_only, size=6, replace=False)
    for g, field in ((gl, "lam"), (gm, "mu")):
        scale = np.abs(g).max()
        assert scale > 0
        for k in sample:
            dK = _unit_stiffness(mesh, model.rho, int(k), field)
            ref = -dt * sum(float(Uf[t] @ (dK @ Pf[t])) for t in range(1, nstep))
            assert abs(g[k] - ref) <= 1e-12 * scale, (field, int(k), g[k], ref)
    # nodes of the PML only receive nothing
    pml_only = np.flatnonzero(~mesh.rd_node.reshape(-1))
    assert np.all(gl[pml_only] == 0.0) and np.all(gm[pml_only] == 0.0)


def test_misfit_closed_form_with_shared_equation():
    # residual of receiver q at step n: a_q + b_q n; receivers 0 and 3 share an equation
    nstep, dt = 37, 2.5e-4
    nrec, ns = 5, 9
    rows = np.array([2, 5, 7, 2, 0], dtype=np.int64)
    eqs = np.array([40, 11, 93, 40, 3], dtype=np.int64)
    a = np.array([0.3, -1.1, 0.25, 0.7, 2.0])
    b = np.array([0.01, 0.02, -0.005, -0.03, 0.0])
    U = np.zeros((nstep + 1, ns))
    n = np.arange(nstep + 1, dtype=np.float64)
    dmeas = -(a[None, :] + b[None, :] * n[:, None])        # u = 0, so u - d = a + b n
    J = O.misfit(U, nstep, dt, rows, dmeas, rec_eq=eqs)

    def s1(A, B):        # sum_{n=1}^{N} (A + B n)^2
        N = nstep
        return N * A * A + A * B * N * (N + 1) + B * B * N * (N + 1) * (2 * N + 1) / 6.0

    groups = {}
    for q in range(nrec):
        ga, gb = groups.get(eqs[q], (0.0, 0.0))
        groups[eqs[q]] = (ga + a[q], gb + b[q])
    ref = 0.5 * dt * sum(s1(A, B) for A, B in groups.values())
    assert abs(J - ref) <= 1e-13 * ref
    # squaring each receiver on its own would give a different value
    alone = 0.5 * dt * sum(s1(a[q], b[q]) for q in range(nrec))
    assert abs(alone - ref) > 1e-3 * ref


def test_misfit_zero_row_is_not_counted():
    nstep, dt = 4, 1e-3
    U = np.zeros((nstep + 1, 2))
    dmeas = np.zeros((nstep + 1, 1))
    dmeas[0, 0] = 5.0                       # step 0 is outside the sum n = NSTEP..1
    assert O.misfit(U, nstep, dt, np.array([1]), dmeas) == 0.0
    dmeas[nstep, 0] = 2.0
    assert O.misfit(U, nstep, dt, np.array([1]), dmeas) == pytest.approx(0.5 * 4.0 * dt, rel=1e-15)


# ------------------------------------------------------------------ TIME_FUNCTION
def _tf(t, lcase, fr=15.0):
    return O.time_function(t, lcase, fr)


def _chopra_period():
    T_f = 2.0 * math.pi / 3.0
    nc = int(10.0 / T_f - 0.5)
    return T_f, nc, 10.0 / (nc + 0.5)


def test_time_function_chopra_cosine_and_derivative():
    T0, nc, T_f = _chopra_period()
    assert nc == 4                           # INTEGER truncation of 10 / (2 pi / 3) - 0.5
    assert _tf(0.0, 4) == 0.0
    # breakpoints T_f / 2 and nc T_f: the pieces meet (values 1 and -1 - (-1)...)
    for tb in (T_f / 2.0, nc * T_f):
        assert abs(_tf(tb - 1e-12, 4) - _tf(tb, 4)) < 1e-9
    assert _tf(T_f / 2.0, 4) == pytest.approx(1.0, abs=1e-15)
    assert _tf(10.0, 4) == pytest.approx(0.0, abs=1e-12)     # ends at zero at t_d
    assert _tf(10.0 + 1e-9, 4) == 0.0 and _tf(10.0 + 1e-9, 5) == 0.0
    # case 5 is the time derivative of case 4 on every piece
    for t in np.linspace(0.01, 9.99, 41):
        h = 1e-6
        fd = (_tf(t + h, 4) - _tf(t - h, 4)) / (2 * h)
        assert abs(_tf(t, 5) - fd) < 1e-6 * max(1.0, abs(fd)), t


def test_time_function_delta_ricker_derivative_heaviside():
    # Greenberg's delta: a Cauchy density of scale 1/k centred at t0: peak k/pi, unit mass
    assert _tf(0.0005, 6) == pytest.approx(10.0 / math.pi, rel=1e-15)
    from scipy.integrate import quad
    mass = quad(lambda t: _tf(t, 6), -np.inf, np.inf, limit=400)[0]
    assert mass == pytest.approx(1.0, rel=1e-8)
    # case 7 is the time derivative of the Ricker pulse (case 2) for t < t_max
    fr = 12.0
    t_max = 6.0 * math.sqrt(6.0) / (2.0 * math.pi * fr)
    for t in np.linspace(1e-3, 0.98 * t_max, 25):
        h = 1e-7
        fd = (_tf(t + h, 2, fr) - _tf(t - h, 2, fr)) / (2 * h)
        assert abs(_tf(t, 7, fr) - fd) < 1e-5 * max(1.0, abs(fd)), t
    assert _tf(t_max, 7, fr) == 0.0
    # smooth Heaviside: linear ramp to 1 at t0 = 5e-4, then constant
    assert _tf(0.00025, 10) == pytest.approx(0.5, rel=1e-15)
    assert _tf(0.0005, 10) == 1.0 and _tf(1.0, 10) == 1.0
    assert _tf(3.0, 9) == 0.0 and _tf(3.0, 8) == 1.0
    assert math.isnan(_tf(0.1, 11)) and math.isnan(_tf(0.1, 0))

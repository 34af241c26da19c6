"""The oracle's assembly and time stepping, pinned independently:

* assembly: a dense numpy assembly from the oracle's element routines with the
  reference's insertion semantics (row-oriented MatSetValuesLocal of the
  column-major Fortran blocks, negative indices dropped) equals the CSR;
* forward: a textbook RK4 of M x3' = -(C x3 + K x2 + G x1) + B f(t) on the
  assembled matrices reproduces the PETSc-sequence restatement;
* adjoint + control: the OtD gradient agrees with central finite differences
  of the misfit (the Gradient_check.F90 pattern), within the OtD/DtO gap;
* energy: with no PML damping (beta0 = 0, alpha0 = 0) the regular-domain
  kinetic + strain energy of the free vibration is conserved to O(dt^4).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from hirensi_amd.mesh import BoxMesh, homogeneous_model
from oracle import oracle as O
from tests.helpers import small_case, random_model, rel_l2


def _dense_assembly(mesh, model):
    orc = O.oracle_elem()
    n = mesh.neq
    K, C, G = (np.zeros((n, n)) for _ in range(3))
    M = np.zeros(n)
    xyz, inod, ID = mesh.xyz(), mesh.inod(), mesh.ID
    pp = mesh.pml_param()

    def add(A, rows, cols, blk):
        for a, r in enumerate(rows):
            if r < 0:
                continue
            for b, c in enumerate(cols):
                if c >= 0 and blk[a, b] != 0.0:
                    A[r, c] += blk[a, b]

    mrd, mpml = [], []
    for e in range(mesh.nel):
        nodes = inod[e]
        xt = xyz[nodes]
        lam, mu = model.lam[nodes], model.mu[nodes]
        nd1 = np.concatenate([ID[nodes, c] for c in range(3)])
        nd2 = np.concatenate([ID[nodes, 3 + c] for c in range(6)])
        if mesh.mtel[e] == 1:
            add(K, nd1, nd1, orc.element_k(xt, lam, mu))
            mrd.append((nd1, np.diag(orc.element_m(xt, model.rho[e], pp, 201))))
        else:
            for A, fm, fn, fa in ((C, 203, 302, 101), (K, 204, 303, 102), (G, 205, 304, 103)):
                add(A, nd1, nd1, orc.element_m(xt, model.rho[e], pp, fm))
                add(A, nd2, nd2, orc.element_en(xt, pp, fn))
                up = orc.element_ea(xt, lam, mu, pp, fa)
                lo = orc.element_ea(xt, lam, mu, pp, fa + 3)
                # Fortran EA(81,162) passed row-oriented as a 162 x 81 block -> A[ND2, ND1] = EA^T
                add(A, nd2, nd1, up.T)
                # transpose(EA_lower) (162 x 81 Fortran) passed as 81 x 162 -> A[ND1, ND2] = -EA_lower
                add(A, nd1, nd2, -lo)
            mpml.append((nd1, np.diag(orc.element_m(xt, model.rho[e], pp, 202))))
            mpml.append((nd2, np.diag(orc.element_en(xt, pp, 301))))
    for idx, v in mrd + mpml:
        for r, val in zip(idx, v):
            if r >= 0:
                M[r] += val
    return K, C, G, M


def test_assembly_matches_dense_restatement():
    mesh = BoxMesh(4, 3, 3, npml=(1, 1, 1, 1, 1))
    model = random_model(mesh, seed=4)
    K, C, G, M = _dense_assembly(mesh, model)
    s = O.System(mesh, model)
    for w, D in enumerate((K, C, G)):
        A = s.csr(w).toarray()
        assert np.array_equal(A, D), w
    assert np.array_equal(s.diag_m(), M)
    # the u-u block of K in the regular domain is symmetric
    u = np.zeros(mesh.neq, bool)
    u[mesh.ID[:, :3][mesh.ID[:, :3] >= 0]] = True
    Kuu = K[np.ix_(u, u)]
    assert np.abs(Kuu - Kuu.T).max() <= 1e-12 * np.abs(Kuu).max()


def _textbook_rk4(sysm, mesh, B, nstep, dt, fr):
    K, C, G = (sysm.csr(w) for w in range(3))
    iM = 1.0 / sysm.diag_m()
    Bt = B * iM

    def rhs(x1, x2, x3, f):
        return x2, x3, -iM * (C @ x3 + K @ x2 + G @ x1) + f * Bt

    n = mesh.neq
    x1, x2, x3 = np.zeros(n), np.zeros(n), np.zeros(n)
    out = [x2[mesh.store_eq].copy()]
    for k in range(nstep):
        T = k * dt
        f0, fh, f1 = (O.time_function(t, 2, fr) for t in (T, T + 0.5 * dt, T + dt))
        a = rhs(x1, x2, x3, f0)
        b = rhs(x1 + 0.5 * dt * a[0], x2 + 0.5 * dt * a[1], x3 + 0.5 * dt * a[2], fh)
        c = rhs(x1 + 0.5 * dt * b[0], x2 + 0.5 * dt * b[1], x3 + 0.5 * dt * b[2], fh)
        d = rhs(x1 + dt * c[0], x2 + dt * c[1], x3 + dt * c[2], f1)
        x1 = x1 + dt / 6 * (a[0] + 2 * b[0] + 2 * c[0] + d[0])
        x2 = x2 + dt / 6 * (a[1] + 2 * b[1] + 2 * c[1] + d[1])
        x3 = x3 + dt / 6 * (a[2] + 2 * b[2] + 2 * c[2] + d[2])
        out.append(x2[mesh.store_eq].copy())
    return np.array(out)


def test_forward_matches_textbook_rk4():
    mesh, model, rec_eq, src_eq, src_val = small_case(nex=6, ney=6, nez=5, npml=2, seed=2)
    s = O.System(mesh, model)
    B = np.zeros(mesh.neq)
    B[src_eq] = src_val
    nstep, dt, fr = 150, 1e-4, 60.0
    U = s.forward(B, nstep, dt, lcase=2, fr=fr)
    Ut = _textbook_rk4(s, mesh, B, nstep, dt, fr)
    assert np.abs(U).max() > 0
    assert rel_l2(U, Ut) < 1e-12


def _misfit_of(mesh, model, B, nstep, dt, fr, rows, dmeas):
    s = O.System(mesh, model)
    U = s.forward(B, nstep, dt, lcase=2, fr=fr)
    return O.misfit(U, nstep, dt, rows, dmeas)


@pytest.mark.slow
def test_adjoint_gradient_vs_finite_differences():
    mesh, model, rec_eq, src_eq, src_val = small_case(nex=8, ney=8, nez=6, npml=2, seed=3)
    nstep, dt, fr = 250, 1e-4, 60.0
    B = np.zeros(mesh.neq)
    B[src_eq] = src_val
    true = random_model(mesh, seed=8)
    st = O.System(mesh, true)
    Ut = st.forward(B, nstep, dt, lcase=2, fr=fr)
    pos = {int(e): r for r, e in enumerate(mesh.store_eq)}
    rows = np.array([pos[int(e)] for e in rec_eq])
    dmeas = Ut[:, rows].copy()
    s = O.System(mesh, model)
    U = s.forward(B, nstep, dt, lcase=2, fr=fr)
    P = s.adjoint(U, nstep, dt, rec_eq, dmeas)
    gl, gm = s.control(U, P, nstep, dt)
    # perturb nodes touched only by regular-domain elements: the gradient sums
    # regular-domain elements only (Update_g skips MTYPE /= 1)
    rd_only = (~mesh.pml_node).reshape(-1)
    rng = np.random.default_rng(12)
    for g, field in ((gl, "lam"), (gm, "mu")):
        d = np.where(rd_only, rng.uniform(-1, 1, mesh.nnode), 0.0)
        base = getattr(model, field)
        h = 1e-3 * np.abs(base).max()
        mp = random_model(mesh, seed=3)
        mm = random_model(mesh, seed=3)
        setattr(mp, field, base + h * d)
        setattr(mm, field, base - h * d)
        fd = (_misfit_of(mesh, mp, B, nstep, dt, fr, rows, dmeas) -
              _misfit_of(mesh, mm, B, nstep, dt, fr, rows, dmeas)) / (2 * h)
        adj = float(np.dot(g, d))
        assert abs(fd) > 0
        assert abs(adj - fd) <= 2e-2 * abs(fd), (field, adj, fd)


def test_energy_conserved_without_pml_damping():
    mesh = BoxMesh(6, 6, 5, npml=(1,) * 5, top_stress_bc="none")
    mesh.alpha0, mesh.beta0 = 0.0, 0.0
    model = homogeneous_model(mesh)
    s = O.System(mesh, model)
    n = mesh.neq
    K = s.csr(0)
    G = s.csr(2)
    # free vibration of the u-u elastic system from a smooth initial displacement
    u = np.zeros(n, bool)
    u[mesh.ID[:, :3][mesh.ID[:, :3] >= 0]] = True
    Kuu = K[u][:, u]
    M = s.diag_m()[u]
    rng = np.random.default_rng(0)
    x = rng.standard_normal(u.sum()) * 1e-3
    v = np.zeros_like(x)
    dt = 1e-5

    def energy(x, v):
        return 0.5 * v @ (M * v) + 0.5 * x @ (Kuu @ x)

    e0 = energy(x, v)

    def f(x, v):
        return v, -(Kuu @ x) / M

    for _ in range(100):
        a = f(x, v)
        b = f(x + 0.5 * dt * a[0], v + 0.5 * dt * a[1])
        c = f(x + 0.5 * dt * b[0], v + 0.5 * dt * b[1])
        d = f(x + dt * c[0], v + dt * c[1])
        x = x + dt / 6 * (a[0] + 2 * b[0] + 2 * c[0] + d[0])
        v = v + dt / 6 * (a[1] + 2 * b[1] + 2 * c[1] + d[1])
    assert abs(energy(x, v) - e0) <= 1e-6 * e0
    # with beta0 = 0 the G blocks (products of betas) vanish
    assert G.nnz == 0 or np.abs(G.data).max() == 0.0


def test_energy_diagnostic_matches_textbook():
    """or_rk4_forward_energy: 0.5 (M_RD x3) . x3 every `every` steps, with M_RD the
    regular-domain elements' lumped mass (element_m flag 201 assembled on its own)
    and x3 from the textbook RK4 of the assembled system."""
    mesh, model, rec_eq, src_eq, src_val = small_case(nex=6, ney=6, nez=5, npml=2, seed=2)
    s = O.System(mesh, model)
    B = np.zeros(mesh.neq)
    B[src_eq] = src_val
    nstep, dt, fr, every = 90, 1e-4, 60.0, 15
    U, E = s.forward_energy(B, nstep, dt, every, lcase=2, fr=fr)
    assert np.array_equal(U, s.forward(B, nstep, dt, lcase=2, fr=fr))
    # M_RD from the element routine
    orc = O.oracle_elem()
    xyz, inod, ID = mesh.xyz(), mesh.inod(), mesh.ID
    mrd = np.zeros(mesh.neq)
    for e in np.flatnonzero(mesh.mtel == 1):
        nodes = inod[e]
        nd1 = np.concatenate([ID[nodes, c] for c in range(3)])
        d = np.diag(orc.element_m(xyz[nodes], model.rho[e], mesh.pml_param(), 201))
        ok = nd1 >= 0
        np.add.at(mrd, nd1[ok], d[ok])
    assert np.allclose(s.diag_m_rd(), mrd, rtol=1e-14, atol=0)
    # x3 of the textbook RK4
    K, C, G = (s.csr(w) for w in range(3))
    iM = 1.0 / s.diag_m()
    Bt = B * iM
    n = mesh.neq
    x1, x2, x3 = np.zeros(n), np.zeros(n), np.zeros(n)

    def rhs(a1, a2, a3, f):
        return a2, a3, -iM * (C @ a3 + K @ a2 + G @ a1) + f * Bt

    ref = []
    for k in range(nstep + 1):
        if k % every == 0:
            ref.append(0.5 * np.dot(mrd * x3, x3))
        if k == nstep:
            break
        T = k * dt
        f0, fh, f1 = (O.time_function(t, 2, fr) for t in (T, T + 0.5 * dt, T + dt))
        a = rhs(x1, x2, x3, f0)
        b = rhs(x1 + 0.5 * dt * a[0], x2 + 0.5 * dt * a[1], x3 + 0.5 * dt * a[2], fh)
        c = rhs(x1 + 0.5 * dt * b[0], x2 + 0.5 * dt * b[1], x3 + 0.5 * dt * b[2], fh)
        d = rhs(x1 + dt * c[0], x2 + dt * c[1], x3 + dt * c[2], f1)
        x1 = x1 + dt / 6 * (a[0] + 2 * b[0] + 2 * c[0] + d[0])
        x2 = x2 + dt / 6 * (a[1] + 2 * b[1] + 2 * c[1] + d[1])
        x3 = x3 + dt / 6 * (a[2] + 2 * b[2] + 2 * c[2] + d[2])
    ref = np.array(ref)
    assert E[0] == 0.0 and ref[1:].min() > 0
    assert rel_l2(E, ref) < 1e-12


def test_petsc_numbering_of_a_node_order():
    mesh = BoxMesh(5, 4, 3, npml=(1, 1, 1, 1, 1))
    assert np.array_equal(mesh.petsc_numbering(np.arange(mesh.nnode)), mesh.ID)
    perm = np.random.default_rng(3).permutation(mesh.nnode)
    ID = mesh.petsc_numbering(perm)
    assert np.array_equal(ID >= 0, mesh.ID >= 0)
    eqs = ID[ID >= 0]
    assert np.array_equal(np.sort(eqs), np.arange(mesh.neq))
    # node-major in the given order, DOFs interleaved
    first = ID[perm].max(axis=1)
    act = first >= 0
    assert np.all(np.diff(first[act]) > 0)
    with pytest.raises(ValueError):
        mesh.petsc_numbering(np.zeros(mesh.nnode, dtype=np.int64))

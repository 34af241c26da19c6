"""Shared set-ups for the parity tests: small boxes, seeded heterogeneous
materials, sources and receivers in the reference's equation numbering."""
import numpy as np

from hirensi_amd.mesh import BoxMesh, Model, homogeneous_model, surface_grid_nodes, receiver_dofs


def random_model(mesh: BoxMesh, seed=0, spread=0.25):
    """Seeded nodal lambda/mu around Vs 200 / Vp 400, rho per element, in the
    reference's scaled units (lambda, mu ~ 80-160 MPa, rho ~ 2e-3); PML nodes
    take the material of their mapped regular-domain node."""
    rng = np.random.default_rng(seed)
    base = homogeneous_model(mesh)
    src = mesh.material_map()
    lam = base.lam * (1.0 + spread * rng.uniform(-1, 1, mesh.nnode))
    mu = base.mu * (1.0 + spread * rng.uniform(-1, 1, mesh.nnode))
    lam, mu = lam[src], mu[src]
    rho = base.rho * (1.0 + 0.05 * rng.uniform(-1, 1, mesh.nel))
    return Model(lam, mu, rho)


def small_case(nex=8, ney=8, nez=6, npml=2, top="reference", seed=0):
    mesh = BoxMesh(nex, ney, nez, npml=(npml,) * 5, top_stress_bc=top)
    model = random_model(mesh, seed)
    hx = 0.5 * (nex - 2 * npml)
    hy = 0.5 * (ney - 2 * npml)
    rec_nodes = surface_grid_nodes(mesh, 3, 3, -hx, hx, -hy, hy)
    rec_eq = receiver_dofs(mesh, rec_nodes)
    src_eq = np.array([mesh.dof(mesh.surface_node(0.0, 0.0), 2)], dtype=np.int64)
    src_val = np.array([1.0])
    return mesh, model, rec_eq, src_eq, src_val


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def oracle_postprocess(sysm, it, gl, gm, lam, mu, method="TV", factor=0.3, wreg=None):
    """Program_Main.F90:416-481 in numpy on the oracle's regularisation and
    gradient-mass vectors: returns g_lambda, g_mu, cost_reg_lambda,
    cost_reg_mu, factor_lambda, factor_mu."""
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


class OracleProblem:
    """The inversion's physics on the CPU oracle (one shot): forward + misfit,
    adjoint + control + sign, regularisation and 1/M_g (test infrastructure)."""

    def __init__(self, mesh, rho, src_eq, src_val, rec_eq, nstep, dt, fr, dis_meas, method="TV", wreg=None):
        from oracle import oracle as O
        self.O = O
        self.mesh, self.rho = mesh, rho
        self.B = np.zeros(mesh.neq)
        np.add.at(self.B, src_eq, src_val)
        self.rec_eq = np.asarray(rec_eq, dtype=np.int64)
        pos = {int(e): r for r, e in enumerate(mesh.store_eq)}
        self.rows = np.array([pos[int(e)] for e in rec_eq])
        self.nstep, self.dt, self.fr = nstep, dt, fr
        self.dis_meas = dis_meas
        self.method, self.wreg = method, wreg
        self.src = mesh.material_map()
        self.last = None

    def extend(self, v):
        return v[self.src]

    def forward_misfit(self, lam, mu):
        sysm = self.O.System(self.mesh, Model(lam, mu, self.rho))
        U = sysm.forward(self.B, self.nstep, self.dt, lcase=2, fr=self.fr)
        self.last = (sysm, U)
        return self.O.misfit(U, self.nstep, self.dt, self.rows, self.dis_meas)

    def gradient(self, it, lam, mu):
        sysm, U = self.last
        P = sysm.adjoint(U, self.nstep, self.dt, self.rec_eq, self.dis_meas)
        gl, gm = sysm.control(U, P, self.nstep, self.dt)
        gl, gm, cl, cm, _, _ = oracle_postprocess(sysm, it, gl, gm, lam, mu, self.method, 0.3, self.wreg)
        return gl, gm, cl, cm

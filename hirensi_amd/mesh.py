"""Structured 27-node hexahedral box mesh with a hybrid displacement-stress PML.

This module produces the mesh data the reference solver reads from its
pre-processing pipeline (ANSYS box mesh -> 20->27 node conversion -> PIC
partitioner), restated for a structured box:

* node coordinates ``XYZ``, connectivity ``INOD`` in the reference's 27-node
  local ordering (``Shape_Functions.F90:124-162``), element groups ``MTEL``
  (1 = regular domain, 2 = PML), and the DOF activity table ``ID``
  (node x 9: u_x u_y u_z s_xx s_yy s_zz s_xy s_xz s_yz);
* equation numbers, node-major with the active DOFs of a node interleaved
  (``step_3_PIC_Mesh_Partitioner/PIC_ReNumbering.F90:270-285``);
* ``PML_PARAM(6,4)`` = [alpha0, beta0, L, x0] per face (x+, x-, y+, y-, z+, z-)
  as built in ``Adjustments.F90:53-84``;
* nodal Lame parameters with the PML extension (PML nodes copy the material
  of their mapped interface node, ``PIC_Heterogeneous_Material.F90:1340-1349``).

DOF activity rules (``step_1_Ansys_Mesh_Generator/Ansys_in_out_hdf5.txt:617-630,
727-830`` and ``step_2_20_To_27/Main.F90:447-521``):

* a node touched by no PML element has its six stress DOFs constrained;
* every node on the outer box sides or bottom has u constrained;
* nodes touched by a PML element have their stresses active, except that on the
  top surface (z = 0) the corner and mid-edge nodes of the original 20-node mesh
  have s_zz, s_xz, s_yz constrained; the face-centre nodes added by the 20->27
  conversion keep all six (``top_stress_bc='reference'``).  ``'none'`` leaves
  every PML stress DOF active.

Lattice convention: node (i, j, k) with i in [0, 2*nex], j in [0, 2*ney],
k in [0, 2*nez]; k = 0 is the bottom, k = 2*nez the free surface z = 0.
Node number n = i + NX*(j + NY*k).  Element (ex, ey, ez) spans lattice
[2ex, 2ex+2] x [2ey, 2ey+2] x [2ez, 2ez+2]; element number
e = ex + nex*(ey + ney*ez).
"""
from __future__ import annotations

import dataclasses
import numpy as np

# local node (xi, eta, zeta) signs in the reference's 27-node order
# (Shape_Functions.F90:129-162)
LOCAL_NODES = np.array([
    (+1, -1, -1), (+1, +1, -1), (-1, +1, -1), (-1, -1, -1),
    (+1, -1, +1), (+1, +1, +1), (-1, +1, +1), (-1, -1, +1),
    (+1, 0, -1), (0, +1, -1), (-1, 0, -1), (0, -1, -1),
    (+1, -1, 0), (+1, +1, 0), (-1, +1, 0), (-1, -1, 0),
    (+1, 0, +1), (0, +1, +1), (-1, 0, +1), (0, -1, +1),
    (0, 0, -1), (+1, 0, 0), (0, +1, 0), (-1, 0, 0), (0, -1, 0), (0, 0, +1),
    (0, 0, 0),
], dtype=np.int64)

# stress DOF slots in the reference ID table (1-based 4..9 -> 0-based 3..8)
S_XX, S_YY, S_ZZ, S_XY, S_XZ, S_YZ = 3, 4, 5, 6, 7, 8
NDOF = 9


@dataclasses.dataclass
class BoxMesh:
    nex: int
    ney: int
    nez: int
    npml: tuple = (4, 4, 4, 4, 4)       # element layers: x-, x+, y-, y+, z-
    h: tuple = (1.0, 1.0, 1.0)
    top_stress_bc: str = "reference"     # 'reference' | 'none'
    alpha0: float = 5.0                  # Adjustments.F90:76
    beta0: float = 400.0                 # Adjustments.F90:77

    def __post_init__(self):
        self.npml = tuple(int(v) for v in self.npml)
        self.h = tuple(float(v) for v in self.h)
        if self.top_stress_bc not in ("reference", "none"):
            raise ValueError("top_stress_bc must be 'reference' or 'none'")
        px0, px1, py0, py1, pz0 = self.npml
        if min(self.npml) < 1:
            # PML_PARAM holds the layer thickness as a divisor (PML_FINDER_3D)
            raise ValueError("every side needs at least one PML element layer")
        if px0 + px1 >= self.nex or py0 + py1 >= self.ney or pz0 >= self.nez:
            raise ValueError("PML layers leave no regular domain")
        self.NX, self.NY, self.NZ = 2 * self.nex + 1, 2 * self.ney + 1, 2 * self.nez + 1
        self.nnode = self.NX * self.NY * self.NZ
        self.nel = self.nex * self.ney * self.nez
        hx, hy, hz = self.h
        self.xmin = -0.5 * self.nex * hx
        self.ymin = -0.5 * self.ney * hy
        self.zmin = -self.nez * hz
        # lattice coordinates per axis (node spacing h/2)
        self.xc = self.xmin + np.arange(self.NX) * (0.5 * hx)
        self.yc = self.ymin + np.arange(self.NY) * (0.5 * hy)
        self.zc = self.zmin + np.arange(self.NZ) * (0.5 * hz)
        # regular-domain element box [lo, hi)
        self.rd_lo = (px0, py0, pz0)
        self.rd_hi = (self.nex - px1, self.ney - py1, self.nez)
        self._build_dofs()

    # ------------------------------------------------------------------ layout
    def node_index(self, i, j, k):
        return i + self.NX * (j + self.NY * k)

    def element_is_rd(self):
        ex = np.arange(self.nex)[:, None, None]
        ey = np.arange(self.ney)[None, :, None]
        ez = np.arange(self.nez)[None, None, :]
        lo, hi = self.rd_lo, self.rd_hi
        rd = ((ex >= lo[0]) & (ex < hi[0]) & (ey >= lo[1]) & (ey < hi[1])
              & (ez >= lo[2]) & (ez < hi[2]))
        # element order e = ex + nex*(ey + ney*ez)  -> array [ez, ey, ex]
        return np.ascontiguousarray(rd.transpose(2, 1, 0)).ravel()

    def _node_touch(self, elem_mask):
        """Boolean lattice [k, j, i]: node touched by an element in ``elem_mask``."""
        m = elem_mask.reshape(self.nez, self.ney, self.nex)
        t = np.zeros((self.NZ, self.NY, self.NX), dtype=bool)
        for dz in range(3):
            for dy in range(3):
                for dx in range(3):
                    t[dz:dz + 2 * self.nez:2, dy:dy + 2 * self.ney:2, dx:dx + 2 * self.nex:2] |= m
        return t

    def _build_dofs(self):
        rd = self.element_is_rd()
        self.mtel = np.where(rd, 1, 2).astype(np.int32)
        self.pml_node = self._node_touch(~rd)          # [k, j, i]
        self.rd_node = self._node_touch(rd)
        act = np.zeros((self.NZ, self.NY, self.NX, NDOF), dtype=bool)
        k = np.arange(self.NZ)[:, None, None]
        j = np.arange(self.NY)[None, :, None]
        i = np.arange(self.NX)[None, None, :]
        boundary = ((i == 0) | (i == self.NX - 1) | (j == 0) | (j == self.NY - 1) | (k == 0))
        boundary = np.broadcast_to(boundary, (self.NZ, self.NY, self.NX))
        act[..., 0:3] = ~boundary[..., None]
        act[..., 3:9] = self.pml_node[..., None]
        if self.top_stress_bc == "reference":
            top = self.NZ - 1
            ii = np.arange(self.NX)[None, :]
            jj = np.arange(self.NY)[:, None]
            original = ~((ii % 2 == 1) & (jj % 2 == 1))      # not a face-centre node
            for s in (S_ZZ, S_XZ, S_YZ):
                act[top, :, :, s] &= ~original
        self.active = act.reshape(self.nnode, NDOF)           # node-major, lattice order
        eq = np.full(self.active.shape, -1, dtype=np.int64)
        eq[self.active] = np.arange(int(self.active.sum()), dtype=np.int64)
        self.ID = eq                                           # 0-based eq number, -1 inactive
        self.neq = int(self.active.sum())
        self.n_u = int(self.active[:, 0:3].sum())
        self.n_s = int(self.active[:, 3:9].sum())
        # stored set: u-DOFs of nodes touching a regular-domain element
        # (PIC_Inversion_Data_Structure.F90:276-301)
        rdn = self.rd_node.reshape(-1)
        self.store_eq = self.ID[rdn][:, 0:3].reshape(-1)
        if np.any(self.store_eq < 0):
            raise ValueError("stored RD displacement DOF is constrained")

    def petsc_numbering(self, node_order):
        """ID table of the equation numbering the reference's partitioner builds
        (PIC_ReNumbering.F90:263-285): nodes taken in ``node_order`` (PETSc node
        order, e.g. grouped by partition), the active DOFs of each node numbered
        consecutively.  Returns (nnode, 9) in lattice order, -1 where constrained."""
        node_order = np.asarray(node_order, dtype=np.int64)
        if node_order.shape != (self.nnode,) or np.any(np.sort(node_order) != np.arange(self.nnode)):
            raise ValueError("node_order must be a permutation of the lattice nodes")
        act = self.active[node_order]                     # PETSc node-major
        eq = np.full(act.shape, -1, dtype=np.int64)
        eq[act] = np.arange(int(act.sum()), dtype=np.int64)
        out = np.full((self.nnode, NDOF), -1, dtype=np.int64)
        out[node_order] = eq
        return out

    # ------------------------------------------------------------- reference data
    def xyz(self):
        X, Y, Z = np.meshgrid(self.xc, self.yc, self.zc, indexing="ij")
        return np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(),
                         Z.transpose(2, 1, 0).ravel()], axis=1)

    def inod(self):
        """Connectivity (nel, 27), 0-based node numbers, reference local ordering."""
        ex, ey, ez = np.meshgrid(np.arange(self.nex), np.arange(self.ney), np.arange(self.nez),
                                 indexing="ij")
        ex = ex.transpose(2, 1, 0).ravel()
        ey = ey.transpose(2, 1, 0).ravel()
        ez = ez.transpose(2, 1, 0).ravel()
        out = np.empty((self.nel, 27), dtype=np.int64)
        for a, (sx, sy, sz) in enumerate(LOCAL_NODES):
            out[:, a] = self.node_index(2 * ex + 1 + sx, 2 * ey + 1 + sy, 2 * ez + 1 + sz)
        return out

    def pml_param(self):
        """PML_PARAM(6,4) as a (6,4) array: rows x+, x-, y+, y-, z+, z-;
        columns alpha0, beta0, L, x0 (Adjustments.F90:53-84)."""
        px0, px1, py0, py1, pz0 = self.npml
        hx, hy, hz = self.h
        xmax, ymax = -self.xmin, -self.ymin
        p = np.zeros((6, 4))
        p[:, 0] = self.alpha0
        p[:, 1] = self.beta0
        p[0, 2], p[0, 3] = px1 * hx, xmax - px1 * hx
        p[1, 2], p[1, 3] = px0 * hx, self.xmin + px0 * hx
        p[2, 2], p[2, 3] = py1 * hy, ymax - py1 * hy
        p[3, 2], p[3, 3] = py0 * hy, self.ymin + py0 * hy
        p[4, 2], p[4, 3] = 1.0, 1.0e300        # no top PML: zone never entered
        p[5, 2], p[5, 3] = pz0 * hz, self.zmin + pz0 * hz
        return p

    def material_map(self):
        """For every node, the regular-domain node whose material it carries:
        itself inside the RD node box, else the nearest node of that box
        (the PML extension of PIC_Heterogeneous_Material.F90:1340-1349)."""
        lo = [2 * v for v in self.rd_lo]
        hi = [2 * v for v in self.rd_hi]
        i = np.clip(np.arange(self.NX), lo[0], hi[0])
        j = np.clip(np.arange(self.NY), lo[1], hi[1])
        k = np.clip(np.arange(self.NZ), lo[2], hi[2])
        K, J, I = np.meshgrid(k, j, i, indexing="ij")
        return self.node_index(I, J, K).ravel()

    def rd_node_box(self):
        """Lattice index ranges (inclusive) of the nodes touching RD elements."""
        return tuple((2 * lo, 2 * hi) for lo, hi in zip(self.rd_lo, self.rd_hi))

    def surface_node(self, x, y):
        """Node index of the surface node at (x, y) (must lie on the lattice)."""
        i = int(round((x - self.xmin) / (0.5 * self.h[0])))
        j = int(round((y - self.ymin) / (0.5 * self.h[1])))
        if not (0 <= i < self.NX and 0 <= j < self.NY):
            raise ValueError("surface point outside the mesh")
        return self.node_index(i, j, self.NZ - 1)

    def dof(self, node, comp):
        e = int(self.ID[node, comp])
        if e < 0:
            raise ValueError("constrained DOF")
        return e


# --------------------------------------------------------------------- materials
def vs_vp_rho_to_lame(vs, vp, rho):
    """Scaled units of the reference (Adjustments.F90:25, instructions.txt:104):
    rho in kg/m^3 * 1e-6, lambda and mu in MPa."""
    rho_s = np.asarray(rho, dtype=np.float64) * 1.0e-6
    mu = rho_s * np.asarray(vs, dtype=np.float64) ** 2
    lam = rho_s * np.asarray(vp, dtype=np.float64) ** 2 - 2.0 * mu
    return lam, mu, rho_s


@dataclasses.dataclass
class Model:
    lam: np.ndarray       # nodal lambda (nnode)
    mu: np.ndarray        # nodal mu (nnode)
    rho: np.ndarray       # per-element density, scaled (nel)


def homogeneous_model(mesh: BoxMesh, vs=200.0, vp=400.0, rho=2000.0) -> Model:
    lam, mu, rho_s = vs_vp_rho_to_lame(vs, vp, rho)
    return Model(np.full(mesh.nnode, float(lam)), np.full(mesh.nnode, float(mu)),
                 np.full(mesh.nel, float(rho_s)))


def layered_model(mesh: BoxMesh, vs_layers, rho_layers, interfaces, anomaly=None,
                  vp_over_vs=2.0) -> Model:
    """Horizontal layers (top first) with interfaces at depths ``interfaces``
    (z values, descending), optional Gaussian Vs anomaly
    ``(amp, radius, (x0, y0, z0))`` applied as vs *= 1 + amp*exp(-r^2/R^2).
    Materials are evaluated at RD nodes and extended into the PML; density is
    per element from the element-centre depth."""
    xyz = mesh.xyz()
    src = mesh.material_map()
    p = xyz[src]
    z = p[:, 2]
    vs = np.full(mesh.nnode, float(vs_layers[-1]))
    for zi, v in zip(reversed(interfaces), reversed(vs_layers[:-1])):
        vs = np.where(z >= zi, float(v), vs)
    if anomaly is not None:
        amp, rad, c = anomaly
        r2 = ((p - np.asarray(c)) ** 2).sum(axis=1)
        vs = vs * (1.0 + amp * np.exp(-r2 / rad ** 2))
    vp = vp_over_vs * vs
    # element-centre density
    ez = np.arange(mesh.nez)
    zc = mesh.zmin + (ez + 0.5) * mesh.h[2]
    rho_z = np.full(mesh.nez, float(rho_layers[-1]))
    for zi, r in zip(reversed(interfaces), reversed(rho_layers[:-1])):
        rho_z = np.where(zc >= zi, float(r), rho_z)
    rho_el = np.repeat(rho_z, mesh.nex * mesh.ney)
    # lambda, mu use the density of the layer the (mapped) node lies in
    rho_n = np.full(mesh.nnode, float(rho_layers[-1]))
    for zi, r in zip(reversed(interfaces), reversed(rho_layers[:-1])):
        rho_n = np.where(z >= zi, float(r), rho_n)
    lam, mu, _ = vs_vp_rho_to_lame(vs, vp, rho_n)
    return Model(lam, mu, rho_el * 1.0e-6)


# ------------------------------------------------------------------- acquisition
def surface_grid_nodes(mesh: BoxMesh, nx, ny, x0, x1, y0, y1):
    xs = np.linspace(x0, x1, nx)
    ys = np.linspace(y0, y1, ny)
    return np.array([mesh.surface_node(x, y) for y in ys for x in xs], dtype=np.int64)


def receiver_dofs(mesh: BoxMesh, nodes):
    """EqDis: 3 displacement equations per receiver node, node-major."""
    return np.array([mesh.dof(n, c) for n in nodes for c in range(3)], dtype=np.int64)


def point_load(mesh: BoxMesh, node, comp=2, amp=1.0):
    """Sparse force vector B (reference numbering): a unit vertical point load."""
    return np.array([mesh.dof(node, comp)], dtype=np.int64), np.array([float(amp)])


# ----------------------------------------------------------------------- configs
def config1():
    """Plain forward: 16x16x8 elements, 1 m, 4-element PML, homogeneous
    Vs 200 / Vp 400 / rho 2000, 15 Hz Ricker vertical point load at the
    centre surface node, dt 1e-4, 2000 steps, 9x9 surface receivers."""
    mesh = BoxMesh(16, 16, 8, npml=(4, 4, 4, 4, 4))
    model = homogeneous_model(mesh)
    rec = surface_grid_nodes(mesh, 9, 9, -4.0, 4.0, -4.0, 4.0)
    src = mesh.surface_node(0.0, 0.0)
    return dict(mesh=mesh, model=model, rec_nodes=rec, src_node=src, dt=1.0e-4,
                nstep=2000, fr=15.0)


def config2(seed=20261202, with_anomaly=True):
    """Forward + adjoint gradient: 64x64x32 elements, 6-element PML, three
    layers Vs 150/250/400 (interfaces z=-8, -18), Vp=2Vs, rho 1900/2000/2100,
    Gaussian -30% Vs anomaly of radius 4 m at seeded depth in [-20, -8],
    15 Hz Ricker at the centre surface node, 4000 steps of 1e-4 s,
    400 receivers on a 20x20 surface grid spanning the regular domain."""
    mesh = BoxMesh(64, 64, 32, npml=(6, 6, 6, 6, 6))
    rng = np.random.default_rng(seed)
    zc = float(rng.uniform(-20.0, -8.0))
    anomaly = (-0.30, 4.0, (0.0, 0.0, zc)) if with_anomaly else None
    model = layered_model(mesh, (150.0, 250.0, 400.0), (1900.0, 2000.0, 2100.0),
                          (-8.0, -18.0), anomaly=anomaly)
    half = 0.5 * (mesh.nex - 12) * mesh.h[0]
    # 20x20 grid spanning the RD surface, snapped to the half-metre lattice
    pts = np.round(np.linspace(-half, half, 20) * 2.0) / 2.0
    rec = np.array([mesh.surface_node(x, y) for y in pts for x in pts], dtype=np.int64)
    src = mesh.surface_node(0.0, 0.0)
    return dict(mesh=mesh, model=model, rec_nodes=rec, src_node=src, dt=1.0e-4,
                nstep=4000, fr=15.0, anomaly_z=zc)


def config3(seed=3, nshot=16):
    """Batched multi-shot gradient: 48x48x24 elements, 6-element PML, the
    layering of config 2 (true model with the Gaussian -30 % Vs anomaly at
    z = -12 m, starting model without it), 16 sources on a 4x4 surface grid
    with Ricker central frequencies drawn from U[10, 20] Hz, 3000 steps of
    1e-4 s, 225 shared receivers on a 15x15 surface grid spanning the regular
    domain."""
    mesh = BoxMesh(48, 48, 24, npml=(6, 6, 6, 6, 6))
    rng = np.random.default_rng(seed)
    frs = rng.uniform(10.0, 20.0, nshot)
    layers = ((150.0, 250.0, 400.0), (1900.0, 2000.0, 2100.0), (-8.0, -18.0))
    true = layered_model(mesh, *layers, anomaly=(-0.30, 4.0, (0.0, 0.0, -12.0)))
    start = layered_model(mesh, *layers)
    half = 0.5 * (mesh.nex - 12) * mesh.h[0]
    pts = np.round(np.linspace(-half, half, 15) * 2.0) / 2.0
    rec = np.array([mesh.surface_node(x, y) for y in pts for x in pts], dtype=np.int64)
    side = int(round(np.sqrt(nshot)))
    sp = np.round(np.linspace(-0.6 * half, 0.6 * half, side) * 2.0) / 2.0
    src = np.array([mesh.surface_node(x, y) for y in sp for x in sp], dtype=np.int64)[:nshot]
    return dict(mesh=mesh, true=true, start=start, rec_nodes=rec, src_nodes=src, frs=frs, dt=1.0e-4,
                nstep=3000)


def ellipsoid_model(mesh: BoxMesh, vs_layers, rho_layers, interfaces, inclusions, vp_over_vs=2.0) -> Model:
    """Layered model (as :func:`layered_model`) with ellipsoidal inclusions
    ``(amp, (ax, ay, az), (x0, y0, z0))``: vs *= 1 + amp inside
    ((x - x0)/ax)^2 + ((y - y0)/ay)^2 + ((z - z0)/az)^2 <= 1."""
    base = layered_model(mesh, vs_layers, rho_layers, interfaces, vp_over_vs=vp_over_vs)
    p = mesh.xyz()[mesh.material_map()]
    z = p[:, 2]
    rho = np.full(mesh.nnode, float(rho_layers[-1]))
    vs = np.full(mesh.nnode, float(vs_layers[-1]))
    for zi, r, v in zip(reversed(interfaces), reversed(rho_layers[:-1]), reversed(vs_layers[:-1])):
        rho = np.where(z >= zi, float(r), rho)
        vs = np.where(z >= zi, float(v), vs)
    for amp, axes, c in inclusions:
        q = (((p - np.asarray(c)) / np.asarray(axes)) ** 2).sum(axis=1)
        vs = np.where(q <= 1.0, vs * (1.0 + amp), vs)
    lam, mu, _ = vs_vp_rho_to_lame(vs, vp_over_vs * vs, rho)
    return Model(lam, mu, base.rho)


def gradient_model(mesh: BoxMesh, vs_top, vs_bottom, rho_layers, interfaces, vp_over_vs=2.0) -> Model:
    """Vs linear in depth from vs_top at the surface to vs_bottom at the bottom
    of the regular domain (constant below), densities of the layers."""
    base = layered_model(mesh, (vs_top,) * len(rho_layers), rho_layers, interfaces, vp_over_vs=vp_over_vs)
    p = mesh.xyz()[mesh.material_map()]
    z = p[:, 2]
    zb = mesh.zmin + mesh.rd_lo[2] * mesh.h[2]
    t = np.clip(z / zb, 0.0, 1.0)
    vs = vs_top + (vs_bottom - vs_top) * t
    rho = np.full(mesh.nnode, float(rho_layers[-1]))
    for zi, r in zip(reversed(interfaces), reversed(rho_layers[:-1])):
        rho = np.where(z >= zi, float(r), rho)
    lam, mu, _ = vs_vp_rho_to_lame(vs, vp_over_vs * vs, rho)
    return Model(lam, mu, base.rho)


def config4(seed=4, nx=96, ny=96, nz=48, npml=8, nrec_side=30, nshot_side=3):
    """Full L-BFGS FWI: 96x96x48 elements (1 m), 8-element PML; true model of four
    layers (Vs 150/220/300/400 m/s, interfaces z = -6, -14, -26 m, rho
    1900/1950/2050/2100) with two ellipsoidal inclusions of +25 % and -25 % Vs
    placed by default_rng(seed); starting model Vs linear from 150 m/s at the
    surface to 450 m/s at the bottom of the regular domain; 9 shots on a 3x3
    surface grid at 12 Hz (Ricker), 900 receivers on a 30x30 surface grid
    spanning the regular domain, 5000 steps of 1e-4 s.  Smaller meshes keep the
    same layout (scaled extents)."""
    mesh = BoxMesh(nx, ny, nz, npml=(npml,) * 5)
    rng = np.random.default_rng(seed)
    hx = 0.5 * (nx - 2 * npml) * mesh.h[0]
    hy = 0.5 * (ny - 2 * npml) * mesh.h[1]
    depth = (nz - npml) * mesh.h[2]
    f = depth / 40.0
    layers = ((150.0, 220.0, 300.0, 400.0), (1900.0, 1950.0, 2050.0, 2100.0), (-6.0 * f, -14.0 * f, -26.0 * f))
    inc = []
    for amp in (0.25, -0.25):
        axes = tuple(rng.uniform(0.08, 0.16, 3) * np.array([hx, hy, depth]))
        c = (rng.uniform(-0.5, 0.5) * hx, rng.uniform(-0.5, 0.5) * hy, -rng.uniform(0.25, 0.7) * depth)
        inc.append((amp, axes, c))
    true = ellipsoid_model(mesh, *layers, inc)
    start = gradient_model(mesh, 150.0, 450.0, layers[1], layers[2])
    pts = np.round(np.linspace(-hx, hx, nrec_side) * 2.0) / 2.0
    ptsy = np.round(np.linspace(-hy, hy, nrec_side) * 2.0) / 2.0
    rec = np.array([mesh.surface_node(x, y) for y in ptsy for x in pts], dtype=np.int64)
    sx = np.round(np.linspace(-0.5 * hx, 0.5 * hx, nshot_side) * 2.0) / 2.0
    sy = np.round(np.linspace(-0.5 * hy, 0.5 * hy, nshot_side) * 2.0) / 2.0
    src = np.array([mesh.surface_node(x, y) for y in sy for x in sx], dtype=np.int64)
    return dict(mesh=mesh, true=true, start=start, rec_nodes=rec, src_nodes=src, fr=12.0, dt=1.0e-4,
                nstep=5000, inclusions=inc)

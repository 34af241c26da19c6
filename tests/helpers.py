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

"""The oracle's central-difference variant (EXPLICIT_Central_Difference_3,
Obsolete.F90:815-995) against what fixes it independently of its code:

* the defining recurrence, written unscaled on the assembled K, C, G and the
  lumped M with the running integral as an explicit sum:
  M (u_r - 2 u_{r-1} + u_{r-2}) + dt^2 K u_{r-1} + dt C (u_{r-1} - u_{r-2})
  + dt^3 G u_{r-1} + dt^2 G dt sum_{k<r} u_k = dt^2 B f(r dt)
  (the oracle builds the scaled, merged K' = -dt^2 K - dt C - dt^3 G and
  running vectors instead);
* the scheme approximates the same ODE as the RK4 solve: receiver traces agree
  with the RK4 oracle at small dt, and the gap shrinks when dt is halved.
"""
import numpy as np

from hirensi_amd.mesh import BoxMesh
from oracle import oracle as O
from tests.helpers import random_model, small_case


def test_cd_oracle_satisfies_the_recurrence():
    mesh = BoxMesh(6, 6, 5, npml=(2, 2, 2, 2, 2))
    model = random_model(mesh, seed=4)
    sysm = O.System(mesh, model)
    K, C, G = sysm.csr(0), sysm.csr(1), sysm.csr(2)
    M = sysm.diag_m()
    B = np.zeros(mesh.neq)
    B[mesh.dof(mesh.surface_node(0.0, 0.0), 2)] = 1.0
    nstep, dt, fr = 120, 2.0e-4, 60.0
    U = sysm.cd_forward(B, nstep, dt, lcase=2, fr=fr, store_eq=np.arange(mesh.neq))
    assert np.array_equal(U[0], np.zeros(mesh.neq)) and np.array_equal(U[1], np.zeros(mesh.neq))
    v = U[1:]                                   # v[k] = u_k, k = 0..nstep-1
    assert np.abs(v).max() > 0
    worst = 0.0
    for r in range(2, nstep):
        his = dt * v[1:r].sum(axis=0)
        f = O.time_function(r * dt, 2, fr)
        lhs = M * (v[r] - 2 * v[r - 1] + v[r - 2])
        terms = [dt * dt * (K @ v[r - 1]), dt * (C @ (v[r - 1] - v[r - 2])), dt ** 3 * (G @ v[r - 1]),
                 dt * dt * (G @ his), -dt * dt * f * B]
        res = lhs + sum(terms)
        scale = np.abs(lhs).max() + max(np.abs(t).max() for t in terms)
        worst = max(worst, np.abs(res).max() / scale)
    assert worst < 1e-12, worst


def test_cd_oracle_approaches_the_rk4_solution():
    mesh, model, rec_eq, src_eq, src_val = small_case(seed=2)
    sysm = O.System(mesh, model)
    B = np.zeros(mesh.neq)
    np.add.at(B, src_eq, src_val)
    pos = {int(e): r for r, e in enumerate(mesh.store_eq)}
    rows = np.array([pos[int(e)] for e in rec_eq])
    T, fr = 0.03, 60.0
    ref = sysm.forward(B, 600, T / 600, lcase=2, fr=fr)[:, rows]      # RK4 at dt = 5e-5
    gaps = []
    for nstep in (150, 300):
        dt = T / nstep
        U = sysm.cd_forward(B, nstep, dt, lcase=2, fr=fr)[:, rows]
        # CD row n holds u_{n-1}: compare u_{n-1} with the RK4 solution at (n - 1) dt
        cd = U[1:]
        rk = ref[:: 600 // nstep][:nstep]
        gaps.append(np.linalg.norm(cd - rk) / np.linalg.norm(rk))
    assert gaps[0] < 0.2 and gaps[1] < gaps[0] / 1.8, gaps

"""Mesh/DOF model rules and the C ABI surface (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from hirensi_amd import mesh as Mh
from hirensi_amd.mesh import BoxMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("dims,npml,counts", [
    ((16, 16, 8), 4, (18513, 46128, 100278, 146406, 7803)),
    ((64, 64, 32), 6, (1081665, 3096768, 3179982, 6276750, 1752975)),
    ((48, 48, 24), 6, (461041, 1299600, 1677390, 2976990, 591519)),
    ((32, 32, 16), 4, (139425, 381024, 518454, 899478, 180075)),
])
def test_dof_counts_of_the_benchmark_meshes(dims, npml, counts):
    m = BoxMesh(*dims, npml=(npml,) * 5, top_stress_bc="none")
    assert (m.nnode, m.n_u, m.n_s, m.neq, len(m.store_eq)) == counts


def test_dof_rules():
    m = BoxMesh(8, 7, 6, npml=(2, 1, 1, 2, 2))
    act = m.active.reshape(m.NZ, m.NY, m.NX, 9)
    # box sides and bottom: u fixed
    assert not act[0, :, :, :3].any() and not act[:, 0, :, :3].any() and not act[:, :, -1, :3].any()
    # top surface: u free away from the sides
    assert act[-1, 1:-1, 1:-1, :3].all()
    # nodes touched only by regular-domain elements carry no stress DOFs
    rd_only = ~m.pml_node
    assert not act[rd_only][:, 3:].any()
    assert act[m.pml_node & (np.arange(m.NZ)[:, None, None] < m.NZ - 1)][:, 3:].all()
    # top surface: corner/mid-edge nodes of PML elements lose s_zz, s_xz, s_yz
    top = act[-1]
    j, i = np.meshgrid(np.arange(m.NY), np.arange(m.NX), indexing="ij")
    face_centre = (i % 2 == 1) & (j % 2 == 1)
    pml_top = m.pml_node[-1]
    for s, on in ((Mh.S_XX, True), (Mh.S_YY, True), (Mh.S_XY, True), (Mh.S_ZZ, False), (Mh.S_XZ, False),
                  (Mh.S_YZ, False)):
        assert (top[pml_top & ~face_centre][:, s] == on).all()
        assert top[pml_top & face_centre][:, s].all()
    # node-major interleaved numbering
    eq = m.ID[m.active]
    assert np.array_equal(eq, np.arange(m.neq))


def test_connectivity_and_material_map():
    m = BoxMesh(5, 4, 3, npml=(1, 1, 1, 1, 1))
    inod, xyz = m.inod(), m.xyz()
    # every element's local node a sits at centre + h/2 * (xi, eta, zeta)
    for e in (0, 7, m.nel - 1):
        c = xyz[inod[e]].mean(axis=0)
        assert np.allclose(xyz[inod[e]] - c, 0.5 * np.asarray(m.h) * Mh.LOCAL_NODES)
    src = m.material_map()
    lo = [2 * v for v in m.rd_lo]
    hi = [2 * v for v in m.rd_hi]
    k, j, i = np.unravel_index(src, (m.NZ, m.NY, m.NX))
    assert (i >= lo[0]).all() and (i <= hi[0]).all() and (j >= lo[1]).all() and (j <= hi[1]).all()
    assert (k >= lo[2]).all()
    # RD-box nodes map to themselves
    box = m.rd_node.reshape(-1)
    assert np.array_equal(src[box], np.nonzero(box)[0])


def test_pml_param_layout():
    m = BoxMesh(16, 16, 8, npml=(4, 4, 4, 4, 4))
    p = m.pml_param()
    assert np.allclose(p[:, 0], 5.0) and np.allclose(p[:, 1], 400.0)
    assert np.allclose(p[[0, 1, 2, 3, 5], 2], 4.0)
    assert np.allclose(p[[0, 1, 2, 3, 5], 3], [4.0, -4.0, 4.0, -4.0, -4.0])


def test_configs_build():
    c = Mh.config1()
    assert c["mesh"].neq == 144390 and len(c["rec_nodes"]) == 81
    c2 = Mh.config2()
    assert len(c2["rec_nodes"]) == 400 and len(set(c2["rec_nodes"].tolist())) == 400
    assert -20.0 <= c2["anomaly_z"] <= -8.0
    lam, mu = c2["model"].lam, c2["model"].mu
    assert lam.min() > 0 and mu.min() > 0
    # Vs = sqrt(mu / rho): three layers, -30 % at the anomaly centre
    m = c2["mesh"]
    zc = c2["anomaly_z"]
    top = m.surface_node(0.0, 0.0)
    assert abs(np.sqrt(mu[top] / 1.9e-3) - 150.0) < 1e-3
    k = int(round((zc - m.zmin) / (0.5 * m.h[2])))
    n = m.node_index(m.NX // 2, m.NY // 2, k)
    rho = 2.0e-3 if m.zc[k] >= -18.0 else 2.1e-3
    layer = 250.0 if m.zc[k] >= -18.0 else 400.0
    r2 = (m.zc[k] - zc) ** 2
    assert abs(np.sqrt(mu[n] / rho) - layer * (1 - 0.3 * np.exp(-r2 / 16.0))) < 1e-9


def _header_functions(name="hirensi.h"):
    txt = open(os.path.join(ROOT, "include", name)).read()
    return sorted(set(re.findall(r"^\s*(?:int|double|const char\*)\s+(hr_\w+)\s*\(", txt, re.M)))


def test_cabi_library_exports_every_declared_symbol():
    from hirensi_amd import native
    lib = ctypes.CDLL(native.LIB_PATH)
    names = _header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(native.EXPORTS)
    nn = _header_functions("hirensi_nn.h")
    assert len(nn) == 12
    for n in nn:
        assert hasattr(lib, n), n
    assert set(nn) == set(native.NN_EXPORTS)
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == ["hirensi.h", "hirensi_nn.h"]


def test_fortran_module_binds_every_declared_symbol():
    src = open(os.path.join(ROOT, "hirensi_amd", "fortran", "hirensi_mod.F90")).read()
    bound = set(re.findall(r'bind\(C, name="(hr_\w+)"\)', src))
    declared = set(_header_functions()) | set(_header_functions("hirensi_nn.h"))
    assert declared <= bound, sorted(declared - bound)


def test_host_time_function_matches_oracle():
    from hirensi_amd import solver as S
    from oracle import oracle as O
    cases = [(2, 15.0), (2, 60.0), (3, 0.0), (1, 5.0), (6, 0.0), (7, 15.0), (7, 40.0), (8, 0.0), (9, 0.0), (10, 0.0)]
    for t in np.linspace(0, 0.2, 37):
        for lc, fr in cases:
            assert S.time_function(t, lc, fr) == O.time_function(t, lc, fr), (t, lc)
    # Chopra's cosine wave over its 10 s duration and past it
    for t in np.linspace(0, 10.5, 211):
        for lc in (4, 5):
            assert S.time_function(t, lc, 0.0) == O.time_function(t, lc, 0.0), (t, lc)
    with pytest.raises(S.HrError):
        S.time_function(0.1, 11, 15.0)

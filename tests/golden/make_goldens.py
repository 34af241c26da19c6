"""Writes tests/golden/elements.npz: inputs and outputs of the reference's own
element routines (SHAPE_3_27, PML_FINDER_3D, ELEMENT_K/M/EN/EA/FS/g_3D,
ELEMENT_R_Mg_3D, ELEMENT_TV_TN_Reg_3D),
compiled from its Fortran sources into oracle/_ref/libhr_ref.so by
oracle/ref.mk.  Run in the build container:

    make -f oracle/ref.mk && python tests/golden/make_goldens.py < /dev/null

(stdin closed: the reference pauses on READ(*,*) after a negative Jacobian).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O           # noqa: E402
from hirensi_amd.mesh import BoxMesh, LOCAL_NODES  # noqa: E402

# (w_top, w_bot, w_Length) of Adjustments.F90:174-176
WREG = np.array([50.0, 1.0, 45.0])
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elements.npz")


def element_xt(center, h=(1.0, 1.0, 1.0), distort=0.0, rng=None):
    xt = np.asarray(center, dtype=np.float64) + 0.5 * np.asarray(h) * LOCAL_NODES
    if distort:
        xt = xt + distort * rng.uniform(-1.0, 1.0, xt.shape)
    return xt


def main():
    r = O.ref_elem()
    rng = np.random.default_rng(20261202)
    mesh = BoxMesh(16, 16, 8, npml=(4, 4, 4, 4, 4))
    pp = mesh.pml_param()
    cases = {
        "rd_cube": element_xt((0.5, 0.5, -0.5)),
        "rd_distorted": element_xt((1.5, -0.5, -2.5), distort=0.03, rng=rng),
        "rd_aniso": element_xt((0.25, 0.5, -0.75), h=(0.5, 1.0, 1.5)),
        "pml_corner": element_xt((-7.5, -7.5, -7.5)),
        "pml_edge": element_xt((5.5, -6.5, -1.5)),
        "pml_face": element_xt((0.5, 0.5, -6.5)),
        "pml_distorted": element_xt((6.5, 0.5, -2.5), distort=0.03, rng=rng),
    }
    data = {"pml_param": pp}
    for name, xt in cases.items():
        lam = rng.uniform(100.0, 200.0, 27)
        mu = rng.uniform(50.0, 100.0, 27)
        eu = rng.standard_normal(81)
        ep = rng.standard_normal(81)
        data[f"{name}/xt"] = xt
        data[f"{name}/lam"] = lam
        data[f"{name}/mu"] = mu
        data[f"{name}/eu"] = eu
        data[f"{name}/ep"] = ep
        data[f"{name}/K"] = r.element_k(xt, lam, mu)
        for f in (201, 202, 203, 204, 205):
            data[f"{name}/M{f}"] = np.diag(r.element_m(xt, 2.0e-3, pp, f)).copy()
        for f in (301, 302, 303, 304):
            data[f"{name}/EN{f}"] = np.diag(r.element_en(xt, pp, f)).copy()
        for f in range(101, 107):
            data[f"{name}/EA{f}"] = r.element_ea(xt, lam, mu, pp, f)
        gm, gl = r.element_g(xt, eu, ep)
        data[f"{name}/g_mu"] = gm
        data[f"{name}/g_lambda"] = gl
        for tag, w in (("", None), ("_w", WREG)):
            er, emg = r.element_r_mg(xt, w)
            data[f"{name}/ER{tag}"] = er
            data[f"{name}/EMg{tag}"] = emg
            rm, rl, cost = r.element_tv(xt, mu, lam, w)
            data[f"{name}/TV_mu{tag}"] = rm
            data[f"{name}/TV_lambda{tag}"] = rl
            data[f"{name}/TV_cost{tag}"] = cost
    top = element_xt((0.5, 0.5, -0.5))
    data["fs/xt"] = top
    data["fs/face5"] = r.element_fs(top, 5)
    data["fs/face6"] = r.element_fs(top, 6)
    xg = rng.uniform([-8.0, -8.0, -8.0], [8.0, 8.0, 0.0], size=(40, 3))
    xg[:4] = [[4.0, 4.0, -4.0], [-4.0, 0.0, -1.0], [8.0, -8.0, -8.0], [0.0, 0.0, 0.0]]
    data["pml/xg"] = xg
    data["pml/pab"] = np.array([r.pml_finder_3d(p, pp) for p in xg])
    pts = np.vstack([np.array([[a, b, c] for c in (-1, 0, 1) for b in (-1, 0, 1) for a in (-1, 0, 1)], float),
                     rng.uniform(-1.0, 1.0, size=(8, 3))])
    data["shape/pts"] = pts
    fn, d = zip(*[r.shape_3_27(*p) for p in pts])
    data["shape/fn"] = np.array(fn)
    data["shape/dfxi"] = np.array(d)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(data)} arrays)")


if __name__ == "__main__":
    main()

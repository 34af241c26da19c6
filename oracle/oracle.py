"""ctypes access to the CPU oracle (``oracle/liboracle.so``) and, where it has
been built, to the reference's own element routines (``oracle/_ref/libhr_ref.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` as the checker; the product never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libhr_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.or_shape_3_27.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.or_pml_finder_3d.argtypes = [_dp, _dp, _dp]
        L.or_element_k.argtypes = [_dp, _dp, _dp, _dp]
        L.or_element_m.argtypes = [_dp, C.c_double, _dp, C.c_int, _dp]
        L.or_element_en.argtypes = [_dp, _dp, C.c_int, _dp]
        L.or_element_ea.argtypes = [_dp, _dp, _dp, _dp, C.c_int, _dp]
        L.or_element_fs.argtypes = [_dp, C.c_int, _dp]
        L.or_element_g.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.or_time_function.argtypes = [C.c_double, C.c_int, C.c_double]
        L.or_time_function.restype = C.c_double
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        L.or_system_create.argtypes = [C.c_int64, C.c_int64, _dp, _ip, _i32p, _ip, C.c_int64,
                                       _dp, _dp, _dp, _dp]
        L.or_system_create.restype = C.c_void_p
        L.or_system_destroy.argtypes = [C.c_void_p]
        L.or_nnz.argtypes = [C.c_void_p, C.c_int]
        L.or_nnz.restype = C.c_int64
        L.or_get_csr.argtypes = [C.c_void_p, C.c_int, _ip, _i32p, _dp]
        L.or_get_diag_m.argtypes = [C.c_void_p, _dp]
        L.or_matmult_add.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
        L.or_rk4_forward.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, C.c_int, C.c_double,
                                     C.c_int64, _ip, _dp]
        L.or_rk4_forward_energy.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, C.c_int, C.c_double,
                                            C.c_int64, _ip, _dp, C.c_int, _dp]
        L.or_get_diag_m_rd.argtypes = [C.c_void_p, _dp]
        L.or_cd_forward.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, C.c_int, C.c_double,
                                    C.c_int64, _ip, _dp]
        L.or_rk4_adjoint.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int64, _ip, _dp,
                                     C.c_int, _ip, _ip, _dp, _dp]
        L.or_control.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int64, _ip, _dp, _dp, _dp, _dp]
        L.or_misfit.argtypes = [C.c_int, C.c_double, C.c_int, _ip, _ip, _dp, C.c_int64, _dp]
        L.or_misfit.restype = C.c_double
        L.or_element_r_mg.argtypes = [_dp, C.c_void_p, _dp, _dp]
        L.or_element_tv.argtypes = [_dp, _dp, _dp, C.c_void_p, _dp, _dp, _dp]
        L.or_gradient_mass.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp]
        L.or_regularization.argtypes = [C.c_void_p, C.c_int, C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """The reference's element routines compiled from its Fortran sources."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/libhr_ref.so is not built")
        R = C.CDLL(REF_PATH)
        R.ref_shape_3_27.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        R.ref_pml_finder_3d.argtypes = [_dp, _dp, _dp]
        R.ref_element_k.argtypes = [_dp, _dp, _dp, _dp]
        R.ref_element_m.argtypes = [_dp, C.c_double, _dp, C.c_int, _dp]
        R.ref_element_en.argtypes = [_dp, _dp, C.c_int, _dp]
        R.ref_element_ea.argtypes = [_dp, _dp, _dp, _dp, C.c_int, _dp]
        R.ref_element_fs.argtypes = [_dp, C.c_int, _dp]
        R.ref_element_g.argtypes = [_dp, _dp, _dp, _dp, _dp]
        R.ref_element_r_mg.argtypes = [_dp, C.c_int, _dp, _dp, _dp]
        R.ref_element_tv.argtypes = [_dp, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp]
        _ref = R
    return _ref


# ---------------------------------------------------------------- element level
# Arrays follow the Fortran layout: XT(3,27), matrices column-major.
def _f(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).T).T  # keep values


def _colmajor(shape):
    return np.zeros(shape[::-1], dtype=np.float64)   # returned .T gives (rows, cols)


class _Elem:
    def __init__(self, L, pre):
        self.L, self.p = L, pre

    def _fn(self, name):
        return getattr(self.L, f"{self.p}_{name}")

    def shape_3_27(self, x1, x2, x3):
        fn = np.zeros(27)
        d = np.zeros(81)
        self._fn("shape_3_27")(x1, x2, x3, fn, d)
        return fn, d.reshape(3, 27).T

    def pml_finder_3d(self, xg, pml_param):
        out = np.zeros(19)
        self._fn("pml_finder_3d")(np.ascontiguousarray(xg, dtype=np.float64),
                                  np.ascontiguousarray(np.asarray(pml_param).T, dtype=np.float64), out)
        return out

    @staticmethod
    def _xt(xt):
        """xt: (27, 3) node coordinates -> Fortran XT(3,27) buffer."""
        return np.ascontiguousarray(np.asarray(xt, dtype=np.float64).reshape(27, 3))

    def element_k(self, xt, lam, mu):
        out = np.zeros(81 * 81)
        self._fn("element_k")(self._xt(xt), np.ascontiguousarray(lam, dtype=np.float64),
                              np.ascontiguousarray(mu, dtype=np.float64), out)
        return out.reshape(81, 81).T

    def element_m(self, xt, rho, pml_param, iflag):
        out = np.zeros(81 * 81)
        self._fn("element_m")(self._xt(xt), float(rho),
                              np.ascontiguousarray(np.asarray(pml_param).T, dtype=np.float64), int(iflag), out)
        return out.reshape(81, 81).T

    def element_en(self, xt, pml_param, iflag):
        out = np.zeros(162 * 162)
        self._fn("element_en")(self._xt(xt), np.ascontiguousarray(np.asarray(pml_param).T, dtype=np.float64),
                               int(iflag), out)
        return out.reshape(162, 162).T

    def element_ea(self, xt, lam, mu, pml_param, iflag):
        out = np.zeros(81 * 162)
        self._fn("element_ea")(self._xt(xt), np.ascontiguousarray(lam, dtype=np.float64),
                               np.ascontiguousarray(mu, dtype=np.float64),
                               np.ascontiguousarray(np.asarray(pml_param).T, dtype=np.float64), int(iflag), out)
        return out.reshape(162, 81).T

    def element_fs(self, xt, nface):
        out = np.zeros(81)
        self._fn("element_fs")(self._xt(xt), int(nface), out)
        return out

    def element_g(self, xt, eu, ep):
        gm = np.zeros(27)
        gl = np.zeros(27)
        self._fn("element_g")(self._xt(xt), np.ascontiguousarray(eu, dtype=np.float64),
                              np.ascontiguousarray(ep, dtype=np.float64), gm, gl)
        return gm, gl


    # Regularisation weighting: wreg = (w_top, w_bot, w_Length) or None for the
    # unweighted scheme (Weighted_Regularization = 'No', Adjustments.F90:174-177).
    def element_r_mg(self, xt, wreg=None):
        er = np.zeros(27 * 27)
        emg = np.zeros(27)
        if self.p == "ref":
            w = np.zeros(3) if wreg is None else np.ascontiguousarray(wreg, dtype=np.float64)
            self._fn("element_r_mg")(self._xt(xt), int(wreg is not None), w, er, emg)
        else:
            w = _opt(wreg)
            self._fn("element_r_mg")(self._xt(xt), _addr(w), er, emg)
        return er.reshape(27, 27).T, emg

    def element_tv(self, xt, e_mu, e_lam, wreg=None):
        rm = np.zeros(27)
        rl = np.zeros(27)
        cost = np.zeros(2)
        em = np.ascontiguousarray(e_mu, dtype=np.float64)
        el = np.ascontiguousarray(e_lam, dtype=np.float64)
        if self.p == "ref":
            w = np.zeros(3) if wreg is None else np.ascontiguousarray(wreg, dtype=np.float64)
            self._fn("element_tv")(self._xt(xt), em, el, int(wreg is not None), w, rm, rl, cost)
        else:
            w = _opt(wreg)
            self._fn("element_tv")(self._xt(xt), em, el, _addr(w), rm, rl, cost)
        return rm, rl, cost           # cost = (lambda part, mu part)


def _opt(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _addr(a):
    return None if a is None else a.ctypes.data


def oracle_elem():
    return _Elem(lib(), "or")


def ref_elem():
    return _Elem(ref(), "ref")


def time_function(t, lcase=2, fr=15.0):
    return lib().or_time_function(float(t), int(lcase), float(fr))


# ---------------------------------------------------------------- system level
class System:
    """Assembled K, C, G (CSR, PETSc insertion semantics) and the lumped mass."""

    def __init__(self, mesh, model, nthreads=0):
        L = lib()
        if nthreads:
            L.or_set_threads(int(nthreads))
        self.mesh = mesh
        self.neq = mesh.neq
        xyz = np.ascontiguousarray(mesh.xyz())
        inod = np.ascontiguousarray(mesh.inod())
        mtel = np.ascontiguousarray(mesh.mtel, dtype=np.int32)
        ID = np.ascontiguousarray(mesh.ID, dtype=np.int64)
        pp = np.ascontiguousarray(mesh.pml_param().T)      # Fortran (6,4) column-major
        self.h = L.or_system_create(mesh.nnode, mesh.nel, xyz, inod, mtel, ID, mesh.neq,
                                    np.ascontiguousarray(model.lam, dtype=np.float64),
                                    np.ascontiguousarray(model.mu, dtype=np.float64),
                                    np.ascontiguousarray(model.rho, dtype=np.float64), pp)

    def close(self):
        if self.h:
            lib().or_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def csr(self, which):
        """which: 0 = K, 1 = C, 2 = G -> scipy.sparse.csr_matrix."""
        import scipy.sparse as sp
        L = lib()
        nnz = L.or_nnz(self.h, which)
        rp = np.zeros(self.neq + 1, dtype=np.int64)
        col = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz)
        L.or_get_csr(self.h, which, rp, col, val)
        return sp.csr_matrix((val, col, rp), shape=(self.neq, self.neq))

    def diag_m(self):
        m = np.zeros(self.neq)
        lib().or_get_diag_m(self.h, m)
        return m

    def diag_m_rd(self):
        """DIAG_M_RD: the lumped mass of the regular-domain elements only."""
        m = np.zeros(self.neq)
        lib().or_get_diag_m_rd(self.h, m)
        return m

    def forward_energy(self, B, nstep, dt, every, lcase=2, fr=15.0):
        """The RK4 forward with the regular-domain energy at steps 0, every, ...:
        returns (U_store, energy)."""
        store_eq = np.ascontiguousarray(self.mesh.store_eq, dtype=np.int64)
        ns = len(store_eq)
        u = np.zeros((nstep + 1) * ns)
        e = np.zeros(nstep // every + 1)
        lib().or_rk4_forward_energy(self.h, np.ascontiguousarray(B, dtype=np.float64), int(nstep), float(dt),
                                    int(lcase), float(fr), ns, store_eq, u, int(every), e)
        return u.reshape(nstep + 1, ns), e

    def gradient_mass(self, wreg=None):
        """Nodal 1/Mg over regular-domain elements (0 where Mg <= 1e-10) and
        the 0/1 restriction mask."""
        img = np.zeros(self.mesh.nnode)
        res = np.zeros(self.mesh.nnode)
        w = _opt(wreg)
        lib().or_gradient_mass(self.h, _addr(w), img, res)
        return img, res

    def regularization(self, method, lam, mu, wreg=None):
        """Unscaled nodal regularisation vectors and costs.  method 'TV' or 'TN'.
        Returns reg_lam, reg_mu, cost_lam, cost_mu."""
        n = self.mesh.nnode
        rl, rm, cost = np.zeros(n), np.zeros(n), np.zeros(2)
        w = _opt(wreg)
        lib().or_regularization(self.h, {"TV": 0, "TN": 1}[method], _addr(w),
                                np.ascontiguousarray(lam, dtype=np.float64),
                                np.ascontiguousarray(mu, dtype=np.float64), rl, rm, cost)
        return rl, rm, cost[0], cost[1]

    def matmult_add(self, which, x, y, transpose=False):
        z = np.zeros(self.neq)
        lib().or_matmult_add(self.h, which, int(transpose), np.ascontiguousarray(x, dtype=np.float64),
                             np.ascontiguousarray(y, dtype=np.float64), z)
        return z

    def forward(self, B, nstep, dt, lcase=2, fr=15.0, store_eq=None):
        store_eq = self.mesh.store_eq if store_eq is None else store_eq
        store_eq = np.ascontiguousarray(store_eq, dtype=np.int64)
        ns = len(store_eq)
        u = np.zeros((nstep + 1) * ns)
        lib().or_rk4_forward(self.h, np.ascontiguousarray(B, dtype=np.float64), int(nstep), float(dt),
                             int(lcase), float(fr), ns, store_eq, u)
        return u.reshape(nstep + 1, ns)              # row n = U_Store(:, n)

    def cd_forward(self, B, nstep, dt, lcase=2, fr=15.0, store_eq=None):
        """EXPLICIT_Central_Difference_3: rows n = 1..nstep hold u at the start
        of step n (row 0 stays zero)."""
        store_eq = self.mesh.store_eq if store_eq is None else store_eq
        store_eq = np.ascontiguousarray(store_eq, dtype=np.int64)
        ns = len(store_eq)
        u = np.zeros((nstep + 1) * ns)
        lib().or_cd_forward(self.h, np.ascontiguousarray(B, dtype=np.float64), int(nstep), float(dt),
                            int(lcase), float(fr), ns, store_eq, u)
        return u.reshape(nstep + 1, ns)

    def adjoint(self, u_store, nstep, dt, rec_eq, dis_meas, store_eq=None):
        store_eq = self.mesh.store_eq if store_eq is None else store_eq
        store_eq = np.ascontiguousarray(store_eq, dtype=np.int64)
        ns = len(store_eq)
        rec_eq = np.ascontiguousarray(rec_eq, dtype=np.int64)
        pos = {int(e): r for r, e in enumerate(store_eq)}
        rec_row = np.array([pos[int(e)] for e in rec_eq], dtype=np.int64)
        p = np.zeros((nstep + 1) * ns)
        lib().or_rk4_adjoint(self.h, int(nstep), float(dt), ns, store_eq,
                             np.ascontiguousarray(u_store.reshape(-1)), len(rec_eq), rec_eq, rec_row,
                             np.ascontiguousarray(dis_meas, dtype=np.float64).reshape(-1), p)
        return p.reshape(nstep + 1, ns)

    def control(self, u_store, p_store, nstep, dt, store_eq=None):
        store_eq = self.mesh.store_eq if store_eq is None else store_eq
        store_eq = np.ascontiguousarray(store_eq, dtype=np.int64)
        gl = np.zeros(self.mesh.nnode)
        gm = np.zeros(self.mesh.nnode)
        lib().or_control(self.h, int(nstep), float(dt), len(store_eq), store_eq,
                         np.ascontiguousarray(u_store.reshape(-1)), np.ascontiguousarray(p_store.reshape(-1)),
                         gl, gm)
        return gl, gm


def misfit(u_store, nstep, dt, rec_rows, dis_meas, rec_eq=None):
    """Compute_Misfit.  rec_eq (default: the rows) orders and groups the
    receivers: those sharing an equation add their residuals before squaring."""
    ns = u_store.shape[1]
    rows = np.ascontiguousarray(rec_rows, dtype=np.int64)
    key = rows if rec_eq is None else np.ascontiguousarray(rec_eq, dtype=np.int64)
    return lib().or_misfit(int(nstep), float(dt), len(rows), rows, key,
                           np.ascontiguousarray(u_store.reshape(-1)), ns,
                           np.ascontiguousarray(dis_meas, dtype=np.float64).reshape(-1))

"""Host-side mirror of the reference's explicit solver interface, on top of the
HIP operator (``libhirensi.so``).

Names follow the reference (step_4_3D_Forward_and_Inverse/):

* :class:`ShellMat` -- the PETSc ``Mat`` calls the solver makes on AK, AC, AG:
  ``MatMultAdd``, ``MatMultTransposeAdd``, ``MatDiagonalScale``, ``MatScale``,
  ``MatZeroEntries`` (Solver_PETSc_Explicit.F90:815-824,1005-1034,1222-1229,
  1323-1352; Program_Main.F90:692-697).
* :func:`time_function` -- TIME_FUNCTION (IO.F90:541-733).
* :func:`explicit_rk4_petsc_pml_3d` -- the forward RK4 solve
  (Solver_PETSc_Explicit.F90:660-1117); :func:`explicit_rk4_petsc_pml_3d_shots`
  advances several shots over one material together.
* :func:`explicit_rk4_petsc_pml_3d_adjoint_otd` -- adjoint, control pass and
  sign flip (Solver_PETSc_Explicit.F90:1125-1522, Program_Main.F90:400-413).
* :func:`compute_misfit` -- Compute_Misfit (Solver_PETSc_Inverse.F90:591-650).
* :func:`explicit_central_difference_3` -- the central-difference variant
  EXPLICIT_Central_Difference_3 (Obsolete.F90:815-995), forward only.
* :func:`gradient_mass`, :func:`regularization`, :func:`regularize_gradient` --
  the gradient post-processing of the inversion loop: Assem_R_Mg_3D and its
  inversion (Assemble_PETSc_3D.F90:670-748, Control_Problem_Operators.F90:1-24),
  the TV / TN regularisation (Assemble_PETSc_3D.F90:755-860) and
  Program_Main.F90:416-481.

Errors raise :class:`HrError` with the library's message, the way the
reference would stop; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .native import (HrError, HR_K, HR_C, HR_G, HR_HISTORY_NONE, HR_HISTORY_STORE, HR_HISTORY_CHECKPOINT,
                     HR_HISTORY_REVOLVE, HR_REG_TV, HR_REG_TN)

__all__ = ["Context", "ShellMat", "time_function", "time_table", "explicit_rk4_petsc_pml_3d",
           "explicit_rk4_petsc_pml_3d_shots",
           "explicit_rk4_petsc_pml_3d_adjoint_otd", "compute_misfit", "explicit_central_difference_3",
           "gradient_mass", "regularization",
           "regularize_gradient", "HrError", "HR_REG_TV", "HR_REG_TN",
           "HR_K", "HR_C", "HR_G", "HR_HISTORY_NONE", "HR_HISTORY_STORE", "HR_HISTORY_CHECKPOINT",
           "HR_HISTORY_REVOLVE", "revolve_schedule"]


def _ptr(a):
    """Address of a host numpy array or of a device tensor (``data_ptr``)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    if isinstance(a, int):
        return a
    raise TypeError(f"unsupported array type {type(a)!r}")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


class Context:
    """One mesh + material on the GPU (``hr_ctx``).

    ``mesh`` is a :class:`hirensi_amd.mesh.BoxMesh`, ``model`` a
    :class:`hirensi_amd.mesh.Model`.  With ``device_ptrs=True`` every array
    passed to the methods must be device memory (e.g. torch tensors on cuda).
    """

    def __init__(self, mesh, model=None, device_ptrs=False, tz=0, ty=0, version=0):
        self.L = native.load()
        self.mesh = mesh
        self.device_ptrs = bool(device_ptrs)
        d = native.MeshDesc()
        d.nex, d.ney, d.nez = mesh.nex, mesh.ney, mesh.nez
        for a in range(5):
            d.npml[a] = mesh.npml[a]
        for a in range(3):
            d.h[a] = mesh.h[a]
        d.top_stress_bc = 1 if mesh.top_stress_bc == "reference" else 0
        d.flags = native.HR_DEVICE_PTRS if device_ptrs else 0
        self._last_nshot = 1
        h = C.c_void_p()
        rc = self.L.hr_ctx_create(C.byref(d), C.byref(h))
        self.h = h
        self._check(rc)
        if tz or ty or version:
            self._check(self.L.hr_set_tiling(self.h, int(ty), int(tz), int(version)))
        neq, nn, ns = C.c_longlong(), C.c_longlong(), C.c_longlong()
        self._check(self.L.hr_get_sizes(self.h, C.byref(neq), C.byref(nn), C.byref(ns)))
        self.neq, self.nnode, self.nstore = neq.value, nn.value, ns.value
        if self.neq != mesh.neq:
            raise HrError(f"equation count {self.neq} differs from the mesh model {mesh.neq}")
        pp = _f64(np.asarray(mesh.pml_param()).T)      # Fortran column-major (6,4)
        self._pml = pp
        self._check(self.L.hr_set_pml(self.h, pp.ctypes.data))
        if model is not None:
            self.set_material(model)

    # ------------------------------------------------------------------ util
    def _check(self, rc):
        if rc != 0:
            msg = self.L.hr_last_error(self.h).decode() if self.h else "context creation failed"
            raise HrError(f"hirensi error {rc}: {msg}")

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.hr_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _host(self, a, conv):
        return a if self.device_ptrs else conv(a)

    # --------------------------------------------------------------- set-up
    def set_material(self, model):
        if self.device_ptrs:
            raise HrError("set_material takes host arrays; use set_material_ptrs with device memory")
        lam, mu, rho = _f64(model.lam), _f64(model.mu), _f64(model.rho)
        self._check(self.L.hr_set_material(self.h, lam.ctypes.data, mu.ctypes.data, rho.ctypes.data))

    def set_material_ptrs(self, lam, mu, rho):
        self._check(self.L.hr_set_material(self.h, _ptr(lam), _ptr(mu), _ptr(rho)))

    def store_eq(self):
        out = np.zeros(self.nstore, dtype=np.int64)
        if self.device_ptrs:
            raise HrError("store_eq returns host data; create a host-pointer context")
        self._check(self.L.hr_get_store_eq(self.h, out.ctypes.data))
        return out

    def diag_m(self):
        out = np.zeros(self.neq)
        if self.device_ptrs:
            raise HrError("diag_m returns host data; create a host-pointer context")
        self._check(self.L.hr_get_diag_m(self.h, out.ctypes.data))
        return out

    def recompute_ratio(self):
        return self.L.hr_recompute_ratio(self.h)

    def tiling(self):
        """(tile y-extent, z-chunk length, stage-kernel version) in use."""
        ty, tz, ver = C.c_int(0), C.c_int(0), C.c_int(0)
        self._check(self.L.hr_get_tiling(self.h, C.addressof(ty), C.addressof(tz), C.addressof(ver)))
        return ty.value, tz.value, ver.value

    def version(self):
        return self.tiling()[2]

    def set_numbering(self, ID):
        """Equation numbering of the caller (``hr_set_numbering``): ``ID`` is
        (nnode, 9) in lattice node order, the equation number of each DOF or -1
        where constrained (the reference's ID_PETSc mapped onto the lattice)."""
        ID = np.ascontiguousarray(ID, dtype=np.int64)
        if ID.shape != (self.nnode, 9):
            raise ValueError("ID must be (nnode, 9)")
        if self.device_ptrs:
            raise HrError("set_numbering takes a host array here; pass device memory to hr_set_numbering directly")
        self._check(self.L.hr_set_numbering(self.h, ID.ctypes.data))

    def surface_force(self, nface):
        """ASSEM_FORCE_PETSC_3D: the (eq, val) entries of B for a unit z traction
        on face ``nface[e]`` (0 none, 1..6) of each regular-domain element."""
        nface = np.ascontiguousarray(nface, dtype=np.int32)
        n = C.c_longlong()
        self._check(self.L.hr_surface_force(self.h, nface.ctypes.data, None, None, C.byref(n)))
        eq = np.zeros(n.value, dtype=np.int64)
        val = np.zeros(n.value)
        self._check(self.L.hr_surface_force(self.h, nface.ctypes.data, eq.ctypes.data, val.ctypes.data, C.byref(n)))
        return eq, val

    def set_energy_interval(self, every):
        """Record the regular-domain energy every ``every`` RK4 steps (0: off)."""
        self._check(self.L.hr_set_energy_interval(self.h, int(every)))

    def energy(self):
        """Energies of the last forward, (nshot, nrow) (host contexts)."""
        n = C.c_int()
        self._check(self.L.hr_get_energy(self.h, None, C.byref(n)))
        if self.device_ptrs:
            raise HrError("energy returns host data; use hr_get_energy with a device buffer")
        out = np.zeros(max(n.value, 0) * self._last_nshot)
        self._check(self.L.hr_get_energy(self.h, out.ctypes.data if out.size else None, C.byref(n)))
        return out.reshape(self._last_nshot, n.value)

    def gradient_accumulate(self, u, p, scale, g_lambda, g_mu):
        """Update_g: g += scale * (lambda, mu) element gradients of u and p
        (neq vectors), in place on the nodal g_lambda, g_mu."""
        if not self.device_ptrs:
            u, p = _f64(u), _f64(p)
            for a in (g_lambda, g_mu):
                if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]):
                    raise ValueError("g_lambda and g_mu must be contiguous float64 arrays (updated in place)")
        self._check(self.L.hr_gradient_accumulate(self.h, _ptr(u), _ptr(p), float(scale), _ptr(g_lambda),
                                                  _ptr(g_mu)))

    def autotune(self, nshot=1, steps=6):
        """Pick the stage-kernel z-chunk length by timing a few RK4 steps per
        candidate (``hr_autotune``); returns (tz, ms per step)."""
        tz, ms = C.c_int(), C.c_double()
        self._check(self.L.hr_autotune(self.h, int(nshot), int(steps), C.byref(tz), C.byref(ms)))
        return tz.value, ms.value

    def synchronize(self):
        """Wait for the work queued on the context's stream (``hr_synchronize``)."""
        self._check(self.L.hr_synchronize(self.h))

    def set_graph(self, steps):
        """Capture ``steps`` RK4 steps (even; 0: eager launches) of the time loops
        into one replayed hipGraph (``hr_set_graph``)."""
        self._check(self.L.hr_set_graph(self.h, int(steps)))

    def graph_launches(self):
        n = C.c_longlong()
        self._check(self.L.hr_get_graph_launches(self.h, C.byref(n)))
        return n.value

    def set_history_budget(self, nbytes):
        """HBM bytes for the HR_HISTORY_REVOLVE history (0: half the free memory)."""
        self._check(self.L.hr_set_history_budget(self.h, float(nbytes)))

    def history_plan(self, nstep, history=HR_HISTORY_REVOLVE, nshot=1, segment=64):
        """(segment length, snapshot slots, HBM bytes, predicted recompute ratio)
        of the history hr_forward would keep; segment length 0 = every slice."""
        L, ns = C.c_int(), C.c_int()
        b, r = C.c_double(), C.c_double()
        self._check(self.L.hr_get_history_plan(self.h, int(nstep), int(nshot), int(history), int(segment),
                                               C.byref(L), C.byref(ns), C.byref(b), C.byref(r)))
        return L.value, ns.value, b.value, r.value


class ShellMat:
    """PETSc ``Mat`` of one operator (K, C or G) as a matrix-free shell."""

    def __init__(self, ctx: Context, which: int):
        self.ctx, self.which = ctx, int(which)

    def mult_add(self, x, y, z=None):
        """MatMultAdd: z = y + A x."""
        return self._apply(0, x, y, z)

    def mult_transpose_add(self, x, y, z=None):
        """MatMultTransposeAdd: z = y + A^T x."""
        return self._apply(1, x, y, z)

    def _apply(self, tr, x, y, z):
        c = self.ctx
        if not c.device_ptrs:
            x, y = _f64(x), _f64(y)
            if z is None:
                z = np.empty(c.neq)
        elif z is None:
            raise HrError("device shells need an output vector")
        c._check(c.L.hr_apply(c.h, self.which, tr, _ptr(x), _ptr(y), _ptr(z)))
        return z

    def scale(self, s):
        """MatScale."""
        self.ctx._check(self.ctx.L.hr_scale(self.ctx.h, self.which, float(s)))

    def diagonal_scale(self, l=None, r=None):
        """MatDiagonalScale(A, l, r); None leaves that side unchanged."""
        c = self.ctx
        if not c.device_ptrs:
            l = None if l is None else _f64(l)
            r = None if r is None else _f64(r)
        c._check(c.L.hr_diagonal_scale(c.h, self.which, _ptr(l), _ptr(r)))

    def zero_entries(self):
        """MatZeroEntries followed by re-assembly: the shell returns to A."""
        self.ctx._check(self.ctx.L.hr_reset_scaling(self.ctx.h, self.which))


def time_function(t, lcase=2, fr=15.0):
    """TIME_FUNCTION (IO.F90:541-733): 1 sine, 2 Ricker of central frequency fr,
    3 Gaussian, 4 Chopra's cosine wave, 5 its derivative, 6 Greenberg's delta,
    7 derivative of the Ricker pulse, 8 one, 9 zero, 10 smooth Heaviside.
    Any other case stops the reference (``ERROR IN TIME_FUNCTION``) and raises
    here."""
    v = native.load().hr_time_function(float(t), int(lcase), float(fr))
    if v != v:
        raise HrError(f"ERROR IN TIME_FUNCTION: no case {lcase}")
    return v


def revolve_schedule(nseg, nsnap):
    """The binomial reverse schedule of ``nseg`` segments with ``nsnap``
    snapshot slots (``hr_revolve_schedule``): (actions, advances), actions as
    (op, a, b) rows, the forward's snapshots first."""
    L = native.load()
    n = C.c_int()
    adv = C.c_longlong()
    if L.hr_revolve_schedule(int(nseg), int(nsnap), 0, None, C.byref(n), C.byref(adv)) != 0:
        raise HrError("bad revolve arguments")
    out = np.zeros(3 * max(n.value, 1), dtype=np.int32)
    if L.hr_revolve_schedule(int(nseg), int(nsnap), n.value, out.ctypes.data, C.byref(n), C.byref(adv)) != 0:
        raise HrError("revolve schedule failed")
    return out[:3 * n.value].reshape(-1, 3), adv.value


def time_table(nstep, dt, lcase=2, fr=15.0):
    """f(T), f(T + dt/2), f(T + dt) with T = ISTEP*dt, for ISTEP = 0..nstep-1
    (Solver_PETSc_Explicit.F90:853,983-995)."""
    out = np.empty((nstep, 3))
    for n in range(nstep):
        T = float(n) * dt
        out[n, 0] = time_function(T, lcase, fr)
        out[n, 1] = time_function(T + 0.5 * dt, lcase, fr)
        out[n, 2] = time_function(T + dt, lcase, fr)
    return out


def explicit_rk4_petsc_pml_3d(ctx: Context, nstep, dt, ftab, src_eq, src_val, rec_eq, history=HR_HISTORY_STORE,
                              segment=64, traces=None):
    """Forward solve from rest; returns the receiver traces, row n = u at the
    receiver equations at the start of step n (n = 0..nstep)."""
    nrec = len(rec_eq)
    if ctx.device_ptrs:
        if traces is None:
            raise HrError("device contexts need a traces buffer")
        fptr, sptr, vptr, rptr = _ptr(ftab), _ptr(src_eq), _ptr(src_val), _ptr(rec_eq)
        nsrc = len(src_eq)
    else:
        ftab = _f64(ftab)
        if ftab.shape[0] != nstep:
            raise ValueError("time table must have nstep rows")
        src_eq, src_val, rec_eq = _i64(src_eq), _f64(src_val), _i64(rec_eq)
        nsrc = len(src_eq)
        if traces is None:
            traces = np.zeros((nstep + 1, nrec))
        fptr, sptr, vptr, rptr = ftab.ctypes.data, src_eq.ctypes.data, src_val.ctypes.data, rec_eq.ctypes.data
    ctx._last_nshot = 1
    ctx._check(ctx.L.hr_forward(ctx.h, int(nstep), float(dt), fptr, nsrc, sptr if nsrc else None,
                                vptr if nsrc else None, nrec, rptr if nrec else None,
                                _ptr(traces) if nrec else None, int(history), int(segment)))
    return traces


def explicit_rk4_petsc_pml_3d_shots(ctx: Context, nstep, dt, ftab, src_shot, src_eq, src_val, rec_eq,
                                    history=HR_HISTORY_STORE, segment=64, traces=None, nshot=None):
    """Forward solves of several shots over one material, advanced together
    (``hr_forward_shots``).  ``ftab`` is (nshot, nstep, 3); source entry e
    belongs to shot ``src_shot[e]``; the receivers are shared.  Returns the
    traces (nshot, nstep + 1, nrec).  The adjoint and misfit calls that follow
    act on all shots (observations shaped like the traces)."""
    nrec = len(rec_eq)
    if ctx.device_ptrs:
        if traces is None or nshot is None:
            raise HrError("device contexts need a traces buffer and nshot")
        ptrs = (_ptr(ftab), _ptr(src_shot), _ptr(src_eq), _ptr(src_val), _ptr(rec_eq))
        nsrc = len(src_eq)
    else:
        ftab = _f64(ftab)
        if ftab.ndim != 3 or ftab.shape[1] != nstep or ftab.shape[2] != 3:
            raise ValueError("time table must be (nshot, nstep, 3)")
        nshot = ftab.shape[0]
        src_shot = np.ascontiguousarray(src_shot, dtype=np.int32)
        src_eq, src_val, rec_eq = _i64(src_eq), _f64(src_val), _i64(rec_eq)
        nsrc = len(src_eq)
        if traces is None:
            traces = np.zeros((nshot, nstep + 1, nrec))
        ptrs = (ftab.ctypes.data, src_shot.ctypes.data, src_eq.ctypes.data, src_val.ctypes.data,
                rec_eq.ctypes.data)
    ctx._last_nshot = int(nshot)
    ctx._check(ctx.L.hr_forward_shots(ctx.h, int(nshot), int(nstep), float(dt), ptrs[0], nsrc,
                                      ptrs[1] if nsrc else None, ptrs[2] if nsrc else None,
                                      ptrs[3] if nsrc else None, nrec, ptrs[4] if nrec else None,
                                      _ptr(traces) if nrec else None, int(history), int(segment)))
    return traces


def explicit_central_difference_3(ctx: Context, nstep, dt, src_eq, src_val, rec_eq, lcase=2, fr=15.0, ftab=None):
    """Central-difference forward solve (``hr_forward_cd``).  ``ftab`` (length
    nstep) holds f(n dt) for n = 1..nstep; by default TIME_FUNCTION(lcase, fr).
    Returns traces (nstep + 1, nrec): row n = u at the start of step n."""
    if ctx.device_ptrs:
        raise HrError("explicit_central_difference_3 takes host arrays; use a host-pointer context")
    if ftab is None:
        ftab = np.array([time_function(n * dt, lcase, fr) for n in range(1, nstep + 1)])
    ftab = _f64(ftab)
    if ftab.shape != (nstep,):
        raise ValueError("ftab must hold nstep values")
    src_eq, src_val, rec_eq = _i64(src_eq), _f64(src_val), _i64(rec_eq)
    nsrc, nrec = len(src_eq), len(rec_eq)
    traces = np.zeros((nstep + 1, nrec))
    ctx._check(ctx.L.hr_forward_cd(ctx.h, int(nstep), float(dt), ftab.ctypes.data, nsrc,
                                   src_eq.ctypes.data if nsrc else None, src_val.ctypes.data if nsrc else None,
                                   nrec, rec_eq.ctypes.data if nrec else None, traces.ctypes.data if nrec else None))
    return traces


def explicit_rk4_petsc_pml_3d_adjoint_otd(ctx: Context, nstep, dt, rec_eq, dis_meas, g_lambda=None, g_mu=None,
                                          with_misfit=True, misfit_out=None):
    """OtD adjoint + control pass + sign flip; returns (g_lambda, g_mu, misfit)
    with nodal gradients (zero outside the regular-domain node box)."""
    nrec = len(rec_eq)
    if ctx.device_ptrs:
        if g_lambda is None or g_mu is None:
            raise HrError("device contexts need gradient buffers")
        mptr = _ptr(misfit_out) if with_misfit else None
        ctx._check(ctx.L.hr_adjoint_gradient(ctx.h, int(nstep), float(dt), nrec, _ptr(rec_eq), _ptr(dis_meas),
                                             _ptr(g_lambda), _ptr(g_mu), mptr))
        return g_lambda, g_mu, misfit_out
    rec_eq, dis_meas = _i64(rec_eq), _f64(dis_meas)
    if g_lambda is None:
        g_lambda = np.zeros(ctx.nnode)
    if g_mu is None:
        g_mu = np.zeros(ctx.nnode)
    m = np.zeros(1)
    ctx._check(ctx.L.hr_adjoint_gradient(ctx.h, int(nstep), float(dt), nrec, rec_eq.ctypes.data,
                                         dis_meas.ctypes.data, g_lambda.ctypes.data, g_mu.ctypes.data,
                                         m.ctypes.data if with_misfit else None))
    return g_lambda, g_mu, float(m[0])


def compute_misfit(ctx: Context, nstep, dt, dis_meas):
    """Compute_Misfit of the traces of the last forward solve."""
    if ctx.device_ptrs:
        raise HrError("compute_misfit returns a host scalar; use a host-pointer context")
    dis_meas = _f64(dis_meas)
    nrec = dis_meas.shape[-1]
    m = np.zeros(1)
    ctx._check(ctx.L.hr_misfit(ctx.h, int(nstep), float(dt), nrec, dis_meas.ctypes.data, m.ctypes.data))
    return float(m[0])


_REG = {"TV": HR_REG_TV, "TN": HR_REG_TN}


def _wreg(wreg):
    """Weighted_Regularization: None ('No') or (w_top, w_bot, w_Length)."""
    if wreg is None:
        return None, None
    w = _f64(wreg)
    if w.shape != (3,):
        raise ValueError("wreg must be (w_top, w_bot, w_Length)")
    return w, w.ctypes.data


def gradient_mass(ctx: Context):
    """1/Mg of the gradient mass over the regular domain and the 0/1
    restriction (Assem_R_Mg_3D + Control_Problem_Operators.F90:12-24)."""
    if ctx.device_ptrs:
        raise HrError("gradient_mass returns host arrays; use a host-pointer context")
    img = np.zeros(ctx.nnode)
    res = np.zeros(ctx.nnode)
    ctx._check(ctx.L.hr_gradient_mass(ctx.h, img.ctypes.data, res.ctypes.data))
    return img, res


def regularization(ctx: Context, lam, mu, method="TV", wreg=None):
    """Unscaled regularisation vectors of (lam, mu) and their costs:
    returns reg_lambda, reg_mu, cost_lambda, cost_mu (Program_Main.F90:419-428)."""
    if ctx.device_ptrs:
        raise HrError("regularization returns host arrays; use a host-pointer context")
    lam, mu = _f64(lam), _f64(mu)
    w, wp = _wreg(wreg)
    rl, rm, cost = np.zeros(ctx.nnode), np.zeros(ctx.nnode), np.zeros(2)
    ctx._check(ctx.L.hr_regularization(ctx.h, _REG[method], wp, lam.ctypes.data, mu.ctypes.data, rl.ctypes.data,
                                       rm.ctypes.data, cost.ctypes.data))
    return rl, rm, float(cost[0]), float(cost[1])


def regularize_gradient(ctx: Context, it, g_lambda, g_mu, lam, mu, method="TV", factor=0.3, wreg=None):
    """Sections IV and V.2-3 of the inversion loop (Program_Main.F90:416-481):
    adds the adaptively scaled regularisation (factor = Regularization_factor_
    continuation, Adjustments.F90:138) unless ``it == 1`` and applies 1/Mg, in
    place on g_lambda, g_mu.  Returns (cost_reg_lambda, cost_reg_mu,
    reg_factor_lambda, reg_factor_mu)."""
    w, wp = _wreg(wreg)
    if ctx.device_ptrs:
        out = np.zeros(4)
        import torch
        dev = torch.zeros(4, dtype=torch.float64, device="cuda")
        ctx._check(ctx.L.hr_regularize_gradient(ctx.h, int(it), _REG[method], float(factor), wp, _ptr(lam), _ptr(mu),
                                                _ptr(g_lambda), _ptr(g_mu), dev.data_ptr(), dev.data_ptr() + 16))
        out[:] = dev.cpu().numpy()
        return tuple(float(v) for v in out)
    for a in (g_lambda, g_mu):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]):
            raise ValueError("g_lambda and g_mu must be contiguous float64 arrays (updated in place)")
    lam, mu = _f64(lam), _f64(mu)
    cost, fac = np.zeros(2), np.zeros(2)
    ctx._check(ctx.L.hr_regularize_gradient(ctx.h, int(it), _REG[method], float(factor), wp, lam.ctypes.data,
                                            mu.ctypes.data, g_lambda.ctypes.data, g_mu.ctypes.data, cost.ctypes.data,
                                            fac.ctypes.data))
    return float(cost[0]), float(cost[1]), float(fac[0]), float(fac[1])

"""Text files the reference reads and writes on either side of the solver.

* HisD -- receiver displacement histories: header of ResultsHeader
  (PTC_Results.F90:99-101), one row per written step from ResultsHis
  (PTC_Results.F90:217-228: the time, then NDim values per history node,
  constrained components as 0), read back as the measured data of the
  inversion by Input_Inv_Measured_Data (PTC_Input.F90:961-1046).
* .Lambda / .Mu -- material snapshots for resuming an inversion: one
  E31.23E3 value per node (Material_Save_Disk.F90:26-44, read by
  Input_Het_Mat_Resume, PTC_Input.F90:1185-1250), written every 10
  iterations into two alternating directories (iterations 10, 30, 50, ... into
  the first, 20, 40, ... into the second) and read back at Iter_begin = 11,
  21, 31, ... (Resume_Inversion_Read_Material.F90).

Numbers are written with the Fortran Ew.dEe edit descriptor (mantissa in
[0.1, 1), d digits, a three-digit exponent, right-justified in w columns; a
trailing X edit descriptor does not extend a record), reproducing what the
compiled Fortran writes character for character.
"""
from __future__ import annotations

import math
import os

import numpy as np

HISD_TITLE = "History of Displacements and stresses "


def fortran_e(x: float, w: int, d: int, e: int = 3) -> str:
    """The Fortran Ew.dEe representation of x (correctly rounded to d digits)."""
    x = float(x)
    if not math.isfinite(x):
        raise ValueError("non-finite value")
    neg = math.copysign(1.0, x) < 0.0
    if x == 0.0:
        digits, exp10 = "0" * d, 0
    else:
        s = f"{abs(x):.{d - 1}e}"                  # d significant digits, d.ddd...e+XX
        mant, ex = s.split("e")
        digits = mant.replace(".", "")
        exp10 = int(ex) + 1                         # 0.ddd... form
    body = f"0.{digits}E{'-' if exp10 < 0 else '+'}{abs(exp10):0{e}d}"
    if neg:
        body = "-" + body
    if len(body) > w:
        return "*" * w
    return body.rjust(w)


def hisd_header(app_nodes, petsc_nodes) -> list[str]:
    """The two header records: title and ('Time',5x,NNDH('App#',I19,1x,'PETSc#',I19,1x))."""
    parts = ["Time", " " * 5]
    for a, p in zip(app_nodes, petsc_nodes):
        parts.append(f"App#{int(a):19d} PETSc#{int(p):19d} ")
    return [HISD_TITLE, "".join(parts).rstrip(" ")]


def hisd_row(time: float, values, ndim: int = 3) -> str:
    """(E14.6E3,2x, NNDH(NDim(E18.10E3,2x),3x)) of one step."""
    v = np.asarray(values, dtype=np.float64).reshape(-1, ndim)
    out = [fortran_e(time, 14, 6), "  "]
    for n in range(v.shape[0]):
        for c in range(ndim):
            out.append(fortran_e(v[n, c], 18, 10))
            out.append("  ")
        out.append("   ")
    return "".join(out).rstrip(" ")


def write_hisd(path, dt, traces, rec_eq=None, app_nodes=None, petsc_nodes=None, ndim=3):
    """Write a HisD file from receiver traces [(nstep + 1) x (NNDH * ndim)]
    (row n: displacement at the start of step n, time n * dt).  Columns whose
    equation number in rec_eq is -1 (constrained) are written as 0."""
    tr = np.asarray(traces, dtype=np.float64)
    nrow, ncol = tr.shape
    if ncol % ndim:
        raise ValueError("trace columns must be NNDH * ndim")
    nndh = ncol // ndim
    if rec_eq is not None:
        tr = np.where(np.asarray(rec_eq)[None, :] < 0, 0.0, tr)
    app = np.arange(1, nndh + 1) if app_nodes is None else app_nodes
    pet = np.arange(nndh) if petsc_nodes is None else petsc_nodes
    with open(path, "w") as f:
        for line in hisd_header(app, pet):
            f.write(line + "\n")
        for n in range(nrow):
            f.write(hisd_row(float(n) * dt, tr[n], ndim) + "\n")


def _fields(line, widths):
    pos, out = 0, []
    for w, skip in widths:
        out.append(line[pos:pos + w])
        pos += w + skip
    return out


def read_hisd(path, nstep, nndh, ndim=3):
    """Input_Inv_Measured_Data: skip the two header records, then read rows
    0..nstep with the row format.  Returns (time[nstep + 1],
    dis_meas[(nstep + 1) x (nndh * ndim)])."""
    widths = [(14, 2)]
    for _ in range(nndh):
        widths += [(18, 2)] * (ndim - 1) + [(18, 5)]
    t = np.zeros(nstep + 1)
    dis = np.zeros((nstep + 1, nndh * ndim))
    with open(path) as f:
        f.readline()
        f.readline()
        for n in range(nstep + 1):
            line = f.readline()
            if not line:
                raise EOFError(f"{path}: {n} rows for {nstep + 1} steps")
            vals = [_fortran_real(s) for s in _fields(line.rstrip("\n"), widths)]
            t[n] = vals[0]
            dis[n] = vals[1:]
    return t, dis


def _fortran_real(s: str) -> float:
    s = s.strip()
    return float(s) if s else 0.0                   # BLANK = 'NULL': an empty field reads as zero


# ----------------------------------------------------------------- material
def write_material(path, values):
    """One (E31.23E3) record per node."""
    with open(path, "w") as f:
        for v in np.asarray(values, dtype=np.float64):
            f.write(fortran_e(v, 31, 23) + "\n")


def read_material(path, nj):
    out = np.zeros(nj)
    with open(path) as f:
        for k in range(nj):
            line = f.readline()
            if not line:
                raise EOFError(f"{path}: {k} records for {nj} nodes")
            out[k] = _fortran_real(line[:31])
    return out


def material_file_names(model_name, size=1, rank=0):
    """TRIM(ModelName)//'_'//Size//'_'//Rank//'.Lambda' and '.Mu'."""
    stem = f"{model_name.strip()}_{int(size)}_{int(rank)}"
    return stem + ".Lambda", stem + ".Mu"


def save_material(model_name, it, lam, mu, dir1, dir2, size=1, rank=0):
    """Material_Save_Disk: at iterations with it % 10 == 0, into dir1 when
    it % 20 != 0, else into dir2.  Returns the directory written, or None."""
    if it % 10 != 0:
        return None
    d = dir2 if it % 20 == 0 else dir1
    fl, fm = material_file_names(model_name, size, rank)
    os.makedirs(d, exist_ok=True)
    write_material(os.path.join(d, fl), lam)
    write_material(os.path.join(d, fm), mu)
    return d


def resume_material(model_name, it_begin, nj, dir1, dir2, size=1, rank=0):
    """Resume_Inversion_Read_Material: only at it_begin = 11, 21, 31, ...; the
    snapshot of iteration it_begin - 1 (dir1 when (it_begin - 1) % 20 != 0)."""
    if (it_begin - 1) % 10 != 0:
        raise ValueError("Inversion can only be resumed at iterations 11, 21, 31, .... Change Iter_begin.")
    d = dir2 if (it_begin - 1) % 20 == 0 else dir1
    fl, fm = material_file_names(model_name, size, rank)
    return read_material(os.path.join(d, fl), nj), read_material(os.path.join(d, fm), nj)

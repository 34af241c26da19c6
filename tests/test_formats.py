"""HisD and material-snapshot text formats (hirensi_amd/formats.py).

* Fortran Ew.dEe output against values written by compiled Fortran
  (tests/golden/fortran_e.txt from tests/golden/fortran_e.f90, ROCm flang):
  character for character, including the round-half-even tie 2**-15 at ten
  digits, carries (9.9999999999995 -> 0.1000000000E+002) and 3-digit exponents;
* whole HisD records (header and rows with the reference's edit descriptors,
  repeat counts written out) and material records against a live flang build;
* round trips: HisD rows keep 10 significant digits, material snapshots are
  exact (23 digits); the snapshot alternation and resume rule of
  Material_Save_Disk.F90 / Resume_Inversion_Read_Material.F90.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from hirensi_amd import formats as F

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def _golden():
    rows = []
    for line in open(os.path.join(HERE, "golden", "fortran_e.txt")):
        hexbits, rest = line.rstrip("\n").split(" ", 1)
        v = struct.unpack(">d", bytes.fromhex(hexbits))[0]
        rows.append((v, rest.split("|")))
    return rows


def test_fortran_e_matches_compiled_fortran():
    rows = _golden()
    assert len(rows) == 40
    for v, (e14, e18, e31) in rows:
        assert F.fortran_e(v, 14, 6) == e14, v
        assert F.fortran_e(v, 18, 10) == e18, v
        assert F.fortran_e(v, 31, 23) == e31, v


@pytest.mark.skipif(not os.path.exists(FLANG), reason="ROCm flang absent")
def test_hisd_and_material_records_match_compiled_fortran(tmp_path):
    rng = np.random.default_rng(3)
    vals = rng.standard_normal(6) * 10.0 ** rng.integers(-12, 3, 6)
    t = 0.0125
    lam = 137.25 + rng.standard_normal()
    lit = lambda x: f"{float(x):.17e}".replace("e", "d")
    prog = f"""
program t
  implicit none
  real(8) :: v(6), tm, a
  integer :: i
  v = [{', '.join(lit(x) for x in vals)}]
  tm = {lit(t)}
  a = {lit(lam)}
  write(*, "('History of Displacements and stresses ')")
  write(*, "('Time',5x,2('App#',I19,1x,'PETSc#',I19,1x))") 101, 7, 2048, 31
  write(*, "(E14.6E3,2x, 2(3(E18.10E3,2x),3x) )") tm, (v(i), i = 1, 6)
  write(*, "(3(E31.23E3,2X) )") a
end program t
"""
    (tmp_path / "t.f90").write_text(prog)
    subprocess.run([FLANG, "t.f90", "-o", "t"], cwd=tmp_path, check=True)
    out = subprocess.run(["./t"], cwd=tmp_path, check=True, capture_output=True, text=True,
                         stdin=subprocess.DEVNULL).stdout.split("\n")
    head = F.hisd_header([101, 2048], [7, 31])
    assert out[0] == head[0] and out[1] == head[1]
    assert out[2] == F.hisd_row(t, vals)
    assert out[3] == F.fortran_e(lam, 31, 23)


def test_hisd_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    nstep, nndh = 25, 4
    tr = rng.standard_normal((nstep + 1, 3 * nndh)) * 1e-6
    rec_eq = np.arange(3 * nndh)
    rec_eq[5] = -1                                   # a constrained component is written as 0
    p = tmp_path / "model.HisD"
    F.write_hisd(p, 1e-4, tr, rec_eq=rec_eq)
    t, dis = F.read_hisd(p, nstep, nndh)
    assert np.allclose(t, np.arange(nstep + 1) * 1e-4, rtol=1e-6, atol=0)
    ref = tr.copy()
    ref[:, 5] = 0.0
    assert np.allclose(dis, ref, rtol=5e-10, atol=0)     # 10 significant digits
    lines = open(p).read().split("\n")
    assert lines[0] == F.HISD_TITLE and lines[1].startswith("Time     App#")
    with pytest.raises(EOFError):
        F.read_hisd(p, nstep + 1, nndh)


def test_material_snapshots_round_trip_and_alternate(tmp_path):
    rng = np.random.default_rng(1)
    lam, mu = rng.uniform(50, 166, 57), rng.uniform(50, 166, 57)
    d1, d2 = str(tmp_path / "r1"), str(tmp_path / "r2")
    assert F.save_material("box", 7, lam, mu, d1, d2) is None
    assert F.save_material("box", 10, lam, mu, d1, d2) == d1
    assert F.save_material("box", 20, 2 * lam, 2 * mu, d1, d2) == d2
    assert sorted(os.listdir(d1)) == ["box_1_0.Lambda", "box_1_0.Mu"]
    l1, m1 = F.resume_material("box", 11, 57, d1, d2)
    assert np.array_equal(l1, lam) and np.array_equal(m1, mu)     # 23 digits: exact
    l2, _ = F.resume_material("box", 21, 57, d1, d2)
    assert np.array_equal(l2, 2 * lam)
    with pytest.raises(ValueError):
        F.resume_material("box", 15, 57, d1, d2)
    first = open(os.path.join(d1, "box_1_0.Lambda")).readline().rstrip("\n")
    assert len(first) == 31 and first == F.fortran_e(lam[0], 31, 23)

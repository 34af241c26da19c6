"""Host check of the nodal PML pull coefficients of the split stage path
(hirensi_amd/csrc/hr_split.h): the side-scaled per-axis stencils and the PML
weight sums against brute-force sums over the PML elements holding each node,
on boxes with regular-domain interfaces on every side (Element_Matrices_3D.F90:618-882
gives the element terms: W(point) DM(p, l) on the u rows, W(node) DM(l, p) on
the s rows); the list-free PML node enumeration and the nodal pass's x-neighbour
rule against a lattice walk."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_split_pull_coefficients_match_element_sums():
    src = os.path.join(HERE, "native", "split_coef_check.cpp")
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "scc")
        subprocess.run(["g++", "-O1", "-std=c++17", src, "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout

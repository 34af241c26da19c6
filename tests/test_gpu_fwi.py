"""Inversion iterations on the HIP path against the same iterations on the CPU
oracle (runs on the GPU).

The product driver (hirensi_amd/fwi.py) with :class:`HipProblem` (HIP forward,
adjoint, regularisation) and the oracle's literal restatement of
Program_Main.F90:379-921 (oracle/fwi_oracle.py) with the oracle's physics take
the same decisions (Armijo cuts) and reach the same models: the per-iteration
misfits and the final lambda, mu agree to 1e-9 relative (FP64; the two sides
differ in summation order only, amplified through the iterations).
"""
import numpy as np
import pytest

from hirensi_amd import fwi
from hirensi_amd import solver as S
from hirensi_amd.mesh import Model
from oracle import fwi_oracle as FO
from tests.helpers import OracleProblem, random_model, rel_l2, small_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["TV", "TN"])
def test_inversion_iterations_match_oracle(method):
    mesh, _, rec_eq, src_eq, src_val = small_case(seed=1)
    nstep, dt, fr = 240, 1.0e-4, 60.0
    true = random_model(mesh, seed=9)
    start = random_model(mesh, seed=2, spread=0.05)
    print(f"fwi case method={method} neq={mesh.neq} nstep={nstep}", flush=True)
    ftab = S.time_table(nstep, dt, 2, fr)[None]
    survey = fwi.Survey(nstep, dt, ftab, np.zeros(len(src_eq), dtype=np.int32), src_eq, src_val, rec_eq)
    ctx = S.Context(mesh, true)
    hp = fwi.HipProblem(ctx, survey, true.rho, fwi.InversionParams(regularization=method))
    survey.dis_meas = hp.observe(true.lam, true.mu)
    op = OracleProblem(mesh, true.rho, src_eq, src_val, rec_eq, nstep, dt, fr, survey.dis_meas[0], method)
    res = (np.ones(mesh.nnode) * mesh.rd_node.reshape(-1)).astype(float)
    lam0, mu0 = start.lam, start.mu
    n_iter = 3
    lp, mp, hist_p = fwi.invert(hp, lam0, mu0, res, n_iter, fwi.InversionParams(regularization=method))
    lo, mo, hist_o = FO.invert(op, lam0, mu0, res, n_iter, FO.Params())
    for a, b in zip(hist_p, hist_o):
        print(f"iter {a['iter']}: misfit {a['misfit']:.10e} (oracle {b['misfit']:.10e}) cuts {a['armijo_counter']}"
              f" cost_reg {a['cost_reg'][0]:.4e} {a['cost_reg'][1]:.4e}", flush=True)
        assert a["armijo_counter"] == b["armijo_counter"]
        assert abs(a["misfit"] - b["misfit"]) <= 1e-9 * abs(b["misfit"])
        for x, y in zip(a["cost_reg"], b["cost_reg"]):
            assert abs(x - y) <= 1e-9 * max(abs(y), 1e-300)
    assert hist_p[-1]["misfit"] < hist_p[0]["misfit_pre"]
    assert rel_l2(lp, lo) <= 1e-9 and rel_l2(mp, mo) <= 1e-9

"""Stage-kernel timing of the C ABI (hr_set_profiling / hr_get_kernel_stats):
every RK4 stage launch is counted, one in seven is timed with events around
its kernels on each stream, and the reported time is the mean timed span
times the launch count."""
import numpy as np
import pytest

from hirensi_amd import solver as S
from tests.helpers import small_case

pytestmark = pytest.mark.gpu


def _stats(ctx):
    n = np.zeros(5, dtype=np.int64)
    ms = np.zeros(5)
    assert ctx.L.hr_get_kernel_stats(ctx.h, n.ctypes.data, ms.ctypes.data) == 0
    return n, ms


def test_stage_timing_counts_every_launch():
    mesh, model, rec_eq, src_eq, src_val = small_case(seed=3)
    print("profiling: forward of 30 steps", flush=True)
    ctx = S.Context(mesh, model)
    nstep, dt = 30, 1e-4
    ftab = S.time_table(nstep, dt, 2, 60.0)
    assert ctx.L.hr_set_profiling(ctx.h, 2) == 0
    assert ctx.L.hr_set_profiling(ctx.h, 1) == 0
    S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, src_val, rec_eq)
    n, ms = _stats(ctx)
    assert list(n[1:]) == [nstep] * 4
    # 120 launches, one in seven timed: every stage kind has timed launches
    assert np.all(np.isfinite(ms)) and np.all(ms[1:] > 0.0)
    assert ctx.L.hr_set_profiling(ctx.h, 0) == 0
    S.explicit_rk4_petsc_pml_3d(ctx, 5, dt, ftab[:5], src_eq, src_val, rec_eq)
    n2, _ = _stats(ctx)
    assert np.array_equal(n2, n)                           # off: nothing counted
    assert ctx.L.hr_set_profiling(ctx.h, 2) == 0
    n3, ms3 = _stats(ctx)
    assert not n3.any() and not ms3.any()

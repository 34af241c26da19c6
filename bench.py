"""Benchmark: one forward + OtD adjoint + lambda/mu gradient solve per step.

Workload: 64 x 64 x 32 elements of 1 m with 6-element PML (6.26 M DOFs incl.
PML stresses), three layers Vs 150/250/400 m/s, Vp = 2 Vs, rho 1900/2000/2100,
15 Hz Ricker at the centre surface node, 4000 RK4 steps of 1e-4 s, 400
surface receivers.  Observed data come from the model with a -30 % Gaussian
Vs anomaly; the gradient is evaluated at the model without it.  Synthetic
data, FP64 throughout, all inputs resident in HBM before the timed region.

Prints one JSON line.  value = DOF-steps/s = neq * (forward + adjoint steps)
per second of the whole forward + adjoint + gradient solve.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--nstep", type=int, default=4000, help="RK4 steps per solve (workload: 4000)")
    p.add_argument("--history", choices=["store", "checkpoint"], default="store")
    p.add_argument("--segment", type=int, default=64)
    p.add_argument("--tz", type=int, default=0)
    p.add_argument("--ty", type=int, default=0)
    p.add_argument("--version", type=int, default=0)
    p.add_argument("--serial-classes", action="store_true", help="v3: run the tile classes one after another")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--ckpt-budget", type=float, default=8.0e9,
                   help="HBM budget (bytes) of the checkpointed leg (binomial revolve history); 0 skips the leg")
    p.add_argument("--cpu-steps", type=int, default=40, help="oracle RK4 steps timed (each direction)")
    p.add_argument("--graph-leg", type=int, default=8,
                   help="also time the gradient with time loops replayed from hipGraphs of this many steps "
                        "(0 skips the leg)")
    p.add_argument("--copy-calib", action="store_true", help="also run a plain 1 GiB device copy (PMC scale)")
    p.add_argument("--shots", type=int, default=0,
                   help="instead of the default line: batched multi-shot gradient on the 48x48x24 survey "
                        "(use with --nstep; shots/s against single-shot solves)")
    p.add_argument("--seq-shots", type=int, default=2, help="single-shot gradients timed for the comparison")
    p.add_argument("--fwi", choices=["config4", "small"], default=None,
                   help="instead of the default line: L-BFGS FWI on config 4 (96x96x48, 9 shots) or its 48x48x24 "
                        "reduction, plain sequential and/or batched + checkpointed (use with --nstep, --fwi-iters)")
    p.add_argument("--fwi-iters", type=int, default=20)
    p.add_argument("--fwi-variant", choices=["both", "batched", "plain"], default="both")
    p.add_argument("--fwi-budget-s", type=float, default=900.0, help="stop a variant after this many seconds")
    p.add_argument("--precond", action="store_true",
                   help="instead of the default line: train the learned gradient preconditioner on solver-made "
                        "pairs (32x32x16) and run L-BFGS with and without it on a 64x64x32 survey")
    p.add_argument("--pc-pairs", type=int, default=200, help="training pairs")
    p.add_argument("--pc-nstep", type=int, default=900, help="RK4 steps of the training solves")
    p.add_argument("--pc-epochs", type=int, default=60)
    p.add_argument("--pc-fwi-nstep", type=int, default=1800, help="RK4 steps of the inversion solves")
    p.add_argument("--pc-fwi-iters", type=int, default=8)
    p.add_argument("--pc-seed", type=int, default=5, help="seed of the inversion's true model")
    p.add_argument("--pc-weights", default="", help="load trained weights (.npz) instead of training")
    p.add_argument("--pc-save", default="", help="save the trained weights (.npz) here")
    p.add_argument("--pc-variants", default="plain,learned",
                   help="inversions to compare: plain (batched shots, stored history), plain-seq (one shot per "
                        "solve, revolve history), learned (batched, stored, learned H0), learned-revolve "
                        "(batched, revolve history, learned H0); the last learned one is compared with the "
                        "first plain one")
    p.add_argument("--fd-check", action="store_true",
                   help="instead of timing: check the lambda/mu gradient against central finite differences "
                        "along 5 seeded smooth directions on this mesh (use with --nstep)")
    return p.parse_args()


def step_bytes(neq, nnode):
    """Algorithmic bytes of one RK4 step by the fixed accounting convention of the
    hot path (reads of the stage inputs, M^-1 and lambda, mu; writes of g_s; a
    separate update pass reading 7 and writing 3 vectors):
    B_step = 8 (36 N + 8 N_node).  A stage launch processes a quarter step."""
    return 8.0 * (36.0 * neq + 8.0 * nnode)


def stage_bytes(neq, nnode):
    """Algorithmic bytes of one launch of the stage kernel for RK4 stage s:
    stage s reads its 2+s input vectors (x1, x2, x3, g1..g_{s-1}), M^-1 and
    lambda, mu; stages 1-3 write g_s; stage 4 writes the updated x1, x2, x3."""
    b = {}
    for s in (1, 2, 3, 4):
        reads = (2 + s) * neq + neq + 2 * nnode
        writes = neq if s < 4 else 3 * neq
        b[s] = 8.0 * (reads + writes)
    return b


def split_stage_bytes(mesh, transpose=False):
    """Algorithmic bytes of one RK4 stage s of the split stage path (the default):
    the nodal pass over the n_p PML nodes reads the 2+s raw inputs of all nine
    components and four node-table doubles, writes the 12 flux-record doubles of
    the y- and z-lines (the x-lines are summed in the pass, from LDS) and the
    nine node-local outputs (stage 4: also x1, x2); the nodal pull reads the
    flux records (each once), the node-local outputs, M^-1 of the u rows, the interface sums
    of the n_i interface nodes (and lambda, mu when transposed) and writes nine outputs; the regular-domain
    kernel reads, per regular-domain-only node, the displacement of the raw inputs
    that enter its stage input (1, 2, 3, 3 vectors; stage 4: all six for the
    update), lambda, mu and M^-1, and writes three outputs (stage 4: nine), and the
    interface sums of the interface nodes."""
    lo, hi = mesh.rd_lo, mesh.rd_hi
    nxy = mesh.NX * mesh.NY
    ax, ay = 2 * (hi[0] - lo[0]) - 1, 2 * (hi[1] - lo[1]) - 1
    n_p = (2 * lo[2] + 1) * nxy + (mesh.NZ - 2 * lo[2] - 1) * (nxy - ax * ay)
    n_r = mesh.nnode - n_p
    # interface nodes: the regular domain's node box less its interior
    bx, by, bz = ax + 2, ay + 2, mesh.NZ - 2 * lo[2]
    n_i = bx * by * bz - ax * ay * (bz - 1)
    b = {}
    for st in (1, 2, 3, 4):
        k1 = n_p * (9 * (2 + st) + 4 + 12 + 9 + (18 if st == 4 else 0))
        k3 = n_p * (12 + 9 + 1 + (2 if transpose else 0) + 9) + 3 * n_i
        nv = 6 if st == 4 else {1: 1, 2: 2, 3: 3}[st]
        k2 = n_r * (3 * nv + 2 + 2 + (9 if st == 4 else 3)) + 3 * n_i
        b[st] = 8.0 * (k1 + k2 + k3)
    return b


def traffic_file(split):
    return os.path.join("profiles", "split_stage_traffic_64x64x32.json" if split else
                        "stage_kernel_traffic_64x64x32.json")


def measured_traffic(split):
    """HBM bytes per stage launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload and stage path (copy-kernel
    calibrated), or None."""
    path = os.path.join(ROOT, traffic_file(split))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def _log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_config1(threads):
    """The oracle on config 1 in full (16x16x8, PML 4, 2000 RK4 steps, forward
    only, the reference's CPU-runnable case), assembly excluded."""
    import numpy as np
    from oracle import oracle as O
    from hirensi_amd.mesh import config1
    c = config1()
    mesh = c["mesh"]
    O.lib().or_set_threads(threads)
    s = O.System(mesh, c["model"])
    B = np.zeros(mesh.neq)
    B[mesh.dof(c["src_node"], 2)] = 1.0
    t0 = time.perf_counter()
    s.forward(B, c["nstep"], c["dt"], lcase=2, fr=c["fr"])
    wall = time.perf_counter() - t0
    s.close()
    return {"value": mesh.neq * c["nstep"] / wall, "unit": "DOF-steps/s", "wall_s": wall, "cores": threads,
            "sample": f"config 1 in full: 16x16x8 el, PML 4, {c['nstep']} RK4 forward steps, neq {mesh.neq}"}


def cpu_baseline(cfg, nsteps):
    """The oracle (CSR + PETSc-sequence RK4, OpenMP) on the same mesh and model,
    timed over `nsteps` forward and `nsteps` adjoint steps; assembly excluded."""
    import numpy as np
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0))
    O.lib().or_set_threads(threads)
    mesh, model = cfg["mesh"], cfg["start"]
    s = O.System(mesh, model)
    B = np.zeros(mesh.neq)
    B[cfg["src_eq"]] = 1.0
    t0 = time.perf_counter()
    U = s.forward(B, nsteps, cfg["dt"], lcase=2, fr=cfg["fr"])
    t1 = time.perf_counter()
    dmeas = np.zeros((nsteps + 1, len(cfg["rec_eq"])))
    s.adjoint(U, nsteps, cfg["dt"], cfg["rec_eq"], dmeas)      # builds the transposes on first use
    t2 = time.perf_counter()
    s.adjoint(U, nsteps, cfg["dt"], cfg["rec_eq"], dmeas)
    t3 = time.perf_counter()
    s.close()
    wall = (t1 - t0) + (t3 - t2)
    return {"value": mesh.neq * 2 * nsteps / wall, "unit": "DOF-steps/s", "cores": threads, "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": f"oracle on the same 64x64x32 mesh and model, {nsteps} forward + {nsteps} adjoint "
                      f"RK4 steps, assembly excluded ({wall:.1f} s)",
            "config1_full": cpu_config1(threads)}


def fd_check(a):
    """Gradient_check.F90 pattern on the benchmark mesh: directional derivatives of
    the misfit along 5 smooth random directions (default_rng(5), unit infinity norm,
    regular-domain-only nodes) by central differences at h = 1e-2, 1e-3, 1e-4 |m|,
    against the OtD gradient.  Prints one JSON line."""
    import numpy as np
    from hirensi_amd import solver as S
    from hirensi_amd.mesh import config2, receiver_dofs
    from tests.test_gpu_fd_gradient import smooth_directions
    _log("fd-check: building the 64x64x32 mesh and models")
    true_cfg, start_cfg = config2(with_anomaly=True), config2(with_anomaly=False)
    mesh = true_cfg["mesh"]
    nstep, dt, fr = a.nstep, true_cfg["dt"], true_cfg["fr"]
    rec_eq = receiver_dofs(mesh, true_cfg["rec_nodes"])
    src_eq = np.array([mesh.dof(true_cfg["src_node"], 2)], dtype=np.int64)
    one = np.array([1.0])
    ftab = S.time_table(nstep, dt, 2, fr)
    ctx = S.Context(mesh, true_cfg["model"])
    _log("fd-check: context ready; observed data")
    dobs = S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, one, rec_eq,
                                       history=S.HR_HISTORY_NONE).copy()
    start = start_cfg["model"]
    _log("fd-check: observed data done; gradient")

    def misfit(model):
        ctx.set_material(model)
        S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, one, rec_eq, history=S.HR_HISTORY_NONE)
        return S.compute_misfit(ctx, nstep, dt, dobs)

    ctx.set_material(start)
    S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, ftab, src_eq, one, rec_eq, history=S.HR_HISTORY_CHECKPOINT,
                                segment=64)
    gl, gm, j0 = S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, rec_eq, dobs)
    _log("fd-check: gradient done; central differences")
    rows = []
    for t, d in enumerate(smooth_directions(mesh, 5)):
        for field, g in (("lambda", gl), ("mu", gm)):
            base = start.lam if field == "lambda" else start.mu
            adj = float(np.dot(g, d))
            errs = []
            for hr in (1e-2, 1e-3, 1e-4):
                h = hr * np.abs(base).max()
                mp = type(start)(start.lam.copy(), start.mu.copy(), start.rho)
                mm = type(start)(start.lam.copy(), start.mu.copy(), start.rho)
                if field == "lambda":
                    mp.lam, mm.lam = base + h * d, base - h * d
                else:
                    mp.mu, mm.mu = base + h * d, base - h * d
                fd = (misfit(mp) - misfit(mm)) / (2.0 * h)
                errs.append(abs(adj - fd) / abs(fd))
                print(f"fd-check direction {t} {field} h/|m| {hr:g}: rel. err. {errs[-1]:.3e}", file=sys.stderr,
                      flush=True)
            rows.append({"direction": t, "field": field, "adjoint": adj, "rel_err": errs})
    print(json.dumps({"metric": "gradient vs central finite differences (64x64x32, PML 6)",
                      "nstep": nstep, "misfit": j0, "recompute_ratio": ctx.recompute_ratio(),
                      "h_over_norm": [1e-2, 1e-3, 1e-4], "checks": rows,
                      "worst_best_h_rel_err": max(min(r["rel_err"]) for r in rows)}), flush=True)
    ctx.close()


def shots_bench(a):
    """Batched multi-shot forward + adjoint + gradient of the 48x48x24 survey:
    all shots advance together (one stage-kernel launch per RK4 stage covers
    every shot), checkpointed history.  The same shots solved one at a time
    through the single-shot path give the comparison (a bounded sample of
    --seq-shots solves, per-shot time extrapolated).  Prints one JSON line."""
    import numpy as np
    import torch
    from hirensi_amd import solver as S
    from hirensi_amd.mesh import config3, receiver_dofs
    cfg = config3(nshot=a.shots)
    mesh, dt = cfg["mesh"], cfg["dt"]
    nstep = a.nstep if a.nstep != 4000 else cfg["nstep"]
    nshot = a.shots
    dev = torch.device("cuda:0")
    rec_eq = receiver_dofs(mesh, cfg["rec_nodes"])
    nrec = len(rec_eq)
    src_eq = np.array([mesh.dof(nd, 2) for nd in cfg["src_nodes"]], dtype=np.int64)
    src_shot = np.arange(nshot, dtype=np.int32)
    ftab = np.stack([S.time_table(nstep, dt, 2, fr) for fr in cfg["frs"]])
    seg = 128

    def T(x, dtype=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=dev)

    ctx = S.Context(mesh, device_ptrs=True)
    d_ftab, d_shot, d_src, d_val, d_rec = (T(ftab), T(src_shot, torch.int32), T(src_eq, torch.int64),
                                           T(np.ones(nshot)), T(rec_eq, torch.int64))
    d_obs = torch.zeros((nshot, nstep + 1, nrec), dtype=torch.float64, device=dev)
    d_tr = torch.zeros_like(d_obs)
    d_gl = torch.zeros(mesh.nnode, dtype=torch.float64, device=dev)
    d_gm = torch.zeros_like(d_gl)
    d_mis = torch.zeros(1, dtype=torch.float64, device=dev)
    tm, sm = cfg["true"], cfg["start"]
    ctx.set_material_ptrs(T(tm.lam), T(tm.mu), T(tm.rho))
    S.explicit_rk4_petsc_pml_3d_shots(ctx, nstep, dt, d_ftab, d_shot, d_src, d_val, d_rec,
                                      history=S.HR_HISTORY_NONE, traces=d_obs, nshot=nshot)
    ctx.set_material_ptrs(T(sm.lam), T(sm.mu), T(sm.rho))
    torch.cuda.synchronize()
    print(f"shots: observed data of {nshot} shots done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    S.explicit_rk4_petsc_pml_3d_shots(ctx, nstep, dt, d_ftab, d_shot, d_src, d_val, d_rec,
                                      history=S.HR_HISTORY_CHECKPOINT, segment=seg, traces=d_tr, nshot=nshot)
    torch.cuda.synchronize()
    print(f"shots: batched forward done ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, d_rec, d_obs, d_gl, d_gm, misfit_out=d_mis)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    print(f"shots: batched gradient done ({t_batch:.1f} s)", file=sys.stderr, flush=True)
    ratio = ctx.recompute_ratio()
    misfit = float(d_mis.item())
    # single shots through the same library, one at a time
    nseq = max(1, min(a.seq_shots, nshot))
    d_tr1 = torch.zeros((nstep + 1, nrec), dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    for q in range(nseq):
        S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, d_ftab[q], d_src[q:q + 1], d_val[q:q + 1], d_rec,
                                    history=S.HR_HISTORY_CHECKPOINT, segment=seg, traces=d_tr1)
        S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, d_rec, d_obs[q], d_gl, d_gm, misfit_out=d_mis)
        torch.cuda.synchronize()
        print(f"shots: single shot {q} done", file=sys.stderr, flush=True)
    t_one = (time.perf_counter() - t0) / nseq
    ctx.close()
    dof_steps = ctx.neq * 2.0 * nstep * nshot
    print(json.dumps({
        "metric": f"batched multi-shot gradient shots/s (48x48x24, {nshot} shots, {nstep} steps, 1 GPU)",
        "value": nshot / t_batch, "unit": "shots/s", "higher_is_better": True, "n_gpus": 1,
        "dtype": "f64", "data": "synthetic (layered model, anomaly in the observed-data model)",
        "config": {"workload": f"fwd+adj+grad of {nshot} shots, 48x48x24 el, PML 6, {nstep} RK4 steps, "
                               f"{nrec // 3} shared receivers, checkpoint segment {seg}",
                   "neq": ctx.neq, "recompute_ratio": ratio, "misfit_sum": misfit},
        "batched_wall_s": t_batch, "batched_dof_steps_per_s": dof_steps / t_batch,
        "single_shot_wall_s": t_one, "single_shot_sample": nseq,
        "sequential_shots_per_s": 1.0 / t_one, "speedup_vs_sequential": t_one * nshot / t_batch}), flush=True)


def fwi_bench(a):
    """L-BFGS FWI (m = 5) of the config-4 survey: the plain form (one shot per
    solve, stored history) against batched shots with the binomially
    checkpointed history, same algorithm and data; misfit and mu model error per
    iteration, wall time, and the batched run's time to reach the plain run's
    final misfit.  Prints one JSON line."""
    import numpy as np
    from hirensi_amd import solver as S
    from hirensi_amd import fwi as F
    from hirensi_amd.mesh import config4
    cfg = config4() if a.fwi == "config4" else config4(nx=48, ny=48, nz=24, npml=6, nrec_side=15)
    mesh = cfg["mesh"]
    nstep = a.nstep if a.nstep != 4000 else cfg["nstep"]
    sv = F.survey_of(cfg, nstep)
    params = F.InversionParams(m_lbfgs=5, bound_lambda_lower=40.0, bound_lambda_upper=1200.0, bound_mu_lower=20.0,
                               bound_mu_upper=600.0, fac_max_update_lambda=10.0, fac_max_update_mu=10.0)
    ctx = S.Context(mesh)
    ctx.set_material(cfg["true"])
    _log(f"fwi {a.fwi}: neq {mesh.neq}, {len(cfg['src_nodes'])} shots, {nstep} steps; observing the true model")
    obs = F.HipProblem(ctx, sv, cfg["true"].rho, log=_log).observe(cfg["true"].lam, cfg["true"].mu)
    sv.dis_meas = obs
    _, restriction = S.gradient_mass(ctx)
    start, true = cfg["start"], cfg["true"]
    runs = {}
    variants = ["plain", "batched"] if a.fwi_variant == "both" else [a.fwi_variant]
    for v in variants:
        if v == "plain":
            prob = F.SequentialHipProblem(ctx, sv, start.rho, params, log=_log)
        else:
            ctx.set_history_budget(0.0)
            prob = F.HipProblem(ctx, sv, start.rho, params, history=S.HR_HISTORY_REVOLVE, log=_log)
        t0 = time.perf_counter()
        inv = F.Inversion(prob, start.lam, start.mu, restriction, params)
        recs = [{"iter": 0, "misfit": inv.misfit, "mu_error": F.model_error(mesh, inv.mu, true.mu),
                 "wall_s": time.perf_counter() - t0}]
        _log(f"fwi {v}: iteration 0 misfit {inv.misfit:.6e} mu error {recs[0]['mu_error']:.4f}")
        stop = ""
        for k in range(a.fwi_iters):
            try:
                r = inv.iterate()
            except F.ArmijoFailure as e:
                stop = str(e)
                break
            recs.append({"iter": r["iter"], "misfit": r["misfit"], "mu_error": F.model_error(mesh, inv.mu, true.mu),
                         "armijo_cuts": r["armijo_counter"], "wall_s": time.perf_counter() - t0})
            _log(f"fwi {v}: iteration {r['iter']} misfit {r['misfit']:.6e} mu error {recs[-1]['mu_error']:.4f} "
                 f"({recs[-1]['wall_s']:.1f} s)")
            if recs[-1]["wall_s"] > a.fwi_budget_s:
                stop = f"time budget of {a.fwi_budget_s} s reached"
                break
        runs[v] = {"records": recs, "forward_solves": prob.nforward, "stopped": stop,
                   "recompute_ratio": ctx.recompute_ratio()}
    out = {"metric": f"FWI time-to-misfit ({a.fwi}: {mesh.nex}x{mesh.ney}x{mesh.nez} el, "
                     f"{len(cfg['src_nodes'])} shots, {nstep} steps, L-BFGS m=5)",
           "unit": "s", "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
           "data": "synthetic (4-layer model with two +-25 % ellipsoids; start: linear Vs 150-450 m/s)",
           "config": {"workload": a.fwi, "neq": mesh.neq, "nstep": nstep, "nshot": len(cfg["src_nodes"]),
                      "nrec": len(cfg["rec_nodes"])}, "runs": runs}
    if "plain" in runs and "batched" in runs:
        target = runs["plain"]["records"][-1]["misfit"]
        hit = [r for r in runs["batched"]["records"] if r["misfit"] <= target]
        tp = runs["plain"]["records"][-1]["wall_s"]
        if hit:
            out["value"] = hit[0]["wall_s"]
            out["plain_time_s"] = tp
            out["speedup"] = tp / hit[0]["wall_s"]
    ctx.close()
    print(json.dumps(out), flush=True)


# variant -> (batched shots, history, learned H0)
PC_VARIANTS = {"plain": (True, 1, False), "plain-seq": (False, 3, False), "learned": (True, 1, True),
               "learned-revolve": (True, 3, True)}


def precond_bench(a):
    """Learned gradient preconditioner: training pairs from the FP64 solver on
    random layered models with Gaussian inclusions (32x32x16 elements, PML 4, 4
    shots at 30 Hz), training with the HIP kernels (BF16 weights, FP32
    accumulation), then L-BFGS (m = 5) on a 64x64x32 survey of the same family
    with the plain H0 and with the learned one.  Reports the training cost, the
    held-out cosine of raw gradient vs network output with the true error, the
    misfit and mu error per iteration, and the iterations and wall time of the
    learned run to reach the plain run's final misfit.  Prints one JSON line."""
    import numpy as np
    from hirensi_amd import fwi as F
    from hirensi_amd import precond as PC
    from hirensi_amd import solver as S
    from hirensi_amd.mesh import BoxMesh
    out = {"metric": "learned-preconditioned L-BFGS: wall time to the plain run's final misfit (64x64x32)",
           "unit": "s", "higher_is_better": False, "n_gpus": 1, "dtype": "f64 solver, bf16/f32 network",
           "data": "synthetic (random layered models with 1-3 Gaussian Vs inclusions)",
           "network": {"layers": len(PC.LAYERS), "params": PC.n_params()}}
    net = PC.ConvNet(seed=0)
    if a.pc_weights:
        net.load(a.pc_weights)
        out["training"] = {"weights": a.pc_weights}
    else:
        t0 = time.perf_counter()
        X, T = PC.generate_pairs(a.pc_pairs, seed=1, nstep=a.pc_nstep, log=_log)
        t_gen = time.perf_counter() - t0
        nval = max(4, a.pc_pairs // 10)
        t1 = time.perf_counter()
        hist = PC.train(net, X[nval:], T[nval:], epochs=a.pc_epochs, batch=8, lr=1e-3, log=_log, log_every=10)
        import torch
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t1
        raw_cos, nn_cos = PC.evaluate(net, X[:nval], T[:nval])
        tr_raw, tr_nn = PC.evaluate(net, X[nval:nval + 20], T[nval:nval + 20])
        _log(f"precond: held-out cosine with the true error: raw gradient {raw_cos:.3f}, network {nn_cos:.3f}")
        out["training"] = {"pairs": a.pc_pairs, "held_out": nval, "pair_nstep": a.pc_nstep,
                           "pair_generation_s": t_gen, "train_s": t_train, "epochs": a.pc_epochs,
                           "loss_first": hist[0], "loss_last": hist[-1], "loss_history": hist,
                           "held_out_cosine_raw": raw_cos, "held_out_cosine_net": nn_cos,
                           "train_cosine_raw": tr_raw, "train_cosine_net": tr_nn}
        if a.pc_save:
            net.save(a.pc_save)
    mesh = BoxMesh(64, 64, 32, npml=(6,) * 5)
    true, start = PC.random_case(mesh, np.random.default_rng(a.pc_seed))
    sv = PC.training_survey(mesh, a.pc_fwi_nstep, nshot_side=2, nrec_side=20)
    ctx = S.Context(mesh)
    _log(f"precond: 64x64x32 survey, {sv.ftab.shape[0]} shots, {a.pc_fwi_nstep} steps; observing the true model")
    sv.dis_meas = F.HipProblem(ctx, sv, true.rho).observe(true.lam, true.mu)
    _, restriction = S.gradient_mass(ctx)
    params = F.InversionParams(m_lbfgs=5, bound_lambda_lower=40.0, bound_lambda_upper=1200.0, bound_mu_lower=20.0,
                               bound_mu_upper=600.0, fac_max_update_lambda=10.0, fac_max_update_mu=10.0)
    runs = {}
    ctx.set_history_budget(a.ckpt_budget)     # the revolve variants' HBM budget (0: half the free memory)
    variants = [v.strip() for v in a.pc_variants.split(",") if v.strip()]
    bad = [v for v in variants if v not in PC_VARIANTS]
    if bad or not any(v.startswith("plain") for v in variants) or not any(v.startswith("learned") for v in variants):
        raise SystemExit(f"--pc-variants: one plain and one learned variant of {sorted(PC_VARIANTS)}")
    for v in variants:
        batched, history, learned = PC_VARIANTS[v]
        if batched:
            prob = F.HipProblem(ctx, sv, start.rho, params, history=history, log=_log)
        else:
            prob = F.SequentialHipProblem(ctx, sv, start.rho, params, log=_log)
        h0 = PC.LearnedH0(net, mesh, restriction) if learned else None
        t0 = time.perf_counter()
        inv = F.Inversion(prob, start.lam, start.mu, restriction, params, h0_map=h0)
        recs = [{"iter": 0, "misfit": inv.misfit, "mu_error": F.model_error(mesh, inv.mu, true.mu),
                 "wall_s": time.perf_counter() - t0}]
        stop = ""
        for _ in range(a.pc_fwi_iters):
            try:
                r = inv.iterate()
            except F.ArmijoFailure as e:
                stop = str(e)
                break
            recs.append({"iter": r["iter"], "misfit": r["misfit"], "mu_error": F.model_error(mesh, inv.mu, true.mu),
                         "armijo_cuts": r["armijo_counter"], "learned_h0": r.get("learned_h0"),
                         "fallback": r.get("fallback"), "wall_s": time.perf_counter() - t0})
            _log(f"precond {v}: iteration {r['iter']} misfit {r['misfit']:.6e} mu error {recs[-1]['mu_error']:.4f}"
                 f"{' (learned step)' if r.get('learned_h0') else ''} ({recs[-1]['wall_s']:.1f} s)")
        runs[v] = {"records": recs, "forward_solves": prob.nforward, "stopped": stop,
                   "shots": "batched" if batched else "one per solve",
                   "history": {S.HR_HISTORY_STORE: "store", S.HR_HISTORY_REVOLVE: "revolve"}[history],
                   "recompute_ratio": ctx.recompute_ratio()}
        if h0 is not None:
            runs[v]["h0_calls"], runs[v]["h0_used"] = h0.calls, h0.used
    ctx.close()
    plain = next(v for v in variants if v.startswith("plain"))
    learned = [v for v in variants if v.startswith("learned")][-1]
    target = runs[plain]["records"][-1]["misfit"]
    hit = [r for r in runs[learned]["records"] if r["misfit"] <= target]
    out["config"] = {"workload": "64x64x32 el, PML 6, 4 shots at 30 Hz, 400 receivers", "nstep": a.pc_fwi_nstep,
                     "iterations": a.pc_fwi_iters, "true_model_seed": a.pc_seed, "plain": plain, "learned": learned}
    out["runs"] = runs
    out["plain_final_misfit"] = target
    out["plain_time_s"] = runs[plain]["records"][-1]["wall_s"]
    out["value"] = hit[0]["wall_s"] if hit else None
    out["learned_iterations_to_plain_misfit"] = hit[0]["iter"] if hit else None
    if hit:
        out["speedup"] = out["plain_time_s"] / hit[0]["wall_s"]
    # the same comparison on the model error (relative L2 to the true mu)
    target_e = runs[plain]["records"][-1]["mu_error"]
    hit_e = [r for r in runs[learned]["records"] if r["mu_error"] <= target_e]
    out["plain_final_mu_error"] = target_e
    out["learned_time_to_plain_mu_error_s"] = hit_e[0]["wall_s"] if hit_e else None
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.precond:
        precond_bench(a)
        return
    if a.fwi:
        fwi_bench(a)
        return
    if a.shots:
        shots_bench(a)
        return
    if a.fd_check:
        fd_check(a)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported (shots would shard across GPUs "
              f"with one all-reduce of the gradient per iteration)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from hirensi_amd import solver as S
    from hirensi_amd.mesh import config2, receiver_dofs

    dev = torch.device("cuda:0")
    _log("building the 64x64x32 mesh and models")
    true_cfg = config2(with_anomaly=True)
    start_cfg = config2(with_anomaly=False)
    mesh = true_cfg["mesh"]
    nstep, dt, fr = a.nstep, true_cfg["dt"], true_cfg["fr"]
    rec_eq = receiver_dofs(mesh, true_cfg["rec_nodes"])
    src_eq = np.array([mesh.dof(true_cfg["src_node"], 2)], dtype=np.int64)
    src_val = np.array([1.0])
    ftab = S.time_table(nstep, dt, 2, fr)
    nrec = len(rec_eq)
    history = S.HR_HISTORY_STORE if a.history == "store" else S.HR_HISTORY_CHECKPOINT

    def T(x, dtype=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=dev)

    version = a.version | (0x100 if a.serial_classes else 0)
    ctx = S.Context(mesh, device_ptrs=True, tz=a.tz, ty=a.ty, version=version)
    d_ftab, d_src, d_val, d_rec = T(ftab), T(src_eq, torch.int64), T(src_val), T(rec_eq, torch.int64)
    d_obs = torch.zeros((nstep + 1, nrec), dtype=torch.float64, device=dev)
    d_tr = torch.zeros_like(d_obs)
    d_gl = torch.zeros(mesh.nnode, dtype=torch.float64, device=dev)
    d_gm = torch.zeros_like(d_gl)
    d_mis = torch.zeros(1, dtype=torch.float64, device=dev)
    # observed data from the true model (untimed set-up)
    tm = true_cfg["model"]
    ctx.set_material_ptrs(T(tm.lam), T(tm.mu), T(tm.rho))
    S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, d_ftab, d_src, d_val, d_rec, history=S.HR_HISTORY_NONE,
                                traces=d_obs)
    sm = start_cfg["model"]
    ctx.set_material_ptrs(T(sm.lam), T(sm.mu), T(sm.rho))
    torch.cuda.synchronize()
    _log("observed data done")
    tuned = None
    if not a.tz:
        # the stage-kernel z-chunk length for this mesh, from a few timed RK4 steps per candidate
        tz, ms = ctx.autotune(nshot=1, steps=6)
        tuned = {"tz": tz, "ms_per_rk4_step": ms}
        _log(f"autotune: z-chunks of {tz} element layers ({ms:.3f} ms per RK4 step)")

    def gradient():
        S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, d_ftab, d_src, d_val, d_rec, history=history,
                                    segment=a.segment, traces=d_tr)
        S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, d_rec, d_obs, d_gl, d_gm, misfit_out=d_mis)

    for w in range(a.warmup):
        gradient()
        torch.cuda.synchronize()
        _log(f"warmup gradient {w + 1}/{a.warmup} done")
    L = ctx.L
    L.hr_set_profiling(ctx.h, 2)
    L.hr_set_profiling(ctx.h, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        gradient()
        if k + 1 == a.steps or (k + 1) % 5 == 0:
            _log(f"timed gradient {k + 1}/{a.steps} queued")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n_l = (np.zeros(5, dtype=np.int64))
    ms_l = np.zeros(5)
    L.hr_get_kernel_stats(ctx.h, n_l.ctypes.data, ms_l.ctypes.data)
    L.hr_set_profiling(ctx.h, 0)
    wall = (t1 - t0) / a.steps
    misfit = float(d_mis.item())
    gnorm = float(torch.linalg.norm(d_gm).item())
    neq = ctx.neq
    dof_steps = neq * 2.0 * nstep
    value = dof_steps / wall
    # roofline of the stage kernel: algorithmic bytes per launch (a quarter of B_step) /
    # average launch duration; the fused kernel's own byte count is reported beside it
    split = ctx.version() == 4
    if split:
        # forward and adjoint stages in equal numbers
        sf, sa = split_stage_bytes(mesh), split_stage_bytes(mesh, transpose=True)
        sb = {st: 0.5 * (sf[st] + sa[st]) for st in (1, 2, 3, 4)}
    else:
        sb = stage_bytes(neq, mesh.nnode)
    launches = int(n_l[1:].sum())
    kern_ms = float(ms_l[1:].sum())
    alg_per_launch = step_bytes(neq, mesh.nnode) / 4.0
    fused_bytes = sum(sb[s] * n_l[s] for s in (1, 2, 3, 4))
    avg_us = kern_ms * 1e3 / launches
    achieved = alg_per_launch * launches / (kern_ms * 1e-3) / 1e9
    tr = measured_traffic(split)
    out = {
        "metric": "fwd+adjoint gradient DOF-steps/s (64x64x32, 4000 steps, 1 GPU)",
        "value": value,
        "unit": "DOF-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded layered model; observed traces from the anomaly model)",
        "config": {"workload": "fwd+adj+grad, 64x64x32 el, PML 6, 4000 RK4 steps, 400 receivers",
                   "neq": neq, "nnode": mesh.nnode, "nstep": nstep, "history": a.history,
                   "recompute_ratio": ctx.recompute_ratio(), "misfit": misfit, "g_mu_norm": gnorm,
                   "gradient_wall_s": wall, "parallelism": "single GPU", "autotune": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr["traffic_bytes_per_launch"] if tr else None,
                     "traffic_source": traffic_file(split) if tr else None,
                     "kernel": ("RK4 stage, split path: nodal PML pass beside the regular-domain tile kernel, "
                                "then the nodal PML pull" if split else
                                "stage kernel (all tile classes of one RK4 stage)"),
                     "launches": launches,
                     "avg_launch_us": avg_us,
                     "alg_bytes_per_launch": alg_per_launch,
                     "alg_bytes_convention": "B_step/4, B_step = 8 (36 N + 8 N_node)",
                     "own_bytes_per_launch": fused_bytes / launches,
                     "own_achieved": fused_bytes / (kern_ms * 1e-3) / 1e9,
                     "stage_kernel_share_of_wall": kern_ms * 1e-3 / (wall * a.steps)},
    }
    if a.graph_leg > 0:
        # the same gradient with its time loops replayed from a captured hipGraph of
        # --graph-leg steps, against eager launches, both without timing events
        L.hr_set_profiling(ctx.h, 0)
        g_ref = d_gm.clone()
        walls = {}
        for mode, k in (("eager", 0), ("graph", a.graph_leg)):
            ctx.set_graph(k)
            gradient()
            torch.cuda.synchronize()
            t0g = time.perf_counter()
            for _ in range(2):
                gradient()
            torch.cuda.synchronize()
            walls[mode] = (time.perf_counter() - t0g) / 2
            _log(f"{mode} gradient without timing events: {walls[mode]:.3f} s")
        same = bool(torch.equal(d_gm, g_ref))
        ctx.set_graph(0)
        out["graph"] = {"steps_per_graph": a.graph_leg, "graph_wall_s": walls["graph"],
                        "eager_wall_s": walls["eager"], "speedup": walls["eager"] / walls["graph"],
                        "value": dof_steps / walls["graph"], "graph_launches": ctx.graph_launches(),
                        "gradient_bitwise_equal_to_eager": same}
    if a.ckpt_budget > 0:
        # the same gradient with the forward history checkpointed in HBM within a budget
        # (binomial revolve schedule); one warm-up (allocates the snapshot pool), one timed
        ctx.set_history_budget(a.ckpt_budget)
        L_, ns_, nb_, ratio_ = ctx.history_plan(nstep, S.HR_HISTORY_REVOLVE)

        def gradient_ckpt():
            S.explicit_rk4_petsc_pml_3d(ctx, nstep, dt, d_ftab, d_src, d_val, d_rec, history=S.HR_HISTORY_REVOLVE,
                                        traces=d_tr)
            S.explicit_rk4_petsc_pml_3d_adjoint_otd(ctx, nstep, dt, d_rec, d_obs, d_gl, d_gm, misfit_out=d_mis)

        gradient_ckpt()
        torch.cuda.synchronize()
        t0c = time.perf_counter()
        gradient_ckpt()
        torch.cuda.synchronize()
        wall_c = time.perf_counter() - t0c
        _log(f"checkpointed gradient done ({wall_c:.2f} s, recompute ratio {ctx.recompute_ratio():.3f})")
        out["checkpointed"] = {"history": "revolve", "budget_bytes": a.ckpt_budget, "history_bytes": nb_,
                               "segment": L_, "snapshots": ns_, "recompute_ratio": ctx.recompute_ratio(),
                               "gradient_wall_s": wall_c, "value": dof_steps / wall_c, "unit": "DOF-steps/s",
                               "store_history_bytes": 8.0 * ctx.nstore * (nstep + 1)}
    if a.copy_calib:
        x = torch.empty(1 << 27, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
    ctx.close()
    if not a.no_cpu_baseline:
        _log("CPU baseline (oracle)")
        cfg = {"mesh": mesh, "start": sm, "src_eq": src_eq, "rec_eq": rec_eq, "dt": dt, "fr": fr}
        out["cpu_baseline"] = cpu_baseline(cfg, a.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

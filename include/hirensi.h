/*
 * hirensi.h - C ABI of the MI355X explicit elastic forward / adjoint operator.
 *
 * Replaces the assembled-matrix path of HiReNSI's explicit solver
 * (step_4_3D_Forward_and_Inverse/, PETSc 3.4 Mat/Vec API):
 *   - the matrices AK, AC, AG assembled by ASSEM_REGULAR_DOMAIN_PETSC_3D and
 *     ASSEM_PML_PETSC_3D (Assemble_PETSc_3D.F90:6-436) and their products
 *     MatMultAdd / MatMultTransposeAdd (Solver_PETSc_Explicit.F90:1005-1034,
 *     1323-1352), with the in-place MatDiagonalScale / MatScale of
 *     Solver_PETSc_Explicit.F90:815-824,1222-1229;
 *   - the RK4 forward loop EXPLICIT_RK4_PETSC_PML_3D
 *     (Solver_PETSc_Explicit.F90:660-1117, called at Solve_Forward.F90:117);
 *   - the OtD adjoint EXPLICIT_RK4_PETSC_PML_3D_Adjoint_OtD
 *     (Solver_PETSc_Explicit.F90:1125-1405), the control pass
 *     EXPLICIT_RK4_PETSC_PML_3D_Control_OtD + Update_g
 *     (Solver_PETSc_Explicit.F90:1413-1522, Assemble_PETSc_3D.F90:556-661)
 *     and the sign flip of Program_Main.F90:404-413;
 *   - Compute_Misfit (Solver_PETSc_Inverse.F90:591-650).
 *
 * Conventions
 *   - Equation numbers are 0-based, node-major with the active DOFs of a node
 *     interleaved (u_x u_y u_z s_xx s_yy s_zz s_xy s_xz s_yz), nodes in
 *     lattice order n = i + NX*(j + NY*k) of the structured box (NX = 2*nex+1).
 *   - All calls return 0 on success, a positive code otherwise; the message
 *     is available from hr_last_error().  No call throws.
 *   - Pointers are host pointers unless the context was created with
 *     HR_DEVICE_PTRS, in which case every array argument is device memory.
 *   - Calls on one context are stream-ordered and not re-entrant.
 */
#ifndef HIRENSI_H
#define HIRENSI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hr_ctx hr_ctx;

#define HR_DEVICE_PTRS 1

/* which operator */
#define HR_K 0
#define HR_C 1
#define HR_G 2

/* history handling of the forward wavefield for the adjoint */
#define HR_HISTORY_NONE 0
#define HR_HISTORY_STORE 1        /* keep u on the regular-domain box for every step */
#define HR_HISTORY_CHECKPOINT 2   /* keep the state every `segment` steps, recompute */
#define HR_HISTORY_REVOLVE 3      /* binomial checkpoints within an HBM budget (hr_set_history_budget) */

/* Structured box mesh: the ANSYS box + PML of the reference pre-processing
 * (step_1_Ansys_Mesh_Generator, step_2_20_To_27/Main.F90:447-521). */
typedef struct hr_mesh_desc {
  int nex, ney, nez;        /* elements per axis, PML included */
  int npml[5];              /* PML element layers (>= 1): x-, x+, y-, y+, z- (top is free) */
  double h[3];              /* element edge lengths */
  int top_stress_bc;        /* 1: s_zz, s_xz, s_yz fixed at original top PML nodes */
  int flags;                /* HR_DEVICE_PTRS */
} hr_mesh_desc;

/* Context and mesh (replaces MatCreateAIJ, PETSc_Global_Objects.F90:46-51). */
int hr_ctx_create(const hr_mesh_desc* desc, hr_ctx** out);
int hr_ctx_destroy(hr_ctx* ctx);
const char* hr_last_error(const hr_ctx* ctx);
int hr_get_sizes(hr_ctx* ctx, long long* neq, long long* nnode, long long* nstore);
/* stored set (U_Store_Numbers_Global, PIC_Inversion_Data_Structure.F90:276-301):
 * equation numbers of the u DOFs of the regular-domain node box, in the order
 * the history is kept (component-major, box lattice order) */
int hr_get_store_eq(hr_ctx* ctx, long long* store_eq);

/* PML_PARAM(6,4), Fortran column-major: [alpha0 beta0 L x0] for faces
 * x+ x- y+ y- z+ z- (Adjustments.F90:53-84, PML.F90:325).  Always host
 * memory (a 24-value parameter block), also under HR_DEVICE_PTRS. */
int hr_set_pml(hr_ctx* ctx, const double* pml_param);

/* Nodal lambda and mu (PMat_Lambda, PMat_Mu) and per-element density
 * (PMAT(MTEL,3)), element order e = ex + nex*(ey + ney*ez).  Replaces the
 * element loops of ASSEM_REGULAR_DOMAIN_PETSC_3D / ASSEM_PML_PETSC_3D. */
int hr_set_material(hr_ctx* ctx, const double* lam, const double* mu, const double* rho_elem);

/* Lumped mass DIAG_M (Solve_Forward.F90:96-97), reference numbering. */
int hr_get_diag_m(hr_ctx* ctx, double* diag_m);

/* MatShell state of operator `which`: A_eff = s * diag(l) * A * diag(r).
 * hr_scale multiplies s (MatScale), hr_diagonal_scale multiplies l and r
 * elementwise (MatDiagonalScale; NULL leaves a side unchanged),
 * hr_reset_scaling restores s = 1, l = r = 1 (MatZeroEntries + reassembly). */
int hr_scale(hr_ctx* ctx, int which, double s);
int hr_diagonal_scale(hr_ctx* ctx, int which, const double* l, const double* r);
int hr_reset_scaling(hr_ctx* ctx, int which);

/* z = y + A_eff x (MatMultAdd) or z = y + A_eff^T x (MatMultTransposeAdd).
 * x, y, z have neq entries; z may alias y.  Host vectors: returns with z
 * written.  HR_DEVICE_PTRS: returns with the product queued on the context's
 * stream (hr_synchronize waits for it). */
int hr_apply(hr_ctx* ctx, int which, int transpose, const double* x, const double* y, double* z);

/* Waits until the work the context has queued is done (device-pointer
 * contexts: hr_apply, hr_diagonal_scale).  No counterpart in the reference
 * (PETSc's host MatMultAdd returns with its result). */
int hr_synchronize(hr_ctx* ctx);

/* TIME_FUNCTION (IO.F90:541-733), every LCASE: 1 sine, 2 Ricker, 3 Gaussian,
 * 4 Chopra cosine, 5 its derivative, 6 Greenberg delta, 7 Ricker derivative,
 * 8 one, 9 zero, 10 smooth Heaviside; NaN for other cases. */
double hr_time_function(double t, int lcase, double fr);

/* Forward RK4 from rest (EXPLICIT_RK4_PETSC_PML_3D).
 *   ftab[3*n + {0,1,2}] = f(n dt), f(n dt + dt/2), f(n dt + dt) for n < nstep
 *   B = sum of src_val at src_eq (force vector before M^-1 scaling)
 *   traces[(nstep+1) x nrec]: u at rec_eq at the start of steps 0..nstep
 *   history: HR_HISTORY_* ; segment: checkpoint interval (HR_HISTORY_CHECKPOINT) */
int hr_forward(hr_ctx* ctx, int nstep, double dt, const double* ftab, int nsrc, const long long* src_eq,
               const double* src_val, int nrec, const long long* rec_eq, double* traces, int history,
               int segment);

/* Batched shots over one material: nshot forward solves of
 * EXPLICIT_RK4_PETSC_PML_3D (Solver_PETSc_Explicit.F90:660-1117) advanced
 * together, one stage-kernel launch per RK4 stage covering every shot.  The
 * reference runs one shot per solve; this entry point has no single
 * counterpart there.  ftab[nshot x nstep x 3]; source entry e belongs to shot
 * src_shot[e]; the receivers are shared; traces[nshot x (nstep+1) x nrec].
 * hr_adjoint_gradient and hr_misfit then act on all shots of this forward
 * (observations shot-major, gradient and misfit summed over shots). */
int hr_forward_shots(hr_ctx* ctx, int nshot, int nstep, double dt, const double* ftab, int nsrc,
                     const int* src_shot, const long long* src_eq, const double* src_val, int nrec,
                     const long long* rec_eq, double* traces, int history, int segment);

/* The reference's central-difference variant EXPLICIT_Central_Difference_3
 * (Obsolete.F90:815-995) on the same operators, one shot:
 *   u_{n+1} = 2 u_n - u_{n-1} - M^-1 [dt^2 K u_n + dt C (u_n - u_{n-1})
 *             + dt^3 G u_n + dt^2 G u_his] + dt^2 f((n+1) dt) M^-1 B,
 *   u_his <- u_his + dt u_{n+1}, from zero initial state.
 *   ftab[nstep]: f(n dt) for n = 1..nstep (TIME_FUNCTION at T = ISTEP dt)
 *   traces[(nstep+1) x nrec]: row n = u at the start of step ISTEP = n
 *   (U_Store(:, ISTEP)); row 0 is the zero initial state (the reference leaves
 *   column 0 unset).  Leaves no history: hr_adjoint_gradient and hr_misfit
 *   need an RK4 forward. */
int hr_forward_cd(hr_ctx* ctx, int nstep, double dt, const double* ftab, int nsrc, const long long* src_eq,
                  const double* src_val, int nrec, const long long* rec_eq, double* traces);

/* OtD adjoint + gradient (Adjoint_OtD, Control_OtD, g <- -g).  Uses the
 * history of the last hr_forward / hr_forward_shots.  dis_meas[(nstep+1) x
 * nrec] measured data per shot, shots one after another.
 * g_lambda, g_mu: nodal (nnode), zero outside the regular-domain node box;
 * the sum over shots.  misfit (optional, may be NULL): Compute_Misfit of the
 * last forward, summed over shots. */
int hr_adjoint_gradient(hr_ctx* ctx, int nstep, double dt, int nrec, const long long* rec_eq,
                        const double* dis_meas, double* g_lambda, double* g_mu, double* misfit);

/* Compute_Misfit of the traces of the last forward against dis_meas (summed over its shots). */
int hr_misfit(hr_ctx* ctx, int nstep, double dt, int nrec, const double* dis_meas, double* misfit);

/* hr_misfit without allocation: the same kernels and summation order (the
 * result equals hr_misfit's bit for bit), with the per-step partial sums in a
 * fixed device array of 1048576 doubles (8 MiB per GPU, held until the process
 * ends) instead of memory the library allocates.  No counterpart in the
 * reference: it is for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  Needs a context
 * created with HR_DEVICE_PTRS (dis_meas and misfit on the device); refuses
 * nshot x (nstep + 1) + nshot > 1048576 with an error.  All calls on one GPU
 * share the array: calls that overlap on one GPU race. */
int hr_misfit_noalloc(hr_ctx* ctx, int nstep, double dt, int nrec, const double* dis_meas, double* misfit);

/* Update_g (Assemble_PETSc_3D.F90:556-661) for one step of the control pass
 * (Solver_PETSc_Explicit.F90:1474-1510): over the regular-domain elements
 * (MTYPE 1), g_lambda += scale * W div u div p and g_mu += scale * W 2 eps(u):eps(p)
 * at every GLL point, i.e. ELEMENT_g_3D assembled with ADD_VALUES and
 * multiplied by Scale_L.  u, p: neq entries in the context's numbering (the
 * displacement DOFs of the regular-domain nodes are read, as the reference
 * scatters only U_Store/P_Store entries); g_lambda, g_mu: nodal, accumulated. */
int hr_gradient_accumulate(hr_ctx* ctx, const double* u, const double* p, double scale, double* g_lambda,
                           double* g_mu);

/* Caller-supplied equation numbering (replaces the lattice order): id[9*n + c]
 * = equation number of component c (u_x u_y u_z s_xx s_yy s_zz s_xy s_xz s_yz)
 * of lattice node n, -1 where constrained -- the reference's ID_PETSc from
 * PIC_ReNumbering.F90:263-285 (nodes in PETSc order, active DOFs interleaved)
 * mapped onto the lattice.  The active set must be the context's; the numbers
 * a permutation of 0..neq-1.  Every later call that takes equation numbers or
 * neq-vectors uses it; resets the MatShell scalings and the last forward. */
int hr_set_numbering(hr_ctx* ctx, const long long* id);

/* Surface traction force (ASSEM_FORCE_PETSC_3D, Assemble_PETSc_3D.F90:444-548,
 * with ELEMENT_FS_3D, Element_Matrices_3D.F90:890-1096): a unit traction in z on
 * face nface[e] of every regular-domain element (ID_BC(IEL,1): 0 none, 1..6 as
 * the reference numbers faces, 6 treated as 5), GLL face quadrature, assembled
 * with ADD_VALUES; constrained DOFs dropped.  nface: host, element order
 * e = ex + nex*(ey + ney*ez).  Output: the nonzero entries of B as (eq, val)
 * pairs in ascending eq (host; pass eq = val = NULL to get *nnz only), ready
 * as the src_eq / src_val of hr_forward. */
int hr_surface_force(hr_ctx* ctx, const int* nface, long long* eq, double* val, long long* nnz);

/* Regular-domain energy (TOTAL_ENERGY_RD_PETSC_EXPLICIT, Obsolete.F90:108-150,
 * called every I_PRINT_STEP steps at Solver_PETSc_Explicit.F90:915-916):
 * 0.5 * (M_RD x3) . x3 -- the reference's AK_RD term is zero because AK_RD is
 * never assembled (Assemble_PETSc_3D.F90:163).  hr_set_energy_interval(every)
 * makes later RK4 forwards record it at steps 0, every, 2 every, ... <= nstep
 * (0: off); hr_get_energy returns the last forward's values, nshot x nrow. */
int hr_set_energy_interval(hr_ctx* ctx, int every);
int hr_get_energy(hr_ctx* ctx, double* energy, int* nrow);

/* Gradient post-processing on the regular domain (Program_Main.F90:416-484).
 * Nodal vectors have nnode entries; only regular-domain elements contribute
 * (MTYPE == 1).  wreg: NULL for Weighted_Regularization = 'No', else the host
 * triple (w_top, w_bot, w_Length) of the weighted scheme
 * (Adjustments.F90:174-177).  Scalar outputs follow HR_DEVICE_PTRS like arrays. */
#define HR_REG_TV 0   /* Assem_TV_TN_Regularization (Assemble_PETSc_3D.F90:755-860) */
#define HR_REG_TN 1   /* MatMult(Reg_PETSC, .) with Reg from Assem_R_Mg_3D */

/* 1/Mg of the gradient mass (Assem_R_Mg_3D, Assemble_PETSc_3D.F90:670-748)
 * as inverted by Control_Problem_Operators.F90:12-24: imgrad = 1/Mg where
 * Mg > 1e-10, else 0; restriction (may be NULL) = 1 there, else 0. */
int hr_gradient_mass(hr_ctx* ctx, double* imgrad, double* restriction);

/* Unscaled regularisation vectors of lam and mu (nodal) and their costs
 * cost[2] = (lambda, mu), may be NULL.  HR_REG_TV: ELEMENT_TV_TN_Reg_3D
 * (Element_Matrices_3D.F90:1899-2075) summed over the regular domain, the
 * cost with the reference's integer-truncated numerator; HR_REG_TN: R m with R
 * of ELEMENT_R_Mg_3D (Element_Matrices_3D.F90:1555-1718), cost 1/2 m.R m
 * (Program_Main.F90:419-428,464-474 before scaling). */
int hr_regularization(hr_ctx* ctx, int method, const double* wreg, const double* lam, const double* mu,
                      double* reg_lam, double* reg_mu, double* cost);

/* Sections IV and V.2-3 of the inversion loop (Program_Main.F90:416-481) on
 * the misfit gradient g_lambda, g_mu (in place, nodal, as returned by
 * hr_adjoint_gradient): regularisation vectors of (lam, mu), adaptive factor
 * f = factor * ||g|| / ||reg|| per field (0 when ||reg|| = 0, where the
 * reference divides by zero), g <- g + f reg unless iter == 1, g <- g / Mg.
 * cost_reg[2] (may be NULL): f * cost per field, 0 when iter == 1;
 * reg_factor[2] (may be NULL): the two factors. */
int hr_regularize_gradient(hr_ctx* ctx, int iter, int method, double factor, const double* wreg, const double* lam,
                           const double* mu, double* g_lambda, double* g_mu, double* cost_reg, double* reg_factor);

/* Forward steps recomputed per adjoint (1.0 with a stored history). */
double hr_recompute_ratio(const hr_ctx* ctx);

/* The forward history for the adjoint.  The reference keeps every regular-domain
 * displacement slice in host memory (U_Store_Mapping, PETSc_Global_Objects_Inversion.F90:71);
 * HR_HISTORY_REVOLVE keeps it in HBM within `bytes`: every slice if they fit,
 * else snapshots of the RK4 state at segment starts placed by the binomial
 * schedule of Griewank & Walther (revolve) plus one segment of slices, chosen
 * for the fewest recomputed steps.  0 (the default) = half of the free device
 * memory.  No counterpart in the reference, which has no recomputation. */
int hr_set_history_budget(hr_ctx* ctx, double bytes);
/* The plan hr_forward would use: segment length (0: every slice kept),
 * snapshot slots, HBM bytes and the predicted recompute ratio. */
int hr_get_history_plan(hr_ctx* ctx, int nstep, int nshot, int history, int segment, int* seglen, int* nsnap,
                        double* bytes, double* ratio);
/* The binomial reverse schedule of nseg segments with nsnap snapshot slots
 * (host only, no context): actions as (op, a, b) triples, the forward's
 * snapshots first (op 3 = snapshot slot a <- segment b), then the reverse sweep
 * (0 zero state, 1 restore slot a, 2 advance segment a -> b, 3 snapshot slot a
 * <- segment b, 4 record segment a, 5 consume segment a); advances = segment
 * advances of the sweep.  Returns 2 when max_actions is too small. */
int hr_revolve_schedule(int nseg, int nsnap, int max_actions, int* actions, int* nactions, long long* advances);

/* Stage-kernel configuration: tile of 4 x ty elements in x-y (ty = 3), tz
 * elements per z-chunk (0: 10 for version 4, whose chunks start at the regular
 * domain's bottom layer; 7 for version 3), kernel version (0 keeps it): 4 (the default) splits
 * each RK4 stage into a nodal pass over the PML nodes, the tiled
 * regular-domain kernel and a nodal pull over the PML nodes; 3 runs the fused
 * tile classes (mixed, regular-domain, PML) for every stage.  Single-operator
 * applies and the central-difference step always use the fused kernel.  The
 * kernels of a stage run concurrently on side streams unless
 * HR_TILING_SERIAL_CLASSES is or-ed into version.  Invalidates the history of
 * the last hr_forward. */
#define HR_TILING_SERIAL_CLASSES 0x100
int hr_set_tiling(hr_ctx* ctx, int ty, int tz, int version);
/* The stage-kernel configuration in use: tile y-extent, z-chunk length and
 * kernel version (no counterpart in the reference; for callers that report it). */
int hr_get_tiling(hr_ctx* ctx, int* ty, int* tz, int* version);

/* Stage-kernel timing (off by default; on = 1, reset = 2).  One stage launch
 * in seven is timed with HIP events recorded right before and after its
 * kernel on each stream it uses (its span: first start to last end).
 * hr_get_kernel_stats fills, per stage kind (0 = single-operator apply,
 * 1..4 = RK4 stage), launches[5] = stage launches while profiling and
 * ms[5] = their kernel time, the mean span of the timed launches times
 * launches. */
int hr_set_profiling(hr_ctx* ctx, int mode);
int hr_get_kernel_stats(hr_ctx* ctx, long long* launches, double* ms);

/* Autotuner of the stage-kernel tiling: times `steps` RK4 steps of nshot shots
 * (zero state, no sources: the same launches and traffic) for each z-chunk
 * length of the tiles (3..14 element layers) and keeps the fastest; returns it
 * and its time per step (ms).  Invalidates the history of the last forward.
 * No counterpart in the reference (a PETSc AIJ product has no tiling). */
int hr_autotune(hr_ctx* ctx, int nshot, int steps, int* tz_best, double* ms_per_step);

/* Captured time loops: with steps > 0 (even), the time loops of hr_forward /
 * hr_forward_shots (history NONE or STORE, no energy diagnostic) and of
 * hr_adjoint_gradient (STORE history) capture `steps` RK4 steps into one
 * hipGraph and replay it; the per-step values come from a device-side step
 * index, and results are bit-identical to the eager launches.  0 (the default):
 * eager launches.  Checkpointed and revolve histories and profiling runs stay
 * eager.  No counterpart in the reference (PETSc calls MatMultAdd per step). */
int hr_set_graph(hr_ctx* ctx, int steps);
/* hipGraph replays since the context was created. */
int hr_get_graph_launches(hr_ctx* ctx, long long* launches);

#ifdef __cplusplus
}
#endif
#endif /* HIRENSI_H */

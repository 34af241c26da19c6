! ISO_C_BINDING interface to the C ABI of include/hirensi.h, for the
! reference's Fortran driver (step_4_3D_Forward_and_Inverse/).  USE HIRENSI
! next to USE PARAMETERS; the calls replace the assembled-matrix path:
!   ASSEM_*_PETSC_3D + MatAssembly      -> hr_ctx_create, hr_set_pml, hr_set_material
!   EXPLICIT_RK4_PETSC_PML_3D           -> hr_forward
!   EXPLICIT_Central_Difference_3       -> hr_forward_cd
!   *_Adjoint_OtD + *_Control_OtD       -> hr_adjoint_gradient
!   MatMultAdd / MatMultTransposeAdd    -> hr_apply (or the MatShell adapter)
!   Assem_R_Mg_3D + iMg inversion       -> hr_gradient_mass
!   Program_Main.F90:416-481            -> hr_regularize_gradient (hr_regularization)
!   (wreg: c_loc of (w_top, w_bot, w_Length), or c_null_ptr when unweighted)
module hirensi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: HR_K = 0, HR_C = 1, HR_G = 2
  integer(c_int), parameter :: HR_DEVICE_PTRS = 1
  integer(c_int), parameter :: HR_HISTORY_NONE = 0, HR_HISTORY_STORE = 1, HR_HISTORY_CHECKPOINT = 2
  integer(c_int), parameter :: HR_HISTORY_REVOLVE = 3
  integer(c_int), parameter :: HR_REG_TV = 0, HR_REG_TN = 1

  type, bind(C) :: hr_mesh_desc
    integer(c_int) :: nex, ney, nez
    integer(c_int) :: npml(5)            ! x-, x+, y-, y+, z-
    real(c_double) :: h(3)
    integer(c_int) :: top_stress_bc
    integer(c_int) :: flags
  end type hr_mesh_desc

  interface
    integer(c_int) function hr_ctx_create(desc, ctx) bind(C, name="hr_ctx_create")
      import :: c_int, c_ptr, hr_mesh_desc
      type(hr_mesh_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: ctx
    end function hr_ctx_create

    integer(c_int) function hr_ctx_destroy(ctx) bind(C, name="hr_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function hr_ctx_destroy

    type(c_ptr) function hr_last_error(ctx) bind(C, name="hr_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function hr_last_error

    integer(c_int) function hr_get_sizes(ctx, neq, nnode, nstore) bind(C, name="hr_get_sizes")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: neq, nnode, nstore
    end function hr_get_sizes

    integer(c_int) function hr_get_store_eq(ctx, store_eq) bind(C, name="hr_get_store_eq")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: store_eq(*)
    end function hr_get_store_eq

    integer(c_int) function hr_set_pml(ctx, pml_param) bind(C, name="hr_set_pml")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: pml_param(6, 4)
    end function hr_set_pml

    integer(c_int) function hr_set_material(ctx, lam, mu, rho_elem) bind(C, name="hr_set_material")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: lam(*), mu(*), rho_elem(*)
    end function hr_set_material

    integer(c_int) function hr_get_diag_m(ctx, diag_m) bind(C, name="hr_get_diag_m")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: diag_m(*)
    end function hr_get_diag_m

    integer(c_int) function hr_scale(ctx, which, s) bind(C, name="hr_scale")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      real(c_double), value :: s
    end function hr_scale

    integer(c_int) function hr_diagonal_scale(ctx, which, l, r) bind(C, name="hr_diagonal_scale")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      type(c_ptr), value :: l, r          ! c_loc(vector) or c_null_ptr
    end function hr_diagonal_scale

    integer(c_int) function hr_reset_scaling(ctx, which) bind(C, name="hr_reset_scaling")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
    end function hr_reset_scaling

    integer(c_int) function hr_apply(ctx, which, transpose, x, y, z) bind(C, name="hr_apply")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: which, transpose
      real(c_double), intent(in) :: x(*), y(*)
      real(c_double), intent(inout) :: z(*)
    end function hr_apply

    ! MatMultAdd(A, x, k, k) passes one vector as both y and z: bind the same
    ! symbol with c_loc arguments (k declared TARGET), which is conforming
    ! Fortran where two array dummies sharing an actual argument are not.
    integer(c_int) function hr_apply_ptr(ctx, which, transpose, x, y, z) bind(C, name="hr_apply")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: which, transpose
      type(c_ptr), value :: x, y, z
    end function hr_apply_ptr

    integer(c_int) function hr_gradient_accumulate(ctx, u, p, scale, g_lambda, g_mu) &
        bind(C, name="hr_gradient_accumulate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*), p(*)
      real(c_double), value :: scale
      real(c_double), intent(inout) :: g_lambda(*), g_mu(*)
    end function hr_gradient_accumulate

    integer(c_int) function hr_set_numbering(ctx, id) bind(C, name="hr_set_numbering")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(in) :: id(9, *)
    end function hr_set_numbering

    integer(c_int) function hr_surface_force(ctx, nface, eq, val, nnz) bind(C, name="hr_surface_force")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: nface(*)
      type(c_ptr), value :: eq, val        ! c_loc of the outputs, or c_null_ptr to size
      integer(c_long_long), intent(out) :: nnz
    end function hr_surface_force

    integer(c_int) function hr_set_energy_interval(ctx, every) bind(C, name="hr_set_energy_interval")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: every
    end function hr_set_energy_interval

    integer(c_int) function hr_get_energy(ctx, energy, nrow) bind(C, name="hr_get_energy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, energy
      integer(c_int), intent(out) :: nrow
    end function hr_get_energy

    integer(c_int) function hr_set_history_budget(ctx, bytes) bind(C, name="hr_set_history_budget")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: bytes
    end function hr_set_history_budget

    integer(c_int) function hr_autotune(ctx, nshot, steps, tz_best, ms_per_step) bind(C, name="hr_autotune")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nshot, steps
      integer(c_int), intent(out) :: tz_best
      real(c_double), intent(out) :: ms_per_step
    end function hr_autotune

    integer(c_int) function hr_get_history_plan(ctx, nstep, nshot, history, segment, seglen, nsnap, bytes, &
        ratio) bind(C, name="hr_get_history_plan")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nstep, nshot, history, segment
      integer(c_int), intent(out) :: seglen, nsnap
      real(c_double), intent(out) :: bytes, ratio
    end function hr_get_history_plan

    integer(c_int) function hr_revolve_schedule(nseg, nsnap, max_actions, actions, nactions, advances) &
        bind(C, name="hr_revolve_schedule")
      import :: c_int, c_long_long
      integer(c_int), value :: nseg, nsnap, max_actions
      integer(c_int), intent(out) :: actions(*)
      integer(c_int), intent(out) :: nactions
      integer(c_long_long), intent(out) :: advances
    end function hr_revolve_schedule

    integer(c_int) function hr_synchronize(ctx) bind(C, name="hr_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function hr_synchronize

    integer(c_int) function hr_set_graph(ctx, steps) bind(C, name="hr_set_graph")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: steps
    end function hr_set_graph

    integer(c_int) function hr_get_graph_launches(ctx, launches) bind(C, name="hr_get_graph_launches")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: launches
    end function hr_get_graph_launches

    ! device pointers only (HR_DEVICE_PTRS contexts)
    integer(c_int) function hr_misfit_noalloc(ctx, nstep, dt, nrec, dis_meas, misfit) &
        bind(C, name="hr_misfit_noalloc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nstep, nrec
      real(c_double), value :: dt
      type(c_ptr), value :: dis_meas, misfit
    end function hr_misfit_noalloc

    ! learned preconditioner kernels (include/hirensi_nn.h): device pointers, a HIP stream
    integer(c_int) function hr_nn_conv3d(x, w, bias, y, cin, cout, nb, d, h, wd, accumulate, stream) &
        bind(C, name="hr_nn_conv3d")
      import :: c_int, c_ptr
      type(c_ptr), value :: x, w, bias, y, stream
      integer(c_int), value :: cin, cout, nb, d, h, wd, accumulate
    end function hr_nn_conv3d

    integer(c_int) function hr_nn_conv3d_wgrad(x, dy, part, nslice, dw, db, cin, cout, nb, d, h, wd, stream) &
        bind(C, name="hr_nn_conv3d_wgrad")
      import :: c_int, c_ptr
      type(c_ptr), value :: x, dy, part, dw, db, stream
      integer(c_int), value :: nslice, cin, cout, nb, d, h, wd
    end function hr_nn_conv3d_wgrad

    integer(c_int) function hr_nn_transpose_weights(w, wt, cin, cout, stream) bind(C, name="hr_nn_transpose_weights")
      import :: c_int, c_ptr
      type(c_ptr), value :: w, wt, stream
      integer(c_int), value :: cin, cout
    end function hr_nn_transpose_weights

    integer(c_int) function hr_nn_leaky_relu(x, n, slope, stream) bind(C, name="hr_nn_leaky_relu")
      import :: c_int, c_ptr, c_long_long, c_float
      type(c_ptr), value :: x, stream
      integer(c_long_long), value :: n
      real(c_float), value :: slope
    end function hr_nn_leaky_relu

    integer(c_int) function hr_nn_leaky_relu_grad(y, dy, n, slope, stream) bind(C, name="hr_nn_leaky_relu_grad")
      import :: c_int, c_ptr, c_long_long, c_float
      type(c_ptr), value :: y, dy, stream
      integer(c_long_long), value :: n
      real(c_float), value :: slope
    end function hr_nn_leaky_relu_grad

    integer(c_int) function hr_nn_avgpool2(x, y, c, d, h, wd, stream) bind(C, name="hr_nn_avgpool2")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: x, y, stream
      integer(c_long_long), value :: c
      integer(c_int), value :: d, h, wd
    end function hr_nn_avgpool2

    integer(c_int) function hr_nn_avgpool2_grad(dy, dx, c, d, h, wd, accumulate, stream) &
        bind(C, name="hr_nn_avgpool2_grad")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: dy, dx, stream
      integer(c_long_long), value :: c
      integer(c_int), value :: d, h, wd, accumulate
    end function hr_nn_avgpool2_grad

    integer(c_int) function hr_nn_upsample2(x, y, c, d, h, wd, stream) bind(C, name="hr_nn_upsample2")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: x, y, stream
      integer(c_long_long), value :: c
      integer(c_int), value :: d, h, wd
    end function hr_nn_upsample2

    integer(c_int) function hr_nn_upsample2_grad(dy, dx, c, d, h, wd, stream) bind(C, name="hr_nn_upsample2_grad")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: dy, dx, stream
      integer(c_long_long), value :: c
      integer(c_int), value :: d, h, wd
    end function hr_nn_upsample2_grad

    integer(c_int) function hr_nn_mse(y, t, n, grad, part, loss, stream) bind(C, name="hr_nn_mse")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: y, t, grad, part, loss, stream
      integer(c_long_long), value :: n
    end function hr_nn_mse

    integer(c_int) function hr_nn_adam(w, g, m, v, wb, n, lr, b1, b2, eps, step, stream) bind(C, name="hr_nn_adam")
      import :: c_int, c_ptr, c_long_long, c_float
      type(c_ptr), value :: w, g, m, v, wb, stream
      integer(c_long_long), value :: n
      real(c_float), value :: lr, b1, b2, eps
      integer(c_int), value :: step
    end function hr_nn_adam

    integer(c_int) function hr_nn_to_bf16(w, wb, n, stream) bind(C, name="hr_nn_to_bf16")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: w, wb, stream
      integer(c_long_long), value :: n
    end function hr_nn_to_bf16

    real(c_double) function hr_time_function(t, lcase, fr) bind(C, name="hr_time_function")
      import :: c_double, c_int
      real(c_double), value :: t, fr
      integer(c_int), value :: lcase
    end function hr_time_function

    integer(c_int) function hr_forward(ctx, nstep, dt, ftab, nsrc, src_eq, src_val, nrec, rec_eq, traces, &
                                       history, segment) bind(C, name="hr_forward")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: nstep, nsrc, nrec, history, segment
      real(c_double), value :: dt
      real(c_double), intent(in) :: ftab(3, *), src_val(*)
      integer(c_long_long), intent(in) :: src_eq(*), rec_eq(*)
      real(c_double), intent(out) :: traces(*)
    end function hr_forward

    integer(c_int) function hr_forward_shots(ctx, nshot, nstep, dt, ftab, nsrc, src_shot, src_eq, src_val, nrec, &
                                             rec_eq, traces, history, segment) bind(C, name="hr_forward_shots")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: nshot, nstep, nsrc, nrec, history, segment
      real(c_double), value :: dt
      real(c_double), intent(in) :: ftab(3, *), src_val(*)
      integer(c_int), intent(in) :: src_shot(*)
      integer(c_long_long), intent(in) :: src_eq(*), rec_eq(*)
      real(c_double), intent(out) :: traces(*)
    end function hr_forward_shots

    integer(c_int) function hr_adjoint_gradient(ctx, nstep, dt, nrec, rec_eq, dis_meas, g_lambda, g_mu, misfit) &
        bind(C, name="hr_adjoint_gradient")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: nstep, nrec
      real(c_double), value :: dt
      integer(c_long_long), intent(in) :: rec_eq(*)
      real(c_double), intent(in) :: dis_meas(*)
      real(c_double), intent(out) :: g_lambda(*), g_mu(*)
      real(c_double), intent(out) :: misfit
    end function hr_adjoint_gradient

    integer(c_int) function hr_misfit(ctx, nstep, dt, nrec, dis_meas, misfit) bind(C, name="hr_misfit")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nstep, nrec
      real(c_double), value :: dt
      real(c_double), intent(in) :: dis_meas(*)
      real(c_double), intent(out) :: misfit
    end function hr_misfit

    integer(c_int) function hr_forward_cd(ctx, nstep, dt, ftab, nsrc, src_eq, src_val, nrec, rec_eq, traces) &
        bind(C, name="hr_forward_cd")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: nstep, nsrc, nrec
      real(c_double), value :: dt
      real(c_double), intent(in) :: ftab(*), src_val(*)
      integer(c_long_long), intent(in) :: src_eq(*), rec_eq(*)
      real(c_double), intent(out) :: traces(*)
    end function hr_forward_cd

    integer(c_int) function hr_gradient_mass(ctx, imgrad, restriction) bind(C, name="hr_gradient_mass")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: imgrad(*), restriction(*)
    end function hr_gradient_mass

    integer(c_int) function hr_regularization(ctx, method, wreg, lam, mu, reg_lam, reg_mu, cost) &
        bind(C, name="hr_regularization")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, wreg
      integer(c_int), value :: method
      real(c_double), intent(in) :: lam(*), mu(*)
      real(c_double), intent(out) :: reg_lam(*), reg_mu(*), cost(2)
    end function hr_regularization

    integer(c_int) function hr_regularize_gradient(ctx, iter, method, factor, wreg, lam, mu, g_lambda, g_mu, &
                                                   cost_reg, reg_factor) bind(C, name="hr_regularize_gradient")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, wreg
      integer(c_int), value :: iter, method
      real(c_double), value :: factor
      real(c_double), intent(in) :: lam(*), mu(*)
      real(c_double), intent(inout) :: g_lambda(*), g_mu(*)
      real(c_double), intent(out) :: cost_reg(2), reg_factor(2)
    end function hr_regularize_gradient

    real(c_double) function hr_recompute_ratio(ctx) bind(C, name="hr_recompute_ratio")
      import :: c_double, c_ptr
      type(c_ptr), value :: ctx
    end function hr_recompute_ratio

    integer(c_int) function hr_set_tiling(ctx, ty, tz, version) bind(C, name="hr_set_tiling")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ty, tz, version
    end function hr_set_tiling

    integer(c_int) function hr_get_tiling(ctx, ty, tz, version) bind(C, name="hr_get_tiling")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: ty, tz, version
    end function hr_get_tiling

    integer(c_int) function hr_set_profiling(ctx, mode) bind(C, name="hr_set_profiling")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function hr_set_profiling

    integer(c_int) function hr_get_kernel_stats(ctx, launches, ms) bind(C, name="hr_get_kernel_stats")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: launches(5)
      real(c_double), intent(out) :: ms(5)
    end function hr_get_kernel_stats
  end interface

contains

  ! Fortran string of hr_last_error
  function hr_error_message(ctx) result(msg)
    type(c_ptr), value :: ctx
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    interface
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
        import :: c_ptr, c_size_t
        type(c_ptr), value :: s
      end function c_strlen
    end interface
    type(c_ptr) :: cp
    integer :: i, n
    cp = hr_last_error(ctx)
    n = int(c_strlen(cp))
    call c_f_pointer(cp, p, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = p(i)
    end do
  end function hr_error_message

end module hirensi

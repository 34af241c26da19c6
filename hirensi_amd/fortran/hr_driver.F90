! A Fortran driver of the C ABI, in the reference driver's call order
! (Program_Main.F90:318-413): mesh and material, forward solve with stored
! history, misfit, adjoint + gradient.  Reads its inputs from a stream file,
! writes traces, gradients and misfit to another; the Python side writes the
! same inputs and compares its own results with these.
!
! Input (stream, native endianness): int32 nex, ney, nez, npml(5), top_rule,
! nstep, nsrc, nrec; float64 h(3), dt; float64 pml_param(6,4), lam(nnode),
! mu(nnode), rho(nel), ftab(3,nstep); int64 src_eq(nsrc); float64 src_val(nsrc);
! int64 rec_eq(nrec); float64 dis_meas(nrec, 0:nstep).
! Output: float64 traces(nrec, 0:nstep), g_lambda(nnode), g_mu(nnode), misfit.
program hr_driver
  use iso_c_binding
  use hirensi
  implicit none
  type(hr_mesh_desc) :: d
  type(c_ptr) :: ctx
  integer(c_int) :: ihead(13), rc, nstep, nsrc, nrec
  real(c_double) :: rhead(4), misfit
  real(c_double) :: pml(6, 4)
  real(c_double), allocatable :: lam(:), mu(:), rho(:), ftab(:, :), src_val(:), dmeas(:), traces(:), gl(:), gm(:)
  integer(c_long_long), allocatable :: src_eq(:), rec_eq(:)
  integer(c_long_long) :: neq, nnode, nstore
  integer :: nel, u
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) ihead, rhead
  d%nex = ihead(1); d%ney = ihead(2); d%nez = ihead(3)
  d%npml = ihead(4:8)
  d%top_stress_bc = ihead(9)
  d%h = rhead(1:3)
  d%flags = 0
  nstep = ihead(10); nsrc = ihead(11); nrec = ihead(12)
  rc = hr_ctx_create(d, ctx)
  call check(rc, 'hr_ctx_create')
  rc = hr_get_sizes(ctx, neq, nnode, nstore)
  call check(rc, 'hr_get_sizes')
  nel = d%nex * d%ney * d%nez
  allocate(lam(nnode), mu(nnode), rho(nel), ftab(3, nstep), src_eq(nsrc), src_val(nsrc), rec_eq(nrec))
  allocate(dmeas(nrec * (nstep + 1)), traces(nrec * (nstep + 1)), gl(nnode), gm(nnode))
  read(u) pml, lam, mu, rho, ftab, src_eq, src_val, rec_eq, dmeas
  close(u)
  rc = hr_set_pml(ctx, pml)
  call check(rc, 'hr_set_pml')
  rc = hr_set_material(ctx, lam, mu, rho)
  call check(rc, 'hr_set_material')
  rc = hr_forward(ctx, nstep, rhead(4), ftab, nsrc, src_eq, src_val, nrec, rec_eq, traces, HR_HISTORY_STORE, 1_c_int)
  call check(rc, 'hr_forward')
  rc = hr_adjoint_gradient(ctx, nstep, rhead(4), nrec, rec_eq, dmeas, gl, gm, misfit)
  call check(rc, 'hr_adjoint_gradient')
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) traces, gl, gm, misfit
  close(u)
  rc = hr_ctx_destroy(ctx)
  write(*, '(A, ES24.16)') 'hr_driver: misfit ', misfit

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      write(*, '(A, A, A, I0, A, A)') 'hr_driver: ', what, ' failed (', code, '): ', hr_error_message(ctx)
      stop 1
    end if
  end subroutine check

end program hr_driver

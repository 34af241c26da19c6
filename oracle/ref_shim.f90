! C-callable entry points around the reference's own element routines
! (Element_Matrices_3D.F90, PML.F90, Shape_Functions.F90, Source1.F90),
! used only to write golden vectors and to check the CPU oracle.
! Each call describes one 27-node element as a one-element mesh:
! XYZ(27,3) = transpose(XT), INOD(:,1) = 1..27, NGP = 3.
module hr_ref_shim
  use iso_c_binding
  use PARAMETERS
  implicit none
  external :: ELEMENT_K_3D, ELEMENT_M_3D, ELEMENT_EN_3D, ELEMENT_EA_3D, ELEMENT_FS_3D
  external :: ELEMENT_g_3D, SHAPE_3_27, PML_FINDER_3D, ELEMENT_R_Mg_3D, ELEMENT_TV_TN_Reg_3D
contains

  subroutine set_globals()
    NJ = 27
    NEL = 1
    NNODE = 27
    NDIM = 3
    NDOF = 9
    NMAT = 2
    NPM = 8
    NINT_Lobatto = 3
    Weighted_Regularization = 'No '
  end subroutine

  subroutine one_element(xt, xyz, inod, ngp, mtel, pmat, mtype, rho)
    real(c_double), intent(in) :: xt(3, 27)
    real(8), intent(out) :: xyz(27, 3), pmat(2, 8)
    integer, intent(out) :: inod(27, 1), ngp(1), mtel(1)
    integer, intent(in) :: mtype
    real(8), intent(in) :: rho
    integer :: i
    call set_globals()
    do i = 1, 27
      xyz(i, :) = xt(:, i)
      inod(i, 1) = i
    end do
    ngp(1) = 3
    mtel(1) = mtype
    pmat = 0.0d0
    pmat(:, 1) = 1.0d0          ! E (overridden by nodal Lame parameters)
    pmat(:, 2) = 0.25d0         ! nu
    pmat(:, 3) = rho            ! density
  end subroutine

  subroutine ref_shape_3_27(x1, x2, x3, fn, dfxi) bind(C, name="ref_shape_3_27")
    real(c_double), value :: x1, x2, x3
    real(c_double), intent(out) :: fn(27), dfxi(27, 3)
    call SHAPE_3_27(x1, x2, x3, fn, dfxi)
  end subroutine

  subroutine ref_pml_finder_3d(xg, pml_param, pab) bind(C, name="ref_pml_finder_3d")
    real(c_double), intent(in) :: xg(3), pml_param(6, 4)
    real(c_double), intent(out) :: pab(19)
    call set_globals()
    call PML_FINDER_3D(xg, pml_param, pab)
  end subroutine

  subroutine ref_element_k(xt, lam, mu, ek) bind(C, name="ref_element_k")
    real(c_double), intent(in) :: xt(3, 27), lam(27), mu(27)
    real(c_double), intent(out) :: ek(81, 81)
    real(8) :: xyz(27, 3), pmat(2, 8)
    integer :: inod(27, 1), ngp(1), mtel(1)
    call one_element(xt, xyz, inod, ngp, mtel, pmat, 1, 1.0d0)
    call ELEMENT_K_3D(xyz, inod, ngp, mtel, pmat, lam, mu, 1, ek)
  end subroutine

  subroutine ref_element_m(xt, rho, pml_param, iflag, em) bind(C, name="ref_element_m")
    real(c_double), intent(in) :: xt(3, 27), pml_param(6, 4)
    real(c_double), value :: rho
    integer(c_int), value :: iflag
    real(c_double), intent(out) :: em(81, 81)
    real(8) :: xyz(27, 3), pmat(2, 8)
    integer :: inod(27, 1), ngp(1), mtel(1), fl
    call one_element(xt, xyz, inod, ngp, mtel, pmat, 1, rho)
    fl = iflag
    call ELEMENT_M_3D(xyz, inod, ngp, mtel, pmat, pml_param, 1, em, fl)
  end subroutine

  subroutine ref_element_en(xt, pml_param, iflag, en) bind(C, name="ref_element_en")
    real(c_double), intent(in) :: xt(3, 27), pml_param(6, 4)
    integer(c_int), value :: iflag
    real(c_double), intent(out) :: en(162, 162)
    real(8) :: xyz(27, 3), pmat(2, 8)
    integer :: inod(27, 1), ngp(1), mtel(1), fl
    call one_element(xt, xyz, inod, ngp, mtel, pmat, 2, 1.0d0)
    fl = iflag
    call ELEMENT_EN_3D(xyz, inod, ngp, mtel, pmat, pml_param, 1, en, fl)
  end subroutine

  subroutine ref_element_ea(xt, lam, mu, pml_param, iflag, ea) bind(C, name="ref_element_ea")
    real(c_double), intent(in) :: xt(3, 27), lam(27), mu(27), pml_param(6, 4)
    integer(c_int), value :: iflag
    real(c_double), intent(out) :: ea(81, 162)
    real(8) :: xyz(27, 3), pmat(2, 8)
    integer :: inod(27, 1), ngp(1), mtel(1), fl
    call one_element(xt, xyz, inod, ngp, mtel, pmat, 2, 1.0d0)
    fl = iflag
    call ELEMENT_EA_3D(xyz, inod, ngp, mtel, pmat, lam, mu, pml_param, 1, ea, fl)
  end subroutine

  subroutine ref_element_fs(xt, nface, efs) bind(C, name="ref_element_fs")
    real(c_double), intent(in) :: xt(3, 27)
    integer(c_int), value :: nface
    real(c_double), intent(out) :: efs(81)
    real(8) :: xyz(27, 3), pmat(2, 8)
    integer :: inod(27, 1), ngp(1), mtel(1), id_bc(1, 9)
    call one_element(xt, xyz, inod, ngp, mtel, pmat, 1, 1.0d0)
    id_bc = 0
    id_bc(1, 1) = nface
    call ELEMENT_FS_3D(xyz, inod, ngp, mtel, pmat, id_bc, 1, efs)
  end subroutine

  subroutine ref_element_g(xt, eu, ep, g_mu, g_lambda) bind(C, name="ref_element_g")
    real(c_double), intent(in) :: xt(3, 27), eu(81), ep(81)
    real(c_double), intent(out) :: g_mu(27), g_lambda(27)
    real(8) :: xyz(27, 3), pmat(2, 8)
    integer :: inod(27, 1), ngp(1), mtel(1)
    call one_element(xt, xyz, inod, ngp, mtel, pmat, 1, 1.0d0)
    call ELEMENT_g_3D(xyz, inod, ngp, 1, eu, ep, g_mu, g_lambda)
  end subroutine

  ! weighted regularisation: wreg = (w_top, w_bot, w_Length) when weighted /= 0
  subroutine set_weighting(weighted, wreg)
    integer(c_int), value :: weighted
    real(c_double), intent(in) :: wreg(3)
    if (weighted /= 0) then
      Weighted_Regularization = 'Yes'
      w_top = wreg(1)
      w_bot = wreg(2)
      w_Length = wreg(3)
    else
      Weighted_Regularization = 'No '
    end if
  end subroutine

  subroutine ref_element_r_mg(xt, weighted, wreg, er, emg) bind(C, name="ref_element_r_mg")
    real(c_double), intent(in) :: xt(3, 27), wreg(3)
    integer(c_int), value :: weighted
    real(c_double), intent(out) :: er(27, 27), emg(27)
    real(8) :: xyz(27, 3), pmat(2, 8)
    integer :: inod(27, 1), ngp(1), mtel(1)
    call one_element(xt, xyz, inod, ngp, mtel, pmat, 1, 1.0d0)
    call set_weighting(weighted, wreg)
    call ELEMENT_R_Mg_3D(xyz, inod, ngp, mtel, 1, er, emg)
  end subroutine

  subroutine ref_element_tv(xt, e_mu, e_lambda, weighted, wreg, reg_mu, reg_lambda, cost) bind(C, name="ref_element_tv")
    real(c_double), intent(in) :: xt(3, 27), e_mu(27), e_lambda(27), wreg(3)
    integer(c_int), value :: weighted
    real(c_double), intent(out) :: reg_mu(27), reg_lambda(27), cost(2)
    real(8) :: xyz(27, 3), pmat(2, 8), c_lam, c_mu
    integer :: inod(27, 1), ngp(1), mtel(1)
    call one_element(xt, xyz, inod, ngp, mtel, pmat, 1, 1.0d0)
    call set_weighting(weighted, wreg)
    call ELEMENT_TV_TN_Reg_3D(xyz, inod, ngp, mtel, 1, e_mu, e_lambda, reg_mu, reg_lambda, c_lam, c_mu)
    cost(1) = c_lam
    cost(2) = c_mu
  end subroutine

end module hr_ref_shim

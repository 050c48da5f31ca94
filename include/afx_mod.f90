! afx_mod.f90 -- Fortran iso_c_binding interface to libafx.so (include/afx.h).
!
! A Fortran caller of the reference (type(global_matrix) arrays, complex(8) vectors,
! 1-based index arrays) passes its arrays straight through: complex(8) has the interleaved
! (re, im) layout the C ABI expects, INU/INL are (0:n) offsets, IAU/IAL hold 1-based ids.
!
!   call afx_cocg13_single_f(...)   replaces   call COCG13_single(...)   (COCG13_single.f90:9-36)
!   call afx_solve_cg_f(...)        replaces   call solveCG(...)         (forward_bxyz_mpi.f90:70-94)
!   call afx_solvepardiso_f(...)    replaces   call solvePARDISO(...)    (solvePARDISO.f90:6-16)
!   afx_solve_pardiso_inv(...)      replaces   call solvePARDISOinv(...) (solvePARDISOinv.f90:5-23)
!   afx_solve_pardiso_jointinv(...) replaces   call solvePARDISOjointinv(...) (solvePARDISOjointinv.f90:5-26)
!   afx_solve(...) on a handle      keeps the system on the device across calls
!   afx_ixyh(...)                   computes   IXYHOUT's seafloor currents (n_ebfem_bxyz_mpi.f90:192-268)
module afx
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: AFX_OK = 0, AFX_ERR_MAXIT = -100, AFX_ERR_BADOPT = 101
  integer(c_int), parameter :: AFX_ERR_BREAKDOWN = -200, AFX_ERR_HIP = -300, AFX_ERR_ARG = -400
  integer(c_int), parameter :: AFX_PC_NONE = 0, AFX_PC_IC = 1, AFX_PC_SSOR = 2, AFX_PC_DIAG = 3
  integer(c_int), parameter :: AFX_FP64 = 0, AFX_MIXED = 1
  integer(c_int), parameter :: AFX_PC_SSOR_BJ = 4, AFX_PC_NODEBLK = 5, AFX_PC_GRAD = 256

  type, bind(C) :: afx_opts
    integer(c_int) :: precond
    real(c_double) :: sigma_diag, sigma, tol
    integer(c_int) :: maxit, nset, precision
    real(c_double) :: inner_tol
    integer(c_int) :: max_outer, chunk, check_true
  end type afx_opts

  type, bind(C) :: afx_stats
    integer(c_int) :: iter, error
    real(c_double) :: resid, true_resid
    integer(c_int) :: outer
  end type afx_stats

  interface
    subroutine afx_default_opts(o) bind(C, name="afx_default_opts")
      import :: afx_opts
      type(afx_opts), intent(out) :: o
    end subroutine

    integer(c_int) function afx_create(n, INU, IAU, INL, IAL, h) bind(C, name="afx_create")
      import :: c_int, c_ptr
      integer(c_int), value :: n
      integer(c_int), intent(in) :: INU(*), IAU(*), INL(*), IAL(*)
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function afx_create_from_n6line(nline, ntet, n6line, h) bind(C, name="afx_create_from_n6line")
      import :: c_int, c_ptr
      integer(c_int), value :: nline, ntet
      integer(c_int), intent(in) :: n6line(*)
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function afx_destroy(h) bind(C, name="afx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function afx_set_values(h, D, AU, AL) bind(C, name="afx_set_values")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: D(*), AU(*), AL(*)
    end function

    integer(c_int) function afx_precond_setup(h, kind, sigma_diag, sigma) bind(C, name="afx_precond_setup")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: kind
      real(c_double), value :: sigma_diag, sigma
    end function

    ! solvePARDISO(doftot, nsr, A, rhs, xout, ip): rhs(doftot,nsr), xout(doftot,nsr) column-major
    integer(c_int) function afx_solve(h, nrhs, B, ldb, X, ldx, opts, stats) bind(C, name="afx_solve")
      import :: c_int, c_ptr, c_double_complex, afx_opts, afx_stats
      type(c_ptr), value :: h
      integer(c_int), value :: nrhs, ldb, ldx
      complex(c_double_complex), intent(in) :: B(*)
      complex(c_double_complex), intent(inout) :: X(*)
      type(afx_opts), intent(in) :: opts
      type(afx_stats), intent(out) :: stats(*)
    end function

    ! several independent meshes in one launch: system q owns rows row_off(q+1)+1 .. row_off(q+2)
    ! (row_off(1) = 0, row_off(m+1) = n); dB, dX are device pointers (n complex(8) each), stats(m)
    integer(c_int) function afx_create_stacked(m, row_off, n, INU, IAU, INL, IAL, h) &
                                               bind(C, name="afx_create_stacked")
      import :: c_int, c_ptr
      integer(c_int), value :: m, n
      integer(c_int), intent(in) :: row_off(*), INU(*), IAU(*), INL(*), IAL(*)
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function afx_solve_stacked_device(h, dB, dX, opts, stats) bind(C, name="afx_solve_stacked_device")
      import :: c_int, c_ptr, afx_opts, afx_stats
      type(c_ptr), value :: h, dB, dX
      type(afx_opts), intent(in) :: opts
      type(afx_stats), intent(out) :: stats(*)
    end function

    integer(c_int) function afx_cocg13_single(NP, NPL, NPU, D, AL, INL, IAL, AU, INU, IAU, B, X, PRECOND, &
                                              SIGMA_DIAG, SIGMA, RESID, ITER, ERROR, NSET) &
                                              bind(C, name="afx_cocg13_single")
      import :: c_int, c_double, c_double_complex, c_char
      integer(c_int), value :: NP, NPL, NPU, NSET
      complex(c_double_complex), intent(inout) :: D(*), AL(*), AU(*), B(*), X(*)
      integer(c_int), intent(in) :: INL(*), IAL(*), INU(*), IAU(*)
      character(kind=c_char), intent(in) :: PRECOND(*)
      real(c_double), value :: SIGMA_DIAG, SIGMA
      real(c_double), intent(inout) :: RESID
      integer(c_int), intent(inout) :: ITER
      integer(c_int), intent(out) :: ERROR
    end function

    integer(c_int) function afx_cocg13_single_release() bind(C, name="afx_cocg13_single_release")
      import :: c_int
    end function

    integer(c_int) function afx_solve_cg(doftot, NPL, NPU, D, AL, INL, IAL, AU, INU, IAU, rhs, X) &
                                         bind(C, name="afx_solve_cg")
      import :: c_int, c_double_complex
      integer(c_int), value :: doftot, NPL, NPU
      complex(c_double_complex), intent(in) :: D(*), AL(*), AU(*), rhs(*)
      integer(c_int), intent(in) :: INL(*), IAL(*), INU(*), IAU(*)
      complex(c_double_complex), intent(out) :: X(*)
    end function

    ! edge end nodes line(2,nline) (g_line%line) and a one-byte Dirichlet mask per edge, for the
    ! nodal correction (AFX_PC_GRAD)
    integer(c_int) function afx_set_nodes(h, node, line, line_bc) bind(C, name="afx_set_nodes")
      import :: c_int, c_ptr, c_signed_char
      type(c_ptr), value :: h
      integer(c_int), value :: node
      integer(c_int), intent(in) :: line(*)
      integer(c_signed_char), intent(in) :: line_bc(*)
    end function

    ! solvePARDISO(doftot, nsr, A, rhs, xout, ip) on A's arrays (see afx_solvepardiso_f)
    integer(c_int) function afx_solve_pardiso(doftot, nsr, INU, IAU, INL, IAL, D, AU, AL, node, line, line_bc, &
                                              rhs, xout, tol) bind(C, name="afx_solve_pardiso")
      import :: c_int, c_double, c_double_complex, c_signed_char
      integer(c_int), value :: doftot, nsr, node
      integer(c_int), intent(in) :: INU(*), IAU(*), INL(*), IAL(*), line(*)
      complex(c_double_complex), intent(in) :: D(*), AU(*), AL(*), rhs(*)
      integer(c_signed_char), intent(in) :: line_bc(*)
      complex(c_double_complex), intent(out) :: xout(*)
      real(c_double), value :: tol
    end function

    ! solvePARDISOinv(doftot, nsr, A, rhs, xout, PTR, ut, iflag_comp, ip, np): PTR(c) and ut(c) as
    ! plain CRS arrays (afx.h); ut_cap entries per component
    integer(c_int) function afx_solve_pardiso_inv(doftot, nsr, INU, IAU, INL, IAL, D, AU, AL, node, line, line_bc, &
                                                  rhs, xout, nobs, iflag_comp, ptr_stack, ptr_base, ptr_item, ptr_val, &
                                                  ut_cap, ut_stack, ut_ntot, ut_item, ut_val, tol) &
                                                  bind(C, name="afx_solve_pardiso_inv")
      import :: c_int, c_double, c_double_complex, c_signed_char
      integer(c_int), value :: doftot, nsr, node, nobs, ut_cap
      integer(c_int), intent(in) :: INU(*), IAU(*), INL(*), IAL(*), line(*), iflag_comp(*)
      integer(c_int), intent(in) :: ptr_stack(*), ptr_base(*), ptr_item(*)
      real(c_double), intent(in) :: ptr_val(*)
      complex(c_double_complex), intent(in) :: D(*), AU(*), AL(*), rhs(*)
      integer(c_signed_char), intent(in) :: line_bc(*)
      complex(c_double_complex), intent(out) :: xout(*), ut_val(*)
      integer(c_int), intent(out) :: ut_stack(*), ut_ntot(*), ut_item(*)
      real(c_double), value :: tol
    end function

    ! solvePARDISOjointinv(ACT, MT, nline, nsr, A, rhs, rhs_mt, xout, xout_mt, PTR, PTR_mt, ut, ut_mt,
    ! iflag_comp, ip, np): act / mt as 0 or 1, PTR_mt(4) and ut_mt(4) in the layout of PTR and ut
    integer(c_int) function afx_solve_pardiso_jointinv(act, mt, doftot, nsr, INU, IAU, INL, IAL, D, AU, AL, node, &
                                                       line, line_bc, rhs, rhs_mt, xout, xout_mt, nobs, iflag_comp, &
                                                       ptr_stack, ptr_base, ptr_item, ptr_val, nobs_mt, ptr_mt_stack, &
                                                       ptr_mt_base, ptr_mt_item, ptr_mt_val, ut_cap, ut_stack, ut_ntot, &
                                                       ut_item, ut_val, ut_mt_cap, ut_mt_stack, ut_mt_ntot, ut_mt_item, &
                                                       ut_mt_val, tol) bind(C, name="afx_solve_pardiso_jointinv")
      import :: c_int, c_double, c_double_complex, c_signed_char
      integer(c_int), value :: act, mt, doftot, nsr, node, nobs, nobs_mt, ut_cap, ut_mt_cap
      integer(c_int), intent(in) :: INU(*), IAU(*), INL(*), IAL(*), line(*), iflag_comp(*)
      integer(c_int), intent(in) :: ptr_stack(*), ptr_base(*), ptr_item(*)
      integer(c_int), intent(in) :: ptr_mt_stack(*), ptr_mt_base(*), ptr_mt_item(*)
      real(c_double), intent(in) :: ptr_val(*), ptr_mt_val(*)
      complex(c_double_complex), intent(in) :: D(*), AU(*), AL(*), rhs(*), rhs_mt(*)
      integer(c_signed_char), intent(in) :: line_bc(*)
      complex(c_double_complex), intent(out) :: xout(*), xout_mt(*), ut_val(*), ut_mt_val(*)
      integer(c_int), intent(out) :: ut_stack(*), ut_ntot(*), ut_item(*), ut_mt_stack(*), ut_mt_ntot(*), ut_mt_item(*)
      real(c_double), value :: tol
    end function

    integer(c_int) function afx_solve_pardiso_release() bind(C, name="afx_solve_pardiso_release")
      import :: c_int
    end function

    ! IXYHOUT's arrays (n_ebfem_bxyz_mpi.f90:192-268) in the reference's own shapes:
    ! xyzobs(3,nobs), exy(2,nobs), vxyz(3,node), fxyz(3,node); coeff_vF as stack(0:nobs) from 0,
    ! item (1-based nodes), val. out(2,4,nobs) holds ixyh_e, ixyh_vF, ixyh, vh_obs for each point.
    integer(c_int) function afx_ixyh(nobs, xyzobs, exy, stack, item, val, node, vxyz, fxyz, sigma, out, h) &
                                     bind(C, name="afx_ixyh")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nobs, node
      real(c_double), intent(in) :: xyzobs(*), val(*), fxyz(*)
      complex(c_double_complex), intent(in) :: exy(*), vxyz(*)
      integer(c_int), intent(in) :: stack(*), item(*)
      real(c_double), value :: sigma
      complex(c_double_complex), intent(out) :: out(*)
      real(c_double), intent(out) :: h(*)
    end function
  end interface

contains

  ! The reference's COCG13_single call, argument for argument (character(*) PRECOND padded
  ! to the 20 characters the C side reads).
  subroutine afx_cocg13_single_f(NP, NPL, NPU, D, AL, INL, IAL, AU, INU, IAU, B, X, PRECOND, &
                                 SIGMA_DIAG, SIGMA, RESID, ITER, ERROR, NSET)
    integer(4), intent(in) :: NP, NPL, NPU, NSET
    complex(8), intent(inout) :: D(NP), AL(NPU), AU(NPU), B(NP), X(NP)
    integer(4), intent(in) :: INL(0:NP), IAL(NPL), INU(0:NP), IAU(NPU)
    character(len=*), intent(in) :: PRECOND
    real(8), intent(in) :: SIGMA_DIAG, SIGMA
    real(8), intent(inout) :: RESID
    integer(4), intent(inout) :: ITER
    integer(4), intent(out) :: ERROR
    character(kind=c_char, len=20) :: pc
    integer(c_int) :: rc
    pc = PRECOND
    rc = afx_cocg13_single(NP, NPL, NPU, D, AL, INL, IAL, AU, INU, IAU, B, X, pc, SIGMA_DIAG, SIGMA, &
                           RESID, ITER, ERROR, NSET)
    if (rc /= AFX_OK .and. rc /= AFX_ERR_MAXIT .and. rc /= AFX_ERR_BADOPT) ERROR = rc
  end subroutine afx_cocg13_single_f

  ! The reference's solvePARDISO call on A's arrays: rhs(doftot,nsr) and xout(doftot,nsr) as
  ! PARDISO takes them, line(2,nline) and line_bc(nline) (GENBCCSEM's logical mask) for the nodal
  ! correction, FP64 COCG to a 1e-12 true residual. ierr = 0, -100 (not reached) or AFX_ERR_*.
  subroutine afx_solvepardiso_f(doftot, nsr, D, AU, AL, INU, IAU, INL, IAL, node, line, line_bc, rhs, xout, ierr)
    integer(4), intent(in) :: doftot, nsr, node
    complex(8), intent(in) :: D(*), AU(*), AL(*)
    integer(4), intent(in) :: INU(0:doftot), IAU(*), INL(0:doftot), IAL(*), line(2,doftot)
    logical, intent(in) :: line_bc(doftot)
    complex(8), intent(in) :: rhs(doftot,nsr)
    complex(8), intent(out) :: xout(doftot,nsr)
    integer(4), intent(out) :: ierr
    integer(c_signed_char), allocatable :: bc(:)
    allocate(bc(doftot))
    bc = 0_c_signed_char
    where (line_bc) bc = 1_c_signed_char
    ierr = afx_solve_pardiso(doftot, nsr, INU, IAU, INL, IAL, D, AU, AL, node, line, bc, rhs, xout, 1.d-12)
    deallocate(bc)
  end subroutine afx_solvepardiso_f

end module afx

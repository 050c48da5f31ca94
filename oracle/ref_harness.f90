! Test-infrastructure driver around the reference's own Fortran routines (built by
! oracle/build_ref.sh with flang from the sources under /root/reference; never shipped).
!
! Reads a driver file on stdin, runs, in the reference's own order:
!   READMESH_TOTAL, READLINE, READCOND, set_iccg_var7_dofn, INITIALIZE, GENMAT, COPY_UL_ICCG12,
!   GENBCCSEM, SET_BC_ICCG, COCG13_single, FINDELEMENT0 + EDGEBASISFUN + FACEBASISFUN per site,
!   MKLINE_V2 + MKN6 (from the mesh) and readvxyz/readfxyz,
! and writes every intermediate array as raw little-endian stream files into the current
! directory (<name>.bin), for the golden-vector generator in tests/golden/.
program ref_harness
use mesh_type
use line_type
use param
use iccg_var_takuto
use fem_util
use fem_edge_util
use readFvxyz
implicit none
type(mesh)          :: g_mesh
type(line_info)     :: g_line, l2
type(param_forward) :: g_param
type(param_source)  :: sparam
type(param_cond)    :: g_cond
type(global_matrix) :: A
character(70)       :: meshfile, vfile, ffile, condfile
character(50)       :: linefile
character(20)       :: precond
real(8)             :: omega, zmin, zmax, xout, yout, resid, sd, sg, sigma_air, v, g_tol
integer(4)          :: nodes, maxit, nset, nobs, i, j, ierr, iter, nline, ntet, iele, dorun
integer(4)          :: nseq, kseq, ios
real(8)             :: bfac
character(8)        :: tag
integer(4), allocatable :: table_dof(:,:)
complex(8), allocatable :: b_vec(:,:), bs(:,:), Avalue_bc(:), X(:), B(:), vxyz(:,:)
real(8),    allocatable :: fxyz(:,:)
logical,    allocatable :: line_bc(:)
integer(4), allocatable :: ibc(:)
real(8)             :: x3(3), a4(4), w(3,6), wfe(3,6), elen(6), elm_xyz(3,4)
integer(4), allocatable :: iobs(:)
real(8),    allocatable :: wobs(:,:,:), wfobs(:,:,:), aobs(:,:)

read(*,'(a)') meshfile
read(*,'(a)') linefile
read(*,'(a)') condfile
read(*,'(a)') vfile
read(*,'(a)') ffile
read(*,*) nodes
read(*,*) omega
read(*,*) zmin, zmax, xout, yout
read(*,*) sigma_air
read(*,'(a)') precond
read(*,*) resid, maxit, nset, sd, sg, dorun
g_tol = resid
read(*,*) nobs

!# inputs through the reference readers
call READMESH_TOTAL(g_mesh, meshfile)
call READLINE(linefile, g_line)
g_cond%condflag  = 1
g_cond%condfile  = condfile
g_cond%sigma_air = sigma_air
call READCOND(g_cond)
g_cond%ntet   = g_mesh%ntet
g_cond%nphys1 = g_mesh%ntet - g_cond%nphys2
g_param%nodes     = nodes
g_param%vxyz_file = vfile
g_param%fxyz_file = ffile
sparam%nsource = 1
nline = g_line%nline
ntet  = g_line%ntet

allocate(vxyz(3,g_mesh%node), fxyz(3,g_mesh%node))
call readfxyz(nodes, g_mesh%node, fxyz, ffile)
call readvxyz(nodes, g_mesh%node, vxyz, vfile)
call dumpc2('vxyz', vxyz, 3, g_mesh%node)
call dumpr2('fxyz', fxyz, 3, g_mesh%node)

!# pattern (n_ebfem_bxyz_mpi.f90:104-107)
allocate(table_dof(nline,1))
do i = 1, nline
  table_dof(i,1) = i
end do
call set_iccg_var7_dofn(1, 6, nline, nline, nline, ntet, g_line%n6line, table_dof, A, 0)
call dumpi('INU', A%INU, nline + 1)
call dumpi('INL', A%INL, nline + 1)
call dumpi('IAU', A%IAU, A%iau_tot)
call dumpi('IAL', A%IAL, A%ial_tot)

!# forward_bxyz steps (forward_bxyz_mpi.f90:43-58)
allocate(b_vec(nline,1), bs(nline,1))
call INITIALIZE(A, b_vec, bs, nline, 1)
call GENMAT(g_mesh, g_line, A, b_vec, omega, sparam, g_param, g_cond)
call dumpc('D_genmat', A%D, nline)
call dumpc('AU_genmat', A%AU, A%iau_tot)
call dumpc('b_genmat', b_vec(:,1), nline)
call COPY_UL_ICCG12(A, 0)
call dumpc('AL_copy', A%AL, A%ial_tot)
allocate(Avalue_bc(nline), line_bc(nline), ibc(nline))
call GENBCCSEM(zmin, zmax, xout, yout, g_line, g_mesh, Avalue_bc, line_bc)
do i = 1, nline
  ibc(i) = 0
  if (line_bc(i)) ibc(i) = 1
end do
call dumpi('line_bc', ibc, nline)
call SET_BC_ICCG(A, nline, 1, b_vec, Avalue_bc, line_bc, 0)
call dumpc('D', A%D, nline)
call dumpc('AU', A%AU, A%iau_tot)
call dumpc('AL', A%AL, A%ial_tot)
call dumpc('b', b_vec(:,1), nline)

!# COCG13_single (COCG13_single.f90:9-341)
if (dorun .eq. 1) then
  allocate(X(nline), B(nline))
  X(:) = (0.d0, 0.d0)
  B(:) = b_vec(:,1)
  ierr = 0
  iter = maxit
  call COCG13_single(nline, A%ial_tot, A%iau_tot, A%D, A%AL, A%INL, A%IAL, A%AU, A%INU, A%IAU, &
                   & B, X, precond, sd, sg, resid, iter, ierr, nset)
  call dumpc('x', X, nline)
  call dumpr('resid', (/resid/), 1)
  call dumpi('iter', (/iter, ierr/), 2)
end if

!# observation-site basis functions (PREPOBSCOEFF's calls, n_ebfem_bxyz_mpi.f90:684-700)
allocate(iobs(max(nobs,1)), wobs(3,6,max(nobs,1)), wfobs(3,6,max(nobs,1)), aobs(4,max(nobs,1)))
do i = 1, nobs
  read(*,*) x3(1:3)
  call FINDELEMENT0(x3, g_mesh, iele, a4)
  do j = 1, 4
    elm_xyz(1:3,j) = g_mesh%xyz(1:3, g_mesh%n4(iele,j))
  end do
  call EDGEBASISFUN(elm_xyz, x3, w, elen, v)
  call FACEBASISFUN(elm_xyz, x3, wfe, v)
  iobs(i) = iele
  wobs(:,:,i) = w
  wfobs(:,:,i) = wfe
  aobs(:,i) = a4
end do
if (nobs .ge. 1) then
  call dumpi('obs_iele', iobs, nobs)
  call dumpr('obs_w', wobs, 18*nobs)
  call dumpr('obs_wf', wfobs, 18*nobs)
  call dumpr('obs_a', aobs, 4*nobs)
end if

!# edge numbering from the mesh (tmtgem/mesh/mkline.f90:19-20)
call MKLINE_V2(l2, g_mesh%node, g_mesh%ntet, 4, g_mesh%n4)
call MKN6(l2, g_mesh%node, g_mesh%ntet, g_mesh%n4)
call dumpi('mk_line', l2%line, 2*l2%nline)
call dumpi('mk_n6line', l2%n6line, 6*g_mesh%ntet)

!# optional sequence of further COCG13_single calls on the same arrays (its SAVEd DD, SCALE and
!# IFLAG carry over, and D/AU/AL keep what earlier calls did to them): per call one line
!# "nset precond bfac", B = bfac * b, X = 0, the driver's tolerance and iteration cap
nseq = 0
read(*,*,iostat=ios) nseq
if (ios .ne. 0) nseq = 0
if (nseq .gt. 0 .and. .not. allocated(X)) allocate(X(nline), B(nline))
do kseq = 1, nseq
  read(*,*) nset, precond, bfac
  X(:) = (0.d0, 0.d0)
  B(:) = b_vec(:,1) * bfac
  ierr = 0
  iter = maxit
  resid = g_tol
  call COCG13_single(nline, A%ial_tot, A%iau_tot, A%D, A%AL, A%INL, A%IAL, A%AU, A%INU, A%IAU, &
                   & B, X, precond, sd, sg, resid, iter, ierr, nset)
  write(tag,'(i0)') kseq
  call dumpc('seq_x'//trim(tag), X, nline)
  call dumpc('seq_b'//trim(tag), B, nline)
  call dumpr('seq_resid'//trim(tag), (/resid/), 1)
  call dumpi('seq_iter'//trim(tag), (/iter, ierr/), 2)
end do

contains
  subroutine dumpi(name, a, n)
    character(*), intent(in) :: name
    integer(4),   intent(in) :: n
    integer(4),   intent(in) :: a(n)
    open(91, file=trim(name)//'.bin', access='stream', form='unformatted', status='replace')
    write(91) a
    close(91)
  end subroutine
  subroutine dumpr(name, a, n)
    character(*), intent(in) :: name
    integer(4),   intent(in) :: n
    real(8),      intent(in) :: a(n)
    open(91, file=trim(name)//'.bin', access='stream', form='unformatted', status='replace')
    write(91) a
    close(91)
  end subroutine
  subroutine dumpr2(name, a, m, n)
    character(*), intent(in) :: name
    integer(4),   intent(in) :: m, n
    real(8),      intent(in) :: a(m,n)
    open(91, file=trim(name)//'.bin', access='stream', form='unformatted', status='replace')
    write(91) a
    close(91)
  end subroutine
  subroutine dumpc(name, a, n)
    character(*), intent(in) :: name
    integer(4),   intent(in) :: n
    complex(8),   intent(in) :: a(n)
    open(91, file=trim(name)//'.bin', access='stream', form='unformatted', status='replace')
    write(91) a
    close(91)
  end subroutine
  subroutine dumpc2(name, a, m, n)
    character(*), intent(in) :: name
    integer(4),   intent(in) :: m, n
    complex(8),   intent(in) :: a(m,n)
    open(91, file=trim(name)//'.bin', access='stream', form='unformatted', status='replace')
    write(91) a
    close(91)
  end subroutine
end program ref_harness

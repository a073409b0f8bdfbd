! Golden-vector harness for the reference TABI-PB build (test infrastructure only).
!
! Links the reference's own compiled objects (src/treecode3d_pb.f, src/main.f90,
! src/readin.f90, src/dgmres_dep.f, src/dgetrf_dep.f, src/var_modules.f90) and
! drives the same steps as the reference program (src/main.f90:39-159), dumping
! the intermediate vectors the reference never writes out:
!   geom.bin   nface, tree-ordered centroids/normals/areas, orderarr, bvct
!   leaves.bin leaf (ibeg,iend) pairs in tree traversal order
!   mv.bin     x (tree order) and y = MATVECmain(x)       (src/main.f90:467-488)
!   ms.bin     r and z = MSOLVEdric(r)                     (src/main.f90:511-529)
!   solve.bin  iter, ierr, err, xvct (tree order), soleng, couleng
!               (DGMRES call src/main.f90:136-137, energy src/main.f90:144-157)
! Run in a directory holding usrdata.in and test_proteins/<fname>.{pqr,vert,face};
! readin deletes the .vert/.face files it reads (src/readin.f90:254-256).
! Optional argument: "nosolve" skips the GMRES solve and the energy.
program ref_harness
  use molecule
  use comdata
  use bicg
  use treecode
  use treecode3d_procedures
  implicit none
  character(100) :: fhead, arg
  real*8 :: bulk_strength, pi, one_over_4pi, kappa2, ptl, soleng, couleng, dist, err, tol
  real*8 :: r0(3), diff(3)
  integer :: meof, iprcd, ndim, maxl, lrgw, jscal, jpre, nrmax, ligw, mlwk, nms, lenw, leniw
  integer :: itol, itmax, iter, ierr, i, iatm, jatm, nl, u
  integer, allocatable :: lb(:), le(:)
  real*8, allocatable :: xin(:), yout(:)
  logical :: dosolve
  external MATVECmain, MSOLVEdric
  common // pi, one_over_4pi

  dosolve = .true.
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    if (trim(arg) == 'nosolve') dosolve = .false.
  end if

  ! usrdata.in: ten positional "key value" lines, same order as src/main.f90:39-50
  open(101, file="usrdata.in")
  read(101,*,iostat=meof) fhead, fname
  read(101,*,iostat=meof) fhead, den
  read(101,*,iostat=meof) fhead, eps0
  read(101,*,iostat=meof) fhead, eps1
  read(101,*,iostat=meof) fhead, bulk_strength
  read(101,*,iostat=meof) fhead, order
  read(101,*,iostat=meof) fhead, maxparnode
  read(101,*,iostat=meof) fhead, theta
  read(101,*,iostat=meof) fhead, iprcd
  read(101,*,iostat=meof) fhead, isurf_type
  close(101)

  ! constants exactly as src/main.f90:55-65 (note the default-real literal in kappa2)
  pi = acos(-1.d0)
  one_over_4pi = 0.25d0/pi
  kappa2 = 8.430325455*bulk_strength/eps1
  kappa = sqrt(kappa2)
  para = 332.0716d0
  eps = eps1/eps0

  call readin
  call form_matrix
  call treecode_initialization

  ndim = 2*nface
  open(newunit=u, file='geom.bin', access='stream', form='unformatted', status='replace')
  write(u) nface, kappa, eps
  write(u) tr_xyz, tr_q, tr_area, orderarr, bvct
  close(u)

  allocate(lb(nface), le(nface))
  nl = 0
  call walk(troot)
  open(newunit=u, file='leaves.bin', access='stream', form='unformatted', status='replace')
  write(u) nl, lb(1:nl), le(1:nl)
  close(u)

  ! a fixed, smooth, sign-changing test vector in tree order
  allocate(xin(ndim), yout(ndim))
  do i = 1, ndim
    xin(i) = sin(0.37d0*i) + 0.5d0*cos(1.3d0*i) + 0.25d0
  end do
  call MATVECmain(ndim, xin, yout)
  open(newunit=u, file='mv.bin', access='stream', form='unformatted', status='replace')
  write(u) ndim, xin, yout
  close(u)

  call MSOLVEdric(ndim, xin, yout, 0, 0, 0, 0.d0, 0, 0.d0, 0)
  open(newunit=u, file='ms.bin', access='stream', form='unformatted', status='replace')
  write(u) ndim, xin, yout
  close(u)

  if (.not. dosolve) stop

  ! GMRES set-up exactly as src/main.f90:55-58,93-115
  itol = 2; itmax = 1000000; tol = 1.d-4
  maxl = 10; lrgw = 1 + ndim*(maxl+6) + maxl*(maxl+3); jscal = 0; jpre = -1
  nrmax = 10; ligw = 20; mlwk = ligw; nms = 0; lenw = 10; leniw = 10
  allocate(sb(ndim), sx(ndim), rgwk(lrgw), igwk(ligw), rwork(lenw), iwork(leniw))
  rgwk = 0.d0; igwk = 0; rwork = 0.d0; iwork = 0
  igwk(1:7) = (/maxl, maxl, jscal, jpre, nrmax, mlwk, nms/)
  xvct = 0.d0
  call DGMRES(ndim, bvct, xvct, MATVECmain, MSOLVEdric, itol, tol, itmax, &
              iter, err, ierr, 0, sb, sx, rgwk, lrgw, igwk, ligw, rwork, iwork)

  ! energies as src/main.f90:144-157
  soleng = 0.d0; couleng = 0.d0
  do iatm = 1, nchr
    r0 = chrpos(:, iatm)
    call potential_molecule(r0, ptl, 0)
    soleng = soleng + atmchr(iatm)*ptl
    do jatm = iatm+1, nchr
      diff = r0 - chrpos(:, jatm)
      dist = sqrt(dot_product(diff, diff))
      couleng = couleng + 1/eps0/dist*atmchr(iatm)*atmchr(jatm)
    end do
  end do
  soleng = soleng*0.5d0*para*4*pi
  couleng = couleng*para
  open(newunit=u, file='solve.bin', access='stream', form='unformatted', status='replace')
  write(u) iter, ierr, err, soleng, couleng, xvct
  close(u)
  print *, 'harness: iter=', iter, ' ierr=', ierr, ' soleng=', soleng

contains

  recursive subroutine walk(p)
    type(tnode), pointer :: p
    integer :: k
    if (p%num_children == 0) then
      nl = nl + 1
      lb(nl) = p%ibeg
      le(nl) = p%iend
    else
      do k = 1, p%num_children
        call walk(p%child(k)%p_to_tnode)
      end do
    end if
  end subroutine walk

end program ref_harness

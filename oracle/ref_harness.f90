! Golden-vector and timing harness for the reference TABI-PB build (test
! infrastructure only).
!
! Links the reference's own compiled objects (src/treecode3d_pb.f, src/main.f90,
! src/readin.f90, src/dgmres_dep.f, src/dgetrf_dep.f, src/var_modules.f90) and
! drives the same steps as the reference program (src/main.f90:39-161), dumping
! the intermediate vectors the reference never writes out:
!   geom.bin   nface, tree-ordered centroids/normals/areas, orderarr, bvct
!   leaves.bin leaf (ibeg,iend) pairs in tree traversal order
!   mv.bin     x (tree order) and y = MATVECmain(x)       (src/main.f90:467-488)
!   ms.bin     r and z = MSOLVE(r): MSOLVEdric (src/main.f90:511-529) when
!              iprcd = 1, else MSOLVEprec (src/main.f90:532-542), as the
!              program binds them (src/main.f90:118-123)
!   solve.bin  iter, ierr, err, soleng, couleng, xvct (tree order)
!              (DGMRES call src/main.f90:136-137, energy src/main.f90:144-157)
!   surface_potential.dat  written by the reference's own output_potential
!              (src/main.f90:226-326) with option "potential"
! Run in a directory holding usrdata.in and the surface files readin expects
! (src/readin.f90:90-246); readin deletes the MSMS .vert/.face files it reads.
!
! Options (any order):
!   nosolve     skip the GMRES solve and the energy
!   nomsolve    skip the MSOLVE dump (large surfaces: the solve refactors every leaf)
!   tol=X       GMRES tolerance (default 1.d-4 as src/main.f90:58)
!   nrmax=K     maximum number of restarts (default 10 as src/main.f90:97)
!   potential   call output_potential after the energy
!   time        print the wall time of every phase
!   sample=K    bounded timing sample instead of dumps: times the reference's own
!               compiled routines on every K-th matvec target and every L-th leaf
!   lstride=L   (default 8) and prints the measured pieces (subroutine time_sample)
program ref_harness
  use molecule
  use comdata
  use bicg
  use treecode
  use treecode3d_procedures
  implicit none
  character(100) :: fhead, arg
  real*8 :: bulk_strength, pi, one_over_4pi, kappa2, ptl, soleng, couleng, dist, err, tol
  real*8 :: r0(3), diff(3), tw0
  integer :: meof, iprcd, ndim, maxl, lrgw, jscal, jpre, nrmax, ligw, mlwk, nms, lenw, leniw
  integer :: itol, itmax, iter, ierr, i, iatm, jatm, nl, u, nsample, ia, lstride
  integer, allocatable :: lb(:), le(:)
  real*8, allocatable :: xin(:), yout(:)
  logical :: dosolve, domsolve, dopot, dotime
  external MATVECmain, MSOLVEdric, MSOLVEprec
  common // pi, one_over_4pi

  dosolve = .true.; domsolve = .true.; dopot = .false.; dotime = .false.
  tol = 1.d-4; nrmax = 10; nsample = 0; lstride = 8
  do ia = 1, command_argument_count()
    call get_command_argument(ia, arg)
    if (trim(arg) == 'nosolve') dosolve = .false.
    if (trim(arg) == 'nomsolve') domsolve = .false.
    if (trim(arg) == 'potential') dopot = .true.
    if (trim(arg) == 'time') dotime = .true.
    if (arg(1:4) == 'tol=') read(arg(5:), *) tol
    if (arg(1:6) == 'nrmax=') read(arg(7:), *) nrmax
    if (arg(1:7) == 'sample=') read(arg(8:), *) nsample
    if (arg(1:8) == 'lstride=') read(arg(9:), *) lstride
  end do

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

  if (nsample > 0) then
    call time_sample(nsample, lstride)
    stop
  end if

  tw0 = wall()
  call readin
  call phase('readin')
  call form_matrix
  call phase('form_matrix')
  call treecode_initialization
  call phase('treecode_initialization')

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
  call phase('MATVECmain')
  open(newunit=u, file='mv.bin', access='stream', form='unformatted', status='replace')
  write(u) ndim, xin, yout
  close(u)

  if (domsolve) then
    if (iprcd == 1) then
      call MSOLVEdric(ndim, xin, yout, 0, 0, 0, 0.d0, 0, 0.d0, 0)
    else
      call MSOLVEprec(ndim, xin, yout, 0, 0, 0, 0.d0, 0, 0.d0, 0)
    end if
    call phase('MSOLVE')
    open(newunit=u, file='ms.bin', access='stream', form='unformatted', status='replace')
    write(u) ndim, xin, yout
    close(u)
  end if

  if (.not. dosolve) stop

  ! GMRES set-up exactly as src/main.f90:55-58,93-115 (tol and NRMAX as given)
  itol = 2; itmax = 1000000
  maxl = 10; lrgw = 1 + ndim*(maxl+6) + maxl*(maxl+3); jscal = 0; jpre = -1
  ligw = 20; mlwk = ligw; nms = 0; lenw = 10; leniw = 10
  allocate(sb(ndim), sx(ndim), rgwk(lrgw), igwk(ligw), rwork(lenw), iwork(leniw))
  rgwk = 0.d0; igwk = 0; rwork = 0.d0; iwork = 0
  igwk(1:7) = (/maxl, maxl, jscal, jpre, nrmax, mlwk, nms/)
  xvct = 0.d0
  if (iprcd == 1) then
    call DGMRES(ndim, bvct, xvct, MATVECmain, MSOLVEdric, itol, tol, itmax, &
                iter, err, ierr, 0, sb, sx, rgwk, lrgw, igwk, ligw, rwork, iwork)
  else
    call DGMRES(ndim, bvct, xvct, MATVECmain, MSOLVEprec, itol, tol, itmax, &
                iter, err, ierr, 0, sb, sx, rgwk, lrgw, igwk, ligw, rwork, iwork)
  end if
  call phase('DGMRES')

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
  call phase('energy')
  open(newunit=u, file='solve.bin', access='stream', form='unformatted', status='replace')
  write(u) iter, ierr, err, soleng, couleng, xvct
  close(u)
  print *, 'harness: iter=', iter, ' ierr=', ierr, ' soleng=', soleng
  if (dopot) then
    call output_potential
    call phase('output_potential')
  end if

contains

  real*8 function wall()
    integer*8 :: cnt, rate
    call system_clock(cnt, rate)
    wall = dble(cnt)/dble(rate)
  end function wall

  subroutine phase(name)
    character(*) :: name
    real*8 :: t
    t = wall()
    if (dotime) then
      write(*,'(a,a,a,f12.4)') 'harness-time ', name, ' s=', t - tw0
      flush(6)
    end if
    tw0 = t
  end subroutine phase

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

  ! Bounded timing sample of one solve, all arithmetic in the reference's own
  ! compiled routines except the leaf-block fill of the preconditioner (below).
  ! Printed lines "harness-sample <piece> <seconds> <units measured> <units total>":
  !   readin       readin (file parsing and surface_area culling), in full
  !   form_geom    form_matrix with no charges (centroids, normals, areas), in full
  !   form_rhs     the right-hand side loop of form_matrix (src/main.f90:597-607):
  !                form_matrix with m charges minus the geometry-only call
  !   tree         treecode_initialization (src/main.f90:329-464), in full
  !   pb_kernel    pb_kernel (src/main.f90:664-694), in full, once per matvec
  !   matvec_first the per-target body of TREE_COMPP_PB (src/treecode3d_pb.f:
  !                1158-1185, COMPP_TREE) for every k-th target, computing the
  !                lazy node moments on the way
  !   matvec       the same targets again with the moments in place
  !   leaf_fill    LEAFMATVECpara's block fill (src/treecode3d_pb.f:1446-1528)
  !                for every l-th leaf; units are sum (2n)^2
  !   leaf_lu      dgetrf + dgetrs of those blocks (src/treecode3d_pb.f:1530-1532);
  !                units are sum (2n)^3
  !   energy       potential_molecule (src/main.f90:614-657) for m charges
  subroutine time_sample(kstride, lstride)
    integer, intent(in) :: kstride, lstride
    integer :: m, nq, ii, kk0, nlv, ilf, nrow, nrow2, ib, inc, nfull, ipass
    integer, allocatable :: ipiv(:)
    real*8 :: t0, t_geom, t_fm, tempx, temp_area, peng(2), tempq(16,2), pre1, pre2
    real*8 :: u_fill_s, u_fill_t, u_lu_s, u_lu_t, t_fill, t_lu
    real*8, allocatable :: xv(:), yv(:), mat(:), r2s(:)

    t0 = wall()
    call readin
    write(*,'(a,f12.5,2i12)') 'harness-sample readin ', wall() - t0, nface, nface
    flush(6)
    nq = nchr
    ! geometry only
    nchr = 0
    t0 = wall()
    call form_matrix
    t_geom = wall() - t0
    write(*,'(a,f12.5,2i12)') 'harness-sample form_geom ', t_geom, nface, nface
    deallocate(tr_xyz, tr_q, bvct, xvct, tchg, schg, tr_area)
    ! geometry plus m charges
    m = max(1, min(nq, int(4.d8/dble(max(nface, 1)))))
    nchr = m
    t0 = wall()
    call form_matrix
    t_fm = wall() - t0
    nchr = nq
    write(*,'(a,f12.5,2i12)') 'harness-sample form_rhs ', max(t_fm - t_geom, 0.d0), m, nq
    flush(6)
    t0 = wall()
    call treecode_initialization
    write(*,'(a,f12.5,2i12)') 'harness-sample tree ', wall() - t0, nface, nface
    flush(6)

    nfull = 2*nface
    allocate(xv(nfull), yv(nfull))
    do ii = 1, nfull
      xv(ii) = sin(0.37d0*ii) + 0.5d0*cos(1.3d0*ii) + 0.25d0
    end do
    t0 = wall()
    call pb_kernel(xv)
    write(*,'(a,f12.5,2i12)') 'harness-sample pb_kernel ', wall() - t0, nface, nface
    ! the per-target loop body of TREE_COMPP_PB for every kstride-th target
    pre1 = 0.5d0*(1.d0 + eps)
    pre2 = 0.5d0*(1.d0 + 1.d0/eps)
    ! pass 1 computes the lazy moments of every node the sample accepts (about
    ! all of them) besides the traversal; pass 2 repeats the same targets with
    ! the moments in place, so it times the per-target work alone
    do ipass = 1, 2
    kk0 = 0
    t0 = wall()
    do ii = 1, numpars, kstride
      tarpos(1) = x(ii); tarpos(2) = y(ii); tarpos(3) = z(ii)
      tarq = tr_q(:, ii)
      tempq = tchg(ii, :, :)
      peng_old(1) = xv(ii)
      peng_old(2) = xv(ii + numpars)
      peng = 0.d0
      tempx = x(ii)
      temp_area = tr_area(ii)
      x(ii) = x(ii) + 100.123456789d0
      tr_area(ii) = 0.d0
      call COMPP_TREE(troot, peng, x, y, z, schg, xv, kappa, theta, numpars, kk, eps, &
                      tempq, der_cof)
      yv(ii) = pre1*peng_old(1) - peng(1)
      yv(numpars + ii) = pre2*peng_old(2) - peng(2)
      x(ii) = tempx
      tr_area(ii) = temp_area
      kk0 = kk0 + 1
    end do
    if (ipass == 1) then
      write(*,'(a,f12.5,2i12)') 'harness-sample matvec_first ', wall() - t0, kk0, numpars
    else
      write(*,'(a,f12.5,2i12)') 'harness-sample matvec ', wall() - t0, kk0, numpars
    end if
    flush(6)
    end do
    call REMOVE_MMT(troot)

    ! leaves: every kstride/8-th leaf (at least one in 64), sizes weighted as the
    ! work scales
    allocate(lb(nface), le(nface))
    nl = 0
    call walk(troot)
    nlv = lstride
    u_fill_s = 0; u_fill_t = 0; u_lu_s = 0; u_lu_t = 0; t_fill = 0; t_lu = 0
    do ilf = 1, nl
      nrow2 = 2*(le(ilf) - lb(ilf) + 1)
      u_fill_t = u_fill_t + dble(nrow2)**2
      u_lu_t = u_lu_t + dble(nrow2)**3
    end do
    do ilf = 1, nl, nlv
      ib = lb(ilf)
      nrow = le(ilf) - lb(ilf) + 1
      nrow2 = 2*nrow
      allocate(mat(nrow2*nrow2), r2s(nrow2), ipiv(nrow2))
      t0 = wall()
      call fill_block(ib, nrow, mat, pre1, pre2)
      t_fill = t_fill + (wall() - t0)
      r2s(1:nrow) = xv(ib:ib + nrow - 1)
      r2s(nrow + 1:nrow2) = xv(ib + nface:ib + nface + nrow - 1)
      t0 = wall()
      call dgetrf(nrow2, nrow2, mat, nrow2, ipiv, inc)
      call dgetrs('N', nrow2, 1, mat, nrow2, ipiv, r2s, nrow2, inc)
      t_lu = t_lu + (wall() - t0)
      u_fill_s = u_fill_s + dble(nrow2)**2
      u_lu_s = u_lu_s + dble(nrow2)**3
      deallocate(mat, r2s, ipiv)
    end do
    write(*,'(a,f12.5,2es16.8)') 'harness-sample leaf_fill ', t_fill, u_fill_s, u_fill_t
    write(*,'(a,f12.5,2es16.8)') 'harness-sample leaf_lu ', t_lu, u_lu_s, u_lu_t
    write(*,'(a,2i12)') 'harness-sample leaves ', (nl + nlv - 1)/nlv, nl
    flush(6)

    ! energy: potential_molecule for m charges (xvct holds the test vector)
    xvct = xv
    m = max(1, min(nq, int(2.d8/dble(max(nface, 1)))))
    t0 = wall()
    do iatm = 1, m
      r0 = chrpos(:, iatm)
      call potential_molecule(r0, ptl, 0)
    end do
    write(*,'(a,f12.5,2i12)') 'harness-sample energy ', wall() - t0, m, nq
    flush(6)
  end subroutine time_sample

  ! The dense 2n x 2n block of one leaf, column-major, as LEAFMATVECpara fills it
  ! (src/treecode3d_pb.f:1446-1528): identity coefficients on the diagonal and
  ! -L * area_j off the diagonal with the reference's index placement.
  subroutine fill_block(ib, nrow, mat, pre1, pre2)
    integer, intent(in) :: ib, nrow
    real*8, intent(out) :: mat(*)
    real*8, intent(in) :: pre1, pre2
    integer :: i, j, nrow2, m1, m2, m3, m4
    real*8 :: t(3), c(3), r(3), v(3), rs, g0, gk, krs, ekrs, ct, ct0, tp1, tp2
    real*8 :: g10, g20, g1, g2, g3, g4
    nrow2 = 2*nrow
    mat(1:nrow2*nrow2) = 0.d0
    do i = ib, ib + nrow - 1
      t = (/x(i), y(i), z(i)/)
      c = tr_q(:, i)
      do j = ib, ib + nrow - 1
        if (j == i) cycle
        r = (/x(j), y(j), z(j)/)
        v = tr_q(:, j)
        rs = sqrt(dot_product(r - t, r - t))
        g0 = one_over_4pi/rs
        krs = kappa*rs
        ekrs = exp(-krs)
        gk = ekrs*g0
        ct = dot_product(v, r - t)/rs
        ct0 = dot_product(c, r - t)/rs
        tp1 = g0/rs
        tp2 = (1.d0 + krs)*ekrs
        g10 = ct0*tp1
        g20 = tp2*g10
        g1 = ct*tp1
        g2 = tp2*g1
        g3 = (dot_product(c, v) - 3.d0*ct0*ct)/rs*tp1
        g4 = tp2*g3 - kappa**2*ct0*ct*gk
        m1 = (j - ib)*nrow2 + i - ib + 1
        m2 = (j - ib)*nrow2 + i - ib + nrow + 1
        m3 = (j - ib + nrow)*nrow2 + i - ib + 1
        m4 = (j - ib + nrow)*nrow2 + i - ib + nrow + 1
        mat(m1) = -(g1 - eps*g2)*tr_area(j)
        mat(m2) = -(g0 - gk)*tr_area(j)
        mat(m3) = -(g4 - g3)*tr_area(j)
        mat(m4) = -(g10 - g20/eps)*tr_area(j)
      end do
      mat((i - ib)*(nrow2 + 1) + 1) = pre1
      mat((i - ib + nrow)*(nrow2 + 1) + 1) = pre2
    end do
  end subroutine fill_block

end program ref_harness

! ISO_C_BINDING interface to libtabi_hip.so (include/tabi_hip.h).
!
! Self-contained: depends on no module of the reference.  A Fortran driver
! calls tabi_hip_setup once with its face geometry and then hands
! MATVECgpu / MSOLVEgpu to SLATEC DGMRES in place of MATVECmain / MSOLVEdric
! (src/main.f90:116-123); both keep the argument lists DGMRES uses
! (CALL MATVEC(N, X, Y), CALL MSOLVE(N, R, Z, RPAR, IPAR); src/dgmres_dep.f:756,776,1736,1741).
module tabi_hip
  use iso_c_binding
  implicit none
  private
  public :: tabi_params_t, tabi_hip_setup, tabi_hip_release, tabi_hip_check
  public :: c_tabi_init, c_tabi_matvec, c_tabi_msolve, c_tabi_finalize
  public :: c_tabi_rhs, c_tabi_energy, c_tabi_vertex_potential

  type, bind(C) :: tabi_params_t
    real(c_double) :: eps, kappa, theta
    integer(c_int32_t) :: order, maxparnode, precision, flags
  end type tabi_params_t

  interface
    integer(c_int) function c_tabi_init(prm, n, xyz, nrm, area, orderarr, nleaf) &
        bind(C, name='tabi_init')
      import :: c_int, c_int32_t, c_double, tabi_params_t
      type(tabi_params_t), intent(in) :: prm
      integer(c_int32_t), value :: n
      real(c_double), intent(in) :: xyz(*), nrm(*), area(*)
      integer(c_int32_t), intent(out) :: orderarr(*)
      integer(c_int32_t), intent(out) :: nleaf
    end function c_tabi_init

    integer(c_int) function c_tabi_matvec(n2, x, y) bind(C, name='tabi_matvec')
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: n2
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function c_tabi_matvec

    integer(c_int) function c_tabi_msolve(n2, r, z) bind(C, name='tabi_msolve')
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: n2
      real(c_double), intent(in) :: r(*)
      real(c_double), intent(out) :: z(*)
    end function c_tabi_msolve

    subroutine c_tabi_finalize() bind(C, name='tabi_finalize')
    end subroutine c_tabi_finalize

    ! right-hand side in tree order (form_matrix's charge loop, src/main.f90:597-607)
    integer(c_int) function c_tabi_rhs(nq, qpos, q, eps_p, b) bind(C, name='tabi_rhs')
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: nq
      real(c_double), intent(in) :: qpos(*), q(*)
      real(c_double), value :: eps_p
      real(c_double), intent(out) :: b(*)
    end function c_tabi_rhs

    ! solvation and Coulomb energies in kcal/mol (src/main.f90:144-157)
    integer(c_int) function c_tabi_energy(nq, qpos, q, eps_p, x, soleng, couleng) &
        bind(C, name='tabi_energy')
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: nq
      real(c_double), intent(in) :: qpos(*), q(*), x(*)
      real(c_double), value :: eps_p
      real(c_double), intent(out) :: soleng, couleng
    end function c_tabi_energy

    ! vertex-averaged potentials (output_potential, src/main.f90:226-326)
    integer(c_int) function c_tabi_vertex_potential(nvert, vpos, faces, x, vpot) &
        bind(C, name='tabi_vertex_potential')
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: nvert
      real(c_double), intent(in) :: vpos(*), x(*)
      integer(c_int32_t), intent(in) :: faces(*)
      real(c_double), intent(out) :: vpot(*)
    end function c_tabi_vertex_potential

    type(c_ptr) function c_tabi_last_error() bind(C, name='tabi_last_error')
      import :: c_ptr
    end function c_tabi_last_error

    integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function c_strlen
  end interface

contains

  ! Same observable behaviour as the reference's error handling (WRITE + STOP,
  ! e.g. src/main.f90:105-110): print the library's message and stop.
  subroutine tabi_hip_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = c_tabi_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    write(*,*) 'tabi_hip: ', what, ' failed: ', msg(1:n)
    stop 1
  end subroutine tabi_hip_check

  ! Build the GPU tree, interaction lists and leaf factors for nface faces given
  ! in ORIGINAL order (xyz(3,nface), nrm(3,nface), area(nface) as form_matrix
  ! forms tr_xyz, tr_q, tr_area).  orderarr(nface) receives the 1-based tree
  ! permutation.
  subroutine tabi_hip_setup(eps, kappa, theta, order, maxparnode, nface, xyz, nrm, area, &
                            orderarr, nleaf)
    real(c_double), intent(in) :: eps, kappa, theta
    integer, intent(in) :: order, maxparnode, nface
    real(c_double), intent(in) :: xyz(3, nface), nrm(3, nface), area(nface)
    integer(c_int32_t), intent(out) :: orderarr(nface)
    integer, intent(out) :: nleaf
    type(tabi_params_t) :: prm
    integer(c_int32_t) :: nl
    prm%eps = eps; prm%kappa = kappa; prm%theta = theta
    prm%order = order; prm%maxparnode = maxparnode; prm%precision = 0; prm%flags = 0
    call tabi_hip_check(c_tabi_init(prm, int(nface, c_int32_t), xyz, nrm, area, orderarr, nl), &
                        'tabi_init')
    nleaf = nl
  end subroutine tabi_hip_setup

  subroutine tabi_hip_release()
    call c_tabi_finalize()
  end subroutine tabi_hip_release

end module tabi_hip

! MATVEC callback: y = A x (replaces MATVECmain, src/main.f90:467-488)
subroutine MATVECgpu(N, XX, bb)
  use tabi_hip
  use iso_c_binding
  implicit none
  integer :: N
  real(c_double) :: XX(N), bb(N)
  call tabi_hip_check(c_tabi_matvec(int(N, c_int32_t), XX, bb), 'tabi_matvec')
end subroutine MATVECgpu

! MSOLVE callback: z = M^-1 r (replaces MSOLVEdric, src/main.f90:511-529); DGMRES
! passes five arguments of which only N, R and Z are used.
subroutine MSOLVEgpu(N, R, Z, RPAR, IPAR)
  use tabi_hip
  use iso_c_binding
  implicit none
  integer :: N
  real(c_double) :: R(N), Z(N)
  real(c_double) :: RPAR(*)
  integer :: IPAR(*)
  call tabi_hip_check(c_tabi_msolve(int(N, c_int32_t), R, Z), 'tabi_msolve')
end subroutine MSOLVEgpu

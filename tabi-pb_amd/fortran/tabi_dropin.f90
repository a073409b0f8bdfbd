! Reference-side hooks for the drop-in build of the TABI-PB driver.
!
! Compiled against the reference's own modules (MOLECULE, COMDATA, TREECODE from
! src/var_modules.f90; TREECODE3D_PROCEDURES from src/treecode3d_pb.f).
!
! tabi_dropin_setup is called right after treecode_initialization
! (src/main.f90:78), when tr_xyz, tr_q and tr_area are already in tree order
! (src/main.f90:447-453): it restores the original face order through
! orderarr, builds the GPU-side tree, interaction lists and leaf factors,
! checks that the GPU permutation equals the reference's orderarr, so every
! later MATVECgpu / MSOLVEgpu vector is in the order DGMRES already uses, and
! forms the right-hand side bvct on the GPU in tree order (the O(nface*nchr)
! charge loop of form_matrix, src/main.f90:597-607, is left out of the
! drop-in build).
!
! tabi_dropin_energy replaces the O(nchr*nface) energy loop
! (src/main.f90:144-157) with the GPU's; it returns the energies in kcal/mol.
subroutine tabi_dropin_setup
  use molecule, only: nface, eps, kappa, nchr, chrpos, atmchr, eps0
  use comdata, only: bvct
  use treecode, only: tr_xyz, tr_q, tr_area, order, maxparnode, theta
  use treecode3d_procedures, only: orderarr
  use tabi_hip
  use iso_c_binding
  implicit none
  real(c_double), allocatable :: oxyz(:, :), onrm(:, :), oarea(:)
  integer(c_int32_t), allocatable :: gorder(:)
  integer :: i, nl

  allocate(oxyz(3, nface), onrm(3, nface), oarea(nface), gorder(nface))
  do i = 1, nface
    oxyz(:, orderarr(i)) = tr_xyz(:, i)
    onrm(:, orderarr(i)) = tr_q(:, i)
    oarea(orderarr(i)) = tr_area(i)
  end do
  call tabi_hip_setup(eps, kappa, theta, order, maxparnode, nface, oxyz, onrm, oarea, &
                      gorder, nl)
  if (any(gorder /= orderarr(1:nface))) then
    write(*,*) 'tabi_hip: GPU tree permutation differs from orderarr'
    stop 1
  end if
  print *, 'tabi_hip: GPU set-up done, number of leaves:', nl
  deallocate(oxyz, onrm, oarea, gorder)
  call tabi_hip_check(c_tabi_rhs(int(nchr, c_int32_t), chrpos, atmchr, eps0, bvct), 'tabi_rhs')
end subroutine tabi_dropin_setup

subroutine tabi_dropin_energy(soleng, couleng)
  use molecule, only: nchr, chrpos, atmchr, eps0
  use comdata, only: xvct
  use tabi_hip
  use iso_c_binding
  implicit none
  real*8, intent(out) :: soleng, couleng
  real(c_double) :: s, c
  call tabi_hip_check(c_tabi_energy(int(nchr, c_int32_t), chrpos, atmchr, eps0, xvct, s, c), &
                      'tabi_energy')
  soleng = s
  couleng = c
end subroutine tabi_dropin_energy

!! dropin_test.f90 -- runs the same ensemble through the reference's own
!! ADV_PARTICLE_ELLP + COLLISION and through the GPU drop-ins GPU_* of
!! gazpart_amd/fortran/gp_dropin.f90, inside one Fortran program, and
!! writes both final states.  TEST INFRASTRUCTURE (tests/test_gpu_fortran.py).
!!
!! usage: dropin_test <input.bin> <output.bin>
!!   input : int32 n, np, nsteps, scheme, mirror; float64 L, visc, rhof,
!!           dtime, emaj, apr, rhop; field(n, n+4, n+4, 6); state(np, 13)
!!   output: float64 state_ref(np, 13), state_gpu(np, 13)
program dropin_test
  use iso_c_binding
  use gpref_driver
  use gp_dropin
  implicit none
  character(len=512) :: fin, fout
  integer(c_int) :: hdr(5)
  real(c_double) :: r(7)
  real(c_double), allocatable :: fld(:,:,:,:), st0(:,:), sref(:,:), sgpu(:,:), fl(:,:), col(:)
  integer(c_int), allocatable :: idp(:), home(:), near(:)
  integer :: n, np, nsteps, s, c, nclose
  real(c_double) :: em(1,1), ap(1), rh(1,1)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) r
  n = hdr(1); np = hdr(2); nsteps = hdr(3)
  allocate(fld(n, n + 4, n + 4, 6), st0(np, 13), sref(np, 13), sgpu(np, 13), fl(np, 6), col(np))
  allocate(idp(np), home(200 * np), near(200 * np))
  read(10) fld
  read(10) st0
  close(10)
  em(1,1) = r(5); ap(1) = r(6); rh(1,1) = r(7)
  call ref_setup(n, r(1), r(2), r(3), r(4), 1.0d0, 0.0d0, hdr(4), 1, 1, em, ap, rh, &
                 [0.0d0], [1.0d0], [0.0d0])
  do c = 1, 6
    call ref_set_field(c, fld(:,:,:,c))
  end do
  idp = 1

  ! the reference itself
  call ref_set_particles(1, np, st0, idp)
  call ref_collision_init(hdr(1:4), r(1:6))
  do s = 1, nsteps
    call ADV_PARTICLE_ELLP(s)
    call ref_collision(1, s, 0.0d0, 200 * np, home, near, nclose)
  end do
  call ref_get_particles(1, np, sref, fl, col)

  ! the GPU drop-ins on the same module state
  call ref_set_particles(1, np, st0, idp)
  GPU_MIRROR = (hdr(5) /= 0)
  call GPU_INIT
  do s = 1, nsteps
    call GPU_ADV_PARTICLE_ELLP(s)
    call GPU_COLLISION(s, 0.0d0)
  end do
  call GPU_SYNC_PART
  call ref_get_particles(1, np, sgpu, fl, col)
  call GPU_FINALIZE

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) sref
  write(11) sgpu
  close(11)
end program dropin_test

!! dropin_test.f90 -- runs the same ensemble through the reference's own
!! ADV_PARTICLE_ELLP + COLLISION and through the GPU drop-ins GPU_* of
!! gazpart_amd/fortran/gp_dropin.f90, inside one Fortran program, and
!! writes both final states (and, with statistics on, both
!! MEAN_TIME_PART_COL sums).  TEST INFRASTRUCTURE (tests/test_gpu_fortran.py).
!!
!! usage: dropin_test <input.bin> <output.bin>
!!   input : int32 n, np, nsteps, scheme, mirror, npoly, fout2 (0: no
!!           statistics), interph (1: INTERPH and GPU_INTERPH of the six
!!           components at the initial positions, before GPU_INIT, as
!!           INITIATION_PARTICLE_VELOCITY calls it), evolve (1: SOLVE_FLUID
!!           = 1 and, from step nsteps/2 + 1 on, the field halved and DTIME
!!           times 0.8 in both runs, as ADV_FLUID and CHECK_CFL change them);
!!           float64 L, visc, rhof, dtime, apr;
!!           float64 emaj(npoly), rhop(npoly); field(n, n+4, n+4, 6);
!!           state(np, 13); int32 idp(np)
!!   output: float64 state_ref(np, 13), state_gpu(np, 13); with fout2 > 0
!!           also MEAN_TIME_PART_COL(1:8, npoly, npoly) of each run; with
!!           interph = 1 also interp_ref(np, 6), interp_gpu(np, 6)
!!   With fout2 > 0 both runs also write the reference's per-pass lines to
!!   fort.NNN in the current directory (the reference's run first).
program dropin_test
  use iso_c_binding
  use gpref_driver
  use gp_dropin
  implicit none
  character(len=512) :: fin, fout
  integer(c_int) :: hdr(9), inum(4)
  real(c_double) :: r(5), rnum(6)
  real(c_double), allocatable :: fld(:,:,:,:), st0(:,:), sref(:,:), sgpu(:,:), fl(:,:), col(:)
  real(c_double), allocatable :: em(:,:), rh(:,:), mref(:,:,:), mgpu(:,:,:)
  real(c_double), allocatable :: iref(:,:), igpu(:,:)
  integer(c_int), allocatable :: idp(:), home(:), near(:)
  integer :: n, np, nsteps, npoly, fout2_, s, c, nclose
  real(kind=8) :: dtime0

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) r
  n = hdr(1); np = hdr(2); nsteps = hdr(3); npoly = hdr(6); fout2_ = hdr(7)
  allocate(em(1, npoly), rh(1, npoly))
  allocate(fld(n, n + 4, n + 4, 6), st0(np, 13), sref(np, 13), sgpu(np, 13), fl(np, 6), col(np))
  allocate(idp(np), home(200 * np), near(200 * np))
  read(10) em, rh
  read(10) fld
  read(10) st0
  read(10) idp
  close(10)
  call ref_setup(n, r(1), r(2), r(3), r(4), 1.0d0, 0.0d0, hdr(4), 1, npoly, em, [r(5)], rh, &
                 [0.0d0], [1.0d0], [0.0d0])
  do c = 1, 6
    call ref_set_field(c, fld(:,:,:,c))
  end do
  dtime0 = DTIME
  if (hdr(9) /= 0) SOLVE_FLUID = 1

  ! INTERPH before any GPU_INIT: the reference's and the drop-in's
  if (hdr(8) /= 0) then
    allocate(iref(np, 6), igpu(np, 6))
    call INTERPH(hdr(4), XMESH, YMESH, ZMESH, UFLU, np, st0(:,1), st0(:,2), st0(:,3), iref(:,1))
    call INTERPH(hdr(4), XMESH, YMESH, ZMESH, VFLU, np, st0(:,1), st0(:,2), st0(:,3), iref(:,2))
    call INTERPH(hdr(4), XMESH, YMESH, ZMESH, WFLU, np, st0(:,1), st0(:,2), st0(:,3), iref(:,3))
    call INTERPH(hdr(4), XMESH, YMESH, ZMESH, VORTICITY_X, np, st0(:,1), st0(:,2), st0(:,3), &
                 iref(:,4))
    call INTERPH(hdr(4), XMESH, YMESH, ZMESH, VORTICITY_Y, np, st0(:,1), st0(:,2), st0(:,3), &
                 iref(:,5))
    call INTERPH(hdr(4), XMESH, YMESH, ZMESH, VORTICITY_Z, np, st0(:,1), st0(:,2), st0(:,3), &
                 iref(:,6))
    call GPU_INTERPH(hdr(4), XMESH, YMESH, ZMESH, UFLU, np, st0(:,1), st0(:,2), st0(:,3), &
                     igpu(:,1))
    call GPU_INTERPH(hdr(4), XMESH, YMESH, ZMESH, VFLU, np, st0(:,1), st0(:,2), st0(:,3), &
                     igpu(:,2))
    call GPU_INTERPH(hdr(4), XMESH, YMESH, ZMESH, WFLU, np, st0(:,1), st0(:,2), st0(:,3), &
                     igpu(:,3))
    call GPU_INTERPH(hdr(4), XMESH, YMESH, ZMESH, VORTICITY_X, np, st0(:,1), st0(:,2), &
                     st0(:,3), igpu(:,4))
    call GPU_INTERPH(hdr(4), XMESH, YMESH, ZMESH, VORTICITY_Y, np, st0(:,1), st0(:,2), &
                     st0(:,3), igpu(:,5))
    call GPU_INTERPH(hdr(4), XMESH, YMESH, ZMESH, VORTICITY_Z, np, st0(:,1), st0(:,2), &
                     st0(:,3), igpu(:,6))
  end if

  ! the reference itself
  call ref_set_particles(1, np, st0, idp)
  call ref_collision_init(inum, rnum)
  if (fout2_ > 0) call ref_cstat_on(fout2_)
  do s = 1, nsteps
    if (hdr(9) /= 0 .and. s == nsteps / 2 + 1) call evolve
    call ADV_PARTICLE_ELLP(s)
    call ref_collision(1, s, 0.0d0, 200 * np, home, near, nclose)
  end do
  call ref_get_particles(1, np, sref, fl, col)
  if (fout2_ > 0) then
    allocate(mref(8, npoly, npoly), mgpu(8, npoly, npoly))
    call ref_cstat_get(mref)
    call ref_cstat_on(fout2_)
  end if

  ! the GPU drop-ins on the same module state
  do c = 1, 6
    call ref_set_field(c, fld(:,:,:,c))
  end do
  DTIME = dtime0
  call ref_set_particles(1, np, st0, idp)
  GPU_MIRROR = (hdr(5) /= 0)
  call GPU_INIT
  do s = 1, nsteps
    if (hdr(9) /= 0 .and. s == nsteps / 2 + 1) call evolve
    call GPU_ADV_PARTICLE_ELLP(s)
    call GPU_COLLISION(s, 0.0d0)
  end do
  call GPU_SYNC_PART
  call ref_get_particles(1, np, sgpu, fl, col)
  call GPU_FINALIZE
  if (fout2_ > 0) call ref_cstat_get(mgpu)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) sref
  write(11) sgpu
  if (fout2_ > 0) then
    write(11) mref
    write(11) mgpu
  end if
  if (hdr(8) /= 0) then
    write(11) iref
    write(11) igpu
  end if
  close(11)

contains

  ! what ADV_FLUID and CHECK_CFL do to the particles' inputs: a new field
  ! and time step (here the field halved, DTIME times 0.8)
  subroutine evolve
    UFLU = 0.5d0 * UFLU; VFLU = 0.5d0 * VFLU; WFLU = 0.5d0 * WFLU
    VORTICITY_X = 0.5d0 * VORTICITY_X; VORTICITY_Y = 0.5d0 * VORTICITY_Y
    VORTICITY_Z = 0.5d0 * VORTICITY_Z
    DTIME = 0.8d0 * dtime0
  end subroutine evolve
end program dropin_test

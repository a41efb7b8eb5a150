!! gp_dropin.f90 -- drop-in GPU versions of the reference's hot-path
!! subroutines, with their argument interfaces unchanged:
!!
!!   call INTERPH(...)            -> call GPU_INTERPH(...)          (same args)
!!   call ADV_PARTICLE_ELLP(NCYCLE) -> call GPU_ADV_PARTICLE_ELLP(NCYCLE)
!!   call COLLISION(NCYCLE, TIME)   -> call GPU_COLLISION(NCYCLE, TIME)
!!
!! Compiled together with the reference's own modules (DNS_DIM, PARAM_PHYS,
!! PARTICLE_PARALLEL, FLUID_VARIABLE, GEOMETRIC_VARIABLE, COLLISION_VARIABLE)
!! and linked with libgazpart_gpu.so.  GPU_INIT is called once after
!! INITIATION_FROZEN_FLOW + FLUIDCOMM, the particle initialisation and
!! COLLISION_INITIATION (main.f90:251-289); GPU_FINALIZE at the end.
!!
!! One GPU context per particle class (the reference collides particles of
!! one class only, collision.f90:76-171).  With GPU_MIRROR = .true. (default)
!! PART is copied to the device before and back after every call, so every
!! other routine of the reference sees the usual PART; with .false. the
!! device copy is authoritative and GPU_SYNC_PART refreshes PART on demand
!! (before statistics or output).  In that resident mode the collision
!! pass of a colliding class is deferred to the next GPU_ADV_PARTICLE_ELLP,
!! which runs it and the next advance in one call without host
!! synchronisation, the advance of the particles outside the pass's contact
!! solves beside the solves (gp_collision_advance); GPU_SYNC_PART runs a
!! deferred collision pass first, so every routine that reads PART after
!! GPU_COLLISION sees it after the pass.  A non-zero status stops the run with
!! the library's message, as the reference's own `stop`s do.  With
!! LEVEL1_STPAR (at GPU_INIT), GPU_COLLISION also writes the level-1
!! collision statistics and adds them to MEAN_TIME_PART_COL as
!! COLLISION_INTERACTION does, from sums formed on the device.
!!
!! When the fluid evolves (SOLVE_FLUID > 0: ADV_FLUID and CHECK_CFL change
!! the field and DTIME every cycle, main.f90:496-542), GPU_ADV_PARTICLE_ELLP
!! copies UFLU..VORTICITY_Z and DTIME to every context before it advances;
!! the spectral supplier on the GPU (gp_dns_bind) avoids that host copy.
!!
!! GPU_INTERPH can be called before GPU_INIT (INITIATION_PARTICLE_VELOCITY
!! does so, main.f90:286, before COLLISION_INITIATION): it uses a context
!! of its own that holds no particles and no field.
module gp_dropin
  use iso_c_binding
  use gazpart_gpu
  use DNS_DIM
  use PARAM_PHYS
  use PARTICLE_PARALLEL
  use FLUID_VARIABLE
  use GEOMETRIC_VARIABLE
  use COLLISION_VARIABLE
  implicit none

  type(c_ptr), allocatable, save :: GPU_CTX(:)
  type(c_ptr), save :: GPU_FIELD_CTX = c_null_ptr
  logical, save :: GPU_MIRROR = .true.
  integer(c_int), save :: GPU_PRECISION = GP_PREC_FP64
  logical, save :: GPU_DEVICE_AHEAD = .false.
  logical, allocatable, save :: GPU_COLL_PENDING(:)
  integer, allocatable, save :: GPU_COLL_NCYCLE(:)
  real(kind=8), allocatable, save :: GPU_COLL_TIME(:)
  real(kind=8), save :: GPU_DTIME_BOUND = -1.0
  ! collision cell rule (gp_params.cell_rule): 0 the reference's, refused
  ! when it would lose pairs; 2 the reference's unchecked; 1 from the
  ! largest semi-axis
  integer(c_int), save :: GPU_CELL_RULE = 0

contains

  subroutine GPU_CHECK(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    if (rc /= GP_OK) then
      write(*,*) 'gazpart_gpu: ', what, ' failed with status ', rc
      if (c_associated(ctx)) write(*,*) trim(gp_error_message(ctx))
      stop 'gazpart_gpu error'
    end if
  end subroutine GPU_CHECK

  !! one context per class: grid, physics, species of the class, field, PART
  subroutine GPU_INIT
    type(gp_params) :: p
    integer :: J, IDP, ns
    integer(c_int) :: rc
    real(c_double), allocatable :: rows(:,:), fld(:,:,:,:)

    if (allocated(GPU_CTX)) call GPU_FINALIZE
    allocate(GPU_CTX(NIG), GPU_COLL_PENDING(NIG), GPU_COLL_NCYCLE(NIG), GPU_COLL_TIME(NIG))
    GPU_COLL_PENDING = .false.
    GPU_COLL_NCYCLE = 0
    GPU_COLL_TIME = 0.0
    allocate(fld(NX, -1:NY+2, -1:NZ+2, 6))
    fld(:,:,:,1) = UFLU; fld(:,:,:,2) = VFLU; fld(:,:,:,3) = WFLU
    fld(:,:,:,4) = VORTICITY_X; fld(:,:,:,5) = VORTICITY_Y; fld(:,:,:,6) = VORTICITY_Z
    do J = 1, NIG
      p%n = NX
      p%interp_scheme = INTERP_SCHEME
      p%precision = GPU_PRECISION
      p%renorm_quat = 0
      p%L = LXMAX
      p%visc = VISC
      p%rhof = RHOF
      p%dtime = DTIME
      p%ecw = ECW
      p%muw = MUW
      p%dpmax = 2.0*maxval(EMAJ_PART(:,:))        ! collision_initiation.f90:50
      p%prefilter = 2
      p%xcd_remap = 0
      p%keep_fluid = 0
      p%timing = 0
      p%graph = 0
      p%brick = 0
      p%lm_seed = 0
      p%lm_waves = 0
      ! resident mode: each collision pass solves its pairs ahead
      p%overlap = 1
      p%cell_rule = GPU_CELL_RULE
      p%tile = 0
      p%reserved = 0
      GPU_CTX(J) = c_null_ptr
      rc = gp_create(p, GPU_CTX(J))
      call GPU_CHECK(rc, GPU_CTX(J), 'gp_create')
      rc = gp_bind_field(GPU_CTX(J), fld)
      call GPU_CHECK(rc, GPU_CTX(J), 'gp_bind_field')
      ns = POLYDISP(J)
      allocate(rows(GP_SP_STRIDE, ns))
      rows = 0.0
      do IDP = 1, ns
        ! the reference's own expressions (fluid_force_ellipsoid.f90:59,70,
        ! calculate_ellipsoid_matrix.f90:33-35, initiation_particle_position.f90:620-622)
        rows(1, IDP) = EMAJ_PART(J,IDP)
        rows(2, IDP) = EMAJ_PART(J,IDP)/APR_PART(J)
        rows(3, IDP) = rows(2, IDP)
        rows(4, IDP) = RHOP(J, IDP)*(4.0/3.0)*PPI*(EMAJ_PART(J,IDP)**3.0)/((APR_PART(J)**(2.0)))
        rows(5, IDP) = (2.0*EMAJ_PART(J,IDP))/(APR_PART(J)**(2.0/3.0))
        rows(6, IDP) = IPXX(J,IDP)
        rows(7, IDP) = IPYY(J,IDP)
        rows(8, IDP) = IPZZ(J,IDP)
        rows(9, IDP) = GRAVITY(J)
        rows(10, IDP) = ECP(J)
        rows(11, IDP) = MUP(J)
        rows(12, IDP) = RHOP(J, IDP)
        rows(13, IDP) = EMAJ_PART(J,IDP)
        rows(14, IDP) = APR_PART(J)
      end do
      rc = gp_set_species(GPU_CTX(J), ns, rows)
      call GPU_CHECK(rc, GPU_CTX(J), 'gp_set_species')
      deallocate(rows)
      ! level-1 collision statistics (MEAN_PART_COL) summed on the device on
      ! the cycles COLLISION_INTERACTION sums them (flags read here, once)
      if (LEVEL1_STPAR .and. FOUT2 > 0) then
        rc = gp_set_collision_stats(GPU_CTX(J), 1_c_int, int(ns, c_int), int(FOUT2, c_int), 1_c_int)
        call GPU_CHECK(rc, GPU_CTX(J), 'gp_set_collision_stats')
      end if
      call GPU_PUT_PART(J)
    end do
    deallocate(fld)
    GPU_DTIME_BOUND = DTIME
    GPU_DEVICE_AHEAD = .false.
  end subroutine GPU_INIT

  !! the evolving field and time step into every context (SOLVE_FLUID > 0)
  subroutine GPU_REBIND_FIELD
    integer :: J
    integer(c_int) :: rc
    real(c_double), allocatable :: fld(:,:,:,:)
    allocate(fld(NX, -1:NY+2, -1:NZ+2, 6))
    fld(:,:,:,1) = UFLU; fld(:,:,:,2) = VFLU; fld(:,:,:,3) = WFLU
    fld(:,:,:,4) = VORTICITY_X; fld(:,:,:,5) = VORTICITY_Y; fld(:,:,:,6) = VORTICITY_Z
    do J = 1, size(GPU_CTX)
      rc = gp_bind_field(GPU_CTX(J), fld)
      call GPU_CHECK(rc, GPU_CTX(J), 'gp_bind_field')
      if (DTIME /= GPU_DTIME_BOUND) then
        rc = gp_set_dtime(GPU_CTX(J), real(DTIME, c_double))
        call GPU_CHECK(rc, GPU_CTX(J), 'gp_set_dtime')
      end if
    end do
    GPU_DTIME_BOUND = DTIME
    deallocate(fld)
  end subroutine GPU_REBIND_FIELD

  !! a deferred collision pass of class J, run now
  subroutine GPU_FLUSH_COLLISION(J)
    integer, intent(in) :: J
    integer(c_int) :: rc
    if (.not. GPU_COLL_PENDING(J)) return
    rc = gp_collision(GPU_CTX(J), int(GPU_COLL_NCYCLE(J), c_int), &
                      real(GPU_COLL_TIME(J), c_double))
    call GPU_CHECK(rc, GPU_CTX(J), 'COLLISION')
    GPU_COLL_PENDING(J) = .false.
  end subroutine GPU_FLUSH_COLLISION

  subroutine GPU_FINALIZE
    integer :: J
    integer(c_int) :: rc
    if (c_associated(GPU_FIELD_CTX)) then
      rc = gp_destroy(GPU_FIELD_CTX)
      GPU_FIELD_CTX = c_null_ptr
    end if
    if (.not. allocated(GPU_CTX)) return
    do J = 1, size(GPU_CTX)
      if (c_associated(GPU_CTX(J))) rc = gp_destroy(GPU_CTX(J))
    end do
    deallocate(GPU_CTX, GPU_COLL_PENDING, GPU_COLL_NCYCLE, GPU_COLL_TIME)
  end subroutine GPU_FINALIZE

  !! PART(1:NPART_LOC(J), J) -> device
  subroutine GPU_PUT_PART(J)
    integer, intent(in) :: J
    integer :: I, np
    integer(c_int) :: rc
    real(c_double), allocatable :: st(:,:)
    integer(c_int), allocatable :: sp(:)
    np = NPART_LOC(J)
    if (np < 1) return
    allocate(st(np, 13), sp(np))
    do I = 1, np
      st(I, 1) = PART(I,J)%XP; st(I, 2) = PART(I,J)%YP; st(I, 3) = PART(I,J)%ZP
      st(I, 4) = PART(I,J)%UP; st(I, 5) = PART(I,J)%VP; st(I, 6) = PART(I,J)%WP
      st(I, 7) = PART(I,J)%ELLQUAT%a; st(I, 8) = PART(I,J)%ELLQUAT%b
      st(I, 9) = PART(I,J)%ELLQUAT%c; st(I, 10) = PART(I,J)%ELLQUAT%d
      st(I, 11) = PART(I,J)%OMEGAX; st(I, 12) = PART(I,J)%OMEGAY; st(I, 13) = PART(I,J)%OMEGAZ
      sp(I) = PART(I,J)%IDP - 1
    end do
    rc = gp_set_particles(GPU_CTX(J), np, st, sp)
    call GPU_CHECK(rc, GPU_CTX(J), 'gp_set_particles')
    deallocate(st, sp)
  end subroutine GPU_PUT_PART

  !! device -> PART(1:NPART_LOC(J), J)
  subroutine GPU_GET_PART(J)
    integer, intent(in) :: J
    integer :: I, np
    integer(c_int) :: rc
    real(c_double), allocatable :: st(:,:)
    np = NPART_LOC(J)
    if (np < 1) return
    call GPU_FLUSH_COLLISION(J)
    allocate(st(np, 13))
    rc = gp_get_particles(GPU_CTX(J), np, st)
    call GPU_CHECK(rc, GPU_CTX(J), 'gp_get_particles')
    do I = 1, np
      PART(I,J)%XP = st(I, 1); PART(I,J)%YP = st(I, 2); PART(I,J)%ZP = st(I, 3)
      PART(I,J)%UP = st(I, 4); PART(I,J)%VP = st(I, 5); PART(I,J)%WP = st(I, 6)
      PART(I,J)%ELLQUAT%a = st(I, 7); PART(I,J)%ELLQUAT%b = st(I, 8)
      PART(I,J)%ELLQUAT%c = st(I, 9); PART(I,J)%ELLQUAT%d = st(I, 10)
      PART(I,J)%OMEGAX = st(I, 11); PART(I,J)%OMEGAY = st(I, 12); PART(I,J)%OMEGAZ = st(I, 13)
    end do
    deallocate(st)
  end subroutine GPU_GET_PART

  !! refresh PART from the device (resident mode)
  subroutine GPU_SYNC_PART
    integer :: J
    if (.not. GPU_DEVICE_AHEAD) return
    do J = 1, NIG
      call GPU_GET_PART(J)
    end do
    GPU_DEVICE_AHEAD = .false.
  end subroutine GPU_SYNC_PART

  !! ADV_PARTICLE_ELLP (adv_particle_new.f90:7), same interface
  subroutine GPU_ADV_PARTICLE_ELLP(NCYCLE)
    integer, intent(in) :: NCYCLE
    integer :: J
    integer(c_int) :: rc
    if (.not. allocated(GPU_CTX)) stop 'GPU_ADV_PARTICLE_ELLP: call GPU_INIT first'
    if (SOLVE_FLUID > 0) call GPU_REBIND_FIELD
    do J = 1, NIG
      if (NPART_LOC(J) < 1) cycle
      if (GPU_COLL_PENDING(J)) then
        ! resident mode: the previous cycle's collision pass and this
        ! advance in one call
        rc = gp_collision_advance(GPU_CTX(J), int(GPU_COLL_NCYCLE(J), c_int), &
                                  real(GPU_COLL_TIME(J), c_double), int(NCYCLE, c_int))
        call GPU_CHECK(rc, GPU_CTX(J), 'COLLISION + ADV_PARTICLE_ELLP')
        GPU_COLL_PENDING(J) = .false.
        cycle
      end if
      if (GPU_MIRROR) call GPU_PUT_PART(J)
      rc = gp_adv_particle_ellp(GPU_CTX(J), int(NCYCLE, c_int))
      call GPU_CHECK(rc, GPU_CTX(J), 'ADV_PARTICLE_ELLP')
      if (GPU_MIRROR) call GPU_GET_PART(J)
    end do
    if (.not. GPU_MIRROR) GPU_DEVICE_AHEAD = .true.
  end subroutine GPU_ADV_PARTICLE_ELLP

  !! COLLISION (collision.f90:7) at NPROC = 1, same interface
  subroutine GPU_COLLISION(NCYCLE, TIME)
    integer, intent(in) :: NCYCLE
    real(kind=8), intent(in) :: TIME
    integer :: J
    integer(c_int) :: rc
    if (.not. allocated(GPU_CTX)) stop 'GPU_COLLISION: call GPU_INIT first'
    do J = 1, NIG
      if (COLLIDEF(J) /= 1 .or. NPART_LOC(J) < 1) cycle
      call GPU_FLUSH_COLLISION(J)
      ! resident mode: run with the next advance (its statistics passes,
      ! which write lines now, stay immediate)
      if (.not. GPU_MIRROR .and. .not. (LEVEL1_STPAR .and. FOUT2 > 0)) then
        GPU_COLL_PENDING(J) = .true.
        GPU_COLL_NCYCLE(J) = NCYCLE
        GPU_COLL_TIME(J) = TIME
        cycle
      end if
      if (GPU_MIRROR) call GPU_PUT_PART(J)
      rc = gp_collision(GPU_CTX(J), int(NCYCLE, c_int), real(TIME, c_double))
      call GPU_CHECK(rc, GPU_CTX(J), 'COLLISION')
      if (GPU_MIRROR) call GPU_GET_PART(J)
      if (LEVEL1_STPAR .and. FOUT2 > 0) then
        if (mod(NCYCLE, FOUT2) == 0) call GPU_COLLISION_STATS(J, TIME)
      end if
    end do
    if (.not. GPU_MIRROR) GPU_DEVICE_AHEAD = .true.
  end subroutine GPU_COLLISION

  !! what COLLISION_INTERACTION does with one pass's MEAN_PART_COL
  !! (collision_interaction.f90:423-453): one line per size pair on the
  !! reference's statistics units, and the STAT_TIME sum MEAN_TIME_PART_COL;
  !! the entries the pass does not form keep the reference's ZERO
  subroutine GPU_COLLISION_STATS(J, TIME)
    integer, intent(in) :: J
    real(kind=8), intent(in) :: TIME
    integer :: ns, IDP1, IDP2, NUMFILE
    integer(c_int) :: rc
    integer(c_int64_t) :: npass     ! (not nstat: that name is DNS_DIM's NSTAT)
    real(c_double), allocatable :: last(:,:,:), tsum(:,:,:)
    real(kind=8), allocatable :: mpc(:,:,:)
    ns = POLYDISP(J)
    allocate(last(8, ns, ns), tsum(8, ns, ns), mpc(NSTAT, POLYDISPMAX, POLYDISPMAX))
    rc = gp_get_collision_stats(GPU_CTX(J), last, tsum, npass)
    call GPU_CHECK(rc, GPU_CTX(J), 'gp_get_collision_stats')
    mpc = ZERO
    do IDP1 = 1, ns
      do IDP2 = 1, ns
        mpc(1:8, IDP1, IDP2) = last(:, IDP2, IDP1)
        NUMFILE = 560 + sum(POLYDISP(1:J)**2) - POLYDISP(J)**2 + (IDP1 - 1)*POLYDISP(J) + IDP2
        write(NUMFILE, 10000) TIME, mpc(1, IDP1, IDP2), mpc(2, IDP1, IDP2), mpc(3, IDP1, IDP2), &
                              mpc(4, IDP1, IDP2)/EMAJ_PART(J, IDP1), mpc(5, IDP1, IDP2), &
                              mpc(6, IDP1, IDP2), mpc(7, IDP1, IDP2), mpc(8, IDP1, IDP2)
      end do
    end do
    if (STAT_TIME) MEAN_TIME_PART_COL = MEAN_TIME_PART_COL + mpc
    deallocate(last, tsum, mpc)
10000 format (30(e17.7))
  end subroutine GPU_COLLISION_STATS

  !! INTERPH (interpH.f90:26), same interface; UAIN is passed to the GPU
  subroutine GPU_INTERPH(INT_SCHEME, XAIN, YAIN, ZAIN, UAIN, NINT, XINT, YINT, ZINT, UINT)
    integer, intent(in) :: INT_SCHEME
    real(kind=8), dimension(ISTART(1):IEND(1)), intent(in) :: XAIN
    real(kind=8), dimension(ISTART(2):IEND(2)), intent(in) :: YAIN
    real(kind=8), dimension(ISTART(3):IEND(3)), intent(in) :: ZAIN
    real(kind=8), dimension(ISTART(1):IEND(1), ISTART(2)-NGHTCELL:IEND(2)+NGHTCELL, &
                           ISTART(3)-NGHTCELL:IEND(3)+NGHTCELL), intent(in) :: UAIN
    integer, intent(in) :: NINT
    real(kind=8), dimension(NINT), intent(in) :: XINT, YINT, ZINT
    real(kind=8), dimension(NINT), intent(out) :: UINT
    integer(c_int) :: rc
    type(gp_params) :: p
    if (INT_SCHEME /= 3 .and. INT_SCHEME /= 4) then
      stop 'GPU_INTERPH: only INT_SCHEME 3 and 4 run on the GPU'
    end if
    if (XAIN(ISTART(1)+1) - XAIN(ISTART(1)) /= DX) stop 'GPU_INTERPH: mesh is not XMESH'
    if (.not. c_associated(GPU_FIELD_CTX)) then
      ! a context for field-argument interpolation only (grid and scheme;
      ! the physics entries are unused but must be valid)
      p%n = NX
      p%interp_scheme = int(INT_SCHEME, c_int)
      p%precision = GP_PREC_FP64
      p%renorm_quat = 0
      p%L = LXMAX
      p%visc = merge(VISC, 1.0d0, VISC > 0.0)
      p%rhof = RHOF
      p%dtime = merge(DTIME, 1.0d0, DTIME > 0.0)
      p%ecw = ECW
      p%muw = MUW
      p%dpmax = 0.0
      p%prefilter = 2
      p%xcd_remap = 0
      p%keep_fluid = 0
      p%timing = 0
      p%graph = 0
      p%brick = 0
      p%lm_seed = 0
      p%lm_waves = 0
      p%overlap = 0
      p%cell_rule = 2
      p%tile = 0
      p%reserved = 0
      rc = gp_create(p, GPU_FIELD_CTX)
      call GPU_CHECK(rc, GPU_FIELD_CTX, 'gp_create')
    end if
    rc = gp_interph_field(GPU_FIELD_CTX, int(INT_SCHEME, c_int), UAIN, int(NINT, c_int), &
                          XINT, YINT, ZINT, UINT)
    call GPU_CHECK(rc, GPU_FIELD_CTX, 'INTERPH')
  end subroutine GPU_INTERPH

end module gp_dropin

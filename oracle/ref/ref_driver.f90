!! ref_driver.f90 -- C-callable driver around the reference's own hot-path
!! subroutines (compiled unmodified from the reference's source/ directory
!! by oracle/ref/Makefile).  TEST INFRASTRUCTURE: used only to write golden
!! fixtures and to pin the C oracle; it is never part of the product.
!!
!! The driver sets the reference's module state (grid, physics, particle
!! classes, fields, particles) the way INIT_RUN / MESHING / READPARAM would
!! for a single-rank frozen-field run, then calls the reference routines.
!! MYID is set to 1 (NPROC = 1) so that the rank-0-only MPI_WTIME timers and
!! prints are skipped; nothing else in the hot path reads MYID.
module gpref_driver
  use iso_c_binding
  use DNS_DIM
  use PARAM_PHYS
  use PARTICLE_PARALLEL
  use FLUID_VARIABLE
  use GEOMETRIC_VARIABLE
  use COLLISION_VARIABLE
  use STATISTICS
  use mod_quaternion
  implicit none

  integer :: NPREF = 0

contains

  !! grid, physics and class tables (nclass classes x nsize sizes)
  subroutine ref_setup(n, L, visc_, rhof_, dtime_, ecw_, muw_, scheme, &
                       nclass, nsize, emaj, apr, rhop_, grav, ecp_, mup_) &
                       bind(C, name='ref_setup')
    integer(c_int), value :: n, scheme, nclass, nsize
    real(c_double), value :: L, visc_, rhof_, dtime_, ecw_, muw_
    real(c_double), intent(in) :: emaj(nclass, nsize), apr(nclass), rhop_(nclass, nsize)
    real(c_double), intent(in) :: grav(nclass), ecp_(nclass), mup_(nclass)
    integer :: i, j, idp

    MYID = 1
    NPROC = 1
    NX = n; NY = n; NZ = n
    ISTART = 1
    IEND(1) = n; IEND(2) = n; IEND(3) = n
    NGHTCELL = 2
    PPI = 4.*atan(1.D0)
    TWOPI = 2.*PPI
    LXMAX = L; LYMAX = L; LZMAX = L
    DX = LXMAX / real(NX-1)
    DY = LYMAX / real(NY-1)
    DZ = LZMAX / real(NZ-1)
    if (allocated(XMESH)) deallocate(XMESH, YMESH, ZMESH)
    allocate(XMESH(n), YMESH(n), ZMESH(n))
    do i = 1, n
      XMESH(i) = (i-1)*DX
      YMESH(i) = DY*(i-1)
      ZMESH(i) = DZ*(i-1)
    end do
    if (allocated(UFLU)) deallocate(UFLU, VFLU, WFLU, VORTICITY_X, VORTICITY_Y, VORTICITY_Z)
    allocate(UFLU(n, -1:n+2, -1:n+2), VFLU(n, -1:n+2, -1:n+2), WFLU(n, -1:n+2, -1:n+2))
    allocate(VORTICITY_X(n, -1:n+2, -1:n+2), VORTICITY_Y(n, -1:n+2, -1:n+2))
    allocate(VORTICITY_Z(n, -1:n+2, -1:n+2))
    UFLU = 0.; VFLU = 0.; WFLU = 0.
    VORTICITY_X = 0.; VORTICITY_Y = 0.; VORTICITY_Z = 0.

    VISC = visc_; RHOF = rhof_
    DTIME = dtime_; ECW = ecw_; MUW = muw_
    INTERP_SCHEME = scheme
    SOLVE_FLUID = 0
    FROZEN_FLOW = 2
    SOLVE_PART = .true.
    WALL_BOUNDARY = .false.
    PERIODICITY = .not. WALL_BOUNDARY
    FOUT2 = 1000000
    STAT_TIME = .false.; LEVEL0_STPAR = .false.; LEVELX_STPAR = .false.
    LEVEL1_STPAR = .false.; LEVEL1_STPDF = .false.

    NIG = nclass
    POLYDISPMAX = nsize
    if (allocated(PARTDEF)) deallocate(PARTDEF, APR_PART, GRAVITY, COLLIDEF, POLYDISP, &
                                       EMAJ_PART, RHOP, IPXX, IPYY, IPZZ, ECP, MUP)
    allocate(PARTDEF(NIG), APR_PART(NIG), GRAVITY(NIG), COLLIDEF(NIG), POLYDISP(NIG))
    allocate(EMAJ_PART(NIG, nsize), RHOP(NIG, nsize))
    allocate(IPXX(NIG, nsize), IPYY(NIG, nsize), IPZZ(NIG, nsize))
    allocate(ECP(NIG), MUP(NIG))
    do j = 1, NIG
      PARTDEF(j) = 2
      COLLIDEF(j) = 1
      POLYDISP(j) = nsize
      APR_PART(j) = apr(j)
      GRAVITY(j) = grav(j)
      ECP(j) = ecp_(j)
      MUP(j) = mup_(j)
      do idp = 1, nsize
        EMAJ_PART(j, idp) = emaj(j, idp)
        RHOP(j, idp) = rhop_(j, idp)
        !! inertia as initiation_particle_position.f90:609-622
        ELL_A = EMAJ_PART(j, idp)
        lambda = APR_PART(j)
        ELL_B = ELL_A/lambda
        ELL_C = ELL_B
        IPXX(j, idp) = (ELL_B**2 + ELL_C**2)/5.0
        IPYY(j, idp) = (ELL_A**2 + ELL_C**2)/5.0
        IPZZ(j, idp) = (ELL_A**2 + ELL_B**2)/5.0
      end do
    end do
  end subroutine ref_setup

  !! one field component (1..6 = u v w wx wy wz), reference layout
  subroutine ref_set_field(comp, arr) bind(C, name='ref_set_field')
    integer(c_int), value :: comp
    real(c_double), intent(in) :: arr(NX, -1:NY+2, -1:NZ+2)
    select case (comp)
    case (1); UFLU = arr
    case (2); VFLU = arr
    case (3); WFLU = arr
    case (4); VORTICITY_X = arr
    case (5); VORTICITY_Y = arr
    case (6); VORTICITY_Z = arr
    end select
  end subroutine ref_set_field

  !! particles of class ig: state(np, 13) = x y z u v w qa qb qc qd ox oy oz
  subroutine ref_set_particles(ig, np, st, idp) bind(C, name='ref_set_particles')
    integer(c_int), value :: ig, np
    real(c_double), intent(in) :: st(np, 13)
    integer(c_int), intent(in) :: idp(np)
    integer :: i
    if (.not. allocated(NPART_LOC)) then
      allocate(NPART_LOC(NIG))
      NPART_LOC = 0
    end if
    if (np > NPREF .or. .not. allocated(PART)) then
      if (allocated(PART)) deallocate(PART)
      NPREF = max(np, 1)
      allocate(PART(NPREF, NIG))
    end if
    NPMAX_LOC = NPREF
    NPART_LOC(ig) = np
    do i = 1, np
      PART(i, ig)%XP = st(i, 1); PART(i, ig)%YP = st(i, 2); PART(i, ig)%ZP = st(i, 3)
      PART(i, ig)%UP = st(i, 4); PART(i, ig)%VP = st(i, 5); PART(i, ig)%WP = st(i, 6)
      PART(i, ig)%ELLQUAT%a = st(i, 7); PART(i, ig)%ELLQUAT%b = st(i, 8)
      PART(i, ig)%ELLQUAT%c = st(i, 9); PART(i, ig)%ELLQUAT%d = st(i, 10)
      PART(i, ig)%OMEGAX = st(i, 11); PART(i, ig)%OMEGAY = st(i, 12)
      PART(i, ig)%OMEGAZ = st(i, 13)
      PART(i, ig)%IDP = idp(i)
      PART(i, ig)%COLOR = 1
    end do
  end subroutine ref_set_particles

  subroutine ref_get_particles(ig, np, st, fl, color) bind(C, name='ref_get_particles')
    integer(c_int), value :: ig, np
    real(c_double), intent(out) :: st(np, 13), fl(np, 6), color(np)
    integer :: i
    do i = 1, np
      st(i, 1) = PART(i, ig)%XP; st(i, 2) = PART(i, ig)%YP; st(i, 3) = PART(i, ig)%ZP
      st(i, 4) = PART(i, ig)%UP; st(i, 5) = PART(i, ig)%VP; st(i, 6) = PART(i, ig)%WP
      st(i, 7) = PART(i, ig)%ELLQUAT%a; st(i, 8) = PART(i, ig)%ELLQUAT%b
      st(i, 9) = PART(i, ig)%ELLQUAT%c; st(i, 10) = PART(i, ig)%ELLQUAT%d
      st(i, 11) = PART(i, ig)%OMEGAX; st(i, 12) = PART(i, ig)%OMEGAY
      st(i, 13) = PART(i, ig)%OMEGAZ
      fl(i, 1) = PART(i, ig)%UFAP; fl(i, 2) = PART(i, ig)%VFAP; fl(i, 3) = PART(i, ig)%WFAP
      fl(i, 4) = PART(i, ig)%VORTXAP; fl(i, 5) = PART(i, ig)%VORTYAP
      fl(i, 6) = PART(i, ig)%VORTZAP
      color(i) = PART(i, ig)%COLOR
    end do
  end subroutine ref_get_particles

  !! INTERPH on one field component
  subroutine ref_interph(scheme, comp, nint, x, y, z, out) bind(C, name='ref_interph')
    integer(c_int), value :: scheme, comp, nint
    real(c_double), intent(in) :: x(nint), y(nint), z(nint)
    real(c_double), intent(out) :: out(nint)
    select case (comp)
    case (1); call INTERPH(scheme, XMESH, YMESH, ZMESH, UFLU, nint, x, y, z, out)
    case (2); call INTERPH(scheme, XMESH, YMESH, ZMESH, VFLU, nint, x, y, z, out)
    case (3); call INTERPH(scheme, XMESH, YMESH, ZMESH, WFLU, nint, x, y, z, out)
    case (4); call INTERPH(scheme, XMESH, YMESH, ZMESH, VORTICITY_X, nint, x, y, z, out)
    case (5); call INTERPH(scheme, XMESH, YMESH, ZMESH, VORTICITY_Y, nint, x, y, z, out)
    case (6); call INTERPH(scheme, XMESH, YMESH, ZMESH, VORTICITY_Z, nint, x, y, z, out)
    end select
  end subroutine ref_interph

  subroutine ref_force(ig, idp, up, vp, wp, uf, vf, wf, q, fd, fl) bind(C, name='ref_force')
    integer(c_int), value :: ig, idp
    real(c_double), value :: up, vp, wp, uf, vf, wf
    real(c_double), intent(in) :: q(4)
    real(c_double), intent(out) :: fd(3), fl(3)
    type(quaternion) :: qq
    real(kind=8), dimension(ndim, 1) :: FDRAG, FLIFT
    qq%a = q(1); qq%b = q(2); qq%c = q(3); qq%d = q(4)
    call FLUID_FORCE_ELLIPSOID(ig, idp, up, vp, wp, uf, vf, wf, qq, FDRAG, FLIFT)
    fd = FDRAG(:, 1); fl = FLIFT(:, 1)
  end subroutine ref_force

  !! FLUID_TORQUE_ELLIPSOID reads the stage-n velocity and the fluid at the
  !! particle from PART(1, ig): st13 and fl6 fill that record first.
  subroutine ref_torque(ig, idp, st13, fl6, ox, oy, oz, q, tp, tr) bind(C, name='ref_torque')
    integer(c_int), value :: ig, idp
    real(c_double), intent(in) :: st13(13), fl6(6), q(4)
    real(c_double), value :: ox, oy, oz
    real(c_double), intent(out) :: tp(3), tr(3)
    type(quaternion) :: qq
    real(kind=8), dimension(ndim, 1) :: TPITCH, TROT
    if (.not. allocated(PART)) then
      NPREF = 1
      allocate(PART(1, NIG))
    end if
    PART(1, ig)%UP = st13(4); PART(1, ig)%VP = st13(5); PART(1, ig)%WP = st13(6)
    PART(1, ig)%UFAP = fl6(1); PART(1, ig)%VFAP = fl6(2); PART(1, ig)%WFAP = fl6(3)
    PART(1, ig)%VORTXAP = fl6(4); PART(1, ig)%VORTYAP = fl6(5); PART(1, ig)%VORTZAP = fl6(6)
    qq%a = q(1); qq%b = q(2); qq%c = q(3); qq%d = q(4)
    call FLUID_TORQUE_ELLIPSOID(1, ig, idp, ox, oy, oz, qq, TPITCH, TROT)
    tp = TPITCH(:, 1); tr = TROT(:, 1)
  end subroutine ref_torque

  subroutine ref_quat(q, ox, oy, oz, dt) bind(C, name='ref_quat')
    real(c_double), intent(inout) :: q(4)
    real(c_double), value :: ox, oy, oz, dt
    type(quaternion) :: qq
    qq%a = q(1); qq%b = q(2); qq%c = q(3); qq%d = q(4)
    call QUATERNION_INTEGRATION(qq, ox, oy, oz, dt)
    q(1) = qq%a; q(2) = qq%b; q(3) = qq%c; q(4) = qq%d
  end subroutine ref_quat

  subroutine ref_wall(ig, idp, x, y, z, q, nrm, arm, delta_x, hit) bind(C, name='ref_wall')
    integer(c_int), value :: ig, idp
    real(c_double), value :: x, y, z
    real(c_double), intent(in) :: q(4)
    real(c_double), intent(out) :: nrm(3), arm(3), delta_x
    integer(c_int), intent(out) :: hit
    type(quaternion) :: qq
    real(kind=8), dimension(ndim, 1) :: wn, ia
    real(kind=8) :: dxw
    logical :: isect
    qq%a = q(1); qq%b = q(2); qq%c = q(3); qq%d = q(4)
    dxw = 0.
    call WALL_BOUNDARY_CONDITION_NEW(1, 1, ig, idp, x, y, z, qq, wn, ia, dxw, isect)
    nrm = wn(:, 1); arm = ia(:, 1); delta_x = dxw
    hit = 0
    if (isect) hit = 1
  end subroutine ref_wall

  subroutine ref_rebound(ig, idp, uvw, omg, q, arm, nrm) bind(C, name='ref_rebound')
    integer(c_int), value :: ig, idp
    real(c_double), intent(inout) :: uvw(3), omg(3)
    real(c_double), intent(in) :: q(4), arm(3), nrm(3)
    type(quaternion) :: qq
    real(kind=8), dimension(ndim, 1) :: ia, wn
    logical :: flag
    qq%a = q(1); qq%b = q(2); qq%c = q(3); qq%d = q(4)
    ia(:, 1) = arm; wn(:, 1) = nrm
    flag = .false.
    call WALL_ELLIPSOID_REBOUND(uvw(1), uvw(2), uvw(3), omg(1), omg(2), omg(3), &
                                qq, ia, wn, ig, idp, 1, flag)
  end subroutine ref_rebound

  subroutine ref_contact(ig, idp1, idp2, b1, q1, b2, q2, flag, e1, e2, cn) &
                         bind(C, name='ref_contact')
    integer(c_int), value :: ig, idp1, idp2
    real(c_double), intent(in) :: b1(3), q1(4), b2(3), q2(4)
    integer(c_int), intent(out) :: flag
    real(c_double), intent(out) :: e1(3), e2(3), cn(3)
    type(quaternion) :: qa, qb
    real(kind=8), dimension(ndim, 1) :: EPT1, EPT2, CNRM
    logical :: f
    qa%a = q1(1); qa%b = q1(2); qa%c = q1(3); qa%d = q1(4)
    qb%a = q2(1); qb%b = q2(2); qb%c = q2(3); qb%d = q2(4)
    f = .false.
    call ellipsoid_contact_detection(ig, idp1, idp2, b1(1), b1(2), b1(3), qa, &
                                     b2(1), b2(2), b2(3), qb, f, EPT1, EPT2, CNRM)
    flag = 0
    if (f) flag = 1
    e1 = EPT1(:, 1); e2 = EPT2(:, 1); cn = CNRM(:, 1)
  end subroutine ref_contact

  !! COLLISION_RESPONSE on two records (13 state values each)
  subroutine ref_response(ig, idp, s1, s2, r1, r2, nrm) bind(C, name='ref_response')
    integer(c_int), value :: ig, idp
    real(c_double), intent(inout) :: s1(13), s2(13)
    real(c_double), intent(in) :: r1(3), r2(3), nrm(3)
    type(PARTTYPE) :: p1, p2
    real(kind=8), dimension(ndim, 1) :: a1, a2, nn
    p1%UP = s1(4); p1%VP = s1(5); p1%WP = s1(6)
    p1%ELLQUAT%a = s1(7); p1%ELLQUAT%b = s1(8); p1%ELLQUAT%c = s1(9); p1%ELLQUAT%d = s1(10)
    p1%OMEGAX = s1(11); p1%OMEGAY = s1(12); p1%OMEGAZ = s1(13)
    p2%UP = s2(4); p2%VP = s2(5); p2%WP = s2(6)
    p2%ELLQUAT%a = s2(7); p2%ELLQUAT%b = s2(8); p2%ELLQUAT%c = s2(9); p2%ELLQUAT%d = s2(10)
    p2%OMEGAX = s2(11); p2%OMEGAY = s2(12); p2%OMEGAZ = s2(13)
    a1(:, 1) = r1; a2(:, 1) = r2; nn(:, 1) = nrm
    call COLLISION_RESPONSE(ig, idp, p1, p2, a1, a2, nn)
    s1(4) = p1%UP; s1(5) = p1%VP; s1(6) = p1%WP
    s1(11) = p1%OMEGAX; s1(12) = p1%OMEGAY; s1(13) = p1%OMEGAZ
    s2(4) = p2%UP; s2(5) = p2%VP; s2(6) = p2%WP
    s2(11) = p2%OMEGAX; s2(12) = p2%OMEGAY; s2(13) = p2%OMEGAZ
  end subroutine ref_response

  !! ADV_PARTICLE_ELLP (main.f90:466)
  subroutine ref_adv(ncycle) bind(C, name='ref_adv')
    integer(c_int), value :: ncycle
    call ADV_PARTICLE_ELLP(ncycle)
  end subroutine ref_adv

  !! COLLISION_INITIATION; returns the collision grid
  subroutine ref_collision_init(inum, rnum) bind(C, name='ref_collision_init')
    integer(c_int), intent(out) :: inum(4)
    real(c_double), intent(out) :: rnum(6)
    if (allocated(COLBOX)) deallocate(COLBOX)
    call COLLISION_INITIATION
    inum(1) = NXBOX; inum(2) = NYBOX; inum(3) = NZBOX; inum(4) = NBOX
    rnum(1) = XMINBOX; rnum(2) = YMINBOX; rnum(3) = ZMINBOX
    rnum(4) = DXBOX; rnum(5) = DYBOX; rnum(6) = DZBOX
  end subroutine ref_collision_init

  subroutine ref_connectivity(nbox_, con, ncon) bind(C, name='ref_connectivity')
    integer(c_int), value :: nbox_
    integer(c_int), intent(out) :: con(13, nbox_), ncon(nbox_)
    integer :: i, k
    do i = 1, nbox_
      ncon(i) = COLBOX(i)%NCON
      con(:, i) = 0
      do k = 1, COLBOX(i)%NCON
        con(k, i) = COLBOX(i)%CON(k)
      end do
    end do
  end subroutine ref_connectivity

  !! COLLISION_DETECTION_CONNECTIVITY only (class ig)
  subroutine ref_detect(ig, nmax, home, near, nclose) bind(C, name='ref_detect')
    integer(c_int), value :: ig, nmax
    integer(c_int), intent(out) :: home(nmax), near(nmax), nclose
    integer :: nf, nc
    nf = NPART_LOC(ig)
    call COLLISION_DETECTION_CONNECTIVITY(nf, PART(1:nf, ig)%XP, PART(1:nf, ig)%YP, &
                                          PART(1:nf, ig)%ZP, nc, nmax, home, near)
    nclose = nc
  end subroutine ref_detect

  !! COLLISION (main.f90:483) at NPROC = 1 for class ig: detection then
  !! interaction (collision.f90:81-171), with caller-sized pair buffers.
  subroutine ref_collision(ig, ncycle, time, nmax, home, near, nclose) &
                           bind(C, name='ref_collision')
    integer(c_int), value :: ig, ncycle, nmax
    real(c_double), value :: time
    integer(c_int), intent(out) :: home(nmax), near(nmax), nclose
    integer :: nf, nc, i
    nf = NPART_LOC(ig)
    do i = 1, nf
      PART(i, ig)%MYIDP = MYID
    end do
    NPNEIGH_COLL = 0
    call COLLISION_DETECTION_CONNECTIVITY(nf, PART(1:nf, ig)%XP, PART(1:nf, ig)%YP, &
                                          PART(1:nf, ig)%ZP, nc, nmax, home, near)
    call COLLISION_INTERACTION(time, ncycle, ig, nc, home, near)
    nclose = nc
  end subroutine ref_collision

  !! pair_depth at a given normal-angle pair x(2)
  subroutine ref_pair_depth(ig, idp1, idp2, x, b1, q1, b2, q2, cv, p1, p2) &
                            bind(C, name='ref_pair_depth')
    integer(c_int), value :: ig, idp1, idp2
    real(c_double), intent(in) :: x(2), b1(3), q1(4), b2(3), q2(4)
    real(c_double), intent(out) :: cv(3), p1(3), p2(3)
    type(quaternion) :: qa, qb
    real(kind=8), dimension(ndim, 1) :: bb1, bb2, c1, pt1, pt2
    qa%a = q1(1); qa%b = q1(2); qa%c = q1(3); qa%d = q1(4)
    qb%a = q2(1); qb%b = q2(2); qb%c = q2(3); qb%d = q2(4)
    bb1(:, 1) = b1; bb2(:, 1) = b2
    call pair_depth(ig, idp1, idp2, x, bb1, qa, bb2, qb, c1, pt1, pt2)
    cv = c1(:, 1); p1 = pt1(:, 1); p2 = pt2(:, 1)
  end subroutine ref_pair_depth

  !! level-1 collision statistics on: every later ref_collision whose cycle
  !! is a multiple of fout2 adds its MEAN_PART_COL (collision_interaction.f90:
  !! 330-441) to MEAN_TIME_PART_COL (:453), zeroed here.  The reference also
  !! writes each such pass to unit 560 + ... (fort.NNN in the current
  !! directory).
  subroutine ref_cstat_on(fout2_) bind(C, name='ref_cstat_on')
    integer(c_int), value :: fout2_
    FOUT2 = fout2_
    NSTAT = max(NSTAT, 8)
    LEVEL1_STPAR = .true.; STAT_TIME = .true.
    LEVEL1_STPDF = .false.; LEVEL0_STPAR = .false.; LEVELX_STPAR = .false.
    if (allocated(MEAN_TIME_PART_COL)) deallocate(MEAN_TIME_PART_COL)
    allocate(MEAN_TIME_PART_COL(NSTAT, POLYDISPMAX, POLYDISPMAX))
    MEAN_TIME_PART_COL = 0.0d0
  end subroutine ref_cstat_on

  !! STAT_PARTFLUID (stat_partfluid.f90) on the current PART, whose fluid
  !! at the particles (UFAP .. VORTZAP) the last ref_adv left: out(1:36,
  !! IDP, J) = MEAN_PARTFLUID(1:36, J, IDP), read back through
  !! MEAN_TIME_PARTFLUID (with STAT_TIME the routine copies it there, :460).
  !! The routine also writes a line per (J, IDP) to unit 540 + ... .
  subroutine ref_partfluid_stats(ncycle, out) bind(C, name='ref_partfluid_stats')
    use mpi
    integer(c_int), value :: ncycle
    real(c_double), intent(out) :: out(36, POLYDISPMAX, NIG)
    integer :: j, idp, ierr
    logical :: st, up
    !! the routine times itself and sums over ranks through MPI: one rank
    call MPI_INITIALIZED(up, ierr)
    if (.not. up) call MPI_INIT(ierr)
    st = STAT_TIME
    NSTAT = max(NSTAT, 36)
    STAT_TIME = .true.
    if (allocated(MEAN_TIME_PARTFLUID)) deallocate(MEAN_TIME_PARTFLUID)
    allocate(MEAN_TIME_PARTFLUID(NSTAT, NIG, POLYDISPMAX))
    MEAN_TIME_PARTFLUID = 0.0d0
    call STAT_PARTFLUID(ncycle, 0.0d0)
    do j = 1, NIG
      do idp = 1, POLYDISPMAX
        out(:, idp, j) = MEAN_TIME_PARTFLUID(1:36, j, idp)
      end do
    end do
    STAT_TIME = st
  end subroutine ref_partfluid_stats

  !! MEAN_TIME_PART_COL(1:8, IDP1, IDP2)
  subroutine ref_cstat_get(out) bind(C, name='ref_cstat_get')
    real(c_double), intent(out) :: out(8, POLYDISPMAX, POLYDISPMAX)
    out = MEAN_TIME_PART_COL(1:8, :, :)
  end subroutine ref_cstat_get

end module gpref_driver

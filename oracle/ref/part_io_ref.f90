!! part_io_ref.f90 -- runs the reference's own writers and readers of its
!! on-disk formats (SAVE_PARTICLE with ISAVEPART = 2, the ISAVEPART = 2
!! branch of INITIATION_PARTICLE_POSITION, SAVE_MPIIO, READ_MPIIO; compiled
!! unmodified from the reference's source/ directory by oracle/ref/Makefile)
!! on one MPI rank.  TEST INFRASTRUCTURE: it checks the library's readers
!! and writers of the same formats (tests/test_file_formats.py) and is never
!! part of the product.
!!
!! Usage (all files in the current directory):
!!   part_io_ref psave <in.raw> <ncycle>   classes from in.raw -> SAVE_PARTICLE
!!   part_io_ref pread <in.raw> <out.raw>  part_pJJ.ini -> out.raw (shapes from in.raw)
!!   part_io_ref fsave <in.raw> <name>     one field component -> SAVE_MPIIO
!!   part_io_ref fread <name> <n> <out.raw> READ_MPIIO -> out.raw
!!   part_io_ref pstat <in.raw> <out.raw>  classes from in.raw -> STAT_PARTICLE
!!                                         (level 1) -> MEAN_PART(1:41, J, IDP)
!!   part_io_ref ppdf <in.raw> <out.raw>   classes from in.raw -> STAT_PARTICLE
!!                                         at NCYCLE = 500 with LEVEL1_STPDF ->
!!                                         MEAN_PART_PDF(1:12, J, IDP, 1:NPDF)
!!                                         (through MEAN_TIME_PART_PDF)
!!   part_io_ref plag <in.raw> <out.raw>   three states per class (cycles 1, 2,
!!                                         3; FOUT2 = 1, NT0(1) = 1) ->
!!                                         STAT_PARTICLE levels 1 and 2 ->
!!                                         RPX/RPY/RPZ_LOC(1:3, J, 1, IDP)
!!
!! in.raw for psave / pread (stream, native endianness): int32 nig, npoly,
!! npart, n; double L; int32 npclass(npoly, nig); double rhop(npoly, nig),
!! emaj(npoly, nig), apr(nig); per class: double state(npart, 13), int32
!! idp(npart).  out.raw of pread: per class: int32 nloc; int32
!! npclass(npoly); double rhop(npoly), emaj(npoly), apr; double
!! state(nloc, 13); int32 idp(nloc).  Field raw files: int32 n, then
!! double f(n, n, n).
program part_io_ref
  use mpi
  use iso_c_binding
  use DNS_DIM
  use PARAM_PHYS
  use PARTICLE_PARALLEL
  use MPI_STRUCTURES
  use STATISTICS
  use gpref_driver
  implicit none
  character(len=256) :: mode, a1, a2
  character(len=40) :: fname
  integer :: mpierr, ncls, npoly, npart, n, j, i, c, ncycle
  real(8) :: L
  integer, allocatable :: ncl(:, :), idp(:)
  real(8), allocatable :: rh(:, :), em(:, :), apr(:), st(:, :), grav(:), ec(:), mu(:)
  real(8), allocatable :: f3(:, :, :)

  call get_command_argument(1, mode)
  call get_command_argument(2, a1)
  call get_command_argument(3, a2)
  call MPI_INIT(mpierr)
  DEBUG = .false.

  select case (trim(mode))
  case ('psave', 'pread', 'pstat', 'ppdf', 'plag')
    open(unit=40, file=trim(a1), access='stream', form='unformatted', status='old')
    read(40) ncls, npoly, npart, n
    read(40) L
    allocate(ncl(npoly, ncls), rh(npoly, ncls), em(npoly, ncls), apr(ncls))
    allocate(grav(ncls), ec(ncls), mu(ncls), st(npart, 13), idp(npart))
    read(40) ncl, rh, em, apr
    grav = 0.0d0; ec = 1.0d0; mu = 0.0d0
    call ref_setup(n, L, 1.0d-2, 1.0d0, 1.0d-3, 1.0d0, 0.0d0, 3, ncls, npoly, &
                   transpose(em), apr, transpose(rh), grav, ec, mu)
    if (allocated(NPCLASS)) deallocate(NPCLASS)
    allocate(NPCLASS(NIG, POLYDISPMAX))
    NPCLASS = transpose(ncl)
    NPART_FULL = npart
    ISAVEPART = 2
    if (trim(mode) == 'psave') then
      do j = 1, ncls
        read(40) st, idp
        call ref_set_particles(j, npart, st, idp)
      end do
      close(40)
      read(a2, *) ncycle
      call SAVE_PARTICLE(ncycle)
    else if (trim(mode) == 'ppdf') then
      do j = 1, ncls
        read(40) st, idp
        call ref_set_particles(j, npart, st, idp)
      end do
      close(40)
      !! one PDF pass (the routine's NCYCLE multiple of 500); MEAN_TIME_PART_PDF
      !! starts at zero, so it holds that pass's MEAN_PART_PDF
      NSTAT = 100
      LEVEL1_STPAR = .false.; LEVEL2_STPAR = .false.; LEVEL3_STPAR = .false.
      LEVEL1_STPDF = .true.; STAT_TIME = .false.
      if (allocated(MEAN_TIME_PART_PDF)) deallocate(MEAN_TIME_PART_PDF)
      allocate(MEAN_TIME_PART_PDF(NSTAT, NIG, POLYDISPMAX, NPDF))
      MEAN_TIME_PART_PDF = 0.0d0
      NEVEN_PDF = 0
      NCYCLEMAX = 1000
      call STAT_PARTICLE(500, 0.0d0)
      open(unit=41, file=trim(a2), access='stream', form='unformatted', status='replace')
      do j = 1, ncls
        do i = 1, npoly
          write(41) MEAN_TIME_PART_PDF(1:12, j, i, 1:NPDF)
        end do
      end do
      close(41)
    else if (trim(mode) == 'plag') then
      !! levels 1 and 2 on three consecutive statistics cycles; the state of
      !! each class is replaced between the calls (PART's UPT0..WPT0 kept)
      NSTAT = 100
      LEVEL1_STPAR = .true.; LEVEL2_STPAR = .true.; LEVEL3_STPAR = .false.
      LEVEL1_STPDF = .false.; STAT_TIME = .false.
      FOUT2 = 1
      NCYCLEMAX = 3
      DIMLGR = 3
      NT0(:) = 1000
      NT0(1) = 1
      if (allocated(TIME_LGR)) deallocate(TIME_LGR, RPX_LOC, RPY_LOC, RPZ_LOC)
      allocate(TIME_LGR(DIMLGR, NBLGRMAX))
      allocate(RPX_LOC(DIMLGR, NIG, NBLGRMAX, POLYDISPMAX))
      allocate(RPY_LOC(DIMLGR, NIG, NBLGRMAX, POLYDISPMAX))
      allocate(RPZ_LOC(DIMLGR, NIG, NBLGRMAX, POLYDISPMAX))
      TIME_LGR = 0.0d0; RPX_LOC = 0.0d0; RPY_LOC = 0.0d0; RPZ_LOC = 0.0d0
      do c = 1, 3
        do j = 1, ncls
          read(40) st, idp
          call ref_set_particles(j, npart, st, idp)
        end do
        call STAT_PARTICLE(c, 0.1d0 * c)
      end do
      close(40)
      open(unit=41, file=trim(a2), access='stream', form='unformatted', status='replace')
      do j = 1, ncls
        do i = 1, npoly
          write(41) RPX_LOC(1:3, j, 1, i), RPY_LOC(1:3, j, 1, i), RPZ_LOC(1:3, j, 1, i)
        end do
      end do
      close(41)
    else if (trim(mode) == 'pstat') then
      do j = 1, ncls
        read(40) st, idp
        call ref_set_particles(j, npart, st, idp)
      end do
      close(40)
      !! one level-1 pass; with STAT_TIME the routine adds MEAN_PART to the
      !! zeroed MEAN_TIME_PART, which is read back at full precision
      NSTAT = 100
      LEVEL1_STPAR = .true.; LEVEL2_STPAR = .false.; LEVEL3_STPAR = .false.
      LEVEL1_STPDF = .false.; STAT_TIME = .true.
      if (allocated(MEAN_TIME_PART)) deallocate(MEAN_TIME_PART)
      allocate(MEAN_TIME_PART(NSTAT, NIG, POLYDISPMAX))
      MEAN_TIME_PART = 0.0d0
      call STAT_PARTICLE(1, 0.0d0)
      open(unit=41, file=trim(a2), access='stream', form='unformatted', status='replace')
      do j = 1, ncls
        do i = 1, npoly
          write(41) MEAN_TIME_PART(1:41, j, i)
        end do
      end do
      close(41)
    else
      close(40)
      !! room for every particle of the file in each class
      do j = 1, ncls
        st = 0.0d0; idp = 1
        call ref_set_particles(j, npart, st, idp)
      end do
      INIT_PART_POSITION = 3
      call INITIATION_PARTICLE_POSITION
      open(unit=41, file=trim(a2), access='stream', form='unformatted', status='replace')
      do j = 1, ncls
        write(41) NPART_LOC(j)
        write(41) (NPCLASS(j, i), i = 1, npoly)
        write(41) (RHOP(j, i), i = 1, npoly), (EMAJ_PART(j, i), i = 1, npoly), APR_PART(j)
        write(41) (PART(i, j)%XP, i = 1, NPART_LOC(j)), (PART(i, j)%YP, i = 1, NPART_LOC(j)), &
                  (PART(i, j)%ZP, i = 1, NPART_LOC(j)), (PART(i, j)%UP, i = 1, NPART_LOC(j)), &
                  (PART(i, j)%VP, i = 1, NPART_LOC(j)), (PART(i, j)%WP, i = 1, NPART_LOC(j)), &
                  (PART(i, j)%ELLQUAT%a, i = 1, NPART_LOC(j)), &
                  (PART(i, j)%ELLQUAT%b, i = 1, NPART_LOC(j)), &
                  (PART(i, j)%ELLQUAT%c, i = 1, NPART_LOC(j)), &
                  (PART(i, j)%ELLQUAT%d, i = 1, NPART_LOC(j)), &
                  (PART(i, j)%OMEGAX, i = 1, NPART_LOC(j)), &
                  (PART(i, j)%OMEGAY, i = 1, NPART_LOC(j)), &
                  (PART(i, j)%OMEGAZ, i = 1, NPART_LOC(j))
        write(41) (PART(i, j)%IDP, i = 1, NPART_LOC(j))
      end do
      close(41)
    end if
  case ('fsave', 'fread')
    if (trim(mode) == 'fsave') then
      open(unit=40, file=trim(a1), access='stream', form='unformatted', status='old')
      read(40) n
    else
      read(a2, *) n
    end if
    allocate(apr(1), rh(1, 1), em(1, 1), grav(1), ec(1), mu(1))
    apr = 2.0d0; rh = 1.0d3; em = 1.0d-2; grav = 0.0d0; ec = 1.0d0; mu = 0.0d0
    call ref_setup(n, 1.0d0, 1.0d-2, 1.0d0, 1.0d-3, 1.0d0, 0.0d0, 3, 1, 1, &
                   em, apr, rh, grav, ec, mu)
    !! rank 0 of one: SAVE_MPIIO writes the header on rank 0 only
    !! (ref_setup's MYID = 1 only skips the hot path's rank-0 timers)
    MYID = 0
    allocate(f3(n, n, n))
    if (trim(mode) == 'fsave') then
      read(40) f3
      close(40)
      UFLU = 0.0d0
      UFLU(1:n, 1:n, 1:n) = f3
      fname = trim(a2)
      call SAVE_MPIIO(UFLU, fname)
    else
      fname = trim(a1)
      UFLU = 0.0d0
      call READ_MPIIO(UFLU, fname)
      call get_command_argument(4, a2)
      f3 = UFLU(1:n, 1:n, 1:n)
      open(unit=41, file=trim(a2), access='stream', form='unformatted', status='replace')
      write(41) n
      write(41) f3
      close(41)
    end if
  case default
    print *, 'usage: part_io_ref psave|pread|fsave|fread ...'
    call MPI_FINALIZE(mpierr)
    stop 2
  end select
  call MPI_FINALIZE(mpierr)
end program part_io_ref

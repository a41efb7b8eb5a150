!! rng_driver.f90 -- writes the sequences of the reference's RAN1 and Gauss
!! (adv_forcing.f90) from a given seed, each from a fresh generator state
!! (RANDOM_VAR's IY = 0, ISET = 0).  TEST INFRASTRUCTURE
!! (tests/golden/make_rng_golden.py).
!!
!! usage: rng_ref <seed> <n> <output.bin>
!!   output: float64 ran1(n), gauss(n)
program rng_driver
  use RANDOM_VAR
  implicit none
  real(kind=8), external :: RAN1, Gauss
  character(len=512) :: a, fout
  integer :: seed, n, i, idum
  real(kind=8), allocatable :: r(:), g(:)

  call get_command_argument(1, a)
  read(a, *) seed
  call get_command_argument(2, a)
  read(a, *) n
  call get_command_argument(3, fout)
  allocate(r(n), g(n))
  IY = 0
  ISET = 0
  idum = seed
  do i = 1, n
    r(i) = RAN1(idum)
  end do
  IY = 0
  ISET = 0
  idum = seed
  do i = 1, n
    g(i) = Gauss(idum)
  end do
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) r
  write(11) g
  close(11)
end program rng_driver

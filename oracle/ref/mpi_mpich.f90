!! The `mpi` module for flang, built from MPICH's own mpif.h as installed in
!! this image (the shipped mpi.mod is gfortran-format and unreadable by
!! flang).  No MPI definition is written here: every constant and interface
!! comes from the installed header.
module mpi
include 'mpif.h'
end module mpi

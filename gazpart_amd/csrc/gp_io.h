// gp_io.h -- the reference's on-disk formats either side of the step (host
// code, no GPU): the per-class particle files of SAVE_PARTICLE / the
// ISAVEPART = 2 restart read of INITIATION_PARTICLE_POSITION, and the
// velocity-component files of SAVE_MPIIO / READ_MPIIO.
//
// Particle file (save_particle.f90:161-185, read back by
// initiation_particle_position.f90:430-514): Fortran unformatted sequential,
// every record framed by its byte length as a 4-byte little-endian integer
// before and after it (the Fortran compilers' default framing):
//   NPART_FULL (int32) | NPCLASS(J,1:npoly) (int32) | RHOP(J,1:npoly) |
//   EMAJ_PART(J,1:npoly) | APR_PART(J) | then 14 records of NPART_FULL
//   doubles: x y z u v w qa qb qc qd omega_x omega_y omega_z idp
// (reals are real(kind=8); idp, the size index, travels as a double).
// Records longer than 2^31 - 1 bytes would need the compilers' split
// records and are refused.
//
// Field file (MPI_structures.f90:434-493 and 571-621): three int32 NX NX NX,
// then NX^3 doubles in Fortran order (x fastest), no framing.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>

namespace gpio {

constexpr int64_t kMaxRecord = 0x7fffffff;

inline bool rec_write(std::FILE* f, const void* p, int64_t n) {
  if (n < 0 || n > kMaxRecord) return false;
  const int32_t m = (int32_t)n;
  if (std::fwrite(&m, 4, 1, f) != 1) return false;
  if (n && std::fwrite(p, 1, (size_t)n, f) != (size_t)n) return false;
  return std::fwrite(&m, 4, 1, f) == 1;
}

// the next record, which must be exactly n bytes long
inline bool rec_read(std::FILE* f, void* p, int64_t n) {
  int32_t m1 = 0, m2 = 0;
  if (std::fread(&m1, 4, 1, f) != 1 || (int64_t)m1 != n) return false;
  if (n && std::fread(p, 1, (size_t)n, f) != (size_t)n) return false;
  return std::fread(&m2, 4, 1, f) == 1 && m2 == m1;
}

// length of the next record, without consuming it
inline bool rec_peek(std::FILE* f, int64_t* n) {
  int32_t m = 0;
  if (std::fread(&m, 4, 1, f) != 1 || m < 0) return false;
  if (std::fseek(f, -4, SEEK_CUR) != 0) return false;
  *n = m;
  return true;
}

struct File {
  std::FILE* f = nullptr;
  File(const char* path, const char* mode) { f = path ? std::fopen(path, mode) : nullptr; }
  ~File() {
    if (f) std::fclose(f);
  }
  // close and report write errors
  bool close() {
    const bool ok = f && std::fclose(f) == 0;
    f = nullptr;
    return ok;
  }
};

}  // namespace gpio

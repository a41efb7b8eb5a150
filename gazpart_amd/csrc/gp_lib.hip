// gp_lib.hip -- kernels and C-ABI of the MI355X Lagrangian ellipsoid step.
//
// Step structure (one HIP stream per context):
//   k_advance      gather the 6 fields at x^n (lag3/lag4 stencil, LDS-free,
//                  interleaved node layout), RK2 + drag/torque + quaternion
//                  + x-walls + periodic wrap, then the particle's collision
//                  cell and its slot in the cell (atomic count) -- fused.
//   scan           cell counts -> cell offsets
//   k_scatter/rank stable counting sort by (cell, particle index): the
//                  reference bins particles in index order
//                  (collision_detection_connectivity.f90:95-118)
//   k_pair_count   candidate pairs per home particle, scan -> list offsets
//                  (the list order of collision_detection_connectivity.f90:
//                  145-207 is cell ascending, home in index order, self pairs,
//                  then the 13 half-shell cells in CON order)
//   k_prefilter    bounding-sphere test; survivors keep their list index
//   k_contact      FP64 Levenberg-Marquardt contact solve + approach test
//   k_greedy_*     one collision per particle in list order + impulses
// Particles are stored cell-sorted: the next advance reads them through the
// sort permutation, so the field gather of neighbouring threads touches
// neighbouring stencils.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gazpart_gpu.h"
#include "gp_contact.h"
#include "gp_device.h"
#include "gp_io.h"
#include "gp_mlp.h"
#include "gp_scan.h"

namespace gp {

constexpr int BLOCK = 256;

// ---------------------------------------------------------------------------
// device data structures
// ---------------------------------------------------------------------------
template <typename VT>
struct PState {
  double *x, *y, *z;
  VT *u, *v, *w;
  double *qa, *qb, *qc, *qd;
  VT *ox, *oy, *oz;
  float *rm; // largest semi-axis of the particle's species, rounded up to float
  int *sp;   // species row
  int *id;   // reference index (0-based)
  int *grp;  // ensemble: particles of different ensembles never collide
};

struct CellGrid {
  int nx, ny, nz, nbox;
  double xmin, ymin, zmin, dx, dy, dz;
};

// bricks of b x b x b collision cells, row-major: the order in which the
// advance visits the particles (a compact set of stencils per workgroup).
// With tile = 1 the bricks are tiles of b x b x b field cells instead (by
// the stencil's base node, floor(x / dx)), each advanced by one workgroup
// that stages the tile's field nodes in LDS (k_gather_tiled).
struct BrickGrid {
  int b, nx, ny, nz, n;
  int tile;     // 1: field tiles
  int nn;       // field nodes per axis (tiles)
  double dx;    // field spacing (tiles)
};

struct ActiveRec {   // a contact that is approaching (flag and NVREL < ZERO)
  long long k;       // index in the candidate list (its order)
  int i, j;          // particle slots
  double r1[3], r2[3], cn[3];
};

struct WorkRec {     // a candidate pair that passed the prefilter
  long long k;
  int i, j;
};

// Solving ahead.  The contact solves of a step are, nearly all of them,
// pairs that were solved in the step before too (near contacts persist for
// many steps).  So every pair that survived the prefilters of step s is
// listed by the slots its two particles take in the next state (the sort of
// step s fixes the visiting order of the next advance, and with it those
// slots), and at the start of step s+1 (its particles all advanced) those
// pairs are solved again from the step-s+1 state, beside that step's sort
// and prefilters; step s+1's filter then sends a pair found in that list,
// (home, near) slots equal, to the results instead of to its own solve.  A
// pair's solve depends only on the two particles' state, so every result is
// the one the step's own solve would give, whatever the list holds.
struct EarlyRec {     // a contact solved ahead
  int i, j;           // storage slots of the step it was solved for
  int approach, nfev;
  double r1[3], r2[3], cn[3];
};

struct MergeRec {     // a surviving pair whose solve was done ahead
  long long k;        // its key in this step's list
  int e, pad;         // its record among the results solved ahead
};

// k_contact's mode for solving ahead: pairs by slot
struct EarlyIn {
  const int* pair = nullptr;                      // home, near slots per pair
  const unsigned long long* n = nullptr;
  EarlyRec* res = nullptr;                        // null: the step's own solve
};

// the filter's side of it: this step's surviving pairs listed (and put in
// a hash table) for the next step, and looked up in the previous step's
struct EarlyTab {
  const unsigned long long* pkeys = nullptr;      // previous step's table
  const int* pvals = nullptr;
  unsigned long long* nkeys = nullptr;            // this step's (cleared)
  int* nvals = nullptr;
  unsigned long long mask = 0;                    // table size - 1
  const int* inv = nullptr;                       // slot -> slot in the next state
  int* pair = nullptr;                            // this step's pairs (next slots)
  int* slot = nullptr;                            // their table slots
  unsigned long long* npair = nullptr;
  MergeRec* merge = nullptr;                      // found in the previous
  unsigned long long* nmerge = nullptr;
  unsigned long long cap = 0;
};

// a contact (flag true) of a statistics step, with the quantities
// collision_interaction.f90:330-345 sums for it when the ordered pass
// reaches it with both particles still free
struct CollStatRec {
  long long k;        // candidate-list order
  int i, j;           // particle slots (HOME, NEAR)
  int bin;            // (class, IDP1, IDP2)
  int approach;       // NVREL < ZERO: a collision when reached
  double depth, nvrel, nrm, theta;
};

// the collision-statistics buffers as the kernels see them; rec == nullptr
// when the statistics are off
struct CStat {
  CollStatRec* rec;
  unsigned long long* cnt;   // 0 records 1 statistics steps
  unsigned long long cap;
  long long* cycle;
  double* acc;               // nbin x 6 sums of the current step
  double* last;              // nbin x 8
  double* tsum;              // nbin x 8
  int fout2, npoly, nbin;
};

__device__ __forceinline__ bool cstat_step(const CStat& cs) {
  return cs.rec && (*cs.cycle % (long long)cs.fout2) == 0;
}

__device__ __forceinline__ int block_index(int xcd_remap) {
  // AMD dispatches consecutive workgroups round-robin over the 8 XCDs;
  // remapping gives each XCD a contiguous range of (cell-sorted) particles
  // so that neighbouring stencils share that XCD's L2.
  // (the launch pads the grid to a multiple of 8 blocks when remapping)
  const int b = blockIdx.x;
  if (!xcd_remap) return b;
  const int per = gridDim.x >> 3;
  return (b & 7) * per + (b >> 3);
}

__device__ __forceinline__ long long node_index(int n4, int i, int j, int k) {
  return ((long long)(k + 1) * n4 + (j + 1)) * n4 + (i + 1);
}

// Lagrange weights of interp_lag3.f90:196-211 / interp_lag4.f90:148-166
template <typename CT, int S>
__device__ __forceinline__ void lweights(CT a, CT* f) {
  if constexpr (S == 4) {
    f[0] = -(CT(1.0 / 6.0) * (a - CT(2)) * (a - CT(1)) * a);
    f[1] = CT(0.5) * (a - CT(2)) * (a - CT(1)) * (a + CT(1));
    f[2] = -(CT(0.5) * (a - CT(2)) * a * (a + CT(1)));
    f[3] = CT(1.0 / 6.0) * (a - CT(1)) * a * (a + CT(1));
  } else {
    f[0] = CT(1.0 / 24.0) * (a - CT(2)) * (a - CT(1)) * a * (a + CT(1));
    f[1] = -(CT(1.0 / 6.0) * (a - CT(2)) * (a - CT(1)) * a * (a + CT(2)));
    f[2] = CT(0.25) * (a - CT(2)) * (a - CT(1)) * (a + CT(1)) * (a + CT(2));
    f[3] = -(CT(1.0 / 6.0) * (a - CT(2)) * a * (a + CT(1)) * (a + CT(2)));
    f[4] = CT(1.0 / 24.0) * (a - CT(1)) * a * (a + CT(1)) * (a + CT(2));
  }
}

// one x-row of S nodes x NC components, contiguous in the interleaved
// layout; NC = 6 rows start 16-byte (FP64) / 8-byte (FP32) aligned and are
// read as double2 / float2, NC = 1 rows element by element
template <typename FT, int S, int NC> struct Row;
template <int S> struct Row<double, S, 6> {
  double v[S * 6];
  __device__ __forceinline__ void load(const double* __restrict__ p) {
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < S * 3; ++i) {
      double2 t = q[i];
      v[2 * i] = t.x; v[2 * i + 1] = t.y;
    }
  }
};
template <int S> struct Row<float, S, 6> {
  float v[S * 6];
  __device__ __forceinline__ void load(const float* __restrict__ p) {
    const float2* q = reinterpret_cast<const float2*>(p);
#pragma unroll
    for (int i = 0; i < S * 3; ++i) {
      float2 t = q[i];
      v[2 * i] = t.x; v[2 * i + 1] = t.y;
    }
  }
};
template <typename FT, int S> struct Row<FT, S, 1> {
  FT v[S];
  __device__ __forceinline__ void load(const FT* __restrict__ p) {
#pragma unroll
    for (int i = 0; i < S; ++i) v[i] = p[i];
  }
};

// INTERP_LAG3 / INTERP_LAG4 of all six fields at once (interp_lag3.f90:
// 119-290, interp_lag4.f90:84-312).  The field carries ghost planes in x as
// well (index period n, the reference's x wrap of interp_lag3.f90:132-154),
// so the stencil is branch-free.  For lag4 the reference's FYP2 row reads
// J+1 (interp_lag4.f90:196,222,248,274,300): that row gets weight FYP1+FYP2.
// the stencil of a position: the reference's 1-based node left of it on
// each axis (i0, j0, k0) and the Lagrange weights.  False outside the
// interpolation range (NaN included).
template <typename CT, int S>
struct Stencil {
  int i0, j0, k0;
  CT fx[S], fz[S], wy[4];
};

template <typename CT, int S>
__device__ __forceinline__ bool stencil_at(int n, double dx, double x, double y, double z,
                                           Stencil<CT, S>& w) {
  if (!(x == x && y == y && z == z)) return false;
  const double tx = x / dx, ty = y / dx, tz = z / dx;
  if (!(tx > -1.0 && ty > -1.0 && tz > -1.0 && tx < n && ty < n && tz < n)) return false;
  w.i0 = (int)tx + 1; w.j0 = (int)ty + 1; w.k0 = (int)tz + 1;
  if (w.i0 < 1 || w.i0 > n || w.j0 < 1 || w.j0 > n || w.k0 < 1 || w.k0 > n) return false;
  CT fy[S];
  lweights<CT, S>(CT((x - (double)(w.i0 - 1) * dx) / dx), w.fx);
  lweights<CT, S>(CT((y - (double)(w.j0 - 1) * dx) / dx), fy);
  lweights<CT, S>(CT((z - (double)(w.k0 - 1) * dx) / dx), w.fz);
  // lag4: rows j0-2..j0+1 (J+2 never read)
  if constexpr (S == 4) {
    w.wy[0] = fy[0]; w.wy[1] = fy[1]; w.wy[2] = fy[2]; w.wy[3] = fy[3];
  } else {
    w.wy[0] = fy[0]; w.wy[1] = fy[1]; w.wy[2] = fy[2]; w.wy[3] = fy[3] + fy[4];
  }
  return true;
}

// first stencil node relative to i0 (and j0, k0): i0-1..i0+2 (lag3),
// i0-2..i0+2 (lag4); rows j0+lo..j0+lo+3 in both
template <int S> constexpr int stencil_lo() { return (S == 4) ? -1 : -2; }

// the weighted sum over the stencil; row(jy, kz) points at the S nodes x NC
// values of row jy of plane kz (the same arithmetic, in the same order,
// wherever the rows are read from)
template <typename FT, typename CT, int S, int NC, typename RowAt>
__device__ __forceinline__ void stencil_sum(const Stencil<CT, S>& w, const RowAt& row,
                                            CT out[NC]) {
  CT acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = CT(0);
#pragma unroll
  for (int kz = 0; kz < S; ++kz) {
    CT sy[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) sy[c] = CT(0);
#pragma unroll
    for (int jy = 0; jy < 4; ++jy) {
      Row<FT, S, NC> r;
      r.load(row(jy, kz));
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        CT sx = 0;
#pragma unroll
        for (int ix = 0; ix < S; ++ix) sx += w.fx[ix] * CT(r.v[ix * NC + c]);
        sy[c] += w.wy[jy] * sx;
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += w.fz[kz] * sy[c];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = acc[c];
}

template <typename FT, typename CT, int S, int NC>
__device__ __forceinline__ void stencil_sum_global(const FT* __restrict__ f, int n,
                                                   const Stencil<CT, S>& w, CT out[NC]) {
  constexpr int lo = stencil_lo<S>();
  const int n4 = n + 4;
  stencil_sum<FT, CT, S, NC>(
      w,
      [&](int jy, int kz) {
        return f + node_index(n4, w.i0 + lo, w.j0 + lo + jy, w.k0 + lo + kz) * NC;
      },
      out);
}

template <typename FT, typename CT, int S, int NC = 6>
__device__ __forceinline__ bool gather6(const FT* __restrict__ f, int n, double dx, double x,
                                        double y, double z, CT out[NC]) {
  Stencil<CT, S> w;
  if (!stencil_at<CT, S>(n, dx, x, y, z, w)) return false;
  stencil_sum_global<FT, CT, S, NC>(f, n, w, out);
  return true;
}

// fluid at the particles from a buffer filled before the advance
// (k_gather_tiled): the six fields of visiting position i at c * np + i.
// Positions outside the interpolation range were reported by that kernel
// and hold zeros, as the direct gather leaves them.
template <typename CT>
struct BufferGather {
  const CT* __restrict__ fl;
  int np, i;
  __device__ __forceinline__ bool operator()(double, double, double, CT out[6]) const {
#pragma unroll
    for (int c = 0; c < 6; ++c) out[c] = fl[(long long)c * np + i];
    return true;
  }
};

// the six fields from the field in HBM
template <typename FT, typename CT, int S>
struct GlobalGather {
  const FT* __restrict__ f;
  int n;
  double dx;
  __device__ __forceinline__ bool operator()(double x, double y, double z, CT out[6]) const {
    return gather6<FT, CT, S>(f, n, dx, x, y, z, out);
  }
};

// the staged field block of k_gather_tiled (dynamic LDS)
extern __shared__ __align__(16) unsigned char tile_lds[];

// the six fields from a tile's staged nodes in LDS: block of Ex x Ey x Ez
// nodes whose first node is the stencil's first node for base node
// (tbx, tby, tbz) (the tile's first cell); positions whose base node lies
// outside the tile read HBM
template <typename FT, typename CT, int S>
struct TileGather {
  const FT* __restrict__ f;
  int n;
  double dx;
  int tbx, tby, tbz, T, Ex, Ey;
  __device__ __forceinline__ bool operator()(double x, double y, double z, CT out[6]) const {
    Stencil<CT, S> w;
    if (!stencil_at<CT, S>(n, dx, x, y, z, w)) return false;
    const int bx = w.i0 - 1 - tbx, by = w.j0 - 1 - tby, bz = w.k0 - 1 - tbz;
    if ((unsigned)bx < (unsigned)T && (unsigned)by < (unsigned)T && (unsigned)bz < (unsigned)T)
      stencil_sum<FT, CT, S, 6>(
          w,
          [&](int jy, int kz) {
            return reinterpret_cast<const FT*>(tile_lds) +
                   (((bz + kz) * Ey + (by + jy)) * Ex + bx) * 6;
          },
          out);
    else
      stencil_sum_global<FT, CT, S, 6>(f, n, w, out);
    return true;
  }
};

// collision cell of a position (collision_detection_connectivity.f90:74-78):
// IEEE (x - xmin) / dx truncated, 1-based, as a 0-based box.  A position
// outside the grid (the reference prints a warning and writes out of
// bounds) raises *oob and is clamped to the nearest box, so that every
// particle keeps a slot in the sort.
__device__ __forceinline__ int axis_cell(double x, double x0, double d, int nmax, bool* oob) {
  const double t = (x - x0) / d;
  if (!(t > -1.0 && t < (double)nmax)) {   // also NaN
    *oob = true;
    return (t >= (double)nmax) ? nmax : 1;
  }
  const int i = (int)t + 1;
  if (i < 1 || i > nmax) {
    *oob = true;
    return i < 1 ? 1 : nmax;
  }
  return i;
}
__device__ __forceinline__ int cell_of(const CellGrid& g, double x, double y, double z,
                                       bool* oob) {
  *oob = false;
  const int i0 = axis_cell(x, g.xmin, g.dx, g.nx, oob);
  const int j0 = axis_cell(y, g.ymin, g.dy, g.ny, oob);
  const int k0 = axis_cell(z, g.zmin, g.dz, g.nz, oob);
  return (i0 - 1) + (j0 - 1) * g.nx + (k0 - 1) * g.nx * g.ny;
}

__device__ __forceinline__ int brick_of(const CellGrid& g, const BrickGrid& bg, int c) {
  const int i = c % g.nx, j = (c / g.nx) % g.ny, k = c / (g.nx * g.ny);
  return i / bg.b + bg.nx * (j / bg.b + bg.ny * (k / bg.b));
}

// field tile along one axis: the stencil's base node floor(x / dx) (as
// stencil_at takes it) over the tile edge; positions outside the
// interpolation range go to tile 0 (their advance reports the error)
__device__ __forceinline__ int tile_axis(double x, const BrickGrid& bg, int nt) {
  const double t = x / bg.dx;
  const int b = (t > -1.0 && t < (double)bg.nn) ? (int)t : 0;
  return min(b / bg.b, nt - 1);
}

// the brick of a particle at (x, y, z) in collision cell c
__device__ __forceinline__ int brick_at(const CellGrid& g, const BrickGrid& bg, int c, double x,
                                        double y, double z) {
  if (bg.tile)
    return tile_axis(x, bg, bg.nx) +
           bg.nx * (tile_axis(y, bg, bg.ny) + bg.ny * tile_axis(z, bg, bg.nz));
  return brick_of(g, bg, c);
}

// float not below the double (the bounding-sphere test stays conservative)
__device__ __forceinline__ float float_up(double r) {
  float f = (float)r;
  if ((double)f < r) f = nextafterf(f, __builtin_huge_valf());
  return f;
}

// the 13-cell half shell of COLLISION_BOX_CONNECTIVITY
// (collision_box_connectivity.f90:127-316) with the periodic wrap of a
// single rank (all 13 present); 0-based boxes in CON order
__device__ __forceinline__ void half_shell(const CellGrid& g, int c, int nb[13]) {
  const int NX = g.nx, NY = g.ny, NZ = g.nz;
  const int i0 = c % NX + 1, j0 = (c / NX) % NY + 1, k0 = c / (NX * NY) + 1;
  int im1 = i0 - 1, ip1 = i0 + 1, jm1 = j0 - 1, jp1 = j0 + 1, km1 = k0 - 1, kp1 = k0 + 1;
  if (im1 < 1) im1 = NX;
  if (ip1 > NX) ip1 = 1;
  if (jm1 < 1) jm1 = NY;
  if (jp1 > NY) jp1 = 1;
  if (km1 < 1) km1 = NZ;
  if (kp1 > NZ) kp1 = 1;
#define B(I, J, K) ((I) - 1 + ((J) - 1) * NX + ((K) - 1) * NX * NY)
  nb[0] = B(ip1, jp1, km1);
  nb[1] = B(ip1, jm1, k0);
  nb[2] = B(ip1, j0, k0);
  nb[3] = B(i0, jp1, k0);
  nb[4] = B(ip1, jp1, k0);
  nb[5] = B(i0, jm1, kp1);
  nb[6] = B(ip1, jm1, kp1);
  nb[7] = B(im1, j0, kp1);
  nb[8] = B(i0, j0, kp1);
  nb[9] = B(ip1, j0, kp1);
  nb[10] = B(im1, jp1, kp1);
  nb[11] = B(i0, jp1, kp1);
  nb[12] = B(ip1, jp1, kp1);
#undef B
}

template <typename VT>
__device__ __forceinline__ void load_state(const PState<VT>& p, int s, double st[13]) {
  st[0] = p.x[s]; st[1] = p.y[s]; st[2] = p.z[s];
  st[3] = (double)p.u[s]; st[4] = (double)p.v[s]; st[5] = (double)p.w[s];
  st[6] = p.qa[s]; st[7] = p.qb[s]; st[8] = p.qc[s]; st[9] = p.qd[s];
  st[10] = (double)p.ox[s]; st[11] = (double)p.oy[s]; st[12] = (double)p.oz[s];
}
template <typename VT>
__device__ __forceinline__ void store_state(const PState<VT>& p, int s, const double st[13]) {
  p.x[s] = st[0]; p.y[s] = st[1]; p.z[s] = st[2];
  p.u[s] = (VT)st[3]; p.v[s] = (VT)st[4]; p.w[s] = (VT)st[5];
  p.qa[s] = st[6]; p.qb[s] = st[7]; p.qc[s] = st[8]; p.qd[s] = st[9];
  p.ox[s] = (VT)st[10]; p.oy[s] = (VT)st[11]; p.oz[s] = (VT)st[12];
}

__device__ __forceinline__ Species load_species(const Species* __restrict__ t, int s) {
  return t[s];
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* p, T v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// reference layout UAIN(1:n,-1:n+2,-1:n+2) x 6 -> interleaved nodes with x
// ghosts, [(k+1)(n+4)+(j+1)](n+4)+(i+1) nodes of 6 values
template <typename FT, int NC = 6>
__global__ void k_field_prep(const double* __restrict__ ref, FT* __restrict__ out, int n) {
  const int n4 = n + 4;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)n4 * n4 * n4;
  if (t >= total) return;
  int ii = (int)(t % n4);               // stored x index = i + 1, i in [-1, n+2]
  int jj = (int)((t / n4) % n4);
  int kk = (int)(t / ((long long)n4 * n4));
  int i = ii - 1;
  int is = (i < 1) ? i + n : (i > n ? i - n : i);   // x ghosts: index period n
  const long long comp = (long long)n * n4 * n4;
  const long long src = (long long)(is - 1) + (long long)jj * n + (long long)kk * n * n4;
#pragma unroll
  for (int c = 0; c < NC; ++c) out[t * NC + c] = (FT)ref[c * comp + src];
}

// the same node layout straight from planar device arrays (velocity and
// vorticity, [3][n][n][n], x fastest, no ghosts; all ghosts index period n),
// the vorticity scaled by vscale: one pass instead of the reference layout
// and a second copy
template <typename FT>
__global__ void k_field_prep_planar(const double* __restrict__ u, const double* __restrict__ w,
                                    double vscale, FT* __restrict__ out, int n) {
  const int n4 = n + 4;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)n4 * n4 * n4;
  if (t >= total) return;
  const int ii = (int)(t % n4), jj = (int)((t / n4) % n4), kk = (int)(t / ((long long)n4 * n4));
  const int i0 = ((ii - 2) % n + n) % n, j0 = ((jj - 2) % n + n) % n, k0 = ((kk - 2) % n + n) % n;
  const long long nr = (long long)n * n * n;
  const long long src = ((long long)k0 * n + j0) * n + i0;
#pragma unroll
  for (int c = 0; c < 3; ++c) out[t * 6 + c] = (FT)u[c * nr + src];
#pragma unroll
  for (int c = 0; c < 3; ++c) out[t * 6 + 3 + c] = (FT)(vscale * w[c * nr + src]);
}

template <typename FT, typename CT, int S>
__global__ __launch_bounds__(BLOCK) void k_interph(const FT* __restrict__ f, int n, double dx,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z, int comp,
                                                   double* __restrict__ out, int nint,
                                                   unsigned* err) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nint) return;
  CT v[6];
  if (!gather6<FT, CT, S>(f, n, dx, x[p], y[p], z[p], v)) {
    atomicOr(err, EB_INTERP_RANGE);
    out[p] = 0.0;
    return;
  }
  out[p] = (double)v[comp];
}

template <typename FT, typename CT, int S>
__global__ __launch_bounds__(BLOCK) void k_interph1(const FT* __restrict__ f, int n, double dx,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z,
                                                    double* __restrict__ out, int nint,
                                                    unsigned* err) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nint) return;
  CT v[1];
  if (!gather6<FT, CT, S, 1>(f, n, dx, x[p], y[p], z[p], v)) {
    atomicOr(err, EB_INTERP_RANGE);
    out[p] = 0.0;
    return;
  }
  out[p] = (double)v[0];
}

// slot in a counter shared by lanes with the same key: one atomic per
// distinct key in the wavefront (lanes of a brick-ordered wavefront mostly
// share one brick, where per-lane atomics on one address would serialise).
// Every lane that needs a slot must call it together (wave-uniform loop).
__device__ __forceinline__ int grouped_slot(int* cnt, int key) {
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(1);
  int slot = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lkey = __shfl(key, leader);
    const unsigned long long same = __ballot(key == lkey) & todo;
    int base = 0;
    if (lane == leader) base = atomicAdd(&cnt[lkey], __popcll(same));
    base = __shfl(base, leader);
    if ((same >> lane) & 1ull) slot = base + __popcll(same & ((1ull << lane) - 1ull));
    todo &= ~same;
  }
  return slot;
}

// one particle: dst[i] = advance(src[s]), its cell and slot (and brick
// when bcnt; grouped = every lane of the wavefront that is still here
// calls together), error bits and wall count
template <typename VT, typename CT, typename Gather>
__device__ __forceinline__ void advance_one(
    const StepParams& P, const Gather& gather, const Species* __restrict__ spec,
    const PState<VT>& src, const PState<VT>& dst, int i, int s, int np, const CellGrid& g,
    int* __restrict__ cell, int* __restrict__ slot, int* __restrict__ cnt, const BrickGrid& bg,
    int* __restrict__ brick, int* __restrict__ bslot, int* __restrict__ bcnt, bool grouped,
    unsigned* err, unsigned long long* nwall, double* __restrict__ flout) {
  double st[13];
  load_state(src, s, st);
  const int sid = src.sp[s];
  const int pid = src.id[s];
  CT fl[6];
  unsigned e = 0;
  if (!gather(st[0], st[1], st[2], fl)) {
    e |= EB_INTERP_RANGE;
    for (int c = 0; c < 6; ++c) fl[c] = CT(0);
  }
  if (flout)
    for (int c = 0; c < 6; ++c) flout[(long long)c * np + pid] = (double)fl[c];
  const Species sp = load_species(spec, sid);
  int wall = 0;
  e |= advance_particle<CT>(P, sp, st, fl, &wall);
  store_state(dst, i, st);
  dst.rm[i] = src.rm[s];
  dst.sp[i] = sid;
  dst.id[i] = pid;
  dst.grp[i] = src.grp[s];
  int c = 0;
  if (cell) {
    bool oob;
    c = cell_of(g, st[0], st[1], st[2], &oob);
    if (oob) e |= EB_OUT_OF_GRID;
    cell[i] = c;
    slot[i] = atomicAdd(&cnt[c], 1);
  }
  if (bcnt) {
    const int b = brick_at(g, bg, c, st[0], st[1], st[2]);
    brick[i] = b;
    bslot[i] = grouped ? grouped_slot(bcnt, b) : atomicAdd(&bcnt[b], 1);
  }
  if (wall) atomicAdd(nwall, 1ull);
  if (e) atomicOr(err, e);
}

// fused gather + advance + wrap + cell binning; dst[i] = advance(src[perm[i]]).
// skip != nullptr: particles flagged there (those of this step's contact
// solves, advanced later by k_advance_listed) are left out.
template <typename FT, typename VT, typename CT, int S>
__global__ __launch_bounds__(BLOCK) void k_advance(
    StepParams P, const FT* __restrict__ f, const Species* __restrict__ spec, PState<VT> src,
    PState<VT> dst, const int* __restrict__ perm, int np, CellGrid g, int* __restrict__ cell,
    int* __restrict__ slot, int* __restrict__ cnt, BrickGrid bg, int* __restrict__ brick,
    int* __restrict__ bslot, int* __restrict__ bcnt, unsigned* err, unsigned long long* nwall,
    double* __restrict__ flout, int xcd, const int* __restrict__ skip) {
  const int i = block_index(xcd) * BLOCK + threadIdx.x;
  if (i >= np) return;
  const int s = perm ? perm[i] : i;
  if (skip && skip[s]) return;
  advance_one<VT, CT>(P, GlobalGather<FT, CT, S>{f, P.n, P.dx}, spec, src, dst, i, s, np, g, cell,
                      slot, cnt, bg, brick, bslot, bcnt, true, err, nwall, flout);
}

// the advance of the field-tile mode, with the fluid at the particles
// already gathered (k_gather_tiled), one body for both of its callers:
// list == nullptr, visiting position i from the (XCD-remapped) grid,
// particle perm[i], skip flags honoured; list != nullptr, the listed
// particles list[0, *nlist) to their visiting positions inv[s], grid-stride.
// (The same routine compiled into two kernels can round differently, the
// compiler choosing fused multiply-adds per kernel: one kernel keeps the
// eager and the pipelined steps of this mode bit-identical.)
template <typename VT, typename CT>
__global__ __launch_bounds__(BLOCK) void k_advance_fl(
    StepParams P, const CT* __restrict__ fl, const Species* __restrict__ spec, PState<VT> src,
    PState<VT> dst, const int* __restrict__ perm, int np, CellGrid g, int* __restrict__ cell,
    int* __restrict__ slot, int* __restrict__ cnt, BrickGrid bg, int* __restrict__ brick,
    int* __restrict__ bslot, int* __restrict__ bcnt, unsigned* err, unsigned long long* nwall,
    double* __restrict__ flout, int xcd, const int* __restrict__ skip,
    const int* __restrict__ list, const unsigned long long* nlist, const int* __restrict__ inv) {
  const bool listed = list != nullptr;
  const long long n = listed ? (long long)min(*nlist, (unsigned long long)np) : (long long)np;
  const long long stride = listed ? (long long)gridDim.x * blockDim.x : n;
  for (long long t = (long long)(listed ? (int)blockIdx.x : block_index(xcd)) * blockDim.x +
                     threadIdx.x;
       t < n; t += stride) {
    int i, s;
    if (listed) {
      s = list[t];
      i = inv[s];
    } else {
      i = (int)t;
      s = perm[i];
      if (skip && skip[s]) continue;
    }
    advance_one<VT, CT>(P, BufferGather<CT>{fl, np, i}, spec, src, dst, i, s, np, g, cell, slot,
                        cnt, bg, brick, bslot, bcnt, !listed, err, nwall, flout);
  }
}

// the field tiles (gp_params.tile): workgroup t (one wavefront) takes the
// particles at visiting positions boff[t]..boff[t+1], those whose stencil
// base node lay in tile t when they were binned.  stage_tile copies the
// block of nodes every such stencil reaches, tile + S - 1 nodes in x and z
// and tile + 3 rows in y, from HBM into LDS (rows of 8-byte words,
// contiguous and coalesced); TileGather then reads the stencils from LDS
// (a particle whose base node is outside the tile, none unless positions
// changed after the binning, reads HBM).
constexpr int TILE_WG = 64;

template <int S>
__host__ __device__ constexpr int tile_extent_x(int T) { return T + S - 1; }
__host__ __device__ constexpr int tile_extent_y(int T) { return T + 3; }

// quotient and remainder by a small divisor d (dividends below 2^22): a
// float reciprocal estimate, corrected by one step (its error is below one)
struct SmallDiv {
  int d;
  float rcp;
  __device__ explicit SmallDiv(int d_) : d(d_), rcp(1.0f / (float)d_) {}
  __device__ __forceinline__ void divmod(int u, int& q, int& r) const {
    q = (int)((float)u * rcp);
    r = u - q * d;
    if (r < 0) { --q; r += d; }
    if (r >= d) { ++q; r -= d; }
  }
};

template <typename FT, typename CT, int S>
__device__ __forceinline__ TileGather<FT, CT, S> stage_tile(const FT* __restrict__ f,
                                                            const StepParams& P,
                                                            const BrickGrid& bg, int t) {
  const int T = bg.b;
  const int tx = t % bg.nx, ty = (t / bg.nx) % bg.ny, tz = t / (bg.nx * bg.ny);
  const int tbx = tx * T, tby = ty * T, tbz = tz * T;
  const int Ex = tile_extent_x<S>(T), Ey = tile_extent_y(T), Ez = Ex;
  // storage index (ghosts included) of the block's first node on each
  // axis: node_index of the stencil's first node for base node tb
  constexpr int off0 = 2 + stencil_lo<S>();
  const int n4 = P.n + 4;
  // nodes past storage index n + 3 exist in no stencil (base nodes stop at
  // n - 1): a last, partial tile stages fewer
  const int ex = min(Ex, n4 - (tbx + off0)), ey = min(Ey, n4 - (tby + off0)),
            ez = min(Ez, n4 - (tbz + off0));
  constexpr int WPN = 6 * (int)sizeof(FT) / 8;   // 8-byte words per node
  const int wpr = ex * WPN;                       // words per staged row
  const int lrow = Ex * WPN;                      // words per LDS row
  unsigned long long* l64 = reinterpret_cast<unsigned long long*>(tile_lds);
  {
    // lane-strided words, TILE_BATCH loads in flight per lane before their
    // LDS writes (branch-free index arithmetic, so that the loads of a
    // batch issue back to back)
    constexpr int TILE_BATCH = 16;
    const SmallDiv dw(wpr), dy(ey);
    const int total = wpr * ey * ez;
    for (int u0 = (int)threadIdx.x; u0 < total; u0 += TILE_WG * TILE_BATCH) {
      unsigned long long v[TILE_BATCH];
      int at[TILE_BATCH];
#pragma unroll
      for (int u = 0; u < TILE_BATCH; ++u) {
        // past the end: reload the last word (straight-line code), not written
        const int idx = u0 + u * TILE_WG, idc = min(idx, total - 1);
        int r, w, kz, jy;
        dw.divmod(idc, r, w);
        dy.divmod(r, kz, jy);
        const long long node =
            ((long long)(tbz + off0 + kz) * n4 + (tby + off0 + jy)) * n4 + (tbx + off0);
        v[u] = reinterpret_cast<const unsigned long long*>(f + node * 6)[w];
        at[u] = idx < total ? (kz * Ey + jy) * lrow + w : -1;
      }
#pragma unroll
      for (int u = 0; u < TILE_BATCH; ++u)
        if (at[u] >= 0) l64[at[u]] = v[u];
    }
  }
  __syncthreads();
  return TileGather<FT, CT, S>{f, P.n, P.dx, tbx, tby, tbz, T, Ex, Ey};
}

// the six fields at every particle of tile t (skipped ones included: a
// contact pass changes velocities, not positions), by visiting position
template <typename FT, typename VT, typename CT, int S>
__global__ __launch_bounds__(TILE_WG) void k_gather_tiled(
    StepParams P, const FT* __restrict__ f, PState<VT> src, const int* __restrict__ trav,
    const int* __restrict__ boff, int np, BrickGrid bg, CT* __restrict__ fl, unsigned* err,
    int xcd) {
  const int t = block_index(xcd);
  if (t >= bg.n) return;
  const int p0 = boff[t], p1 = boff[t + 1];
  if (p0 >= p1) return;
  const TileGather<FT, CT, S> gather = stage_tile<FT, CT, S>(f, P, bg, t);
  for (int i = p0 + (int)threadIdx.x; i < p1; i += TILE_WG) {
    const int s = trav[i];
    CT v[6];
    if (!gather(src.x[s], src.y[s], src.z[s], v)) {
      atomicOr(err, EB_INTERP_RANGE);
#pragma unroll
      for (int c = 0; c < 6; ++c) v[c] = CT(0);
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) fl[(long long)c * np + i] = v[c];
  }
}

template <typename FT, int S>
static size_t tile_lds_bytes(int T) {
  const size_t e = (size_t)tile_extent_x<S>(T);
  return e * e * (size_t)tile_extent_y(T) * 6 * sizeof(FT);
}

// the particles left out above: storage indices in list[0, *nlist), each
// to its visiting position inv[s].  A few thousand particles: launched as
// one-wavefront workgroups so they spread over as many CUs as possible (each
// lane's stencil gather is dozens of scattered loads, and a CU's address
// path, not HBM, limits a handful of full workgroups).
constexpr int ADVL_BLOCK = 64;

template <typename FT, typename VT, typename CT, int S>
__global__ __launch_bounds__(ADVL_BLOCK) void k_advance_listed(
    StepParams P, const FT* __restrict__ f, const Species* __restrict__ spec, PState<VT> src,
    PState<VT> dst, const int* __restrict__ list, const unsigned long long* nlist,
    const int* __restrict__ inv, int np, CellGrid g,
    int* __restrict__ cell, int* __restrict__ slot, int* __restrict__ cnt, BrickGrid bg,
    int* __restrict__ brick, int* __restrict__ bslot, int* __restrict__ bcnt, unsigned* err,
    unsigned long long* nwall, double* __restrict__ flout) {
  const long long nl = (long long)min(*nlist, (unsigned long long)np);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nl;
       t += (long long)gridDim.x * blockDim.x) {
    const int s = list[t];
    advance_one<VT, CT>(P, GlobalGather<FT, CT, S>{f, P.n, P.dx}, spec, src, dst, inv[s], s, np, g,
                        cell, slot, cnt, bg, brick, bslot, bcnt, false, err, nwall, flout);
  }
}

// cell binning of the current positions (when no advance preceded)
template <typename VT>
__global__ __launch_bounds__(BLOCK) void k_cells(PState<VT> p, int np, CellGrid g,
                                                 int* __restrict__ cell, int* __restrict__ slot,
                                                 int* __restrict__ cnt, BrickGrid bg,
                                                 int* __restrict__ brick, int* __restrict__ bslot,
                                                 int* __restrict__ bcnt, unsigned* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  bool oob;
  int c = cell_of(g, p.x[i], p.y[i], p.z[i], &oob);
  if (oob) atomicOr(err, EB_OUT_OF_GRID);
  cell[i] = c;
  slot[i] = atomicAdd(&cnt[c], 1);
  if (bcnt) {
    const int b = brick_at(g, bg, c, p.x[i], p.y[i], p.z[i]);
    brick[i] = b;
    bslot[i] = grouped_slot(bcnt, b);
  }
}

// visiting order of the next advance: particles grouped by brick (any
// order inside a brick; results do not depend on the storage order)
// Clears folded into kernels that run at that point anyway, each saving a
// launch on the step's critical path: int arrays to zero (bin counts),
// counters to zero, and flag[list[t]] = 0 for the listed particles.  Every
// thread of the grid takes part (grid-stride); members left null do nothing.
struct SideClear {
  int* zi = nullptr;
  long long nzi = 0;
  int* zi2 = nullptr;
  long long nzi2 = 0;
  unsigned long long* zc = nullptr;
  int nzc = 0;
  unsigned long long* zc2 = nullptr;
  int nzc2 = 0;
  unsigned long long* zl = nullptr;          // keys[slot[t]] = 0 for t < *nslot
  const int* slot = nullptr;
  const unsigned long long* nslot = nullptr;
  unsigned long long nslot_cap = 0;
  const int* list = nullptr;
  const unsigned long long* nlist = nullptr;
  int* flag = nullptr;
  int np = 0;
};

__device__ __forceinline__ void side_clear(const SideClear& sc) {
  const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long t = t0; t < sc.nzi; t += st) sc.zi[t] = 0;
  for (long long t = t0; t < sc.nzi2; t += st) sc.zi2[t] = 0;
  if (sc.zc)
    for (long long t = t0; t < sc.nzc; t += st) sc.zc[t] = 0ull;
  if (sc.zc2)
    for (long long t = t0; t < sc.nzc2; t += st) sc.zc2[t] = 0ull;
  if (sc.zl) {
    const long long ns = (long long)min(*sc.nslot, sc.nslot_cap);
    for (long long t = t0; t < ns; t += st) {
      const int h = sc.slot[t];
      if (h >= 0) sc.zl[h] = 0ull;
    }
  }
  if (sc.list) {
    const long long nl = (long long)min(*sc.nlist, (unsigned long long)sc.np);
    for (long long t = t0; t < nl; t += st) sc.flag[sc.list[t]] = 0;
  }
}

__global__ __launch_bounds__(BLOCK) void k_brick_scatter(const int* __restrict__ brick,
                                                         const int* __restrict__ bslot,
                                                         const int* __restrict__ boff,
                                                         int* __restrict__ trav,
                                                         int* __restrict__ inv, int np) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int pos = boff[brick[i]] + bslot[i];
  trav[pos] = i;
  inv[i] = pos;
}

// the order key of every particle, (ensemble, reference index), at its
// cell's slot: the ranking then compares a cell's keys from consecutive
// words instead of two dependent loads per member
__device__ __forceinline__ unsigned long long order_key(int g, int id) {
  return ((unsigned long long)(unsigned)g << 32) | (unsigned long long)(unsigned)id;
}

__global__ __launch_bounds__(BLOCK) void k_scatter(const int* __restrict__ cell,
                                                   const int* __restrict__ slot,
                                                   const int* __restrict__ off,
                                                   const int* __restrict__ id,
                                                   const int* __restrict__ grp,
                                                   unsigned long long* __restrict__ tkey, int np,
                                                   SideClear sc) {
  side_clear(sc);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int c = cell[i];
  if (c >= 0) tkey[off[c] + slot[i]] = order_key(grp ? grp[i] : 0, id[i]);
}

// stable order inside each cell: by ensemble, then reference index (the
// reference bins particles in index order); grp == nullptr: one ensemble
__global__ __launch_bounds__(BLOCK) void k_rank(const int* __restrict__ cell,
                                                const int* __restrict__ off,
                                                const int* __restrict__ cnt,
                                                const unsigned long long* __restrict__ tkey,
                                                const int* __restrict__ id,
                                                const int* __restrict__ grp,
                                                const double* __restrict__ px,
                                                const double* __restrict__ py,
                                                const double* __restrict__ pz,
                                                const float* __restrict__ rm,
                                                int* __restrict__ sorted,
                                                float4* __restrict__ pos4,
                                                int* __restrict__ csort,
                                                int* __restrict__ gsort,
                                                int* __restrict__ inv, int np, int xcd,
                                                SideClear sc) {
  side_clear(sc);
  int i = block_index(xcd) * BLOCK + threadIdx.x;
  if (i >= np) return;
  int c = cell[i];
  const int o = off[c], m = cnt[c];
  int r = 0;
  if (m > 1) {
    // its rank among the cell's keys (alone in its cell, the common case at
    // low density, it has rank 0 and reads none)
    const unsigned long long me = order_key(grp ? grp[i] : 0, id[i]);
    for (int t = 0; t < m; ++t) r += (tkey[o + t] < me) ? 1 : 0;
  }
  sorted[o + r] = i;
  if (inv) inv[i] = o + r;
  // cell-ordered copy of what the bounding-sphere test reads: a cell's
  // members are contiguous here
  pos4[o + r] = make_float4((float)px[i], (float)py[i], (float)pz[i], rm[i]);
  csort[o + r] = c;
  // and of the ensemble: the member ranges of one ensemble are then found
  // without a dependent load per member
  if (grp) gsort[o + r] = grp[i];
}

// members of cell range [o, o+m) (sorted by ensemble) in ensemble g, from
// the ensemble of every sorted position
__device__ __forceinline__ void group_range(const int* __restrict__ gsort, int o, int m, int g,
                                            int* lo, int* hi) {
  int a = o, b = o + m;
  while (a < b && gsort[a] < g) ++a;
  int e = a;
  while (e < b && gsort[e] == g) ++e;
  *lo = a;
  *hi = e;
}

// candidate pairs emitted for the home particle at sorted position p
__global__ __launch_bounds__(BLOCK) void k_pair_count(CellGrid g, const int* __restrict__ sorted,
                                                      const int* __restrict__ cell,
                                                      const int* __restrict__ off,
                                                      const int* __restrict__ cnt,
                                                      const int* __restrict__ gsort,
                                                      int* __restrict__ pc, int nsorted) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nsorted) return;
  const int hi = sorted[p];
  const int c = cell[hi];
  int nb[13];
  half_shell(g, c, nb);
  int k;
  if (!gsort) {
    k = cnt[c] - 1 - (p - off[c]);
#pragma unroll
    for (int v = 0; v < 13; ++v) k += cnt[nb[v]];
  } else {
    const int gr = gsort[p];
    int lo, up;
    group_range(gsort, off[c], cnt[c], gr, &lo, &up);
    k = up - 1 - p;
    for (int v = 0; v < 13; ++v) {
      group_range(gsort, off[nb[v]], cnt[nb[v]], gr, &lo, &up);
      k += up - lo;
    }
  }
  pc[p] = k;
}

// materialised HOME/NEAR (1-based reference indices) for the API
__global__ __launch_bounds__(BLOCK) void k_pair_emit(CellGrid g, const int* __restrict__ sorted,
                                                     const int* __restrict__ cell,
                                                     const int* __restrict__ off,
                                                     const int* __restrict__ cnt,
                                                     const long long* __restrict__ poff,
                                                     const int* __restrict__ id,
                                                     const int* __restrict__ gsort,
                                                     int* __restrict__ home, int* __restrict__ nr,
                                                     int nsorted) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nsorted) return;
  const int hi = sorted[p];
  const int c = cell[hi];
  const int h = id[hi] + 1;
  long long k = poff[p];
  int lo = off[c], up = off[c] + cnt[c];
  const int gr = gsort ? gsort[p] : 0;
  if (gsort) group_range(gsort, off[c], cnt[c], gr, &lo, &up);
  for (int t = p + 1; t < up; ++t) { home[k] = h; nr[k] = id[sorted[t]] + 1; ++k; }
  int nb[13];
  half_shell(g, c, nb);
  for (int v = 0; v < 13; ++v) {
    lo = off[nb[v]];
    up = lo + cnt[nb[v]];
    if (gsort) group_range(gsort, off[nb[v]], cnt[nb[v]], gr, &lo, &up);
    for (int t = lo; t < up; ++t) { home[k] = h; nr[k] = id[sorted[t]] + 1; ++k; }
  }
}

template <typename VT>
__device__ __forceinline__ void min_image(const PState<VT>& p, int i, int j, double L,
                                          double bi[3], double bj[3]) {
  bi[0] = p.x[i]; bi[1] = p.y[i]; bi[2] = p.z[i];
  double vk[3] = {p.x[j] - bi[0], p.y[j] - bi[1], p.z[j] - bi[2]};
  for (int c = 0; c < 3; ++c) {
    if (fabs(vk[c]) > L / 2.0) vk[c] = vk[c] - copysign(L, vk[c]);
    bj[c] = bi[c] + vk[c];
  }
}

// bounding-sphere prefilter over each home's candidates; survivors go to
// the LM work list.  Their list order is kept in the key k = (q << 32) +
// loc, q the home's sorted position and loc the candidate's index among
// the home's candidates: the candidate list is ordered by (q, loc), so k
// orders pairs as their list index poff[q] + loc would, without the pair
// counts and their scan.  The candidates of the workgroup are summed into
// *nclose (the list's length).  Exact: a contact needs
// the support point of one ellipsoid inside the other, which is impossible
// when the centres are farther apart than the sum of the largest semi-axes.
// The test runs on float copies of the centres (rmax rounded up to float);
// tol, 1e-5 of the largest coordinate magnitude, exceeds their rounding
// (a few float ulps of that magnitude) by far, so every dropped pair is
// farther apart than the sum; pairs it keeps beyond it go to the exact
// disjointness test and the solve, which flag them 0.
// Survivors are staged in LDS and appended with one global atomic per
// workgroup (appends from every lane to one counter would serialise);
// LDS overflow, in dense suspensions, appends directly.
constexpr int PF_STAGE = 1024;

// GROUPED: several ensembles (gsort given), a variant of its own so that
// the plain test keeps its registers
template <bool GROUPED>
__global__ __launch_bounds__(BLOCK) void k_prefilter(
    CellGrid g, const float4* __restrict__ pos4, const int* __restrict__ sorted,
    const int* __restrict__ csort, const int* __restrict__ off, const int* __restrict__ cnt,
    const int* __restrict__ gsort, double L, float tol, int use_filter, WorkRec* __restrict__ work,
    unsigned long long* nwork, unsigned long long cap, unsigned long long* nclose, unsigned* err,
    int nsorted, int xcd) {
  __shared__ WorkRec stage[PF_STAGE];
  __shared__ int nstage;
  __shared__ unsigned long long gbase, ncand;
  if (threadIdx.x == 0) { nstage = 0; ncand = 0; }
  __syncthreads();
  const int q = block_index(xcd) * BLOCK + threadIdx.x;
  long long k = (long long)q << 32;
  if (q < nsorted) {
    const int hi = sorted[q];
    const int c = csort[q];
    const float4 h = pos4[q];
    const float Lf = (float)L, Lh = 0.5f * Lf, ri = h.w + tol;
    auto test = [&](const float4 o, int t) {
      float dx = o.x - h.x, dy = o.y - h.y, dz = o.z - h.z;
      if (fabsf(dx) > Lh) dx -= copysignf(Lf, dx);
      if (fabsf(dy) > Lh) dy -= copysignf(Lf, dy);
      if (fabsf(dz) > Lh) dz -= copysignf(Lf, dz);
      const float rs = ri + o.w;
      if (!use_filter || dx * dx + dy * dy + dz * dz <= rs * rs * 1.0001f) {
        WorkRec r;
        r.k = k; r.i = hi; r.j = sorted[t];
        const int sidx = atomicAdd(&nstage, 1);
        if (sidx < PF_STAGE) {
          stage[sidx] = r;
        } else {
          const unsigned long long w = atomicAdd(nwork, 1ull);
          if (w < cap) work[w] = r;
          else atomicOr(err, EB_ACTIVE_OVERFLOW);
        }
      }
      ++k;
    };
    // the 14 member ranges in list order (own cell after q, then the half
    // shell); the first member of every range is fetched before any test
    // (cells hold ~0-2 members, so this puts nearly every load of the
    // thread in flight at once instead of one range after another)
    int slo[14], sup[14];
    // a cell's range is [off[c], off[c + 1]): two words of one line
    slo[0] = q + 1;
    sup[0] = off[c + 1];
    int nb[13];
    half_shell(g, c, nb);
#pragma unroll
    for (int v = 0; v < 13; ++v) {
      slo[v + 1] = off[nb[v]];
      sup[v + 1] = off[nb[v] + 1];
    }
    if (!GROUPED) {
      float4 first[14];
#pragma unroll
      for (int r = 0; r < 14; ++r)
        if (slo[r] < sup[r]) first[r] = pos4[slo[r]];
#pragma unroll
      for (int r = 0; r < 14; ++r) {
        if (slo[r] < sup[r]) {
          test(first[r], slo[r]);
          for (int t = slo[r] + 1; t < sup[r]; ++t) test(pos4[t], t);
        }
      }
    } else {
      // several ensembles: a cell's members are sorted by ensemble, so the
      // home's ensemble's members form one run [glo, ghi) in each range,
      // found member by member over the 14 ranges at once (the loads of one
      // depth in flight together); positions are read for those runs only
      const int gr = gsort[q];
      int glo[14], ghi[14], len = 0;
#pragma unroll
      for (int r = 0; r < 14; ++r) {
        glo[r] = sup[r];
        ghi[r] = sup[r];
        len = max(len, sup[r] - slo[r]);
      }
      unsigned open = (1u << 14) - 1u;   // ranges whose run has not ended
      for (int d = 0; d < len && open; ++d) {
#pragma unroll
        for (int r = 0; r < 14; ++r) {
          const int t = slo[r] + d;
          if (((open >> r) & 1u) && t < sup[r]) {
            const int gt = gsort[t];
            if (gt >= gr && glo[r] == sup[r]) glo[r] = t;
            if (gt > gr) {
              ghi[r] = t;
              open &= ~(1u << r);
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 14; ++r)
        for (int t = glo[r]; t < ghi[r]; ++t) test(pos4[t], t);
    }
  }
  unsigned long long nc = (unsigned long long)(k - ((long long)q << 32));
  for (int o = 32; o > 0; o >>= 1) nc += __shfl_xor(nc, o);
  if ((threadIdx.x & 63) == 0 && nc) atomicAdd(&ncand, nc);
  __syncthreads();
  if (threadIdx.x == 0 && ncand) atomicAdd(nclose, ncand);
  const int ns = min(nstage, PF_STAGE);
  if (ns == 0) return;
  if (threadIdx.x == 0) gbase = atomicAdd(nwork, (unsigned long long)ns);
  __syncthreads();
  for (int t = threadIdx.x; t < ns; t += blockDim.x) {
    const unsigned long long w = gbase + (unsigned long long)t;
    if (w < cap) work[w] = stage[t];
    else atomicOr(err, EB_ACTIVE_OVERFLOW);
  }
}

// open-addressing table of pair keys (home slot + 1) << 32 | (near slot + 1),
// 0 = empty; sized at least twice the pairs it holds, so probes end
__device__ __forceinline__ unsigned long long pair_key(int i, int j) {
  return ((unsigned long long)(unsigned)(i + 1) << 32) | (unsigned long long)(unsigned)(j + 1);
}
__device__ __forceinline__ unsigned long long pair_hash(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}
__device__ __forceinline__ int table_put(unsigned long long* keys, int* vals,
                                         unsigned long long mask, unsigned long long key, int v) {
  unsigned long long h = pair_hash(key) & mask;
  for (unsigned long long probe = 0; probe <= mask; ++probe) {
    const unsigned long long old = atomicCAS(&keys[h], 0ull, key);
    if (old == 0ull || old == key) {
      vals[h] = v;
      return (int)h;
    }
    h = (h + 1) & mask;
  }
  return -1;
}
__device__ __forceinline__ int table_get(const unsigned long long* keys, const int* vals,
                                         unsigned long long mask, unsigned long long key) {
  unsigned long long h = pair_hash(key) & mask;
  for (unsigned long long probe = 0; probe <= mask; ++probe) {
    const unsigned long long k = keys[h];
    if (k == key) return vals[h];
    if (k == 0ull) return -1;
    h = (h + 1) & mask;
  }
  return -1;
}

// one slot per lane that asks, one atomic per wavefront
__device__ __forceinline__ unsigned long long wave_append(bool want, unsigned long long* count) {
  const int lane = threadIdx.x & 63;
  const unsigned long long m = __ballot(want);
  if (!m) return 0;
  const int leader = __ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  return base + __popcll(m & ((1ull << lane) - 1ull));
}

// exact disjointness test (Perram-Wertheim) on the bounding-sphere
// survivors; pairs not proven disjoint go to the LM list in any order (the
// list index k keeps the reference order).  use_pw = 0 passes every pair.
// Quaternions must be unit to 1e-9 for the test to stand for the margin
// function of the flag; others go to the LM.
template <typename VT>
__global__ __launch_bounds__(BLOCK) void k_pw_filter(
    PState<VT> p, const Species* __restrict__ spec, const WorkRec* __restrict__ work,
    const unsigned long long* nwork, unsigned long long cap, double L, int use_pw,
    WorkRec* __restrict__ lm, unsigned long long* nlm, int* __restrict__ flag,
    int* __restrict__ plist, unsigned long long* nplist, SideClear sc, EarlyTab et) {
  side_clear(sc);
  const unsigned long long nw = min(*nwork, cap);
  const int lane = threadIdx.x & 63;
  // the loop bound is uniform over the workgroup, so every lane of a
  // wavefront takes part in each ballot
  for (unsigned long long b0 = (unsigned long long)blockIdx.x * blockDim.x; b0 < nw;
       b0 += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned long long w = b0 + threadIdx.x;
    bool pass = false;
    WorkRec r{};
    if (w < nw) {
      r = work[w];
      pass = true;
      if (use_pw) {
        double bi[3], bj[3];
        min_image(p, r.i, r.j, L, bi, bj);
        const double qi[4] = {p.qa[r.i], p.qb[r.i], p.qc[r.i], p.qd[r.i]};
        const double qj[4] = {p.qa[r.j], p.qb[r.j], p.qc[r.j], p.qd[r.j]};
        const double ni = qi[0] * qi[0] + qi[1] * qi[1] + qi[2] * qi[2] + qi[3] * qi[3];
        const double nj = qj[0] * qj[0] + qj[1] * qj[1] + qj[2] * qj[2] + qj[3] * qj[3];
        if (fabs(ni - 1.0) < 1.0e-9 && fabs(nj - 1.0) < 1.0e-9) {
          const Species si = spec[p.sp[r.i]], sj = spec[p.sp[r.j]];
          const Sym3 E1 = ell_matrix(qi, si.a, si.b, si.c);
          const Sym3 E2 = ell_matrix(qj, sj.a, sj.b, sj.c);
          const double rr[3] = {bj[0] - bi[0], bj[1] - bi[1], bj[2] - bi[2]};
          pass = !pw_disjoint(E1, E2, rr);
        }
      }
    }
    // solving ahead: list the pair for the next step, and take it out of
    // this step's solves when the previous step listed it
    bool solve = pass;
    if (et.nkeys) {
      int found = -1, ni = 0, nj = 0;
      if (pass) {
        found = table_get(et.pkeys, et.pvals, et.mask, pair_key(r.i, r.j));
        ni = et.inv[r.i];
        nj = et.inv[r.j];
      }
      const unsigned long long e = wave_append(pass, et.npair);
      if (pass && e < et.cap) {
        et.pair[2 * e] = ni;
        et.pair[2 * e + 1] = nj;
        et.slot[e] = table_put(et.nkeys, et.nvals, et.mask, pair_key(ni, nj), (int)e);
      }
      const unsigned long long g = wave_append(found >= 0, et.nmerge);
      if (found >= 0 && g < et.cap) {
        et.merge[g] = MergeRec{r.k, found, 0};
        solve = false;
      }
    }
    // one atomic per wavefront: ballot + prefix
    const unsigned long long m = __ballot(solve);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(nlm, (unsigned long long)__popcll(m));
      base = __shfl(base, leader);
      if (solve) lm[base + __popcll(m & ((1ull << lane) - 1ull))] = r;
    }
    // flag (once) and list the particles of the solves, whose advance of the
    // next step waits for this step's collisions
    if (flag) {
      const bool ci = pass && atomicExch(&flag[r.i], 1) == 0;
      const bool cj = pass && atomicExch(&flag[r.j], 1) == 0;
      const unsigned long long mi = __ballot(ci), mj = __ballot(cj);
      const int ni = __popcll(mi), nt = ni + __popcll(mj);
      if (nt) {
        const int leader = __ffsll((long long)(mi | mj)) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(nplist, (unsigned long long)nt);
        base = __shfl(base, leader);
        const unsigned long long below = (1ull << lane) - 1ull;
        if (ci) plist[base + __popcll(mi & below)] = r.i;
        if (cj) plist[base + ni + __popcll(mj & below)] = r.j;
      }
    }
  }
}

// one statistics record per contact (collision_interaction.f90:318-345):
// relative velocity (vx, vy, vz) = VREL, (dx, dy, dz) = EPT1 - EPT2
__device__ __noinline__ void cstat_record(const CStat* __restrict__ csp, long long k, int i, int j,
                                          int s1, int s2, double nvrel, double vx, double vy,
                                          double vz, double dx, double dy, double dz,
                                          unsigned* err) {
  const CStat& cs = *csp;
  if (!cstat_step(cs)) return;
  CollStatRec r;
  r.k = k; r.i = i; r.j = j;
  r.bin = s1 * cs.npoly + s2 % cs.npoly;   // (class, IDP1) row s1, then IDP2
  if (r.bin >= cs.nbin) return;
  r.approach = nvrel < ZERO_R;
  r.depth = sqrt(dx * dx + dy * dy + dz * dz);
  r.nvrel = nvrel;
  r.nrm = sqrt(vx * vx + vy * vy + vz * vz);
  r.theta = acos(nvrel / r.nrm);
  const unsigned long long a = atomicAdd(cs.cnt, 1ull);
  if (a < cs.cap) cs.rec[a] = r;
  else atomicOr(err, EB_ACTIVE_OVERFLOW);
}

// LM contact solve on the work list + NVREL test
// (collision_interaction.f90:221-348); approaching contacts are appended.
// One pair per lane and one wavefront per workgroup, launched with one
// wavefront per SIMD.  Pairs are dealt lane-major (pair = lane * waves +
// wave), so a short list leaves few active lanes in every wavefront rather
// than a few full ones: a wavefront runs the union of its lanes' iteration
// paths, and the kernel lasts as long as its slowest wavefront.
// The solve is capped at 256 registers (amdgpu_waves_per_eu(2)): uncapped,
// a one-wavefront-per-SIMD kernel gets up to 512 and the compiler takes
// ~356, which leaves no SIMD room for a wavefront of the next step's
// advance (~192) running beside it on the second stream; the ~100 values
// it then spills to scratch cost the latency-bound solve little.
constexpr int LM_BLOCK = 64;

template <typename VT>
__global__ __launch_bounds__(LM_BLOCK) __attribute__((amdgpu_waves_per_eu(2))) void k_contact(
    PState<VT> p, const Species* __restrict__ spec, const WorkRec* __restrict__ work,
    const unsigned long long* nwork, unsigned long long cap, double L,
    ActiveRec* __restrict__ act, unsigned long long* nact, unsigned long long acap,
    unsigned long long* nfev_sum, unsigned long long* ncontact, unsigned* err,
    const double* __restrict__ seeds, SeedGate gate, unsigned long long* nfallback,
    const CStat* __restrict__ csp, EarlyIn ein) {
  // ein.res: solving ahead (pairs by slot, results to ein.res;
  // the same solve in the same kernel, so the same bits)
  const bool ahead = ein.res != nullptr;
  const unsigned long long nw = min(ahead ? *ein.n : *nwork, cap);
  const unsigned long long waves = gridDim.x;
  const int lane = threadIdx.x;
  unsigned long long nfev = 0, nfb = 0;
  // the loop bound is uniform over the wavefront (w - lane * waves is), so
  // the ballot below sees every lane
  for (unsigned long long w0 = blockIdx.x; w0 < nw; w0 += waves * LM_BLOCK) {
    const unsigned long long w = w0 + (unsigned long long)lane * waves;
    bool approach = false;
    ActiveRec ar;
    if (w < nw) {
      WorkRec r;
      if (ahead) {
        r.k = 0;
        r.i = ein.pair[2 * w];
        r.j = ein.pair[2 * w + 1];
      } else {
        r = work[w];
      }
      double bi[3], bj[3];
      min_image(p, r.i, r.j, L, bi, bj);
      const double qi[4] = {p.qa[r.i], p.qb[r.i], p.qc[r.i], p.qd[r.i]};
      const double qj[4] = {p.qa[r.j], p.qb[r.j], p.qc[r.j], p.qd[r.j]};
      const Species si = spec[p.sp[r.i]], sj = spec[p.sp[r.j]];
      // network seed when given, kept only through the gate (else the
      // centre-line solve decides)
      const double* sd = (seeds && seeds[2 * w] == seeds[2 * w]) ? seeds + 2 * w : nullptr;
      int nf = 0;
      bool fb = false;
      const ContactOut o = contact_detect_gated(bi, qi, si, bj, qj, sj, sd, gate, nf, fb);
      nfev += (unsigned long long)nf;
      nfb += fb ? 1ull : 0ull;
      if (o.flag) {
        // velocities of the deepest points (collision_interaction.f90:269-324)
        double r1[3] = {o.e1[0] - bi[0], o.e1[1] - bi[1], o.e1[2] - bi[2]};
        double r2[3] = {o.e2[0] - bj[0], o.e2[1] - bj[1], o.e2[2] - bj[2]};
        Rot<double> R1(qi[0], qi[1], qi[2], qi[3]), R2(qj[0], qj[1], qj[2], qj[3]);
        double w1[3], w2[3];
        R1.mul((double)p.ox[r.i], (double)p.oy[r.i], (double)p.oz[r.i], w1[0], w1[1], w1[2]);
        R2.mul((double)p.ox[r.j], (double)p.oy[r.j], (double)p.oz[r.j], w2[0], w2[1], w2[2]);
        double v1[3] = {(double)p.u[r.i] + (w1[1] * r1[2] - w1[2] * r1[1]),
                        (double)p.v[r.i] + (w1[2] * r1[0] - w1[0] * r1[2]),
                        (double)p.w[r.i] + (w1[0] * r1[1] - w1[1] * r1[0])};
        double v2[3] = {(double)p.u[r.j] + (w2[1] * r2[2] - w2[2] * r2[1]),
                        (double)p.v[r.j] + (w2[2] * r2[0] - w2[0] * r2[2]),
                        (double)p.w[r.j] + (w2[0] * r2[1] - w2[1] * r2[0])};
        double nvrel = (v2[0] - v1[0]) * o.cn[0] + (v2[1] - v1[1]) * o.cn[1] +
                       (v2[2] - v1[2]) * o.cn[2];
        if (nvrel < ZERO_R) {
          approach = true;
          ar.k = r.k; ar.i = r.i; ar.j = r.j;
          for (int c = 0; c < 3; ++c) { ar.r1[c] = r1[c]; ar.r2[c] = r2[c]; ar.cn[c] = o.cn[c]; }
        }
        // statistics records out of line: the solve's registers stay as
        // they are (the stores run only on statistics steps)
        if (csp)
          cstat_record(csp, r.k, r.i, r.j, p.sp[r.i], p.sp[r.j], nvrel, v2[0] - v1[0],
                       v2[1] - v1[1], v2[2] - v1[2], o.e1[0] - o.e2[0], o.e1[1] - o.e2[1],
                       o.e1[2] - o.e2[2], err);
      }
      if (ahead) {
        EarlyRec& e = ein.res[w];
        e.i = r.i;
        e.j = r.j;
        e.approach = approach ? 1 : 0;
        e.nfev = nf;
        for (int c = 0; c < 3; ++c) {
          e.r1[c] = approach ? ar.r1[c] : 0.0;
          e.r2[c] = approach ? ar.r2[c] : 0.0;
          e.cn[c] = approach ? ar.cn[c] : 0.0;
        }
        approach = false;
      }
    }
    const unsigned long long m = __ballot(approach);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      unsigned long long base = 0;
      if (lane == leader) {
        base = atomicAdd(nact, (unsigned long long)__popcll(m));
        atomicAdd(ncontact, (unsigned long long)__popcll(m));
      }
      base = __shfl(base, leader);
      if (approach) {
        const unsigned long long a = base + __popcll(m & ((1ull << lane) - 1ull));
        if (a < acap) act[a] = ar;
        else atomicOr(err, EB_ACTIVE_OVERFLOW);
      }
    }
  }
  // one atomic per wavefront for the evaluation and fallback counts
  for (int o = 32; o > 0; o >>= 1) {
    nfev += __shfl_xor(nfev, o);
    nfb += __shfl_xor(nfb, o);
  }
  if (ahead) return;   // counted when the results are used
  if (lane == 0 && nfev) atomicAdd(nfev_sum, nfev);
  if (lane == 0 && nfb) atomicAdd(nfallback, nfb);
}

// the surviving pairs of this step that were solved ahead: their active
// contacts (with this step's list keys) join those of the step's own solve,
// and their evaluations and pairs are counted
__global__ __launch_bounds__(BLOCK) void k_merge_early(
    const MergeRec* __restrict__ merge, const unsigned long long* nmerge, unsigned long long cap,
    const EarlyRec* __restrict__ res, ActiveRec* __restrict__ act, unsigned long long* nact,
    unsigned long long acap, unsigned long long* nfev_sum, unsigned long long* ncontact,
    unsigned long long* nlm, unsigned* err) {
  const unsigned long long n = min(*nmerge, cap);
  unsigned long long nfev = 0, npair = 0;
  for (unsigned long long b0 = (unsigned long long)blockIdx.x * blockDim.x; b0 < n;
       b0 += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned long long t = b0 + threadIdx.x;
    bool approach = false;
    ActiveRec ar;
    if (t < n) {
      const MergeRec mr = merge[t];
      const EarlyRec& e = res[mr.e];
      nfev += (unsigned long long)e.nfev;
      ++npair;
      if (e.approach) {
        approach = true;
        ar.k = mr.k; ar.i = e.i; ar.j = e.j;
        for (int c = 0; c < 3; ++c) { ar.r1[c] = e.r1[c]; ar.r2[c] = e.r2[c]; ar.cn[c] = e.cn[c]; }
      }
    }
    const unsigned long long a = wave_append(approach, nact);
    if (approach) {
      if (a < acap) act[a] = ar;
      else atomicOr(err, EB_ACTIVE_OVERFLOW);
    }
    const unsigned long long ma = __ballot(approach);
    if ((threadIdx.x & 63) == 0 && ma) atomicAdd(ncontact, (unsigned long long)__popcll(ma));
  }
  for (int o = 32; o > 0; o >>= 1) {
    nfev += __shfl_xor(nfev, o);
    npair += __shfl_xor(npair, o);
  }
  if ((threadIdx.x & 63) == 0 && npair) {
    atomicAdd(nfev_sum, nfev);
    atomicAdd(nlm, npair);
  }
}

// network forward for the 64 lanes of a one-wavefront workgroup: features
// -> seed angles.  Every lane calls it (valid or not).
__device__ __forceinline__ bool wave_seed(const MlpDev& W, bool valid, const float f[12],
                                          const double q1[4], double x[2], float (*sF)[16],
                                          float (*sY)[4]) {
  const int lane = threadIdx.x;
  for (int c = 0; c < 12; ++c) sF[lane][c] = valid ? f[c] : 0.f;
  for (int c = 12; c < 16; ++c) sF[lane][c] = 0.f;
  __syncthreads();
  const int rq = 4 * (lane >> 4);
  for (int tile = 0; tile < 4; ++tile) {
    const f32x4 xv = *reinterpret_cast<const f32x4*>(&sF[16 * tile + (lane & 15)][rq]);
    MlpAct a;
    const f32x4 y = mlp_forward_tile(W, to_bf(xv), a, lane);
    if (lane < 16)
      for (int c = 0; c < 4; ++c) sY[16 * tile + lane][c] = y[c];
  }
  __syncthreads();
  const bool ok =
      valid && seed_from_output(q1, sY[lane][0], sY[lane][1], sY[lane][2], sY[lane][3], x);
  __syncthreads();
  return ok;
}

// network seeds for the LM list (NaN: use the centre line)
template <typename VT>
__global__ __launch_bounds__(64) void k_mlp_seed(PState<VT> p, const Species* __restrict__ spec,
                                                 const WorkRec* __restrict__ work,
                                                 const unsigned long long* nwork,
                                                 unsigned long long cap, double L, MlpDev W,
                                                 double* __restrict__ seeds) {
  __shared__ __attribute__((aligned(16))) float sF[64][16];
  __shared__ float sY[64][4];
  const unsigned long long nw = min(*nwork, cap);
  for (unsigned long long w0 = (unsigned long long)blockIdx.x * 64; w0 < nw;
       w0 += (unsigned long long)gridDim.x * 64) {
    const unsigned long long w = w0 + threadIdx.x;
    const bool valid = w < nw;
    float f[12];
    double qi[4] = {1.0, 0.0, 0.0, 0.0};
    if (valid) {
      const WorkRec r = work[w];
      double bi[3], bj[3];
      min_image(p, r.i, r.j, L, bi, bj);
      qi[0] = p.qa[r.i]; qi[1] = p.qb[r.i]; qi[2] = p.qc[r.i]; qi[3] = p.qd[r.i];
      const double qj[4] = {p.qa[r.j], p.qb[r.j], p.qc[r.j], p.qd[r.j]};
      const Species si = spec[p.sp[r.i]], sj = spec[p.sp[r.j]];
      const double ai[3] = {si.a, si.b, si.c}, aj[3] = {sj.a, sj.b, sj.c};
      pair_features(bi, qi, ai, bj, qj, aj, f);
    }
    double x[2];
    const bool ok = wave_seed(W, valid, f, qi, x, sF, sY);
    if (valid) {
      seeds[2 * w] = ok ? x[0] : __builtin_nan("");
      seeds[2 * w + 1] = ok ? x[1] : __builtin_nan("");
    }
  }
}

// synthetic pairs solved from both seeds (the network's gain):
// per pair the evaluation counts, flags, fallback and |cn_ref . cn_net|
struct SeedEvalOut {
  int* nfev_ref;
  int* nfev_net;
  unsigned char* flag_ref;
  unsigned char* flag_net;
  unsigned char* fallback;
  float* cn_dot;
};
__global__ __launch_bounds__(64) void k_mlp_eval(unsigned long long seed, long long n, MlpDev W,
                                                 SeedGate gate, SeedEvalOut out) {
  __shared__ __attribute__((aligned(16))) float sF[64][16];
  __shared__ float sY[64][4];
  for (long long k0 = (long long)blockIdx.x * 64; k0 < n; k0 += (long long)gridDim.x * 64) {
    const long long k = k0 + threadIdx.x;
    const bool valid = k < n;
    PairPose pp{};
    float f[12];
    if (valid) {
      pp = sample_pair(seed, k);
      pair_features(pp.b1, pp.q1, pp.a1, pp.b2, pp.q2, pp.a2, f);
    } else {
      pp.q1[0] = 1.0;
    }
    double x[2];
    const bool ok = wave_seed(W, valid, f, pp.q1, x, sF, sY);
    if (!valid) continue;
    const Species s1 = pose_species(pp.a1), s2 = pose_species(pp.a2);
    const ContactOut r = contact_detect(pp.b1, pp.q1, s1, pp.b2, pp.q2, s2, nullptr);
    int nf = 0;
    bool fbb = false;
    const ContactOut o =
        contact_detect_gated(pp.b1, pp.q1, s1, pp.b2, pp.q2, s2, ok ? x : nullptr, gate, nf, fbb);
    const unsigned char fb = fbb ? 1 : 0;
    out.nfev_ref[k] = r.nfev;
    out.nfev_net[k] = nf;
    out.flag_ref[k] = (unsigned char)r.flag;
    out.flag_net[k] = (unsigned char)o.flag;
    out.fallback[k] = (unsigned char)(fb | (ok ? 0 : 2));
    out.cn_dot[k] = (float)fabs(r.cn[0] * o.cn[0] + r.cn[1] * o.cn[1] + r.cn[2] * o.cn[2]);
  }
}

constexpr int GREEDY_BLOCK = 1024;
constexpr int GREEDY_GRID = 256;

// One collision per particle in list order (collision_interaction.f90:190-
// 207, 397-398).  Contact flags depend only on positions and orientations,
// and a particle's velocity changes only once it is used, so the reference
// loop equals a greedy matching of the approaching contacts in list order.
// Rounds: every undecided contact that holds the lowest list index at both
// of its particles is accepted; contacts touching an accepted particle are
// dropped.  Then the accepted (disjoint) pairs exchange impulses.
// The first round runs over many workgroups (k_greedy_mark / _accept /
// _drop); the rare later rounds in one workgroup (k_greedy_rounds, which
// returns at once when nothing is undecided); then k_greedy_impulse.
// status: 0 undecided, 1 accepted, 2 dropped.  mark[] (per particle,
// LLONG_MAX when idle) and used[] are back at their idle values on exit.
__global__ __launch_bounds__(BLOCK) void k_greedy_mark(const ActiveRec* __restrict__ act,
                                                       const unsigned long long* nact,
                                                       unsigned long long acap,
                                                       long long* mark) {
  const long long na = (long long)min(*nact, acap);
  for (long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x; a < na;
       a += (long long)gridDim.x * blockDim.x) {
    atomicMin(&mark[act[a].i], act[a].k);
    atomicMin(&mark[act[a].j], act[a].k);
  }
}

__global__ __launch_bounds__(BLOCK) void k_greedy_accept(const ActiveRec* __restrict__ act,
                                                         const unsigned long long* nact,
                                                         unsigned long long acap,
                                                         const long long* __restrict__ mark,
                                                         int* __restrict__ used,
                                                         unsigned char* __restrict__ status) {
  const long long na = (long long)min(*nact, acap);
  for (long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x; a < na;
       a += (long long)gridDim.x * blockDim.x) {
    const ActiveRec& r = act[a];
    if (mark[r.i] == r.k && mark[r.j] == r.k) {
      // accepted pairs are disjoint: each particle is written by one pair
      status[a] = 1;
      used[r.i] = 1;
      used[r.j] = 1;
    } else {
      status[a] = 0;
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_greedy_drop(const ActiveRec* __restrict__ act,
                                                       const unsigned long long* nact,
                                                       unsigned long long acap,
                                                       long long* __restrict__ mark,
                                                       const int* __restrict__ used,
                                                       unsigned char* __restrict__ status,
                                                       unsigned long long* left) {
  const long long na = (long long)min(*nact, acap);
  for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < na;
       b0 += (long long)gridDim.x * blockDim.x) {
    const long long a = b0 + threadIdx.x;
    bool undecided = false;
    if (a < na) {
      const ActiveRec& r = act[a];
      if (status[a] == 0) {
        if (used[r.i] || used[r.j]) status[a] = 2;
        else undecided = true;
      }
      mark[r.i] = LLONG_MAX;
      mark[r.j] = LLONG_MAX;
    }
    const unsigned long long m = __ballot(undecided);
    if (m && (threadIdx.x & 63) == __ffsll((long long)m) - 1)
      atomicAdd(left, (unsigned long long)__popcll(m));
  }
}

__device__ __forceinline__ long long ld_relaxed(long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_relaxed(int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// rounds 2, 3, ... over the contacts still undecided (one workgroup)
__global__ __launch_bounds__(GREEDY_BLOCK) void k_greedy_rounds(
    const ActiveRec* __restrict__ act, const unsigned long long* nact, unsigned long long acap,
    long long* mark, int* used, unsigned char* __restrict__ status,
    const unsigned long long* left, unsigned long long* rounds_out) {
  if (*left == 0) {
    if (threadIdx.x == 0) *rounds_out += 1ull;
    return;
  }
  const long long na = (long long)min(*nact, acap);
  __shared__ int s_left;
  int rounds = 1;
  for (;;) {
    if (threadIdx.x == 0) s_left = 0;
    for (long long a = threadIdx.x; a < na; a += blockDim.x) {
      if (status[a] != 0) continue;
      atomicMin(&mark[act[a].i], act[a].k);
      atomicMin(&mark[act[a].j], act[a].k);
    }
    __syncthreads();
    for (long long a = threadIdx.x; a < na; a += blockDim.x) {
      if (status[a] != 0) continue;
      const long long k = act[a].k;
      if (ld_relaxed(&mark[act[a].i]) == k && ld_relaxed(&mark[act[a].j]) == k) {
        status[a] = 3;   // accepted this round (still resets its marks below)
        atomicExch(&used[act[a].i], 1);
        atomicExch(&used[act[a].j], 1);
      }
    }
    __syncthreads();
    for (long long a = threadIdx.x; a < na; a += blockDim.x) {
      const unsigned char st = status[a];
      if (st == 1 || st == 2) continue;
      if (st == 3) {
        status[a] = 1;
      } else if (ld_relaxed(&used[act[a].i]) || ld_relaxed(&used[act[a].j])) {
        status[a] = 2;
      } else {
        atomicAdd(&s_left, 1);
      }
      atomicExch((unsigned long long*)&mark[act[a].i], (unsigned long long)LLONG_MAX);
      atomicExch((unsigned long long*)&mark[act[a].j], (unsigned long long)LLONG_MAX);
    }
    __syncthreads();
    ++rounds;
    const int nleft = s_left;
    __syncthreads();
    if (nleft == 0) break;
  }
  if (threadIdx.x == 0) *rounds_out += (unsigned long long)rounds;
}

// impulses on the accepted, disjoint pairs; used[] back to 0
template <typename VT>
__global__ __launch_bounds__(BLOCK) void k_greedy_impulse(
    PState<VT> p, const Species* __restrict__ spec, const ActiveRec* __restrict__ act,
    const unsigned long long* nact, unsigned long long acap, int* __restrict__ used,
    const unsigned char* __restrict__ status, unsigned long long* ncoll, unsigned* err,
    int* __restrict__ ccount, long long* __restrict__ acc_key) {
  const long long na = (long long)min(*nact, acap);
  for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < na;
       b0 += (long long)gridDim.x * blockDim.x) {
    const long long a = b0 + threadIdx.x;
    bool acc = false;
    if (a < na) {
      const ActiveRec r = act[a];
      used[r.i] = 0;
      used[r.j] = 0;
      if (status[a] == 1) {
        acc = true;
        // for the collision statistics: the list position of the collision
        // that used each particle (k_cstat_reset restores the idle value)
        if (acc_key) {
          acc_key[r.i] = r.k;
          acc_key[r.j] = r.k;
        }
        double u1[3] = {(double)p.u[r.i], (double)p.v[r.i], (double)p.w[r.i]};
        double o1[3] = {(double)p.ox[r.i], (double)p.oy[r.i], (double)p.oz[r.i]};
        double u2[3] = {(double)p.u[r.j], (double)p.v[r.j], (double)p.w[r.j]};
        double o2[3] = {(double)p.ox[r.j], (double)p.oy[r.j], (double)p.oz[r.j]};
        const double q1[4] = {p.qa[r.i], p.qb[r.i], p.qc[r.i], p.qd[r.i]};
        const double q2[4] = {p.qa[r.j], p.qb[r.j], p.qc[r.j], p.qd[r.j]};
        const Species s1 = spec[p.sp[r.i]];
        if (!collision_response(s1, q1, q2, u1, o1, u2, o2, r.r1, r.r2, r.cn))
          atomicOr(err, EB_KE_COLLISION);
        p.u[r.i] = (VT)u1[0]; p.v[r.i] = (VT)u1[1]; p.w[r.i] = (VT)u1[2];
        p.ox[r.i] = (VT)o1[0]; p.oy[r.i] = (VT)o1[1]; p.oz[r.i] = (VT)o1[2];
        p.u[r.j] = (VT)u2[0]; p.v[r.j] = (VT)u2[1]; p.w[r.j] = (VT)u2[2];
        p.ox[r.j] = (VT)o2[0]; p.oy[r.j] = (VT)o2[1]; p.oz[r.j] = (VT)o2[2];
        // accepted pairs are disjoint: plain increments
        ccount[p.id[r.i]] += 1;
        ccount[p.id[r.j]] += 1;
      }
    }
    const unsigned long long m = __ballot(acc);
    if (m && (threadIdx.x & 63) == __ffsll((long long)m) - 1)
      atomicAdd(ncoll, (unsigned long long)__popcll(m));
  }
}

// Level-1 collision statistics (collision_interaction.f90:330-345, 409,
// 423-453).  The reference's ordered pass reaches a contact with both
// particles free exactly when no collision earlier in the list used either
// of them: with acc_key[p] the list position of the collision that used p
// (LLONG_MAX for none), a contact k is reached iff acc_key >= k at both
// (equal when it is that collision).  Reached contacts add to their bin
// (IDP1, IDP2): count, overlap depth, NVREL, |VREL|, angle; reached
// approaching ones are the collisions.  Sums go through LDS per workgroup,
// then one global atomic per bin and quantity.
constexpr int CS_NQ = 6;            // count, collisions, depth, nvrel, |vrel|, theta
constexpr int CS_LDS_BINS = 512;    // bins summed through LDS (24 KiB)

__global__ __launch_bounds__(BLOCK) void k_cstat_accum(CStat cs,
                                                       const long long* __restrict__ acc_key) {
  __shared__ double sacc[CS_LDS_BINS * CS_NQ];
  const bool lds = cs.nbin <= CS_LDS_BINS;
  const long long n = (long long)min(*cs.cnt, cs.cap);
  if (lds) {
    for (int t = threadIdx.x; t < cs.nbin * CS_NQ; t += blockDim.x) sacc[t] = 0.0;
    __syncthreads();
  }
  for (long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x; a < n;
       a += (long long)gridDim.x * blockDim.x) {
    const CollStatRec r = cs.rec[a];
    if (acc_key[r.i] < r.k || acc_key[r.j] < r.k) continue;
    const double q[CS_NQ] = {1.0, r.approach ? 1.0 : 0.0, r.depth, r.nvrel, r.nrm, r.theta};
    double* dst = lds ? sacc + r.bin * CS_NQ : cs.acc + (size_t)r.bin * CS_NQ;
    for (int c = 0; c < CS_NQ; ++c) atomicAdd(dst + c, q[c]);
  }
  if (lds) {
    __syncthreads();
    for (int t = threadIdx.x; t < cs.nbin * CS_NQ; t += blockDim.x)
      if (sacc[t] != 0.0) atomicAdd(cs.acc + t, sacc[t]);
  }
}

// acc_key back to idle over the accepted collisions
__global__ __launch_bounds__(BLOCK) void k_cstat_reset(const ActiveRec* __restrict__ act,
                                                       const unsigned long long* nact,
                                                       unsigned long long acap,
                                                       const unsigned char* __restrict__ status,
                                                       long long* __restrict__ acc_key) {
  const long long na = (long long)min(*nact, acap);
  for (long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x; a < na;
       a += (long long)gridDim.x * blockDim.x) {
    if (status[a] != 1) continue;
    acc_key[act[a].i] = LLONG_MAX;
    acc_key[act[a].j] = LLONG_MAX;
  }
}

// MEAN_PART_COL of the step (collision_interaction.f90:423-441) into last[]
// and added to tsum[] (:453, the STAT_TIME sum); the sums start again from
// the reference's ZERO (1e-33, :171), so that a bin no contact reached reads
// ZERO in 1, 2, 4, 6-8 and 1 in 5, as the reference's does.  The record
// count is cleared and the cycle number advanced every step.  One workgroup.
__global__ __launch_bounds__(BLOCK) void k_cstat_final(CStat cs, double dtime) {
  const long long cyc = *cs.cycle;
  const bool on = (cyc % (long long)cs.fout2) == 0;
  __syncthreads();
  if (on) {
    for (int b = threadIdx.x; b < cs.nbin; b += blockDim.x) {
      double* a = cs.acc + (size_t)b * CS_NQ;
      const double m[8] = {a[0], a[1], 2.0 * a[1] / dtime, a[2], a[2] / a[0], a[3], a[4], a[5]};
      for (int c = 0; c < 8; ++c) {
        cs.last[(size_t)b * 8 + c] = m[c];
        cs.tsum[(size_t)b * 8 + c] += m[c];
      }
      for (int c = 0; c < CS_NQ; ++c) a[c] = ZERO_R;   // MEAN_PART_COL = ZERO (:171)
    }
  }
  if (threadIdx.x == 0) {
    if (on) cs.cnt[1] += 1ull;
    cs.cnt[0] = 0ull;
    *cs.cycle = cyc + 1;
  }
}

// contact detection on caller-given pairs (API / tests)
template <typename VT>
__global__ __launch_bounds__(BLOCK) void k_contact_batch(
    PState<VT> p, const Species* __restrict__ spec, const int* __restrict__ home,
    const int* __restrict__ nr, const int* __restrict__ pos_of_id, long long npairs, double L,
    int np, unsigned char* flag, double* e1, double* e2, double* cn, int* nfev, unsigned* err,
    const double* __restrict__ seeds, SeedGate gate, unsigned long long* nfallback) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs) return;
  const int hi = home[t] - 1, ni = nr[t] - 1;
  if (hi < 0 || hi >= np || ni < 0 || ni >= np) {
    atomicOr(err, EB_OUT_OF_GRID);
    return;
  }
  const int i = pos_of_id[hi], j = pos_of_id[ni];
  double bi[3], bj[3];
  min_image(p, i, j, L, bi, bj);
  const double qi[4] = {p.qa[i], p.qb[i], p.qc[i], p.qd[i]};
  const double qj[4] = {p.qa[j], p.qb[j], p.qc[j], p.qd[j]};
  // the step's own solve: the network seed through its gate when given
  const double* sd = (seeds && seeds[2 * t] == seeds[2 * t]) ? seeds + 2 * t : nullptr;
  int nf = 0;
  bool fb = false;
  const ContactOut o =
      contact_detect_gated(bi, qi, spec[p.sp[i]], bj, qj, spec[p.sp[j]], sd, gate, nf, fb);
  if (fb) atomicAdd(nfallback, 1ull);
  flag[t] = (unsigned char)o.flag;
  for (int c = 0; c < 3; ++c) {
    e1[t * 3 + c] = o.e1[c];
    e2[t * 3 + c] = o.e2[c];
    cn[t * 3 + c] = o.cn[c];
  }
  nfev[t] = nf;
}

// the work list of a batch of 1-based (HOME, NEAR) pairs (slots of the
// resident particles), for the network's seeds
__global__ __launch_bounds__(BLOCK) void k_work_of_pairs(const int* __restrict__ home,
                                                         const int* __restrict__ nr,
                                                         const int* __restrict__ pos_of_id,
                                                         long long npairs, int np,
                                                         WorkRec* __restrict__ work,
                                                         unsigned* err) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs) return;
  const int hi = home[t] - 1, ni = nr[t] - 1;
  WorkRec r;
  r.k = t;
  if (hi < 0 || hi >= np || ni < 0 || ni >= np) {
    atomicOr(err, EB_OUT_OF_GRID);
    r.i = r.j = 0;
  } else {
    r.i = pos_of_id[hi];
    r.j = pos_of_id[ni];
  }
  work[t] = r;
}

// the approaching contacts of the last collision pass: list key, 1-based
// reference indices of both particles, resolved by the one-collision pass
template <typename VT>
__global__ __launch_bounds__(BLOCK) void k_contacts_out(PState<VT> p,
                                                        const ActiveRec* __restrict__ act,
                                                        const unsigned char* __restrict__ status,
                                                        long long na, long long* key, int* home,
                                                        int* nr, unsigned char* resolved) {
  const long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= na) return;
  key[a] = act[a].k;
  home[a] = p.id[act[a].i] + 1;
  nr[a] = p.id[act[a].j] + 1;
  resolved[a] = status[a] == 1 ? 1 : 0;
}

template <typename VT>
__global__ void k_set_groups(PState<VT> p, const int* __restrict__ g, int np) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) p.grp[i] = g[p.id[i]];
}

template <typename VT>
__global__ void k_pos_of_id(PState<VT> p, int* pos, int np) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) pos[p.id[i]] = i;
}

template <typename VT>
__global__ void k_gather_out(PState<VT> p, double* out, int np) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double st[13];
  load_state(p, i, st);
  const int id = p.id[i];
  for (int c = 0; c < 13; ++c) out[(long long)c * np + id] = st[c];
}

template <typename VT>
__global__ void k_scatter_in(PState<VT> p, const double* in, const int* sp,
                             const Species* __restrict__ spec, int np) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double st[13];
  for (int c = 0; c < 13; ++c) st[c] = in[(long long)c * np + i];
  store_state(p, i, st);
  p.sp[i] = sp[i];
  p.rm[i] = float_up(spec[sp[i]].rmax);
  p.id[i] = i;
  p.grp[i] = 0;
}

// per-particle copy of the species' largest semi-axis after a new table
template <typename VT>
__global__ void k_refresh_rm(PState<VT> p, const Species* __restrict__ spec, int np) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) p.rm[i] = float_up(spec[p.sp[i]].rmax);
}

}  // namespace gp

// ===========================================================================
// host side
// ===========================================================================
using namespace gp;

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
};

constexpr int NCNT_HOST = 13;   // device counters mirrored to the host (16 allocated)

struct gp_ctx {
  gp_params prm{};
  StepParams sp{};
  hipStream_t st = nullptr;
  int device = 0;
  std::string err;
  bool mixed = false;
  // field
  void* field = nullptr;
  bool have_field = false;
  // species
  Species* spec = nullptr;
  int nspec = 0;
  double max_emaj = 0.0;
  double max_semi = 0.0;        // largest semi-axis of any species
  // particles
  int np = 0;
  void* pbuf[2] = {nullptr, nullptr};
  int cur = 0;
  int* perm = nullptr;          // order in which the next advance reads cur
  int ngroups = 1;              // ensembles sharing the context
  bool perm_identity = true;
  int max_sp = -1;              // largest species index of the particles
  int lm_waves = 1024;          // LM launch: one wavefront per SIMD
  int nsimd = 1024;             // SIMDs of the device (4 per CU)
  int lm_waves_last = 0;        // width of the last contact-solve launch issued
  // brick visiting order of the advance (gp_params.brick)
  BrickGrid bg{};
  bool use_bricks = false;
  int* brick = nullptr;
  int* bslot = nullptr;
  int* bcnt = nullptr;
  int* boff = nullptr;
  int* trav = nullptr;
  float4* pos4 = nullptr;       // (x, y, z, rmax) as float in cell (sorted) order
  int* csort = nullptr;         // cell of each sorted position
  int* gsort = nullptr;         // ensemble of each sorted position (several ensembles)
  // pipelined steps (gp_params.overlap)
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // field tiles: the fluid at the particles of the next advance, by
  // visiting position (6 x np of the compute type), gathered before it;
  // flpre_valid: the last advance launch filled it (the listed advance of
  // the same step reads it, after ev_gather when it was filled on st2)
  void* flpre = nullptr;
  bool flpre_valid = false, flpre_st2 = false;
  hipEvent_t ev_gather = nullptr;
  int* inlm = nullptr;          // per particle: in a contact solve of this step
  int* plist = nullptr;         // those particles (storage indices)
  int* inv = nullptr;           // storage index -> visiting position of the next advance
  int* ccount = nullptr;        // collisions per particle (reference index)
  // contact-seed network (fp32 masters, Adam moments, bf16 copies)
  float* mlp_th = nullptr;
  float* mlp_m = nullptr;
  float* mlp_v = nullptr;
  float* mlp_grad = nullptr;
  short* mlp_wb = nullptr;      // w1 | w2 | w3 | w2t | w3t (bf16)
  float* mlp_bias = nullptr;    // b1 | b2 | b3
  float* mlp_part = nullptr;    // per-wavefront partial gradients
  float* mlp_lpart = nullptr;
  float* mlp_loss = nullptr;
  int mlp_step = 0;
  bool mlp_ready = false;
  double* seeds = nullptr;      // 2 per LM-list entry
  // acceptance gate of a network-seeded solve (gp_seed_net_gate)
  SeedGate gate{SEED_GATE_MAXFEV, 0.0};
  // the approaching-contact list (act, act_status) belongs to the particles
  // of buffer cur until the next advance (gp_get_contacts)
  bool contacts_valid = false;
  // graph-mode gp_step: the lists were sized with headroom by an eager step
  // (steps may then run without host synchronisation)
  bool lists_sized = false;
  // the last collision pass was a pipelined gp_collision_advance: its solve
  // flags (inlm) are set for the particles of plist only
  bool pipe_prev = false;
  int64_t batch_fallbacks = 0;  // gated seeded solves redone in the last gp_contact_batch
  double* scratch = nullptr;    // this device's copy of the gp_interph_noalloc staging
  // level-2 statistics: u - mean at the origin of each correlation slot,
  // 3 x np doubles by reference index (allocated when a slot starts)
  double* lag_u0[GP_NLGR] = {};
  int lag_np = 0;
  // collision grid
  CellGrid g{};
  int* cell = nullptr;
  int* slot = nullptr;
  int* cnt = nullptr;
  int* off = nullptr;
  unsigned long long* tkey = nullptr;   // order key per cell slot (sort)
  int* sorted = nullptr;
  int* pcount = nullptr;
  long long* poff = nullptr;
  void* scan_scr = nullptr;
  bool cells_valid = false;
  long long* mark = nullptr;
  int* used = nullptr;
  int* pos_of_id = nullptr;
  // work lists
  WorkRec* work = nullptr;    // bounding-sphere survivors
  WorkRec* lmwork = nullptr;  // pairs left for the LM (same capacity)
  unsigned long long work_cap = 0;
  ActiveRec* act = nullptr;
  unsigned char* act_status = nullptr;
  unsigned long long act_cap = 0;
  // counters (device)
  unsigned* d_err = nullptr;
  unsigned long long* d_cnt = nullptr;   // 0 nwall 1 nwork 2 nact 3 nfev 4 ncontact 5 ncoll 6 rounds 7 nlm 8 undecided
                                         // 9 seed fallbacks 10 particles in solves 11 candidates
                                         // 12 pairs solved ahead 13, 14 lists to solve ahead
  unsigned long long* h_cnt = nullptr;   // pinned mirror of the first NCNT_HOST
  long long* h_total = nullptr;
  double* flout = nullptr;
  gp_stats stats{};
  long long last_nclose = 0;
  // stage timing: ev[0] before the advance, ev[1] after it, ev[2] after the
  // pair-count scan, ev[3] after the LM, ev[4] after the greedy pass
  bool timing = false;
  hipEvent_t ev[6] = {};   // ev[5]: before the contact-solve kernel
  bool ev_pending = false;
  double t_stage[4] = {0, 0, 0, 0};
  double t_solve = 0.0;     // the contact-solve kernel alone (ev[5] -> ev[3])
  long long ntimed = 0;
  // captured two-step graph (gp_params.graph) and the buffer configuration
  // it was captured in
  hipGraphExec_t gexec = nullptr;
  int g_cur = -1;
  int* g_perm = nullptr;
  bool g_perm_identity = false;
  // level-1 collision statistics (off while cs_fout2 == 0): records of the
  // step's contacts, per-bin sums (6 per bin), the last statistics step's
  // MEAN_PART_COL (8 per bin) and their running sum (8 per bin)
  int cs_fout2 = 0, cs_nclass = 0, cs_npoly = 0, cs_nbin = 0;
  CollStatRec* cs_rec = nullptr;
  unsigned long long cs_cap = 0;
  double* cs_buf = nullptr;              // acc | last | tsum
  unsigned long long* cs_cnt = nullptr;  // 0 records 1 statistics steps
  long long* cs_cycle = nullptr;         // cycle number of the next collision pass
  CStat* cs_dev = nullptr;               // device copy of cstat() for k_contact
  // the species rows as given (host), and whether flout holds the fluid at
  // the particles of an advance of the current particles
  std::vector<double> sp_rows;
  bool fluid_valid = false;
  // set around a sort inside pipelined steps: its scatter clears the
  // previous step's solve flags
  bool clear_plist = false;
  // solving ahead (pipelined steps, gp_params.overlap = 1): per step parity
  // p, the filter lists its surviving pairs in ep_pair[p] (count d_cnt[13 +
  // p]) with the hash table ep_keys/ep_vals[p], and looks pairs up in the
  // other parity's, solved from this step's state by k_contact on st3 into
  // ep_res; ep_merge takes the pairs found (count d_cnt[12])
  hipStream_t st3 = nullptr;
  hipEvent_t ev_efork = nullptr, ev_edone = nullptr;
  unsigned long long* ep_keys[2] = {nullptr, nullptr};
  int* ep_vals[2] = {nullptr, nullptr};
  int* ep_pair[2] = {nullptr, nullptr};
  int* ep_slot[2] = {nullptr, nullptr};
  EarlyRec* ep_res = nullptr;
  MergeRec* ep_merge = nullptr;
  unsigned long long ep_tab = 0;   // table size (a power of two)
  unsigned long long ep_cap = 0;   // pairs per list
  int ep_par = 0;
  bool ep_step = false;            // this step solves ahead
  unsigned long long last_nlm = 0; // last known count of contact solves
};

// ---------------------------------------------------------------------------
// Staging of gp_interph_noalloc: a file-scope device array, one copy per GPU
// (held until the process ends).  Host code never uses its name as an
// address: each device's copy is looked up once, with that device current,
// when the first context on it is created, and kernels and copies take that
// address.  Every call on a GPU shares its copy.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[4 * (size_t)GP_INTERPH_NOALLOC_MAX];
constexpr int GP_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch[GP_MAX_DEVICES] = {};

static int scratch_init(gp_ctx* ctx) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (ctx->device < 0 || ctx->device >= GP_MAX_DEVICES) {
    ctx->err = "device index beyond the staging table";
    return GP_ERR_HIP;
  }
  if (!g_scratch[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return GP_ERR_HIP;
    }
    g_scratch[ctx->device] = (double*)p;
  }
  ctx->scratch = g_scratch[ctx->device];
  return GP_OK;
}

static void drop_graph(gp_ctx* ctx) {
  if (ctx->gexec) hipGraphExecDestroy(ctx->gexec);
  ctx->gexec = nullptr;
}

// bf16 block offsets (in shorts) and bias offsets of the seed network
constexpr int MW_W1 = 0, MW_W2 = MW_W1 + 64 * 16, MW_W3 = MW_W2 + 64 * 64,
              MW_W2T = MW_W3 + 16 * 64, MW_W3T = MW_W2T + 64 * 64, MW_N = MW_W3T + 64 * 16;
constexpr int MB_B1 = 0, MB_B2 = 64, MB_B3 = 128, MB_N = 144;
constexpr int MLP_PARTS = 512;   // wavefronts of the training kernel

static MlpDev mlp_dev(gp_ctx* ctx) {
  MlpDev d;
  d.w1 = ctx->mlp_wb + MW_W1; d.w2 = ctx->mlp_wb + MW_W2; d.w3 = ctx->mlp_wb + MW_W3;
  d.w2t = ctx->mlp_wb + MW_W2T; d.w3t = ctx->mlp_wb + MW_W3T;
  d.b1 = ctx->mlp_bias + MB_B1; d.b2 = ctx->mlp_bias + MB_B2; d.b3 = ctx->mlp_bias + MB_B3;
  return d;
}

#define HIPCHK(x)                                                               \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      ctx->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " +  \
                 __FILE__ + ":" + std::to_string(__LINE__);                    \
      return GP_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

static inline unsigned nblk(long long n, int b = BLOCK) {
  return (unsigned)((n + b - 1) / b);
}

template <typename VT>
static PState<VT> pstate(gp_ctx* ctx, int which) {
  PState<VT> s;
  char* b = (char*)ctx->pbuf[which];
  const size_t n = (size_t)ctx->np;
  auto dd = [&](int k) { return (double*)(b); };
  (void)dd;
  double* d = (double*)b;
  s.x = d; s.y = d + n; s.z = d + 2 * n;
  s.qa = d + 3 * n; s.qb = d + 4 * n; s.qc = d + 5 * n; s.qd = d + 6 * n;
  VT* v = (VT*)(d + 7 * n);
  s.u = v; s.v = v + n; s.w = v + 2 * n; s.ox = v + 3 * n; s.oy = v + 4 * n; s.oz = v + 5 * n;
  s.rm = (float*)(v + 6 * n);
  int* ip = (int*)(s.rm + n);
  s.sp = ip; s.id = ip + n; s.grp = ip + 2 * n;
  return s;
}

static size_t pstate_bytes(int np, bool mixed) {
  const size_t n = (size_t)np;
  return n * 7 * 8 + n * 6 * (mixed ? 4 : 8) + n * 4 + n * 3 * 4 + 256;
}

template <typename T>
static int dalloc(gp_ctx* ctx, T** p, size_t count) {
  if (*p) hipFree(*p);
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    ctx->err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
    return GP_ERR_NOMEM;
  }
  return GP_OK;
}

static int map_err(gp_ctx* ctx, unsigned e) {
  if (!e) return GP_OK;
  int code = GP_ERR_STATE;
  if (e & EB_NAN_DRAG) code = GP_ERR_NAN_DRAG;
  else if (e & EB_NAN_OMEGA) code = GP_ERR_NAN_OMEGA;
  else if (e & EB_KE_COLLISION) code = GP_ERR_KE_COLLISION;
  else if (e & EB_KE_WALL) code = GP_ERR_KE_WALL;
  else if (e & EB_OUT_OF_GRID) code = GP_ERR_OUT_OF_GRID;
  else if (e & (EB_OVERFLOW | EB_ACTIVE_OVERFLOW)) code = GP_ERR_NCLOSE_OVERFLOW;
  else if (e & EB_INTERP_RANGE) code = GP_ERR_INTERP_RANGE;
  char buf[128];
  snprintf(buf, sizeof buf, "kernel error bits 0x%x (status %d)", e, code);
  ctx->err = buf;
  return code;
}

// reads and clears the device error word (synchronises the stream)
static int check_err(gp_ctx* ctx) {
  unsigned e = 0;
  HIPCHK(hipMemcpyAsync(&e, ctx->d_err, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (e) {
    unsigned z = 0;
    HIPCHK(hipMemcpy(ctx->d_err, &z, sizeof z, hipMemcpyHostToDevice));
  }
  return map_err(ctx, e);
}

// ---------------------------------------------------------------------------
// STAT_PARTICLE, level 1 (stat_particle.f90:169-450): per species row (one
// per class J and size index IDP) the 41 entries of MEAN_PART(:, J, IDP).
// Pass 1 sums the counts, velocities and angular velocities (body frame and
// rotated to the lab frame as transform_basis.f90:43-77 does); pass 2 the
// products of the fluctuations about those means, with the body-frame and
// lab-frame inertia; a last pass forms entries 10, 39, 40, 41 from the sums
// and divides everything by a count (entry 23 is that count).  Species rows
// are class-major, npoly size indices per class.  The reference divides
// every class's entries by NPM(IDP), which after its class loop holds the
// counts of the last class (stat_particle.f90:436-449); ref_norm = 1 keeps
// that, ref_norm = 0 divides each row by its own count.  Sums over
// at most PST_LDS species go through LDS first, one global atomic per entry
// and workgroup; larger tables take global atomics directly.  The order of
// the floating-point sums differs from the reference's particle order.
// ---------------------------------------------------------------------------
constexpr int NSTP = 41;
constexpr int PST_LDS = 8;

template <typename VT>
__device__ __forceinline__ void pst_load(const PState<VT>& p, int i, double u[3], double w[3],
                                         double g[3], Rot<double>& R) {
  u[0] = (double)p.u[i]; u[1] = (double)p.v[i]; u[2] = (double)p.w[i];
  w[0] = (double)p.ox[i]; w[1] = (double)p.oy[i]; w[2] = (double)p.oz[i];
  R = Rot<double>(p.qa[i], p.qb[i], p.qc[i], p.qd[i]);
  R.mul(w[0], w[1], w[2], g[0], g[1], g[2]);
}

template <typename VT, int PASS>
__global__ __launch_bounds__(BLOCK) void k_pstat(PState<VT> p, const Species* __restrict__ spec,
                                                 int np, int nspec, double* acc,
                                                 const double* __restrict__ mean) {
  __shared__ double s_acc[PST_LDS * NSTP];
  const bool lds = nspec <= PST_LDS;
  if (lds) {
    for (int t = threadIdx.x; t < nspec * NSTP; t += blockDim.x) s_acc[t] = 0.0;
    __syncthreads();
  }
  auto add = [&](int sp, int k, double v) {
    if (lds) atomicAdd(&s_acc[sp * NSTP + k], v);
    else atomicAdd(&acc[(size_t)sp * NSTP + k], v);
  };
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
    const int sp = p.sp[i];
    double u[3], w[3], g[3];
    Rot<double> R(1.0, 0.0, 0.0, 0.0);
    pst_load(p, i, u, w, g, R);
    if (PASS == 1) {
      add(sp, 22, 1.0);
      for (int c = 0; c < 3; ++c) {
        add(sp, c, u[c]);
        add(sp, 10 + c, w[c]);
        add(sp, 13 + c, g[c]);
      }
    } else {
      const double* m = mean + (size_t)sp * 10;
      const double uf[3] = {u[0] - m[0], u[1] - m[1], u[2] - m[2]};
      const double of[3] = {w[0] - m[3], w[1] - m[4], w[2] - m[5]};
      const double gf[3] = {g[0] - m[6], g[1] - m[7], g[2] - m[8]};
      const Species S = spec[sp];
      // R diag(I) R^T as matmul(R, matmul(diag(I), transpose(R)))
      const double ib[3] = {S.ixx, S.iyy, S.izz};
      const double r[3][3] = {{R.r00, R.r01, R.r02}, {R.r10, R.r11, R.r12}, {R.r20, R.r21, R.r22}};
      auto il = [&](int a, int b) {
        return r[a][0] * (ib[0] * r[b][0]) + r[a][1] * (ib[1] * r[b][1]) + r[a][2] * (ib[2] * r[b][2]);
      };
      add(sp, 3, uf[0] * uf[0]); add(sp, 4, uf[1] * uf[1]); add(sp, 5, uf[2] * uf[2]);
      add(sp, 6, uf[0] * uf[1]); add(sp, 7, uf[0] * uf[2]); add(sp, 8, uf[1] * uf[2]);
      add(sp, 16, of[0] * of[0]); add(sp, 17, of[1] * of[1]); add(sp, 18, of[2] * of[2]);
      add(sp, 19, of[0] * of[1]); add(sp, 20, of[0] * of[2]); add(sp, 21, of[1] * of[2]);
      add(sp, 23, S.ixx * of[0] * of[0]); add(sp, 24, S.iyy * of[1] * of[1]);
      add(sp, 25, S.izz * of[2] * of[2]);
      add(sp, 26, gf[0] * gf[0]); add(sp, 27, gf[1] * gf[1]); add(sp, 28, gf[2] * gf[2]);
      add(sp, 29, gf[0] * gf[1]); add(sp, 30, gf[0] * gf[2]); add(sp, 31, gf[1] * gf[2]);
      add(sp, 32, il(0, 0) * gf[0] * gf[0]); add(sp, 33, il(1, 1) * gf[1] * gf[1]);
      add(sp, 34, il(2, 2) * gf[2] * gf[2]); add(sp, 35, il(0, 1) * gf[0] * gf[1]);
      add(sp, 36, il(0, 2) * gf[0] * gf[2]); add(sp, 37, il(1, 2) * gf[1] * gf[2]);
    }
  }
  if (lds) {
    __syncthreads();
    for (int t = threadIdx.x; t < nspec * NSTP; t += blockDim.x)
      if (s_acc[t] != 0.0) atomicAdd(&acc[t], s_acc[t]);
  }
}

// per row: the means for pass 2 (u v w, body omega, lab omega) and the count
__global__ void k_pstat_mean(const double* __restrict__ acc, double* __restrict__ mean,
                             int nspec) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nspec) return;
  const double* a = acc + (size_t)s * NSTP;
  const int src[9] = {0, 1, 2, 10, 11, 12, 13, 14, 15};
  for (int c = 0; c < 9; ++c) mean[(size_t)s * 10 + c] = a[src[c]] / a[22];
  mean[(size_t)s * 10 + 9] = a[22];
}

// entries 10, 39, 40, 41 from the sums, then the division by the count
__global__ void k_pstat_final(const Species* __restrict__ spec, double* acc,
                              const double* __restrict__ mean, int nspec, int npoly,
                              int ref_norm) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nspec) return;
  double* a = acc + (size_t)s * NSTP;
  const Species S = spec[s];
  a[9] = 0.5 * (a[3] + a[4] + a[5]);
  a[38] = 0.5 * (S.ixx * a[16] + S.iyy * a[17] + S.izz * a[18]);
  a[39] = 0.5 * (a[23] + a[24] + a[25]);
  a[40] = a[9] + a[39];
  const int nrow = ref_norm ? nspec - npoly + s % npoly : s;
  const double n = mean[(size_t)nrow * 10 + 9];
  for (int k = 0; k < NSTP; ++k) a[k] = a[k] / n;
  a[22] = n;
}

// ---------------------------------------------------------------------------
// STAT_PARTFLUID (stat_partfluid.f90:113-381): per particle, from its state
// and the fluid velocity and vorticity at it of the last advance (UFAP ..
// VORTZAP), the particle Reynolds numbers, the Zastawny et al. (2012)
// coefficients with the fixed fits of that routine (ellipsoid 2, a/b =
// 1.25, :63-100), drag and lift (rotated to the lab frame), pitching and
// rotation torques (body frame), relaxation times, incidence angle, the
// lab-frame major axis and the velocities; summed per species row from the
// reference's ZERO, divided by the count NPM(IDP) (the last class's with
// ref_norm, as :362-381 leaves it).  Per-species constants (PMASS, dequiv,
// TAUPF, GRAVITY) come from the host: pfc[4 * s ..].
// ---------------------------------------------------------------------------
constexpr int NPFS = 36;

struct PfFit {   // stat_partfluid.f90:63-100
  static constexpr double a0 = 1.95, a1 = 18.12, a2 = 1.023, a3 = 4.26, a4 = 0.384, a5 = 21.52,
                          a6 = 0.990, a7 = 2.86, a8 = 0.260;
  static constexpr double b1 = 0.083, b2 = -0.21, b3 = 1.582, b4 = 0.851, b5 = 1.842, b6 = -0.802,
                          b7 = -0.006, b8 = 0.874, b9 = 0.009, b10 = 0.570;
  static constexpr double c1 = 0.935, c2 = 0.146, c3 = -0.469, c4 = 0.145, c5 = 0.116, c6 = 0.748,
                          c7 = 0.041, c8 = 0.221, c9 = 0.657, c10 = 0.044;
  static constexpr double r1 = 0.573, r2 = -0.154, r3 = 116.61, r4 = 1.0;
  static constexpr double r1_ = 1.244, r2_ = 0.239, r3_ = 378.12, r4_ = 0.789;
};

// the 36 entries of one particle (index k - 1 of MEAN_PARTFLUID(k, J, IDP))
__device__ __forceinline__ void partfluid_entries(const Rot<double>& R, const double up[3],
                                                  const double om[3], const double uf[6],
                                                  const double* c, double visc, double rhof,
                                                  double e[NPFS]) {
  using F = PfFit;
  const double pm = c[0], dq = c[1];
  const double ur[3] = {uf[0] - up[0], uf[1] - up[1], uf[2] - up[2]};
  const double urm = sqrt(ur[0] * ur[0] + ur[1] * ur[1] + ur[2] * ur[2]);
  const double rep = (urm * dq) / visc;
  double u[3];
  R.mulT(ur[0], ur[1], ur[2], u[0], u[1], u[2]);   // by conj_q: to the body frame
  const double s23 = sqrt(u[1] * u[1] + u[2] * u[2]);
  const double phi = fabs(atan(s23 / u[0]));
  const double um = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  const double sp = sin(phi), cp = cos(phi);
  const double cd0 = F::a1 / pow(rep, F::a2) + F::a3 / pow(rep, F::a4);
  const double cd90 = F::a5 / pow(rep, F::a6) + F::a7 / pow(rep, F::a8);
  const double cd = cd0 + (cd90 - cd0) * pow(sp, F::a0);
  const double cl = (F::b1 / pow(rep, F::b2) + F::b3 / pow(rep, F::b4)) *
                    pow(sp, F::b5 + F::b6 * pow(rep, F::b7)) *
                    pow(cp, F::b8 + F::b9 * pow(rep, F::b10));
  const double k = (1.0 / 2.0) * (rhof / pm) * (1.0 / 4.0) * (M_PI * (dq * dq));
  double fd[3], fl[3];
  for (int a = 0; a < 3; ++a) fd[a] = k * cd * um * u[a];
  fl[0] = k * cl * (um * um) * sp * copysign(1.0, -u[0]);
  fl[1] = k * cl * (um * um) * cp * (u[1] / s23);
  fl[2] = k * cl * (um * um) * cp * (u[2] / s23);
  double fdl[3], fll[3];
  R.mul(fd[0], fd[1], fd[2], fdl[0], fdl[1], fdl[2]);
  R.mul(fl[0], fl[1], fl[2], fll[0], fll[1], fll[2]);
  e[0] = rep; e[1] = cd; e[2] = cl;
  for (int a = 0; a < 3; ++a) { e[3 + a] = fdl[a]; e[6 + a] = fll[a]; }
  e[9] = sqrt(fdl[0] * fdl[0] + fdl[1] * fdl[1] + fdl[2] * fdl[2]);
  e[10] = sqrt(fll[0] * fll[0] + fll[1] * fll[1] + fll[2] * fll[2]);
  double cu[3];
  R.mulT(uf[3], uf[4], uf[5], cu[0], cu[1], cu[2]);
  const double orl[3] = {0.5 * cu[0] - om[0], 0.5 * cu[1] - om[1], 0.5 * cu[2] - om[2]};
  const double orm = sqrt(orl[0] * orl[0] + orl[1] * orl[1] + orl[2] * orl[2]);
  const double oyz = sqrt(orl[2] * orl[2] + orl[1] * orl[1]);
  const double rer = (orm * (dq * dq)) / visc;
  const double ctp = (F::c1 / pow(rep, F::c2) + F::c3 / pow(rep, F::c4)) *
                     pow(sp, F::c5 + F::c6 * pow(rep, F::c7)) *
                     pow(cp, F::c8 + F::c9 * pow(rep, F::c10));
  const double ctr = F::r1 * pow(rer, F::r2) + F::r3 / pow(rer, F::r4);
  const double ctr2 = F::r1_ * pow(rer, F::r2_) + F::r3_ / pow(rer, F::r4_);
  const double kp = (1.0 / 4.0) * (rhof / pm) * (1.0 / 4.0) * M_PI * (dq * dq * dq) * ctp *
                    (um * um);
  const double tp[3] = {0.0, kp * (fabs(u[2]) / s23) * copysign(1.0, u[0] * u[2]),
                        kp * (fabs(u[1]) / s23) * copysign(1.0, u[0] * u[1])};
  const double h = dq / 2.0, h5 = h * h * h * h * h;
  const double kr = (rhof / pm) / 2.0 * h5;
  const double tr[3] = {ctr * kr * fabs(orl[0]) * orl[0], ctr2 * kr * oyz * orl[1],
                        ctr2 * kr * oyz * orl[2]};
  e[11] = rer; e[12] = ctp; e[13] = ctr;
  for (int a = 0; a < 3; ++a) { e[14 + a] = tp[a]; e[17 + a] = tr[a]; }
  e[20] = sqrt(tp[0] * tp[0] + tp[1] * tp[1] + tp[2] * tp[2]);
  e[21] = sqrt(tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2]);
  e[22] = c[2];   // TAUPF
  e[23] = c[3];   // GRAVITY
  e[24] = phi;
  e[25] = R.r00; e[26] = R.r10; e[27] = R.r20;   // R (1, 0, 0)
  e[28] = (0.5 * rhof * cd * (M_PI / 4.0) * (dq * dq) * um) / pm;
  for (int a = 0; a < 3; ++a) { e[29 + a] = up[a]; e[32 + a] = om[a]; }
  e[35] = 1.0;
}

template <typename VT>
__global__ __launch_bounds__(BLOCK) void k_pfstat(PState<VT> p, const double* __restrict__ fl,
                                                  const double* __restrict__ pfc, int np,
                                                  int nspec, double visc, double rhof,
                                                  double* acc) {
  __shared__ double s_acc[PST_LDS * NPFS];
  const bool lds = nspec <= PST_LDS;
  if (lds) {
    for (int t = threadIdx.x; t < nspec * NPFS; t += blockDim.x) s_acc[t] = 0.0;
    __syncthreads();
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
    const int sp = p.sp[i];
    const long long id = p.id[i];
    const Rot<double> R(p.qa[i], p.qb[i], p.qc[i], p.qd[i]);
    const double up[3] = {(double)p.u[i], (double)p.v[i], (double)p.w[i]};
    const double om[3] = {(double)p.ox[i], (double)p.oy[i], (double)p.oz[i]};
    double uf[6];
    for (int c = 0; c < 6; ++c) uf[c] = fl[(long long)c * np + id];
    double e[NPFS];
    partfluid_entries(R, up, om, uf, pfc + 4 * (size_t)sp, visc, rhof, e);
    double* dst = lds ? s_acc + sp * NPFS : acc + (size_t)sp * NPFS;
    for (int k = 0; k < NPFS; ++k) atomicAdd(dst + k, e[k]);
  }
  if (lds) {
    __syncthreads();
    for (int t = threadIdx.x; t < nspec * NPFS; t += blockDim.x)
      if (s_acc[t] != 0.0) atomicAdd(&acc[t], s_acc[t]);
  }
}

// division by NPM(IDP) (stat_partfluid.f90:378-381); counts in acc[.. + 35]
__global__ void k_pfstat_final(const double* __restrict__ acc, double* __restrict__ out,
                               int nspec, int npoly, int ref_norm) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nspec) return;
  const int nrow = ref_norm ? nspec - npoly + s % npoly : s;
  const double n = acc[(size_t)nrow * NPFS + 35];   // ZERO + the count
  for (int k = 0; k < NPFS; ++k) out[(size_t)s * NPFS + k] = acc[(size_t)s * NPFS + k] / n;
  out[(size_t)s * NPFS + 35] = n;
}

static void grid_init(gp_ctx* ctx) {
  // COLLISION_INITIATION (collision_initiation.f90:50-102), INIT_GRID = 2,
  // NPROC = 1: DELTA_COLL = ZERO; the y/z range spans L + DY.
  const double L = ctx->prm.L, dy = L / (double)(ctx->prm.n - 1);
  // gp_params.cell_rule 1: from the largest semi-axis of any species
  const double reach = ctx->prm.cell_rule == 1 ? ctx->max_semi : ctx->max_emaj;
  const double dpmax = ctx->prm.dpmax > 0.0 ? ctx->prm.dpmax : 2.0 * reach;
  double dxbox = 2.0 * dpmax;
  int nx = (int)(L / dxbox);
  if (nx < 1) nx = 1;
  dxbox = L / (double)nx;
  const double delta = ZERO_R;
  const double yend = dy * (double)(ctx->prm.n - 1);
  const double ymin = 0.0 - delta, ymax = yend + dy + delta;
  CellGrid& g = ctx->g;
  g.nx = nx;
  g.xmin = 0.0;
  g.ymin = ymin;
  g.zmin = ymin;
  g.ny = (int)((ymax - ymin) / dxbox);
  g.nz = (int)((ymax - ymin) / dxbox);
  if (g.ny < 1) g.ny = 1;
  if (g.nz < 1) g.nz = 1;
  g.dx = dxbox;
  g.dy = (ymax - ymin) / (double)g.ny;
  g.dz = (ymax - ymin) / (double)g.nz;
  g.nbox = g.nx * g.ny * g.nz;
  // brick visiting order: b cells a side (0 or < 0: reference cell order)
  const int b = ctx->prm.brick == 0 ? -1 : ctx->prm.brick;
  ctx->use_bricks = b > 0;
  BrickGrid& bg = ctx->bg;
  bg.b = b > 0 ? b : 1;
  bg.nx = (g.nx + bg.b - 1) / bg.b;
  bg.ny = (g.ny + bg.b - 1) / bg.b;
  bg.nz = (g.nz + bg.b - 1) / bg.b;
  bg.n = bg.nx * bg.ny * bg.nz;
  bg.tile = 0;
  if (ctx->prm.tile > 0) {
    // field tiles: base nodes 0..n-1 on each axis
    const int T = ctx->prm.tile, nn = ctx->prm.n;
    ctx->use_bricks = true;
    bg.tile = 1;
    bg.b = T;
    bg.nn = nn;
    bg.dx = ctx->sp.dx;
    bg.nx = bg.ny = bg.nz = (nn + T - 1) / T;
    bg.n = bg.nx * bg.ny * bg.nz;
  }
}

// LDS of one k_gather_tiled workgroup for tile edge T
static size_t tile_bytes(const gp_params* p, int T) {
  if (p->precision == GP_PREC_MIXED)
    return p->interp_scheme == 3 ? tile_lds_bytes<float, 4>(T) : tile_lds_bytes<float, 5>(T);
  return p->interp_scheme == 3 ? tile_lds_bytes<double, 4>(T) : tile_lds_bytes<double, 5>(T);
}
constexpr size_t TILE_LDS_MAX = 65536;

static int ensure_grid_buffers(gp_ctx* ctx) {
  const CellGrid& g = ctx->g;
  int rc;
  if ((rc = dalloc(ctx, &ctx->cnt, (size_t)g.nbox))) return rc;
  if ((rc = dalloc(ctx, &ctx->off, (size_t)g.nbox + 1))) return rc;
  if ((rc = dalloc(ctx, &ctx->bcnt, (size_t)ctx->bg.n))) return rc;
  if ((rc = dalloc(ctx, &ctx->boff, (size_t)ctx->bg.n + 1))) return rc;
  const int64_t sc = std::max(std::max(scan_scratch(g.nbox), scan_scratch((int64_t)ctx->np + 1)),
                              scan_scratch(ctx->bg.n)) + 16;
  if (ctx->scan_scr) hipFree(ctx->scan_scr);
  ctx->scan_scr = nullptr;
  HIPCHK(hipMalloc(&ctx->scan_scr, (size_t)sc * sizeof(long long)));
  return GP_OK;
}

extern "C" {

int gp_create(const gp_params* p, gp_ctx** out) {
  if (!p || !out) return GP_ERR_ARG;
  *out = nullptr;
  if (p->n < 5 || (p->interp_scheme != 3 && p->interp_scheme != 4) ||
      (p->precision != GP_PREC_FP64 && p->precision != GP_PREC_MIXED) || !(p->L > 0.0) ||
      !(p->dtime > 0.0) || !(p->visc > 0.0) || p->cell_rule < 0 || p->cell_rule > 2 ||
      p->graph < 0 || p->graph > 2 || p->overlap < 0 || p->overlap > 2 || p->tile < 0 ||
      (p->tile > 0 && tile_bytes(p, p->tile) > TILE_LDS_MAX))
    return GP_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return GP_ERR_HIP;
  gp_ctx* ctx = new gp_ctx();
  ctx->prm = *p;
  ctx->mixed = p->precision == GP_PREC_MIXED;
  ctx->gate.cos_keep = std::cos(SEED_GATE_KEEP);
  hipGetDevice(&ctx->device);
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
            hipSuccess || ncu < 1)
      ncu = 256;
    ctx->nsimd = 4 * ncu;
    ctx->lm_waves = p->lm_waves > 0 ? p->lm_waves : ctx->nsimd;   // 0: one per SIMD until a solve count is known
  }
  StepParams& s = ctx->sp;
  s.n = p->n;
  s.scheme = p->interp_scheme;
  s.renorm = p->renorm_quat;
  s.prefilter = p->prefilter;
  s.L = p->L;
  s.dx = p->L / (double)(p->n - 1);   // meshing.f90:43
  s.visc = p->visc;
  s.rhof = p->rhof;
  s.dtime = p->dtime;
  s.ecw = p->ecw;
  s.muw = p->muw;
  // default-priority streams: the main one and the one a pipelined advance
  // runs on beside the contact solves
  if (hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->st2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_gather, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->st3, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_efork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_edone, hipEventDisableTiming) != hipSuccess) {
    delete ctx;
    return GP_ERR_HIP;
  }
  if (hipMalloc((void**)&ctx->d_err, sizeof(unsigned)) != hipSuccess ||
      hipMalloc((void**)&ctx->d_cnt, 16 * sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_cnt, 16 * sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_total, 4 * sizeof(long long)) != hipSuccess) {
    delete ctx;
    return GP_ERR_NOMEM;
  }
  hipMemset(ctx->d_err, 0, sizeof(unsigned));
  hipMemset(ctx->d_cnt, 0, 16 * sizeof(unsigned long long));
  ctx->timing = p->timing != 0 && p->graph == 0;
  if (ctx->timing)
    for (auto& e : ctx->ev)
      if (hipEventCreate(&e) != hipSuccess) {
        delete ctx;
        return GP_ERR_HIP;
      }
  // this device's copy of the gp_interph_noalloc staging (device current)
  if (scratch_init(ctx) != GP_OK) {
    gp_destroy(ctx);
    return GP_ERR_HIP;
  }
  hipDeviceSynchronize();
  *out = ctx;
  return GP_OK;
}

int gp_destroy(gp_ctx* ctx) {
  if (!ctx) return GP_ERR_ARG;
  hipStreamSynchronize(ctx->st);
  hipStreamSynchronize(ctx->st2);
  hipStreamSynchronize(ctx->st3);
  void* bufs[] = {ctx->field, ctx->spec, ctx->pbuf[0], ctx->pbuf[1], ctx->perm, ctx->cell,
                  ctx->slot, ctx->cnt, ctx->off, ctx->tkey, ctx->sorted, ctx->pcount,
                  ctx->poff, ctx->scan_scr, ctx->mark, ctx->used, ctx->pos_of_id,
                  ctx->work, ctx->lmwork, ctx->brick, ctx->bslot, ctx->bcnt, ctx->boff, ctx->trav,
                  ctx->pos4, ctx->csort, ctx->gsort, ctx->inlm, ctx->plist, ctx->inv, ctx->ccount, ctx->mlp_th, ctx->mlp_m, ctx->mlp_v, ctx->mlp_grad, ctx->mlp_wb,
                  ctx->mlp_bias, ctx->mlp_part, ctx->mlp_lpart, ctx->mlp_loss, ctx->seeds,
                  ctx->act, ctx->act_status, ctx->d_err, ctx->d_cnt, ctx->flout, ctx->flpre,
                  ctx->ep_keys[0], ctx->ep_keys[1], ctx->ep_vals[0], ctx->ep_vals[1],
                  ctx->ep_pair[0], ctx->ep_pair[1], ctx->ep_slot[0], ctx->ep_slot[1],
                  ctx->ep_res, ctx->ep_merge,
                  ctx->cs_rec, ctx->cs_buf, ctx->cs_cnt, ctx->cs_cycle, ctx->cs_dev};
  for (void* b : bufs)
    if (b) hipFree(b);
  for (double* b : ctx->lag_u0)
    if (b) hipFree(b);
  for (auto& e : ctx->ev)
    if (e) hipEventDestroy(e);
  drop_graph(ctx);
  if (ctx->h_cnt) hipHostFree(ctx->h_cnt);
  if (ctx->h_total) hipHostFree(ctx->h_total);
  hipStreamSynchronize(ctx->st2);
  hipEventDestroy(ctx->ev_fork);
  hipEventDestroy(ctx->ev_join);
  hipEventDestroy(ctx->ev_gather);
  hipStreamSynchronize(ctx->st3);
  hipEventDestroy(ctx->ev_efork);
  hipEventDestroy(ctx->ev_edone);
  hipStreamDestroy(ctx->st3);
  hipStreamDestroy(ctx->st2);
  hipStreamDestroy(ctx->st);
  delete ctx;
  return GP_OK;
}

const char* gp_last_error_string(const gp_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int gp_sync(gp_ctx* ctx) {
  if (!ctx) return GP_ERR_ARG;
  return check_err(ctx);
}

// the reference-layout field dref (device) -> the context's node layout;
// the node buffer is allocated once (its size depends only on n and the
// precision), so captured step graphs stay valid across re-binds
static int bind_field_dev(gp_ctx* ctx, const double* dref) {
  const int n = ctx->prm.n, n4 = n + 4;
  const size_t nodes = (size_t)n4 * n4 * n4;
  const size_t esz = ctx->mixed ? sizeof(float) : sizeof(double);
  if (!ctx->field) HIPCHK(hipMalloc(&ctx->field, nodes * 6 * esz + 256));
  if (ctx->mixed)
    hipLaunchKernelGGL(k_field_prep<float>, dim3(nblk((long long)nodes)), dim3(BLOCK), 0, ctx->st,
                       dref, (float*)ctx->field, n);
  else
    hipLaunchKernelGGL(k_field_prep<double>, dim3(nblk((long long)nodes)), dim3(BLOCK), 0,
                       ctx->st, dref, (double*)ctx->field, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->have_field = true;
  return GP_OK;
}

int gp_bind_field(gp_ctx* ctx, const double* field) {
  if (!ctx || !field) return GP_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st2));
  const int n = ctx->prm.n, n4 = n + 4;
  const size_t nref = (size_t)6 * n * n4 * n4;
  double* dref = nullptr;
  HIPCHK(hipMalloc((void**)&dref, nref * sizeof(double)));
  HIPCHK(hipMemcpyAsync(dref, field, nref * sizeof(double), hipMemcpyHostToDevice, ctx->st));
  const int rc = bind_field_dev(ctx, dref);
  hipFree(dref);
  return rc;
}

int gp_bind_field_planar(gp_ctx* ctx, int n_arrays, const double* uvw, const double* vort,
                         double vort_scale) {
  if (!ctx || !uvw || !vort) return GP_ERR_ARG;
  if (n_arrays != ctx->prm.n) {
    ctx->err = "field arrays of n = " + std::to_string(n_arrays) + " for a context of n = " +
               std::to_string(ctx->prm.n);
    return GP_ERR_ARG;
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st2));
  const int n = ctx->prm.n, n4 = n + 4;
  const size_t nodes = (size_t)n4 * n4 * n4;
  const size_t esz = ctx->mixed ? sizeof(float) : sizeof(double);
  if (!ctx->field) HIPCHK(hipMalloc(&ctx->field, nodes * 6 * esz + 256));
  if (ctx->mixed)
    hipLaunchKernelGGL(k_field_prep_planar<float>, dim3(nblk((long long)nodes)), dim3(BLOCK), 0,
                       ctx->st, uvw, vort, vort_scale, (float*)ctx->field, n);
  else
    hipLaunchKernelGGL(k_field_prep_planar<double>, dim3(nblk((long long)nodes)), dim3(BLOCK), 0,
                       ctx->st, uvw, vort, vort_scale, (double*)ctx->field, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->have_field = true;
  return GP_OK;
}

int gp_bind_field_device(gp_ctx* ctx, const double* field) {
  if (!ctx || !field) return GP_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st2));
  return bind_field_dev(ctx, field);
}

int gp_set_species(gp_ctx* ctx, int nspec, const double* rows) {
  if (!ctx || nspec < 1 || !rows) return GP_ERR_ARG;
  if (ctx->np > 0 && ctx->max_sp >= nspec) {
    ctx->err = "the particles use species index " + std::to_string(ctx->max_sp) +
               " beyond the new table of " + std::to_string(nspec) + " rows";
    return GP_ERR_ARG;
  }
  hipStreamSynchronize(ctx->st);
  drop_graph(ctx);
  std::vector<Species> h((size_t)nspec);
  double emax = 0.0;
  for (int s = 0; s < nspec; ++s) {
    const double* r = rows + (size_t)s * GP_SP_STRIDE;
    Species& d = h[(size_t)s];
    d.a = r[GP_SP_A]; d.b = r[GP_SP_B]; d.c = r[GP_SP_C];
    d.pmass = r[GP_SP_PMASS]; d.deq = r[GP_SP_DEQ];
    d.ixx = r[GP_SP_IXX]; d.iyy = r[GP_SP_IYY]; d.izz = r[GP_SP_IZZ];
    d.grav = r[GP_SP_GRAV]; d.ecp = r[GP_SP_ECP]; d.mup = r[GP_SP_MUP];
    d.rmax = std::max(d.a, std::max(d.b, d.c));
    if (!(d.a > 0 && d.b > 0 && d.c > 0 && d.pmass > 0 && d.deq > 0)) {
      ctx->err = "species row " + std::to_string(s) + " has a non-positive size or mass";
      return GP_ERR_ARG;
    }
    emax = std::max(emax, r[GP_SP_EMAJ] > 0 ? r[GP_SP_EMAJ] : d.a);
  }
  int rc;
  if ((rc = dalloc(ctx, &ctx->spec, (size_t)nspec))) return rc;
  HIPCHK(hipMemcpy(ctx->spec, h.data(), h.size() * sizeof(Species), hipMemcpyHostToDevice));
  ctx->nspec = nspec;
  ctx->max_emaj = emax;
  ctx->max_semi = 0.0;
  for (const Species& d : h) ctx->max_semi = std::max(ctx->max_semi, d.rmax);
  ctx->sp_rows.assign(rows, rows + (size_t)nspec * GP_SP_STRIDE);
  grid_init(ctx);
  // two particles in contact are less than the sum of their largest
  // semi-axes apart; the half shell of cells finds every pair closer than a
  // cell width, so the reference rule is exact while no semi-axis exceeds
  // half a cell (it fails for oblate species with APR < 0.5)
  if (ctx->prm.cell_rule == 0 &&
      ctx->max_semi > 0.5 * std::min(ctx->g.dx, std::min(ctx->g.dy, ctx->g.dz))) {
    ctx->err = "largest semi-axis " + std::to_string(ctx->max_semi) +
               " exceeds half the collision cell " +
               std::to_string(0.5 * std::min(ctx->g.dx, std::min(ctx->g.dy, ctx->g.dz))) +
               ": the reference cell rule would lose pairs (gp_params.cell_rule 1 sizes the "
               "cells from the largest semi-axis, 2 keeps the reference rule unchecked)";
    return GP_ERR_CELL_RULE;
  }
  if (ctx->np > 0) {
    if ((rc = ensure_grid_buffers(ctx))) return rc;
    if (ctx->mixed)
      hipLaunchKernelGGL(k_refresh_rm<float>, dim3(nblk(ctx->np)), dim3(BLOCK), 0, ctx->st,
                         pstate<float>(ctx, ctx->cur), ctx->spec, ctx->np);
    else
      hipLaunchKernelGGL(k_refresh_rm<double>, dim3(nblk(ctx->np)), dim3(BLOCK), 0, ctx->st,
                         pstate<double>(ctx, ctx->cur), ctx->spec, ctx->np);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    // the grid may have changed: the next advance starts from storage order
    ctx->perm_identity = true;
  }
  ctx->cells_valid = false;
  return GP_OK;
}

int gp_set_particles(gp_ctx* ctx, int np, const double* state, const int* species) {
  if (!ctx || np < 1 || !state || !species) return GP_ERR_ARG;
  hipStreamSynchronize(ctx->st);
  hipStreamSynchronize(ctx->st3);
  drop_graph(ctx);
  ctx->contacts_valid = false;
  ctx->lists_sized = false;
  ctx->pipe_prev = false;
  // pairs listed for solving ahead name slots of the old set (a smaller
  // set has fewer): lists and tables emptied together (a table entry names
  // a list position)
  if (ctx->d_cnt) HIPCHK(hipMemset(ctx->d_cnt + 13, 0, 2 * sizeof(unsigned long long)));
  for (int b = 0; b < 2; ++b)
    if (ctx->ep_keys[b])
      HIPCHK(hipMemset(ctx->ep_keys[b], 0, (size_t)ctx->ep_tab * sizeof(unsigned long long)));
  if (!ctx->spec) {
    ctx->err = "gp_set_species must precede gp_set_particles";
    return GP_ERR_STATE;
  }
  int max_sp = 0;
  for (int i = 0; i < np; ++i) {
    if (species[i] < 0 || species[i] >= ctx->nspec) {
      ctx->err = "species index out of range at particle " + std::to_string(i);
      return GP_ERR_ARG;
    }
    max_sp = std::max(max_sp, species[i]);
  }
  ctx->np = np;
  ctx->max_sp = max_sp;
  const size_t pb = pstate_bytes(np, ctx->mixed);
  for (int k = 0; k < 2; ++k) {
    if (ctx->pbuf[k]) hipFree(ctx->pbuf[k]);
    ctx->pbuf[k] = nullptr;
    HIPCHK(hipMalloc(&ctx->pbuf[k], pb));
  }
  int rc;
  if ((rc = dalloc(ctx, &ctx->perm, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->cell, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->slot, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->tkey, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->sorted, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->pcount, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->poff, (size_t)np + 1))) return rc;
  if ((rc = dalloc(ctx, &ctx->mark, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->used, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->pos_of_id, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->brick, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->bslot, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->trav, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->pos4, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->csort, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->gsort, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->inlm, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->plist, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->inv, (size_t)np))) return rc;
  if ((rc = dalloc(ctx, &ctx->ccount, (size_t)np))) return rc;
  HIPCHK(hipMemset(ctx->inlm, 0, (size_t)np * sizeof(int)));
  HIPCHK(hipMemset(ctx->ccount, 0, (size_t)np * sizeof(int)));
  if (ctx->prm.tile > 0) {
    if (ctx->mixed) {
      if ((rc = dalloc(ctx, (float**)&ctx->flpre, (size_t)np * 6))) return rc;
    } else {
      if ((rc = dalloc(ctx, (double**)&ctx->flpre, (size_t)np * 6))) return rc;
    }
  }
  ctx->flpre_valid = false;
  if (ctx->prm.keep_fluid) {
    if ((rc = dalloc(ctx, &ctx->flout, (size_t)np * 6))) return rc;
  }
  if ((rc = ensure_grid_buffers(ctx))) return rc;
  hipLaunchKernelGGL(k_fill<long long>, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st, ctx->mark,
                     (long long)LLONG_MAX, (long long)np);
  hipLaunchKernelGGL(k_fill<int>, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st, ctx->used, 0,
                     (long long)np);
  double* din = nullptr;
  int* dsp = nullptr;
  HIPCHK(hipMalloc((void**)&din, (size_t)np * 13 * sizeof(double)));
  HIPCHK(hipMalloc((void**)&dsp, (size_t)np * sizeof(int)));
  HIPCHK(hipMemcpyAsync(din, state, (size_t)np * 13 * sizeof(double), hipMemcpyHostToDevice,
                        ctx->st));
  HIPCHK(hipMemcpyAsync(dsp, species, (size_t)np * sizeof(int), hipMemcpyHostToDevice, ctx->st));
  ctx->cur = 0;
  if (ctx->mixed)
    hipLaunchKernelGGL(k_scatter_in<float>, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st,
                       pstate<float>(ctx, 0), din, dsp, ctx->spec, np);
  else
    hipLaunchKernelGGL(k_scatter_in<double>, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st,
                       pstate<double>(ctx, 0), din, dsp, ctx->spec, np);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->st));
  hipFree(din);
  hipFree(dsp);
  ctx->perm_identity = true;
  ctx->cells_valid = false;
  ctx->ngroups = 1;
  ctx->fluid_valid = false;
  return GP_OK;
}

int gp_set_groups(gp_ctx* ctx, int np, const int* group) {
  if (!ctx || !group || np != ctx->np) return GP_ERR_ARG;
  int gmax = 0;
  for (int i = 0; i < np; ++i) {
    if (group[i] < 0) {
      ctx->err = "negative ensemble index at particle " + std::to_string(i);
      return GP_ERR_ARG;
    }
    gmax = std::max(gmax, group[i]);
  }
  hipStreamSynchronize(ctx->st);
  drop_graph(ctx);
  int* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, (size_t)np * sizeof(int)));
  HIPCHK(hipMemcpyAsync(d, group, (size_t)np * sizeof(int), hipMemcpyHostToDevice, ctx->st));
  if (ctx->mixed)
    hipLaunchKernelGGL(k_set_groups<float>, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st,
                       pstate<float>(ctx, ctx->cur), d, np);
  else
    hipLaunchKernelGGL(k_set_groups<double>, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st,
                       pstate<double>(ctx, ctx->cur), d, np);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->st));
  hipFree(d);
  ctx->ngroups = gmax + 1;
  ctx->cells_valid = false;
  return GP_OK;
}

int gp_get_particles(gp_ctx* ctx, int np, double* state) {
  if (!ctx || !state || np != ctx->np) return GP_ERR_ARG;
  double* dout = nullptr;
  HIPCHK(hipMalloc((void**)&dout, (size_t)np * 13 * sizeof(double)));
  if (ctx->mixed)
    hipLaunchKernelGGL(k_gather_out<float>, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st,
                       pstate<float>(ctx, ctx->cur), dout, np);
  else
    hipLaunchKernelGGL(k_gather_out<double>, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st,
                       pstate<double>(ctx, ctx->cur), dout, np);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(state, dout, (size_t)np * 13 * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  hipFree(dout);
  return check_err(ctx);
}

int gp_get_fluid_at_particles(gp_ctx* ctx, int np, double* fl) {
  if (!ctx || !fl || np != ctx->np || !ctx->flout) return GP_ERR_ARG;
  HIPCHK(hipMemcpyAsync(fl, ctx->flout, (size_t)np * 6 * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return GP_OK;
}

// the INTERPH kernel on nint points staged at d (x | y | z | out)
static int launch_interph(gp_ctx* ctx, int scheme, int comp, double* d, int nint) {
  const int n = ctx->prm.n;
  const double dx = ctx->sp.dx;
  const double *x = d, *y = d + nint, *z = d + 2 * (size_t)nint;
  double* o = d + 3 * (size_t)nint;
  if (ctx->mixed) {
    if (scheme == 3)
      hipLaunchKernelGGL((k_interph<float, float, 4>), dim3(nblk(nint)), dim3(BLOCK), 0, ctx->st,
                         (const float*)ctx->field, n, dx, x, y, z, comp, o, nint, ctx->d_err);
    else
      hipLaunchKernelGGL((k_interph<float, float, 5>), dim3(nblk(nint)), dim3(BLOCK), 0, ctx->st,
                         (const float*)ctx->field, n, dx, x, y, z, comp, o, nint, ctx->d_err);
  } else {
    if (scheme == 3)
      hipLaunchKernelGGL((k_interph<double, double, 4>), dim3(nblk(nint)), dim3(BLOCK), 0,
                         ctx->st, (const double*)ctx->field, n, dx, x, y, z, comp, o, nint,
                         ctx->d_err);
    else
      hipLaunchKernelGGL((k_interph<double, double, 5>), dim3(nblk(nint)), dim3(BLOCK), 0,
                         ctx->st, (const double*)ctx->field, n, dx, x, y, z, comp, o, nint,
                         ctx->d_err);
  }
  HIPCHK(hipGetLastError());
  return GP_OK;
}

static int interph_args(gp_ctx* ctx, int scheme, int nint, int comp) {
  if (!ctx || nint < 0 || comp < 0 || comp > 5 || (scheme != 3 && scheme != 4)) return GP_ERR_ARG;
  if (!ctx->have_field) {
    ctx->err = "no field bound";
    return GP_ERR_STATE;
  }
  return GP_OK;
}

// points in, staging at d, result out (synchronous)
static int interph_staged(gp_ctx* ctx, int scheme, int nint, const double* x, const double* y,
                          const double* z, int comp, double* out, double* d) {
  HIPCHK(hipMemcpyAsync(d, x, (size_t)nint * 8, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(d + nint, y, (size_t)nint * 8, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(d + 2 * (size_t)nint, z, (size_t)nint * 8, hipMemcpyHostToDevice, ctx->st));
  int rc = launch_interph(ctx, scheme, comp, d, nint);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(out, d + 3 * (size_t)nint, (size_t)nint * 8, hipMemcpyDeviceToHost,
                        ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return GP_OK;
}

int gp_interph(gp_ctx* ctx, int scheme, int nint, const double* x, const double* y,
               const double* z, int comp, double* out) {
  int rc = interph_args(ctx, scheme, nint, comp);
  if (rc) return rc;
  if (nint == 0) return GP_OK;
  double* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, (size_t)nint * 4 * sizeof(double)));
  rc = interph_staged(ctx, scheme, nint, x, y, z, comp, out, d);
  hipFree(d);
  if (rc) return rc;
  return check_err(ctx);
}

int gp_interph_noalloc(gp_ctx* ctx, int scheme, int nint, const double* x, const double* y,
                       const double* z, int comp, double* out) {
  int rc = interph_args(ctx, scheme, nint, comp);
  if (rc) return rc;
  if (nint > GP_INTERPH_NOALLOC_MAX) {
    ctx->err = "gp_interph_noalloc: " + std::to_string(nint) + " points exceed the staging of " +
               std::to_string(GP_INTERPH_NOALLOC_MAX) + " (use gp_interph)";
    return GP_ERR_ARG;
  }
  if (nint == 0) return GP_OK;
  if (!ctx->scratch) {
    ctx->err = "no staging address for this device";
    return GP_ERR_STATE;
  }
  // every point's result is written (zero where the stencil leaves UAIN,
  // with the error set), so nothing of a previous call is read
  if ((rc = interph_staged(ctx, scheme, nint, x, y, z, comp, out, ctx->scratch))) return rc;
  return check_err(ctx);
}

int gp_interph_field(gp_ctx* ctx, int scheme, const double* uain, int nint, const double* x,
                     const double* y, const double* z, double* out) {
  if (!ctx || !uain || nint < 0 || (scheme != 3 && scheme != 4)) return GP_ERR_ARG;
  if (nint == 0) return GP_OK;
  const int n = ctx->prm.n, n4 = n + 4;
  const size_t nref = (size_t)n * n4 * n4, nodes = (size_t)n4 * n4 * n4;
  char* d = nullptr;
  const size_t bytes = (nref + nodes + 4 * (size_t)nint) * sizeof(double) + 256;
  HIPCHK(hipMalloc((void**)&d, bytes));
  double* dref = (double*)d;
  double* dfld = dref + nref;
  double* dx_ = dfld + nodes;
  double* dy_ = dx_ + nint;
  double* dz_ = dy_ + nint;
  double* dout = dz_ + nint;
  HIPCHK(hipMemcpyAsync(dref, uain, nref * 8, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(dx_, x, (size_t)nint * 8, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(dy_, y, (size_t)nint * 8, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(dz_, z, (size_t)nint * 8, hipMemcpyHostToDevice, ctx->st));
  hipLaunchKernelGGL((k_field_prep<double, 1>), dim3(nblk((long long)nodes)), dim3(BLOCK), 0,
                     ctx->st, dref, dfld, n);
  if (scheme == 3)
    hipLaunchKernelGGL((k_interph1<double, double, 4>), dim3(nblk(nint)), dim3(BLOCK), 0, ctx->st,
                       dfld, n, ctx->sp.dx, dx_, dy_, dz_, dout, nint, ctx->d_err);
  else
    hipLaunchKernelGGL((k_interph1<double, double, 5>), dim3(nblk(nint)), dim3(BLOCK), 0, ctx->st,
                       dfld, n, ctx->sp.dx, dx_, dy_, dz_, dout, nint, ctx->d_err);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, (size_t)nint * 8, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  hipFree(d);
  return check_err(ctx);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// step pieces (asynchronous on ctx->st)
// ---------------------------------------------------------------------------
// adds the stage times of the last timed step (waits for its last event)
static void collect_timing(gp_ctx* ctx) {
  if (!ctx->timing || !ctx->ev_pending) return;
  if (hipEventSynchronize(ctx->ev[4]) != hipSuccess) return;
  for (int k = 0; k < 4; ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]) == hipSuccess)
      ctx->t_stage[k] += ms;
  }
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ctx->ev[5], ctx->ev[3]) == hipSuccess) ctx->t_solve += ms;
  ctx->ntimed++;
  ctx->ev_pending = false;
}
static void mark(gp_ctx* ctx, int k) {
  if (ctx->timing) hipEventRecord(ctx->ev[k], ctx->st);
}
static void zero_bins(gp_ctx* ctx, hipStream_t st) {
  hipLaunchKernelGGL(k_fill<int>, dim3(nblk(ctx->g.nbox)), dim3(BLOCK), 0, st, ctx->cnt, 0,
                     (long long)ctx->g.nbox);
  if (ctx->use_bricks)
    hipLaunchKernelGGL(k_fill<int>, dim3(nblk(ctx->bg.n)), dim3(BLOCK), 0, st, ctx->bcnt, 0,
                       (long long)ctx->bg.n);
}

// the advance kernel from buffer cur to 1 - cur in the visiting order perm
// (nullptr: storage order); skip: particles left for k_advance_listed
// blocks of BLOCK over n items; a multiple of 8 when workgroups are remapped
// to XCDs (block_index)
static unsigned xcd_grid(const gp_ctx* ctx, int n) {
  unsigned nb = nblk(n);
  return ctx->prm.xcd_remap ? (nb + 7) / 8 * 8 : nb;
}

template <typename VT>
static void launch_advance_kernel(gp_ctx* ctx, hipStream_t st, const int* perm, const int* skip,
                                  bool bin) {
  const int np = ctx->np;
  if (ctx->flout) ctx->fluid_valid = true;
  auto ps = pstate<VT>(ctx, ctx->cur), pd = pstate<VT>(ctx, 1 - ctx->cur);
  int* bcnt = bin && ctx->use_bricks ? ctx->bcnt : nullptr;
  int* cell = bin ? ctx->cell : nullptr;
  const int xcd = ctx->prm.xcd_remap;
  ctx->flpre_valid = false;
  if (ctx->bg.tile && perm == ctx->trav) {
    // the fluid at every particle first, one workgroup per field tile (the
    // visiting order is the tiles'), then the advance at full occupancy
    unsigned nt = (unsigned)ctx->bg.n;
    if (xcd) nt = (nt + 7) / 8 * 8;
    const size_t lds = tile_bytes(&ctx->prm, ctx->bg.b);
    unsigned nb = nblk(np);
    if (xcd) nb = (nb + 7) / 8 * 8;
#define GP_GATHER(FT, CT, S)                                                                    \
  hipLaunchKernelGGL((k_gather_tiled<FT, VT, CT, S>), dim3(nt), dim3(TILE_WG), lds, st,        \
                     ctx->sp, (const FT*)ctx->field, ps, ctx->trav, ctx->boff, np, ctx->bg,     \
                     (CT*)ctx->flpre, ctx->d_err, xcd)
#define GP_ADVP(CT)                                                                             \
  hipLaunchKernelGGL((k_advance_fl<VT, CT>), dim3(nb), dim3(BLOCK), 0, st, ctx->sp,            \
                     (const CT*)ctx->flpre, ctx->spec, ps, pd, perm, np, ctx->g, cell, ctx->slot, \
                     ctx->cnt, ctx->bg, ctx->brick, ctx->bslot, bcnt, ctx->d_err,                \
                     ctx->d_cnt + 0, ctx->flout, xcd, skip, (const int*)nullptr,                \
                     (const unsigned long long*)nullptr, (const int*)nullptr)
    if (ctx->mixed) {
      if (ctx->prm.interp_scheme == 3) GP_GATHER(float, float, 4);
      else GP_GATHER(float, float, 5);
    } else {
      if (ctx->prm.interp_scheme == 3) GP_GATHER(double, double, 4);
      else GP_GATHER(double, double, 5);
    }
    // the listed advance on the main stream reads the same buffer
    ctx->flpre_st2 = st != ctx->st;
    if (ctx->flpre_st2) (void)hipEventRecord(ctx->ev_gather, st);
    if (ctx->mixed) GP_ADVP(float);
    else GP_ADVP(double);
#undef GP_GATHER
#undef GP_ADVP
    ctx->flpre_valid = true;
    return;
  }
  unsigned nb = nblk(np);
  if (xcd) nb = (nb + 7) / 8 * 8;
  const dim3 gr(nb), bl(BLOCK);
#define GP_ADV(FT, CT, S)                                                                        \
  hipLaunchKernelGGL((k_advance<FT, VT, CT, S>), gr, bl, 0, st, ctx->sp, (const FT*)ctx->field, \
                     ctx->spec, ps, pd, perm, np, ctx->g, cell, ctx->slot, ctx->cnt, ctx->bg,    \
                     ctx->brick, ctx->bslot, bcnt, ctx->d_err, ctx->d_cnt + 0, ctx->flout, xcd,  \
                     skip)
  if (ctx->mixed) {
    if (ctx->prm.interp_scheme == 3) GP_ADV(float, float, 4);
    else GP_ADV(float, float, 5);
  } else {
    if (ctx->prm.interp_scheme == 3) GP_ADV(double, double, 4);
    else GP_ADV(double, double, 5);
  }
#undef GP_ADV
}

// the advance of the particles listed by k_pw_filter (after the collisions)
template <typename VT>
static void launch_advance_listed(gp_ctx* ctx, hipStream_t st) {
  const int np = ctx->np;
  auto ps = pstate<VT>(ctx, ctx->cur), pd = pstate<VT>(ctx, 1 - ctx->cur);
  int* bcnt = ctx->use_bricks ? ctx->bcnt : nullptr;
  const dim3 gr(1024), bl(ADVL_BLOCK);
  if (ctx->flpre_valid) {
    // field tiles: the fluid gathered with the tiles of this advance (on the
    // second stream), the same advance kernel as the tiles' own
    if (ctx->flpre_st2) (void)hipStreamWaitEvent(st, ctx->ev_gather, 0);
#define GP_ADVF(CT)                                                                              \
  hipLaunchKernelGGL((k_advance_fl<VT, CT>), gr, bl, 0, st, ctx->sp, (const CT*)ctx->flpre,    \
                     ctx->spec, ps, pd, (const int*)nullptr, np, ctx->g, ctx->cell, ctx->slot,  \
                     ctx->cnt, ctx->bg, ctx->brick, ctx->bslot, bcnt, ctx->d_err,                \
                     ctx->d_cnt + 0, ctx->flout, 0, (const int*)nullptr, ctx->plist,            \
                     ctx->d_cnt + 10, ctx->inv)
    if (ctx->mixed) GP_ADVF(float);
    else GP_ADVF(double);
#undef GP_ADVF
    return;
  }
#define GP_ADVL(FT, CT, S)                                                                       \
  hipLaunchKernelGGL((k_advance_listed<FT, VT, CT, S>), gr, bl, 0, st, ctx->sp,                 \
                     (const FT*)ctx->field, ctx->spec, ps, pd, ctx->plist, ctx->d_cnt + 10,     \
                     ctx->inv, np, ctx->g, ctx->cell, ctx->slot, ctx->cnt, ctx->bg,              \
                     ctx->brick, ctx->bslot, bcnt, ctx->d_err, ctx->d_cnt + 0, ctx->flout)
  if (ctx->mixed) {
    if (ctx->prm.interp_scheme == 3) GP_ADVL(float, float, 4);
    else GP_ADVL(float, float, 5);
  } else {
    if (ctx->prm.interp_scheme == 3) GP_ADVL(double, double, 4);
    else GP_ADVL(double, double, 5);
  }
#undef GP_ADVL
}

template <typename VT>
static int launch_advance(gp_ctx* ctx, bool bin) {
  if (bin) zero_bins(ctx, ctx->st);
  const int* perm = ctx->perm_identity ? nullptr : (ctx->use_bricks ? ctx->trav : ctx->perm);
  collect_timing(ctx);
  mark(ctx, 0);
  launch_advance_kernel<VT>(ctx, ctx->st, perm, nullptr, bin);
  HIPCHK(hipGetLastError());
  mark(ctx, 1);
  ctx->cur = 1 - ctx->cur;
  ctx->contacts_valid = false;
  ctx->perm_identity = true;
  ctx->cells_valid = bin;
  return GP_OK;
}

template <typename VT>
static int launch_bin(gp_ctx* ctx) {
  const int np = ctx->np;
  hipLaunchKernelGGL(k_fill<int>, dim3(nblk(ctx->g.nbox)), dim3(BLOCK), 0, ctx->st, ctx->cnt, 0,
                     (long long)ctx->g.nbox);
  if (ctx->use_bricks)
    hipLaunchKernelGGL(k_fill<int>, dim3(nblk(ctx->bg.n)), dim3(BLOCK), 0, ctx->st, ctx->bcnt, 0,
                       (long long)ctx->bg.n);
  hipLaunchKernelGGL(k_cells<VT>, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st,
                     pstate<VT>(ctx, ctx->cur), np, ctx->g, ctx->cell, ctx->slot, ctx->cnt,
                     ctx->bg, ctx->brick, ctx->bslot, ctx->use_bricks ? ctx->bcnt : nullptr,
                     ctx->d_err);
  HIPCHK(hipGetLastError());
  ctx->cells_valid = true;
  return GP_OK;
}

// counting sort (asynchronous); offsets: also the candidate counts per home
// and their offsets (the list length before the prefilter, and the
// materialised lists), which the steps themselves do not need
template <typename VT>
static int launch_sort(gp_ctx* ctx, bool offsets = true) {
  const int np = ctx->np;
  const CellGrid& g = ctx->g;
  int rc;
  if (!ctx->cells_valid)
    if ((rc = launch_bin<VT>(ctx))) return rc;
  scan_exclusive_int(ctx->cnt, ctx->off, g.nbox, (int*)ctx->scan_scr, ctx->st);
  auto ps = pstate<VT>(ctx, ctx->cur);
  // the scatter also clears the previous pipelined step's solve flags (its
  // advances are done); the ranking zeroes the step's counters 1..11
  SideClear sc_flags, sc_cnt;
  if (ctx->clear_plist) {
    sc_flags.list = ctx->plist;
    sc_flags.nlist = ctx->d_cnt + 10;
    sc_flags.flag = ctx->inlm;
    sc_flags.np = np;
  }
  sc_cnt.zc = ctx->d_cnt + 1;
  sc_cnt.nzc = 12;
  if (ctx->ep_step) {
    // this step's list and table for solving ahead: the scatter empties the
    // table slots its last filling used, the ranking then zeroes the count
    const int b = ctx->ep_par;
    sc_flags.zl = ctx->ep_keys[b];
    sc_flags.slot = ctx->ep_slot[b];
    sc_flags.nslot = ctx->d_cnt + 13 + b;
    sc_flags.nslot_cap = ctx->ep_cap;
    sc_cnt.zc2 = ctx->d_cnt + 13 + b;
    sc_cnt.nzc2 = 1;
  }
  const int* grp = ctx->ngroups > 1 ? ps.grp : nullptr;
  hipLaunchKernelGGL(k_scatter, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st, ctx->cell, ctx->slot,
                     ctx->off, ps.id, grp, ctx->tkey, np, sc_flags);
  hipLaunchKernelGGL(k_rank, dim3(xcd_grid(ctx, np)), dim3(BLOCK), 0, ctx->st, ctx->cell,
                     ctx->off, ctx->cnt, ctx->tkey, ps.id, grp, ps.x, ps.y, ps.z, ps.rm,
                     ctx->sorted, ctx->pos4, ctx->csort, ctx->gsort,
                     ctx->use_bricks ? nullptr : ctx->inv, np, ctx->prm.xcd_remap, sc_cnt);
  if (offsets) {
    hipLaunchKernelGGL(k_pair_count, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st, g, ctx->sorted,
                       ctx->cell, ctx->off, ctx->cnt, grp ? ctx->gsort : nullptr, ctx->pcount,
                       np);
    scan_exclusive<int, long long>(ctx->pcount, ctx->poff, np, (long long*)ctx->scan_scr, ctx->st);
  }
  if (ctx->use_bricks) {
    scan_exclusive_int(ctx->bcnt, ctx->boff, ctx->bg.n, (int*)ctx->scan_scr, ctx->st);
    hipLaunchKernelGGL(k_brick_scatter, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st, ctx->brick,
                       ctx->bslot, ctx->boff, ctx->trav, ctx->inv, np);
  }
  HIPCHK(hipGetLastError());
  mark(ctx, 2);
  return GP_OK;
}

// the total candidate count of the last sort (synchronises the stream)
static int read_nclose(gp_ctx* ctx) {
  HIPCHK(hipMemcpyAsync(ctx->h_total, ctx->poff + ctx->np, sizeof(long long),
                        hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->last_nclose = ctx->h_total[0];
  return GP_OK;
}

template <typename VT>
static int launch_sort_count(gp_ctx* ctx) {
  int rc;
  if ((rc = launch_sort<VT>(ctx))) return rc;
  return read_nclose(ctx);
}

// work / active lists able to hold `need` prefilter survivors.  Kernels
// already queued may still use the current lists (the eager step that sizes
// them): both streams are drained before anything is freed.
static int sync_cstat(gp_ctx* ctx);

static int ensure_lists(gp_ctx* ctx, unsigned long long need) {
  int rc;
  const unsigned long long acap_need =
      std::max<unsigned long long>(std::max(need, ctx->work_cap) / 4, 4096);
  if (need > ctx->work_cap || acap_need > ctx->act_cap) {
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st2));
  }
  if (need > ctx->work_cap) {
    drop_graph(ctx);
    const unsigned long long cap = need + need / 4 + 1024;
    if ((rc = dalloc(ctx, &ctx->work, (size_t)cap))) return rc;
    if ((rc = dalloc(ctx, &ctx->lmwork, (size_t)cap))) return rc;
    if ((rc = dalloc(ctx, &ctx->seeds, (size_t)cap * 2))) return rc;
    ctx->work_cap = cap;
  }
  // solving ahead: lists of an eighth of the work capacity (the survivors
  // of both filters are a small part of the bounding-sphere survivors; a
  // pair past the capacity is simply solved in its own step), tables of
  // twice that (a power of two), keys and list counts cleared
  const unsigned long long ecap =
      std::min(ctx->work_cap, std::max<unsigned long long>(ctx->work_cap / 8, 65536));
  if (ecap > ctx->ep_cap) {
    HIPCHK(hipStreamSynchronize(ctx->st3));
    drop_graph(ctx);
    unsigned long long t = 1024;
    while (t < 2 * ecap) t <<= 1;
    for (int b = 0; b < 2; ++b) {
      if ((rc = dalloc(ctx, &ctx->ep_keys[b], (size_t)t))) return rc;
      if ((rc = dalloc(ctx, &ctx->ep_vals[b], (size_t)t))) return rc;
      if ((rc = dalloc(ctx, &ctx->ep_pair[b], (size_t)ecap * 2))) return rc;
      if ((rc = dalloc(ctx, &ctx->ep_slot[b], (size_t)ecap))) return rc;
      HIPCHK(hipMemset(ctx->ep_keys[b], 0, (size_t)t * sizeof(unsigned long long)));
    }
    if ((rc = dalloc(ctx, &ctx->ep_res, (size_t)ecap))) return rc;
    if ((rc = dalloc(ctx, &ctx->ep_merge, (size_t)ecap))) return rc;
    HIPCHK(hipMemset(ctx->d_cnt + 12, 0, 3 * sizeof(unsigned long long)));
    ctx->ep_tab = t;
    ctx->ep_cap = ecap;
  }
  const unsigned long long acap = std::max<unsigned long long>(ctx->work_cap / 4, 4096);
  if (acap > ctx->act_cap) {
    drop_graph(ctx);
    if ((rc = dalloc(ctx, &ctx->act, (size_t)acap))) return rc;
    if ((rc = dalloc(ctx, &ctx->act_status, (size_t)acap))) return rc;
    ctx->act_cap = acap;
  }
  // statistics records: at most one per contact solve
  if (ctx->cs_fout2 > 0 && ctx->work_cap > ctx->cs_cap) {
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st2));
    drop_graph(ctx);
    if ((rc = dalloc(ctx, &ctx->cs_rec, (size_t)ctx->work_cap))) return rc;
    ctx->cs_cap = ctx->work_cap;
    return sync_cstat(ctx);
  }
  return GP_OK;
}

static CStat cstat(gp_ctx* ctx) {
  CStat cs{};
  if (ctx->cs_fout2 <= 0 || !ctx->cs_rec) return cs;
  cs.rec = ctx->cs_rec;
  cs.cnt = ctx->cs_cnt;
  cs.cap = ctx->cs_cap;
  cs.cycle = ctx->cs_cycle;
  cs.acc = ctx->cs_buf;
  cs.last = cs.acc + (size_t)ctx->cs_nbin * CS_NQ;
  cs.tsum = cs.last + (size_t)ctx->cs_nbin * 8;
  cs.fout2 = ctx->cs_fout2;
  cs.npoly = ctx->cs_npoly;
  cs.nbin = ctx->cs_nbin;
  return cs;
}

// the device copy k_contact reads (both streams idle)
static int sync_cstat(gp_ctx* ctx) {
  const CStat cs = cstat(ctx);
  HIPCHK(hipMemcpy(ctx->cs_dev, &cs, sizeof(cs), hipMemcpyHostToDevice));
  return GP_OK;
}

// slack of the float bounding-sphere test: 1e-5 of the largest coordinate
// magnitude of the collision grid (the float rounding of a centre is 6e-8
// of it)
static float prefilter_tol(const gp_ctx* ctx) {
  const CellGrid& g = ctx->g;
  const double ext = std::max({1.0, std::fabs(ctx->prm.L), std::fabs(g.xmin),
                               std::fabs(g.ymin), std::fabs(g.zmin),
                               std::fabs(g.xmin + g.dx * g.nx), std::fabs(g.ymin + g.dy * g.ny),
                               std::fabs(g.zmin + g.dz * g.nz)});
  return (float)(1.0e-5 * ext);
}

// prefilters on the sorted lists (asynchronous; the counters 1..11 were
// zeroed by the sort); list_lm: also flag and list the particles of the
// contact solves (for a pipelined next advance); clear_bins: the PW filter
// also zeroes the bin counts and the wall count for the next advance (the
// prefilter was their last reader)
template <typename VT>
static int launch_contacts_filter(gp_ctx* ctx, bool list_lm, bool clear_bins = false) {
  const int np = ctx->np;
  unsigned long long* c = ctx->d_cnt;
  auto ps = pstate<VT>(ctx, ctx->cur);
  if (ctx->ngroups > 1)
    hipLaunchKernelGGL(k_prefilter<true>, dim3(xcd_grid(ctx, np)), dim3(BLOCK), 0, ctx->st,
                       ctx->g, ctx->pos4, ctx->sorted, ctx->csort, ctx->off, ctx->cnt,
                       ctx->gsort, ctx->prm.L, prefilter_tol(ctx), ctx->prm.prefilter,
                       ctx->work, c + 1, ctx->work_cap, c + 11, ctx->d_err, np,
                       ctx->prm.xcd_remap);
  else
    hipLaunchKernelGGL(k_prefilter<false>, dim3(xcd_grid(ctx, np)), dim3(BLOCK), 0, ctx->st,
                       ctx->g, ctx->pos4, ctx->sorted, ctx->csort, ctx->off, ctx->cnt,
                       (const int*)nullptr, ctx->prm.L, prefilter_tol(ctx), ctx->prm.prefilter,
                       ctx->work, c + 1, ctx->work_cap, c + 11, ctx->d_err, np,
                       ctx->prm.xcd_remap);
  // the survivor count stays on the device: a fixed grid, grid-stride inside
  const unsigned fgrid = (unsigned)std::min<unsigned long long>((ctx->work_cap + BLOCK - 1) / BLOCK,
                                                                 2048ull);
  SideClear sc;
  if (clear_bins) {
    sc.zi = ctx->cnt;
    sc.nzi = ctx->g.nbox;
    if (ctx->use_bricks) {
      sc.zi2 = ctx->bcnt;
      sc.nzi2 = ctx->bg.n;
    }
    sc.zc = c;
    sc.nzc = 1;
  }
  EarlyTab et;
  if (ctx->ep_step) {
    const int b = ctx->ep_par;
    et.pkeys = ctx->ep_keys[1 - b];
    et.pvals = ctx->ep_vals[1 - b];
    et.nkeys = ctx->ep_keys[b];
    et.nvals = ctx->ep_vals[b];
    et.mask = ctx->ep_tab - 1;
    et.inv = ctx->inv;
    et.pair = ctx->ep_pair[b];
    et.slot = ctx->ep_slot[b];
    et.npair = c + 13 + b;
    et.merge = ctx->ep_merge;
    et.nmerge = c + 12;
    et.cap = ctx->ep_cap;
  }
  hipLaunchKernelGGL(k_pw_filter<VT>, dim3(fgrid), dim3(BLOCK), 0, ctx->st, ps, ctx->spec,
                     ctx->work, c + 1, ctx->work_cap, ctx->prm.L, ctx->prm.prefilter >= 2 ? 1 : 0,
                     ctx->lmwork, c + 7, list_lm ? ctx->inlm : nullptr, ctx->plist, c + 10, sc,
                     et);
  HIPCHK(hipGetLastError());
  return GP_OK;
}

// contact solves + one-collision pass (asynchronous)
template <typename VT>
static int launch_contacts_solve(gp_ctx* ctx) {
  unsigned long long* c = ctx->d_cnt;
  auto ps = pstate<VT>(ctx, ctx->cur);
  // one wavefront per workgroup: the LM is latency-bound (one pair per lane),
  // so the pairs are spread over as many CUs as possible
  const bool net = ctx->prm.lm_seed == 1;
  mark(ctx, 5);
  if (net)
    hipLaunchKernelGGL(k_mlp_seed<VT>, dim3(ctx->lm_waves), dim3(64), 0, ctx->st, ps, ctx->spec,
                       ctx->lmwork, c + 7, ctx->work_cap, ctx->prm.L, mlp_dev(ctx), ctx->seeds);
  const CStat cs = cstat(ctx);
  ctx->lm_waves_last = ctx->lm_waves;   // a captured graph keeps this width
  hipLaunchKernelGGL(k_contact<VT>, dim3(ctx->lm_waves), dim3(LM_BLOCK), 0, ctx->st, ps, ctx->spec,
                     ctx->lmwork, c + 7, ctx->work_cap, ctx->prm.L, ctx->act, c + 2, ctx->act_cap,
                     c + 3, c + 4, ctx->d_err, net ? ctx->seeds : nullptr, ctx->gate, c + 9,
                     cs.rec ? ctx->cs_dev : nullptr, EarlyIn{});
  if (ctx->ep_step) {
    // the pairs solved ahead, once that solve is done
    HIPCHK(hipStreamWaitEvent(ctx->st, ctx->ev_edone, 0));
    const unsigned mgrid =
        (unsigned)std::min<unsigned long long>((ctx->ep_cap + BLOCK - 1) / BLOCK, 1024ull);
    hipLaunchKernelGGL(k_merge_early, dim3(mgrid), dim3(BLOCK), 0, ctx->st, ctx->ep_merge,
                       c + 12, ctx->ep_cap, ctx->ep_res, ctx->act, c + 2, ctx->act_cap, c + 3,
                       c + 4, c + 7, ctx->d_err);
  }
  mark(ctx, 3);
  hipLaunchKernelGGL(k_greedy_mark, dim3(GREEDY_GRID), dim3(BLOCK), 0, ctx->st, ctx->act, c + 2,
                     ctx->act_cap, ctx->mark);
  hipLaunchKernelGGL(k_greedy_accept, dim3(GREEDY_GRID), dim3(BLOCK), 0, ctx->st, ctx->act,
                     c + 2, ctx->act_cap, ctx->mark, ctx->used, ctx->act_status);
  hipLaunchKernelGGL(k_greedy_drop, dim3(GREEDY_GRID), dim3(BLOCK), 0, ctx->st, ctx->act, c + 2,
                     ctx->act_cap, ctx->mark, ctx->used, ctx->act_status, c + 8);
  hipLaunchKernelGGL(k_greedy_rounds, dim3(1), dim3(GREEDY_BLOCK), 0, ctx->st, ctx->act, c + 2,
                     ctx->act_cap, ctx->mark, ctx->used, ctx->act_status, c + 8, c + 6);
  hipLaunchKernelGGL(k_greedy_impulse<VT>, dim3(GREEDY_GRID), dim3(BLOCK), 0, ctx->st, ps,
                     ctx->spec, ctx->act, c + 2, ctx->act_cap, ctx->used, ctx->act_status, c + 5,
                     ctx->d_err, ctx->ccount, cs.rec ? ctx->mark : nullptr);
  if (cs.rec) {
    // the greedy pass leaves mark[] idle (LLONG_MAX): it carries acc_key here
    const unsigned sgrid = (unsigned)std::min<unsigned long long>((cs.cap + BLOCK - 1) / BLOCK, 64ull);
    hipLaunchKernelGGL(k_cstat_accum, dim3(sgrid), dim3(BLOCK), 0, ctx->st, cs, ctx->mark);
    hipLaunchKernelGGL(k_cstat_reset, dim3(GREEDY_GRID), dim3(BLOCK), 0, ctx->st, ctx->act, c + 2,
                       ctx->act_cap, ctx->act_status, ctx->mark);
    hipLaunchKernelGGL(k_cstat_final, dim3(1), dim3(BLOCK), 0, ctx->st, cs, ctx->prm.dtime);
  }
  HIPCHK(hipGetLastError());
  mark(ctx, 4);
  ctx->ev_pending = ctx->timing;
  // the next advance reads the particles in brick order (or cell order)
  if (!ctx->use_bricks) std::swap(ctx->perm, ctx->sorted);
  ctx->perm_identity = false;
  ctx->cells_valid = false;
  return GP_OK;
}

template <typename VT>
static int launch_contacts(gp_ctx* ctx) {
  int rc;
  if ((rc = launch_contacts_filter<VT>(ctx, false))) return rc;
  return launch_contacts_solve<VT>(ctx);
}

// COLLISION with the list sizes read back from the device
template <typename VT>
static int launch_collide(gp_ctx* ctx) {
  int rc;
  if ((rc = launch_sort_count<VT>(ctx))) return rc;
  // prefilter survivors <= candidates
  if ((rc = ensure_lists(ctx, (unsigned long long)std::max<long long>(ctx->last_nclose, 1))))
    return rc;
  if ((rc = launch_contacts<VT>(ctx))) return rc;
  ctx->contacts_valid = true;
  ctx->pipe_prev = false;
  return GP_OK;
}

// solving ahead, at the start of a pipelined step (every particle advanced):
// the pairs the previous step's filter listed, solved from this step's state
// on the third stream beside the step's sort and filters.  Only for the
// plain centre-line solve without statistics records (a pair solved ahead
// leaves no record, and the seeded solve keeps its own seeds).
static bool early_on(const gp_ctx* ctx) {
  // latency-bound solves only (fewer than one per lane of one wavefront per
  // SIMD, the last count known): there the step waits on its slowest solve,
  // and solving ahead takes that wait off the sort and filters.  Throughput-
  // bound solves gain nothing from the overlap and pay the second pass.
  return ctx->prm.overlap == 1 && ctx->prm.lm_seed == 0 && ctx->cs_fout2 <= 0 && ctx->ep_cap > 0 &&
         !ctx->timing && ctx->last_nlm < (unsigned long long)ctx->nsimd * 64ull;
}

template <typename VT>
static int launch_early(gp_ctx* ctx) {
  ctx->ep_step = early_on(ctx);
  if (!ctx->ep_step) return GP_OK;
  const int b = 1 - ctx->ep_par;   // the previous step's list
  auto ps = pstate<VT>(ctx, ctx->cur);
  HIPCHK(hipEventRecord(ctx->ev_efork, ctx->st));
  HIPCHK(hipStreamWaitEvent(ctx->st3, ctx->ev_efork, 0));
  EarlyIn ein;
  ein.pair = ctx->ep_pair[b];
  ein.n = ctx->d_cnt + 13 + b;
  ein.res = ctx->ep_res;
  unsigned long long* c = ctx->d_cnt;
  hipLaunchKernelGGL(k_contact<VT>, dim3(ctx->lm_waves), dim3(LM_BLOCK), 0, ctx->st3, ps,
                     ctx->spec, ctx->lmwork, c + 7, ctx->ep_cap, ctx->prm.L, ctx->act, c + 2,
                     ctx->act_cap, c + 3, c + 4, ctx->d_err, (const double*)nullptr, ctx->gate,
                     c + 9, (const CStat*)nullptr, ein);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev_edone, ctx->st3));
  return GP_OK;
}

// end of a step that solved ahead: the next one lists into the other parity
static void end_early(gp_ctx* ctx) {
  if (ctx->ep_step) ctx->ep_par ^= 1;
  ctx->ep_step = false;
}

// COLLISION of the advanced particles, with the next ADV_PARTICLE_ELLP of
// every particle outside the contact solves on the second stream beside the
// solves (as inside pipelined gp_step), then the solves' particles advanced
// after the one-collision pass.  Lists sized (no host synchronisation).
template <typename VT>
static int launch_collide_advance(gp_ctx* ctx) {
  int rc;
  if ((rc = launch_early<VT>(ctx))) return rc;
  if (!ctx->pipe_prev)
    hipLaunchKernelGGL(k_fill<int>, dim3(nblk(ctx->np)), dim3(BLOCK), 0, ctx->st, ctx->inlm, 0,
                       (long long)ctx->np);
  // the sort clears the previous pipelined pass's flags
  ctx->clear_plist = ctx->pipe_prev;
  rc = launch_sort<VT>(ctx, false);
  ctx->clear_plist = false;
  if (rc) return rc;
  if ((rc = launch_contacts_filter<VT>(ctx, true, true))) return rc;
  const int* order = ctx->use_bricks ? ctx->trav : ctx->sorted;
  HIPCHK(hipEventRecord(ctx->ev_fork, ctx->st));
  HIPCHK(hipStreamWaitEvent(ctx->st2, ctx->ev_fork, 0));
  launch_advance_kernel<VT>(ctx, ctx->st2, order, ctx->inlm, true);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev_join, ctx->st2));
  if ((rc = launch_contacts_solve<VT>(ctx))) return rc;
  launch_advance_listed<VT>(ctx, ctx->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamWaitEvent(ctx->st, ctx->ev_join, 0));
  ctx->cur = 1 - ctx->cur;
  ctx->perm_identity = true;
  ctx->cells_valid = true;
  ctx->contacts_valid = false;
  ctx->pipe_prev = true;
  end_early(ctx);
  return GP_OK;
}

// k full steps without host synchronisation (lists already sized).  With
// gp_params.overlap (and no stage timers) the advance of step s + 1 of every
// particle outside step s's contact solves runs on a second stream beside
// those solves (it reads and writes none of their particles); the solves'
// particles are advanced after the one-collision pass, then the next sort
// waits for both.
template <typename VT>
static int launch_steps(gp_ctx* ctx, int k) {
  int rc;
  const bool pipe = ctx->prm.overlap && !ctx->timing;
  unsigned long long* c = ctx->d_cnt;
  ctx->pipe_prev = false;
  if (pipe)
    hipLaunchKernelGGL(k_fill<int>, dim3(nblk(ctx->np)), dim3(BLOCK), 0, ctx->st, ctx->inlm, 0,
                       (long long)ctx->np);
  hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(1), 0, ctx->st, c, 0ull, 1ll);
  if ((rc = launch_advance<VT>(ctx, true))) return rc;
  for (int s = 0; s < k; ++s) {
    const bool next = s + 1 < k;
    // every particle is advanced: the previous step's pairs solved ahead
    if (pipe && (rc = launch_early<VT>(ctx))) return rc;
    // from the second step on, the sort clears the previous step's flags
    ctx->clear_plist = pipe && s > 0;
    rc = launch_sort<VT>(ctx, false);
    ctx->clear_plist = false;
    if (rc) return rc;
    if (!(pipe && next)) {
      if ((rc = launch_contacts<VT>(ctx))) return rc;
      end_early(ctx);
      if (next) {
        hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(1), 0, ctx->st, c, 0ull, 1ll);
        if ((rc = launch_advance<VT>(ctx, true))) return rc;
      }
      continue;
    }
    // the prefilter is the last reader of this step's bins: the PW filter
    // clears them (and the wall count) on the main stream, before both
    // advances
    if ((rc = launch_contacts_filter<VT>(ctx, true, true))) return rc;
    // fork: the visiting order of the next advance is this sort's order
    const int* order = ctx->use_bricks ? ctx->trav : ctx->sorted;
    HIPCHK(hipEventRecord(ctx->ev_fork, ctx->st));
    HIPCHK(hipStreamWaitEvent(ctx->st2, ctx->ev_fork, 0));
    launch_advance_kernel<VT>(ctx, ctx->st2, order, ctx->inlm, true);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev_join, ctx->st2));
    if ((rc = launch_contacts_solve<VT>(ctx))) return rc;
    launch_advance_listed<VT>(ctx, ctx->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamWaitEvent(ctx->st, ctx->ev_join, 0));
    // the flags were read by the advance beside the solves: the next step's
    // sort clears them (and the next gp_step starts from a full clear)
    ctx->cur = 1 - ctx->cur;
    ctx->perm_identity = true;
    ctx->cells_valid = true;
    end_early(ctx);
  }
  return GP_OK;
}

// steps per captured graph (even: the buffer flips cancel)
static int graph_unit(const gp_ctx* ctx) { return ctx->prm.overlap ? 8 : 2; }

// captures graph_unit() steps into ctx->gexec
template <typename VT>
static int capture_graph(gp_ctx* ctx) {
  drop_graph(ctx);
  ctx->g_cur = ctx->cur;
  ctx->g_perm = ctx->perm;
  ctx->g_perm_identity = ctx->perm_identity;
  HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
  int rc = launch_steps<VT>(ctx, graph_unit(ctx));
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->st, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    ctx->err = std::string("stream capture failed: ") + hipGetErrorString(e);
    return GP_ERR_HIP;
  }
  e = hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    ctx->gexec = nullptr;
    ctx->err = std::string("graph instantiation failed: ") + hipGetErrorString(e);
    return GP_ERR_HIP;
  }
  return GP_OK;
}

// Width of the contact-solve launch when gp_params.lm_waves = 0, from the
// last known count of solves: while they leave lanes idle (fewer than one
// per lane of one wavefront per SIMD) the solve is latency-bound and one
// wavefront per SIMD is fastest; with several per lane it is throughput-
// bound and more wavefronts hide each other's latency (2 per SIMD fit the
// register budget; 4 queue behind them).  No effect on results.
static void choose_lm_waves(gp_ctx* ctx, unsigned long long nlm) {
  ctx->last_nlm = nlm;
  if (ctx->prm.lm_waves > 0) return;
  const unsigned long long lanes = (unsigned long long)ctx->nsimd * 64ull;
  ctx->lm_waves = ctx->nsimd * (nlm >= 4 * lanes ? 4 : nlm >= lanes ? 2 : 1);
}

template <typename VT>
static bool graph_matches(const gp_ctx* ctx) {
  return ctx->gexec && ctx->g_cur == ctx->cur && ctx->g_perm == ctx->perm &&
         ctx->g_perm_identity == ctx->perm_identity;
}

template <typename VT>
static int graph_steps(gp_ctx* ctx, int nsteps) {
  int rc;
  const int unit = graph_unit(ctx);
  const bool direct = ctx->prm.graph == 2;
  if (!ctx->lists_sized) {
    // one eager step sizes the lists; headroom for the candidate count to grow
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(1), 0, ctx->st, ctx->d_cnt, 0ull,
                       1ll);
    if ((rc = launch_advance<VT>(ctx, true))) return rc;
    if ((rc = launch_collide<VT>(ctx))) return rc;
    --nsteps;
    const unsigned long long need = 2ull * (unsigned long long)ctx->last_nclose + 4096ull;
    if ((rc = ensure_lists(ctx, need))) return rc;
    // the solve's launch width is fixed from here on: from this step's count
    HIPCHK(hipMemcpyAsync(ctx->h_cnt, ctx->d_cnt, NCNT_HOST * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    choose_lm_waves(ctx, ctx->h_cnt[7]);
    drop_graph(ctx);
    ctx->lists_sized = true;
  }
  // gp_params.graph = 2, or fewer steps than a captured unit (a caller
  // stepping once per call, as the reference's main loop does): launched
  // directly, without capture and without host synchronisation
  if (direct || nsteps < unit) {
    if (nsteps > 0 && (rc = launch_steps<VT>(ctx, nsteps))) return rc;
    ctx->contacts_valid = false;
    return GP_OK;
  }
  if (!graph_matches<VT>(ctx) && ctx->gexec) {
    // an odd number of steps since the capture left the buffers swapped:
    // one direct step swaps them back
    if ((rc = launch_steps<VT>(ctx, 1))) return rc;
    --nsteps;
  }
  if (!graph_matches<VT>(ctx) && nsteps >= unit)
    if ((rc = capture_graph<VT>(ctx))) return rc;
  for (; nsteps >= unit; nsteps -= unit) HIPCHK(hipGraphLaunch(ctx->gexec, ctx->st));
  if (nsteps > 0 && (rc = launch_steps<VT>(ctx, nsteps))) return rc;
  ctx->contacts_valid = false;   // pipelined steps reorder the particles
  return GP_OK;
}

static void accumulate_stats(gp_ctx* ctx) {
  // called after a stream sync
  unsigned long long* h = ctx->h_cnt;
  ctx->stats.nahead = (int64_t)std::min(h[12], ctx->ep_cap);
  ctx->stats.nwall = (int64_t)h[0];
  ctx->stats.nprefilter = (int64_t)h[1];
  ctx->stats.nlm = (int64_t)h[7];
  ctx->stats.seed_fallbacks = (int64_t)h[9];
  ctx->stats.ncontact = (int64_t)h[4];
  ctx->stats.ncollision = (int64_t)h[5];
  ctx->stats.lm_nfev = (int64_t)h[3];
  ctx->stats.lm_waves = ctx->lm_waves_last;
  choose_lm_waves(ctx, h[7]);
  ctx->stats.greedy_rounds = (int64_t)h[6];
  // the candidate count the last prefilter saw (the sizing sort's before any)
  ctx->stats.nclose = h[11] ? (int64_t)h[11] : ctx->last_nclose;
}

extern "C" {

static int ready(gp_ctx* ctx) {
  if (!ctx) return GP_ERR_ARG;
  if (!ctx->have_field || !ctx->spec || ctx->np < 1) {
    ctx->err = "field, species and particles must be set first";
    return GP_ERR_STATE;
  }
  return GP_OK;
}

int gp_adv_particle_ellp(gp_ctx* ctx, int ncycle) {
  (void)ncycle;
  int rc = ready(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(1), 0, ctx->st, ctx->d_cnt, 0ull,
                     1ll);
  rc = ctx->mixed ? launch_advance<float>(ctx, true) : launch_advance<double>(ctx, true);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(ctx->h_cnt, ctx->d_cnt, NCNT_HOST * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, ctx->st));
  rc = check_err(ctx);
  ctx->stats.nwall = (int64_t)ctx->h_cnt[0];
  return rc;
}

int gp_collision_detection(gp_ctx* ctx, int64_t nclosemax, int* home, int* near_,
                           int64_t* nclose) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (!nclose || nclosemax < 0 || (nclosemax > 0 && (!home || !near_))) return GP_ERR_ARG;
  rc = ctx->mixed ? launch_sort_count<float>(ctx) : launch_sort_count<double>(ctx);
  if (rc) return rc;
  const long long total = ctx->last_nclose;
  *nclose = total;
  if ((rc = check_err(ctx))) return rc;
  if (total > nclosemax) {
    ctx->err = "NCLOSE exceeds NCLOSEMAX";
    return GP_ERR_NCLOSE_OVERFLOW;
  }
  if (total == 0) return GP_OK;
  int *dh = nullptr, *dn = nullptr;
  HIPCHK(hipMalloc((void**)&dh, (size_t)total * sizeof(int)));
  HIPCHK(hipMalloc((void**)&dn, (size_t)total * sizeof(int)));
  const int* id = ctx->mixed ? pstate<float>(ctx, ctx->cur).id : pstate<double>(ctx, ctx->cur).id;
  hipLaunchKernelGGL(k_pair_emit, dim3(nblk(ctx->np)), dim3(BLOCK), 0, ctx->st, ctx->g,
                     ctx->sorted, ctx->cell, ctx->off, ctx->cnt, ctx->poff, id,
                     ctx->ngroups > 1 ? ctx->gsort : nullptr, dh, dn, ctx->np);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(home, dh, (size_t)total * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipMemcpyAsync(near_, dn, (size_t)total * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  hipFree(dh);
  hipFree(dn);
  return GP_OK;
}

int gp_contact_batch(gp_ctx* ctx, int64_t npairs, const int* home, const int* near_,
                     unsigned char* flag, double* ept1, double* ept2, double* cnormal,
                     int* nfev) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (npairs < 0 || (npairs > 0 && (!home || !near_))) return GP_ERR_ARG;
  if (npairs == 0) return GP_OK;
  const size_t n = (size_t)npairs;
  char* d = nullptr;
  const size_t bytes = n * (2 * sizeof(int) + 1 + 9 * sizeof(double) + sizeof(int)) + 1024;
  HIPCHK(hipMalloc((void**)&d, bytes));
  double* de1 = (double*)d;
  double* de2 = de1 + 3 * n;
  double* dcn = de2 + 3 * n;
  int* dh = (int*)(dcn + 3 * n);
  int* dn = dh + n;
  int* dnf = dn + n;
  unsigned char* df = (unsigned char*)(dnf + n);
  HIPCHK(hipMemcpyAsync(dh, home, n * sizeof(int), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(dn, near_, n * sizeof(int), hipMemcpyHostToDevice, ctx->st));
  // with lm_seed = 1 the pairs are solved as the step solves them: network
  // seeds for the whole batch, then the gated seeded solve
  const bool net = ctx->prm.lm_seed == 1 && ctx->mlp_ready;
  double* dseed = nullptr;
  WorkRec* dwork = nullptr;
  unsigned long long* dnw = nullptr;
  if (net) {
    const size_t sb = n * (2 * sizeof(double) + sizeof(WorkRec)) + 64;
    HIPCHK(hipMalloc((void**)&dseed, sb));
    dwork = (WorkRec*)(dseed + 2 * n);
    dnw = (unsigned long long*)(dwork + n);
    HIPCHK(hipMemcpyAsync(dnw, &n, sizeof(unsigned long long), hipMemcpyHostToDevice, ctx->st));
  }
#define GP_BATCH(VT)                                                                              \
  {                                                                                               \
    auto ps = pstate<VT>(ctx, ctx->cur);                                                          \
    hipLaunchKernelGGL(k_pos_of_id<VT>, dim3(nblk(ctx->np)), dim3(BLOCK), 0, ctx->st, ps,         \
                       ctx->pos_of_id, ctx->np);                                                  \
    if (net) {                                                                                    \
      hipLaunchKernelGGL(k_work_of_pairs, dim3(nblk(npairs)), dim3(BLOCK), 0, ctx->st, dh, dn,    \
                         ctx->pos_of_id, (long long)npairs, ctx->np, dwork, ctx->d_err);          \
      hipLaunchKernelGGL(k_mlp_seed<VT>, dim3(ctx->nsimd), dim3(64), 0, ctx->st, ps, ctx->spec,   \
                         dwork, dnw, (unsigned long long)n, ctx->prm.L, mlp_dev(ctx), dseed);     \
    }                                                                                             \
    hipLaunchKernelGGL(k_contact_batch<VT>, dim3(nblk(npairs)), dim3(BLOCK), 0, ctx->st, ps,      \
                       ctx->spec, dh, dn, ctx->pos_of_id, (long long)npairs, ctx->prm.L, ctx->np, \
                       df, de1, de2, dcn, dnf, ctx->d_err, dseed, ctx->gate, ctx->d_cnt + 9);     \
  }
  if (net) hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(1), 0, ctx->st,
                              ctx->d_cnt + 9, 0ull, 1ll);
  if (ctx->mixed) GP_BATCH(float) else GP_BATCH(double)
#undef GP_BATCH
  HIPCHK(hipGetLastError());
  if (flag) HIPCHK(hipMemcpyAsync(flag, df, n, hipMemcpyDeviceToHost, ctx->st));
  if (ept1) HIPCHK(hipMemcpyAsync(ept1, de1, 3 * n * 8, hipMemcpyDeviceToHost, ctx->st));
  if (ept2) HIPCHK(hipMemcpyAsync(ept2, de2, 3 * n * 8, hipMemcpyDeviceToHost, ctx->st));
  if (cnormal) HIPCHK(hipMemcpyAsync(cnormal, dcn, 3 * n * 8, hipMemcpyDeviceToHost, ctx->st));
  if (nfev) HIPCHK(hipMemcpyAsync(nfev, dnf, n * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
  if (net)
    HIPCHK(hipMemcpyAsync(ctx->h_cnt + 9, ctx->d_cnt + 9, sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  hipFree(d);
  if (dseed) hipFree(dseed);
  if (net) ctx->batch_fallbacks = (int64_t)ctx->h_cnt[9];
  return check_err(ctx);
}

int gp_get_contacts(gp_ctx* ctx, int64_t nmax, int* home, int* near_, unsigned char* resolved,
                    int64_t* n) {
  if (!ctx || !n || nmax < 0 || (nmax > 0 && (!home || !near_))) return GP_ERR_ARG;
  *n = 0;
  if (!ctx->contacts_valid || !ctx->act) {
    ctx->err = "no contact list: call after gp_collision or an eager gp_step, before the next "
               "advance";
    return GP_ERR_STATE;
  }
  HIPCHK(hipMemcpyAsync(ctx->h_cnt + 2, ctx->d_cnt + 2, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  const unsigned long long na = std::min<unsigned long long>(ctx->h_cnt[2], ctx->act_cap);
  *n = (int64_t)na;
  if (na == 0 || nmax == 0) return GP_OK;
  const size_t m = (size_t)na;
  long long* dk = nullptr;
  HIPCHK(hipMalloc((void**)&dk, m * (sizeof(long long) + 2 * sizeof(int) + 1)));
  int* dh = (int*)(dk + m);
  int* dn = dh + m;
  unsigned char* dr = (unsigned char*)(dn + m);
  if (ctx->mixed)
    hipLaunchKernelGGL(k_contacts_out<float>, dim3(nblk((long long)m)), dim3(BLOCK), 0, ctx->st,
                       pstate<float>(ctx, ctx->cur), ctx->act, ctx->act_status, (long long)m, dk,
                       dh, dn, dr);
  else
    hipLaunchKernelGGL(k_contacts_out<double>, dim3(nblk((long long)m)), dim3(BLOCK), 0, ctx->st,
                       pstate<double>(ctx, ctx->cur), ctx->act, ctx->act_status, (long long)m, dk,
                       dh, dn, dr);
  std::vector<long long> k(m);
  std::vector<int> hh(m), nn(m);
  std::vector<unsigned char> rr(m);
  HIPCHK(hipMemcpyAsync(k.data(), dk, m * sizeof(long long), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipMemcpyAsync(hh.data(), dh, m * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipMemcpyAsync(nn.data(), dn, m * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipMemcpyAsync(rr.data(), dr, m, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  hipFree(dk);
  // candidate-list order: the list keys order the pairs as the reference's
  // loop meets them
  std::vector<size_t> ord(m);
  for (size_t a = 0; a < m; ++a) ord[a] = a;
  std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return k[a] < k[b]; });
  const size_t out = std::min<size_t>(m, (size_t)nmax);
  for (size_t a = 0; a < out; ++a) {
    home[a] = hh[ord[a]];
    near_[a] = nn[ord[a]];
    if (resolved) resolved[a] = rr[ord[a]];
  }
  return check_err(ctx);
}

int gp_seed_net_gate(gp_ctx* ctx, int maxfev, double keep_angle) {
  if (!ctx || maxfev < 1 || maxfev > 300 || !(keep_angle >= 0.0) || !(keep_angle <= M_PI))
    return GP_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  drop_graph(ctx);
  ctx->gate.maxfev = maxfev;
  ctx->gate.cos_keep = std::cos(keep_angle);
  return GP_OK;
}

int gp_collision(gp_ctx* ctx, int ncycle, double time) {
  (void)time;
  int rc = ready(ctx);
  if (rc) return rc;
  // the statistics gate of collision_interaction.f90:330 on this cycle
  if (ctx->cs_fout2 > 0)
    hipLaunchKernelGGL(k_fill<long long>, dim3(1), dim3(1), 0, ctx->st, ctx->cs_cycle,
                       (long long)ncycle, 1ll);
  rc = ctx->mixed ? launch_collide<float>(ctx) : launch_collide<double>(ctx);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(ctx->h_cnt, ctx->d_cnt, NCNT_HOST * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, ctx->st));
  rc = check_err(ctx);
  accumulate_stats(ctx);
  return rc;
}

int gp_collision_advance(gp_ctx* ctx, int ncycle, double time, int ncycle_next) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (ctx->cs_fout2 > 0 || ctx->timing) {
    // statistics passes and stage timers: one after the other, eagerly
    if ((rc = gp_collision(ctx, ncycle, time))) return rc;
    return gp_adv_particle_ellp(ctx, ncycle_next);
  }
  if (!ctx->lists_sized) {
    // an eager pass sizes the lists with headroom; later calls then run
    // without host synchronisation
    rc = ctx->mixed ? launch_collide<float>(ctx) : launch_collide<double>(ctx);
    if (rc) return rc;
    const unsigned long long need = 2ull * (unsigned long long)ctx->last_nclose + 4096ull;
    if ((rc = ensure_lists(ctx, need))) return rc;
    HIPCHK(hipMemcpyAsync(ctx->h_cnt, ctx->d_cnt, NCNT_HOST * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    choose_lm_waves(ctx, ctx->h_cnt[7]);
    accumulate_stats(ctx);
    ctx->lists_sized = true;
    if ((rc = check_err(ctx))) return rc;
    return gp_adv_particle_ellp(ctx, ncycle_next);
  }
  rc = ctx->mixed ? launch_collide_advance<float>(ctx) : launch_collide_advance<double>(ctx);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(ctx->h_cnt, ctx->d_cnt, NCNT_HOST * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, ctx->st));
  return GP_OK;
}

int gp_set_dtime(gp_ctx* ctx, double dtime) {
  if (!ctx || !(dtime > 0.0)) return GP_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st2));
  // the step's kernels take DTIME by value: a captured graph is re-captured
  if (dtime != ctx->prm.dtime) drop_graph(ctx);
  ctx->prm.dtime = dtime;
  ctx->sp.dtime = dtime;
  return GP_OK;
}

int gp_set_collision_stats(gp_ctx* ctx, int nclass, int npoly, int fout2, int ncycle) {
  if (!ctx) return GP_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st2));
  drop_graph(ctx);
  void* bufs[] = {ctx->cs_rec, ctx->cs_buf, ctx->cs_cnt, ctx->cs_cycle, ctx->cs_dev};
  for (void* b : bufs)
    if (b) hipFree(b);
  ctx->cs_rec = nullptr; ctx->cs_buf = nullptr; ctx->cs_cnt = nullptr; ctx->cs_cycle = nullptr;
  ctx->cs_dev = nullptr;
  ctx->cs_cap = 0;
  ctx->cs_fout2 = 0;
  if (fout2 <= 0) return GP_OK;
  if (nclass < 1 || npoly < 1 || (long long)nclass * npoly * npoly > GP_CSTAT_MAX_BINS) {
    ctx->err = "collision statistics: need nclass, npoly >= 1 and nclass * npoly^2 <= " +
               std::to_string(GP_CSTAT_MAX_BINS) + " bins";
    return GP_ERR_ARG;
  }
  const int nbin = nclass * npoly * npoly;
  int rc;
  if ((rc = dalloc(ctx, &ctx->cs_buf, (size_t)nbin * (CS_NQ + 16)))) return rc;
  if ((rc = dalloc(ctx, &ctx->cs_cnt, 2))) return rc;
  if ((rc = dalloc(ctx, &ctx->cs_cycle, 1))) return rc;
  const unsigned long long cap = std::max<unsigned long long>(ctx->work_cap, 4096ull);
  if ((rc = dalloc(ctx, &ctx->cs_rec, (size_t)cap))) return rc;
  ctx->cs_cap = cap;
  // sums and their running sum from ZERO (collision_interaction.f90:171,
  // collision_initiation.f90:172), the last pass's values from 0
  std::vector<double> init((size_t)nbin * (CS_NQ + 16), 0.0);
  for (size_t t = 0; t < (size_t)nbin * CS_NQ; ++t) init[t] = ZERO_R;
  for (size_t t = (size_t)nbin * (CS_NQ + 8); t < init.size(); ++t) init[t] = ZERO_R;
  HIPCHK(hipMemcpy(ctx->cs_buf, init.data(), init.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(ctx->cs_cnt, 0, 2 * sizeof(unsigned long long)));
  const long long cyc = ncycle;
  HIPCHK(hipMemcpy(ctx->cs_cycle, &cyc, sizeof(cyc), hipMemcpyHostToDevice));
  ctx->cs_nclass = nclass;
  ctx->cs_npoly = npoly;
  ctx->cs_nbin = nbin;
  ctx->cs_fout2 = fout2;
  if ((rc = dalloc(ctx, &ctx->cs_dev, 1))) return rc;
  return sync_cstat(ctx);
}

int gp_get_collision_stats(gp_ctx* ctx, double* last, double* time_sum, int64_t* nstat) {
  if (!ctx) return GP_ERR_ARG;
  if (ctx->cs_fout2 <= 0) {
    ctx->err = "collision statistics are off (gp_set_collision_stats)";
    return GP_ERR_STATE;
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  const size_t nb = (size_t)ctx->cs_nbin;
  const double* dlast = ctx->cs_buf + nb * CS_NQ;
  if (last) HIPCHK(hipMemcpy(last, dlast, nb * 8 * sizeof(double), hipMemcpyDeviceToHost));
  if (time_sum)
    HIPCHK(hipMemcpy(time_sum, dlast + nb * 8, nb * 8 * sizeof(double), hipMemcpyDeviceToHost));
  if (nstat) {
    unsigned long long h[2];
    HIPCHK(hipMemcpy(h, ctx->cs_cnt, sizeof(h), hipMemcpyDeviceToHost));
    *nstat = (int64_t)h[1];
  }
  return check_err(ctx);
}

int gp_step(gp_ctx* ctx, int nsteps) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (nsteps <= 0) return GP_OK;
  if (ctx->prm.graph) {
    rc = ctx->mixed ? graph_steps<float>(ctx, nsteps) : graph_steps<double>(ctx, nsteps);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->h_cnt, ctx->d_cnt, NCNT_HOST * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, ctx->st));
    return GP_OK;
  }
  for (int s = 0; s < nsteps; ++s) {
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(1), 0, ctx->st, ctx->d_cnt, 0ull,
                       1ll);
    rc = ctx->mixed ? launch_advance<float>(ctx, true) : launch_advance<double>(ctx, true);
    if (rc) return rc;
    rc = ctx->mixed ? launch_collide<float>(ctx) : launch_collide<double>(ctx);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpyAsync(ctx->h_cnt, ctx->d_cnt, NCNT_HOST * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, ctx->st));
  return GP_OK;
}

int gp_get_stats(gp_ctx* ctx, gp_stats* out) {
  if (!ctx || !out) return GP_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  collect_timing(ctx);
  accumulate_stats(ctx);
  ctx->stats.t_advance_ms = ctx->t_stage[0];
  ctx->stats.t_sort_ms = ctx->t_stage[1];
  ctx->stats.t_contact_ms = ctx->t_stage[2];
  ctx->stats.t_greedy_ms = ctx->t_stage[3];
  ctx->stats.t_solve_ms = ctx->t_solve;
  ctx->stats.ntimed = ctx->ntimed;
  *out = ctx->stats;
  return GP_OK;
}

int gp_reset_timers(gp_ctx* ctx) {
  if (!ctx) return GP_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->st));
  collect_timing(ctx);
  for (double& t : ctx->t_stage) t = 0.0;
  ctx->t_solve = 0.0;
  ctx->ntimed = 0;
  return GP_OK;
}

int gp_get_collision_counts(gp_ctx* ctx, int np, int* counts, int reset) {
  if (!ctx || !counts || np != ctx->np || np < 1) return GP_ERR_ARG;
  HIPCHK(hipMemcpyAsync(counts, ctx->ccount, (size_t)np * sizeof(int), hipMemcpyDeviceToHost,
                        ctx->st));
  if (reset)
    hipLaunchKernelGGL(k_fill<int>, dim3(nblk(np)), dim3(BLOCK), 0, ctx->st, ctx->ccount, 0,
                       (long long)np);
  HIPCHK(hipStreamSynchronize(ctx->st));
  return check_err(ctx);
}

int gp_particle_stats(gp_ctx* ctx, int nspec, int npoly, int ref_norm, double* out) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (!out || nspec != ctx->nspec || npoly < 1 || nspec % npoly) return GP_ERR_ARG;
  double *acc = nullptr, *mean = nullptr;
  const size_t na = (size_t)nspec * NSTP, nm = (size_t)nspec * 10;
  if (hipMalloc((void**)&acc, na * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&mean, nm * sizeof(double)) != hipSuccess) {
    if (acc) hipFree(acc);
    ctx->err = "particle statistics: out of device memory";
    return GP_ERR_NOMEM;
  }
  const unsigned grid = std::max(1u, std::min(nblk(ctx->np), 2048u));
  const unsigned sgrid = nblk(nspec);
  hipMemsetAsync(acc, 0, na * sizeof(double), ctx->st);
  if (ctx->mixed) {
    auto ps = pstate<float>(ctx, ctx->cur);
    hipLaunchKernelGGL((k_pstat<float, 1>), dim3(grid), dim3(BLOCK), 0, ctx->st, ps, ctx->spec,
                       ctx->np, nspec, acc, (const double*)nullptr);
    hipLaunchKernelGGL(k_pstat_mean, dim3(sgrid), dim3(BLOCK), 0, ctx->st, acc, mean, nspec);
    hipLaunchKernelGGL((k_pstat<float, 2>), dim3(grid), dim3(BLOCK), 0, ctx->st, ps, ctx->spec,
                       ctx->np, nspec, acc, (const double*)mean);
  } else {
    auto ps = pstate<double>(ctx, ctx->cur);
    hipLaunchKernelGGL((k_pstat<double, 1>), dim3(grid), dim3(BLOCK), 0, ctx->st, ps,
                       ctx->spec, ctx->np, nspec, acc, (const double*)nullptr);
    hipLaunchKernelGGL(k_pstat_mean, dim3(sgrid), dim3(BLOCK), 0, ctx->st, acc, mean, nspec);
    hipLaunchKernelGGL((k_pstat<double, 2>), dim3(grid), dim3(BLOCK), 0, ctx->st, ps,
                       ctx->spec, ctx->np, nspec, acc, (const double*)mean);
  }
  hipLaunchKernelGGL(k_pstat_final, dim3(sgrid), dim3(BLOCK), 0, ctx->st, ctx->spec, acc,
                     (const double*)mean, nspec, npoly, ref_norm ? 1 : 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, acc, na * sizeof(double), hipMemcpyDeviceToHost,
                                          ctx->st);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->st);
  hipFree(acc);
  hipFree(mean);
  if (e != hipSuccess) {
    ctx->err = std::string("particle statistics: ") + hipGetErrorString(e);
    return GP_ERR_HIP;
  }
  return check_err(ctx);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// STAT_PARTICLE's PDFs (LEVEL1_STPDF, stat_particle.f90:536-645): per species
// row 12 histograms of GP_NPDF bins, bin = int((value - min) / width) + 1
// clamped to [1, GP_NPDF]: u, v, w in [-2, 2]; body spin in [-2000, 2000];
// lab spin R(q) omega in [-2000, 2000]; R(q) e_x (the body x axis in the lab
// frame) in [-1, 1].  R and R omega are formed with the reference's
// operation order and no contraction, so the bins are the reference's.
// Counts go to LDS per workgroup when the rows are few, else to global
// memory.
// ---------------------------------------------------------------------------
constexpr int PDF_Q = 12;
constexpr int PDF_LDS_ROWS = 10;   // 10 x 12 x 100 counters = 48 KB

__device__ __forceinline__ int pdf_bin(double v, double lo, double width) {
  double t = __ddiv_rn(__dsub_rn(v, lo), width);
  // int() truncates toward zero; out-of-range values clamp to the end bins
  if (!(t < 1.0e9)) t = 1.0e9;
  if (!(t > -1.0e9)) t = -1.0e9;
  int b = (int)t + 1;
  return b < 1 ? 0 : (b > GP_NPDF ? GP_NPDF - 1 : b - 1);
}

template <typename VT>
__device__ __forceinline__ void pdf_bins(const PState<VT>& p, int i, int b[PDF_Q]) {
  const double du = 4.0 / (double)GP_NPDF, dw = 4000.0 / (double)GP_NPDF,
               dp = 2.0 / (double)GP_NPDF;
  const double u[3] = {(double)p.u[i], (double)p.v[i], (double)p.w[i]};
  const double w[3] = {(double)p.ox[i], (double)p.oy[i], (double)p.oz[i]};
  const double a = p.qa[i], qb = p.qb[i], c = p.qc[i], d = p.qd[i];
  // R of transform_basis.f90:44-54
  double R[3][3];
  R[0][0] = __dsub_rn(1.0, __dmul_rn(2.0, __dadd_rn(__dmul_rn(c, c), __dmul_rn(d, d))));
  R[0][1] = __dmul_rn(2.0, __dsub_rn(__dmul_rn(qb, c), __dmul_rn(a, d)));
  R[0][2] = __dmul_rn(2.0, __dadd_rn(__dmul_rn(qb, d), __dmul_rn(a, c)));
  R[1][0] = __dmul_rn(2.0, __dadd_rn(__dmul_rn(qb, c), __dmul_rn(a, d)));
  R[1][1] = __dsub_rn(1.0, __dmul_rn(2.0, __dadd_rn(__dmul_rn(qb, qb), __dmul_rn(d, d))));
  R[1][2] = __dmul_rn(2.0, __dsub_rn(__dmul_rn(c, d), __dmul_rn(a, qb)));
  R[2][0] = __dmul_rn(2.0, __dsub_rn(__dmul_rn(qb, d), __dmul_rn(a, c)));
  R[2][1] = __dmul_rn(2.0, __dadd_rn(__dmul_rn(c, d), __dmul_rn(a, qb)));
  R[2][2] = __dsub_rn(1.0, __dmul_rn(2.0, __dadd_rn(__dmul_rn(qb, qb), __dmul_rn(c, c))));
  for (int k = 0; k < 3; ++k) {
    b[k] = pdf_bin(u[k], -2.0, du);
    b[3 + k] = pdf_bin(w[k], -2000.0, dw);
    const double g = __dadd_rn(__dadd_rn(__dmul_rn(R[k][0], w[0]), __dmul_rn(R[k][1], w[1])),
                               __dmul_rn(R[k][2], w[2]));
    b[6 + k] = pdf_bin(g, -2000.0, dw);
    b[9 + k] = pdf_bin(R[k][0], -1.0, dp);
  }
}

template <typename VT>
__global__ __launch_bounds__(BLOCK) void k_ppdf(PState<VT> p, int np, int nspec,
                                                unsigned* __restrict__ hist) {
  __shared__ unsigned sh[PDF_LDS_ROWS * PDF_Q * GP_NPDF];
  const bool lds = nspec <= PDF_LDS_ROWS;
  const int nh = nspec * PDF_Q * GP_NPDF;
  if (lds)
    for (int k = threadIdx.x; k < nh; k += blockDim.x) sh[k] = 0u;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
    int b[PDF_Q];
    pdf_bins(p, i, b);
    const int base = p.sp[i] * PDF_Q * GP_NPDF;
    for (int k = 0; k < PDF_Q; ++k) {
      if (lds) atomicAdd(&sh[base + k * GP_NPDF + b[k]], 1u);
      else atomicAdd(&hist[base + k * GP_NPDF + b[k]], 1u);
    }
  }
  if (!lds) return;
  __syncthreads();
  for (int k = threadIdx.x; k < nh; k += blockDim.x)
    if (sh[k]) atomicAdd(&hist[k], sh[k]);
}

// ---------------------------------------------------------------------------
// STAT_PARTICLE level 2 (stat_particle.f90:737-826): the particle velocity
// fluctuation at the start of correlation slot NLGR, u - MEAN_PART(1:3, J,
// IDP), kept per particle (by reference index, so that it survives the
// reordering of the steps); at a later statistics cycle the sums over each
// species row of (u - MEAN_PART(1:3)) * u_t0 / NPM(IDP), the increments of
// RPX/RPY/RPZ_LOC(NTLGR, J, NLGR, IDP), formed per particle as the routine
// forms them and summed in another order.
// ---------------------------------------------------------------------------
template <typename VT>
__global__ __launch_bounds__(BLOCK) void k_lag_origin(PState<VT> p, int np,
                                                      const double* __restrict__ mean,
                                                      double* __restrict__ u0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int s = p.sp[i], id = p.id[i];
  u0[id] = (double)p.u[i] - mean[3 * s];
  u0[np + id] = (double)p.v[i] - mean[3 * s + 1];
  u0[2 * (size_t)np + id] = (double)p.w[i] - mean[3 * s + 2];
}

constexpr int LAG_LDS_ROWS = 1024;

template <typename VT>
__global__ __launch_bounds__(BLOCK) void k_lag_corr(PState<VT> p, int np, int nspec,
                                                    const double* __restrict__ mean,
                                                    const double* __restrict__ npm,
                                                    const double* __restrict__ u0,
                                                    double* __restrict__ out) {
  __shared__ double sh[3 * LAG_LDS_ROWS];
  const bool lds = nspec <= LAG_LDS_ROWS;
  if (lds)
    for (int k = threadIdx.x; k < 3 * nspec; k += blockDim.x) sh[k] = 0.0;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
    const int s = p.sp[i], id = p.id[i];
    const double v[3] = {(double)p.u[i], (double)p.v[i], (double)p.w[i]};
    for (int c = 0; c < 3; ++c) {
      const double t = ((v[c] - mean[3 * s + c]) * u0[(size_t)c * np + id]) / npm[s];
      if (lds) atomicAdd(&sh[3 * s + c], t);
      else atomicAdd(&out[3 * s + c], t);
    }
  }
  if (!lds) return;
  __syncthreads();
  for (int k = threadIdx.x; k < 3 * nspec; k += blockDim.x)
    if (sh[k] != 0.0) atomicAdd(&out[k], sh[k]);
}

extern "C" {

int gp_partfluid_stats(gp_ctx* ctx, int nspec, int npoly, int ref_norm, double* out) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (!out || nspec != ctx->nspec || npoly < 1 || nspec % npoly) return GP_ERR_ARG;
  if (!ctx->flout || !ctx->fluid_valid) {
    ctx->err = "particle-fluid statistics need gp_params.keep_fluid = 1 and an advance of the "
               "current particles";
    return GP_ERR_STATE;
  }
  // per species: PMASS, dequiv (stat_partfluid.f90:124, 135), TAUPF with the
  // elongation factor (:300-314), GRAVITY
  std::vector<double> pfc((size_t)nspec * 4);
  for (int s = 0; s < nspec; ++s) {
    const double* r = ctx->sp_rows.data() + (size_t)s * GP_SP_STRIDE;
    const double rhop = r[GP_SP_RHOP], emaj = r[GP_SP_EMAJ], apr = r[GP_SP_APR];
    double* c = pfc.data() + (size_t)s * 4;
    c[0] = rhop * (4.0 / 3.0) * M_PI * std::pow(emaj, 3.0) / std::pow(apr, 2.0);
    c[1] = 2.0 * emaj / std::pow(apr, 2.0 / 3.0);
    const double minax = emaj / apr;
    const double fbeta = apr == 1.0 ? 1.0
                                    : (apr * std::log(apr + std::sqrt(apr * apr - 1.0))) /
                                          std::sqrt(apr * apr - 1.0);
    c[2] = (2.0 / 9.0) * (rhop * (minax * minax)) / (ctx->prm.rhof * ctx->prm.visc) * fbeta;
    c[3] = r[GP_SP_GRAV];
  }
  double *acc = nullptr, *dout = nullptr, *dpfc = nullptr;
  const size_t na = (size_t)nspec * NPFS;
  if (hipMalloc((void**)&acc, na * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&dout, na * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&dpfc, pfc.size() * sizeof(double)) != hipSuccess) {
    if (acc) hipFree(acc);
    if (dout) hipFree(dout);
    ctx->err = "particle-fluid statistics: out of device memory";
    return GP_ERR_NOMEM;
  }
  hipError_t e = hipMemcpyAsync(dpfc, pfc.data(), pfc.size() * sizeof(double),
                                hipMemcpyHostToDevice, ctx->st);
  // MEAN_PARTFLUID_LOC(:,:,:) = ZERO (:114)
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk((long long)na)), dim3(BLOCK), 0, ctx->st, acc,
                     ZERO_R, (long long)na);
  const unsigned grid = std::max(1u, std::min(nblk(ctx->np), 2048u));
  if (ctx->mixed)
    hipLaunchKernelGGL(k_pfstat<float>, dim3(grid), dim3(BLOCK), 0, ctx->st,
                       pstate<float>(ctx, ctx->cur), (const double*)ctx->flout,
                       (const double*)dpfc, ctx->np, nspec, ctx->prm.visc, ctx->prm.rhof, acc);
  else
    hipLaunchKernelGGL(k_pfstat<double>, dim3(grid), dim3(BLOCK), 0, ctx->st,
                       pstate<double>(ctx, ctx->cur), (const double*)ctx->flout,
                       (const double*)dpfc, ctx->np, nspec, ctx->prm.visc, ctx->prm.rhof, acc);
  hipLaunchKernelGGL(k_pfstat_final, dim3(nblk(nspec)), dim3(BLOCK), 0, ctx->st,
                     (const double*)acc, dout, nspec, npoly, ref_norm ? 1 : 0);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, dout, na * sizeof(double), hipMemcpyDeviceToHost, ctx->st);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->st);
  hipFree(acc);
  hipFree(dout);
  hipFree(dpfc);
  if (e != hipSuccess) {
    ctx->err = std::string("particle-fluid statistics: ") + hipGetErrorString(e);
    return GP_ERR_HIP;
  }
  return check_err(ctx);
}

int gp_particle_pdf(gp_ctx* ctx, int nspec, double* out) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (!out || nspec != ctx->nspec) return GP_ERR_ARG;
  const size_t nh = (size_t)nspec * PDF_Q * GP_NPDF;
  unsigned* hist = nullptr;
  if (hipMalloc((void**)&hist, nh * sizeof(unsigned)) != hipSuccess) {
    ctx->err = "particle PDFs: out of device memory";
    return GP_ERR_NOMEM;
  }
  HIPCHK(hipMemsetAsync(hist, 0, nh * sizeof(unsigned), ctx->st));
  const unsigned grid = std::max(1u, std::min(nblk(ctx->np), 1024u));
  if (ctx->mixed)
    hipLaunchKernelGGL(k_ppdf<float>, dim3(grid), dim3(BLOCK), 0, ctx->st,
                       pstate<float>(ctx, ctx->cur), ctx->np, nspec, hist);
  else
    hipLaunchKernelGGL(k_ppdf<double>, dim3(grid), dim3(BLOCK), 0, ctx->st,
                       pstate<double>(ctx, ctx->cur), ctx->np, nspec, hist);
  std::vector<unsigned> h(nh);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), hist, nh * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->st);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->st);
  hipFree(hist);
  if (e != hipSuccess) {
    ctx->err = std::string("particle PDFs: ") + hipGetErrorString(e);
    return GP_ERR_HIP;
  }
  // the routine's histograms start at ZERO (stat_particle.f90:164)
  for (size_t k = 0; k < nh; ++k) out[k] = (double)h[k] + ZERO_R;
  return check_err(ctx);
}

int gp_lagrangian_origin(gp_ctx* ctx, int slot, int nspec, const double* mean) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (slot < 0 || slot >= GP_NLGR || !mean || nspec != ctx->nspec) return GP_ERR_ARG;
  const size_t n3 = 3 * (size_t)ctx->np;
  if (ctx->lag_np != ctx->np) {
    for (double*& b : ctx->lag_u0)
      if (b) { hipFree(b); b = nullptr; }
    ctx->lag_np = ctx->np;
  }
  if (!ctx->lag_u0[slot] && hipMalloc((void**)&ctx->lag_u0[slot], n3 * sizeof(double)) != hipSuccess) {
    ctx->lag_u0[slot] = nullptr;
    ctx->err = "Lagrangian origin: out of device memory";
    return GP_ERR_NOMEM;
  }
  double* dm = nullptr;
  HIPCHK(hipMalloc((void**)&dm, 3 * (size_t)nspec * sizeof(double)));
  HIPCHK(hipMemcpyAsync(dm, mean, 3 * (size_t)nspec * sizeof(double), hipMemcpyHostToDevice,
                        ctx->st));
  if (ctx->mixed)
    hipLaunchKernelGGL(k_lag_origin<float>, dim3(nblk(ctx->np)), dim3(BLOCK), 0, ctx->st,
                       pstate<float>(ctx, ctx->cur), ctx->np, (const double*)dm, ctx->lag_u0[slot]);
  else
    hipLaunchKernelGGL(k_lag_origin<double>, dim3(nblk(ctx->np)), dim3(BLOCK), 0, ctx->st,
                       pstate<double>(ctx, ctx->cur), ctx->np, (const double*)dm,
                       ctx->lag_u0[slot]);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->st));
  hipFree(dm);
  return check_err(ctx);
}

int gp_lagrangian_corr(gp_ctx* ctx, int slot, int nspec, const double* mean, const double* npm,
                       double* out) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (slot < 0 || slot >= GP_NLGR || !mean || !npm || !out || nspec != ctx->nspec)
    return GP_ERR_ARG;
  if (!ctx->lag_u0[slot] || ctx->lag_np != ctx->np) {
    ctx->err = "Lagrangian slot " + std::to_string(slot) + " has no origin (gp_lagrangian_origin)";
    return GP_ERR_STATE;
  }
  const size_t ns3 = 3 * (size_t)nspec;
  double* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, (2 * ns3 + nspec) * sizeof(double)));
  double *dm = d, *dn = d + ns3, *dout = d + ns3 + nspec;
  HIPCHK(hipMemcpyAsync(dm, mean, ns3 * sizeof(double), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(dn, npm, (size_t)nspec * sizeof(double), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemsetAsync(dout, 0, ns3 * sizeof(double), ctx->st));
  const unsigned grid = std::max(1u, std::min(nblk(ctx->np), 1024u));
  if (ctx->mixed)
    hipLaunchKernelGGL(k_lag_corr<float>, dim3(grid), dim3(BLOCK), 0, ctx->st,
                       pstate<float>(ctx, ctx->cur), ctx->np, nspec, (const double*)dm,
                       (const double*)dn, (const double*)ctx->lag_u0[slot], dout);
  else
    hipLaunchKernelGGL(k_lag_corr<double>, dim3(grid), dim3(BLOCK), 0, ctx->st,
                       pstate<double>(ctx, ctx->cur), ctx->np, nspec, (const double*)dm,
                       (const double*)dn, (const double*)ctx->lag_u0[slot], dout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, ns3 * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  hipFree(d);
  return check_err(ctx);
}

int gp_collision_grid(gp_ctx* ctx, int* nbox_xyz, double* box_min_size) {
  if (!ctx || !nbox_xyz || !box_min_size) return GP_ERR_ARG;
  nbox_xyz[0] = ctx->g.nx; nbox_xyz[1] = ctx->g.ny; nbox_xyz[2] = ctx->g.nz;
  nbox_xyz[3] = ctx->g.nbox;
  box_min_size[0] = ctx->g.xmin; box_min_size[1] = ctx->g.ymin; box_min_size[2] = ctx->g.zmin;
  box_min_size[3] = ctx->g.dx; box_min_size[4] = ctx->g.dy; box_min_size[5] = ctx->g.dz;
  return GP_OK;
}

}  // extern "C"


// ===========================================================================
// contact-seed network
// ===========================================================================
static int mlp_alloc(gp_ctx* ctx) {
  if (ctx->mlp_th) return GP_OK;
  int rc;
  if ((rc = dalloc(ctx, &ctx->mlp_th, (size_t)MLP_NPAD))) return rc;
  if ((rc = dalloc(ctx, &ctx->mlp_m, (size_t)MLP_NPAD))) return rc;
  if ((rc = dalloc(ctx, &ctx->mlp_v, (size_t)MLP_NPAD))) return rc;
  if ((rc = dalloc(ctx, &ctx->mlp_grad, (size_t)MLP_NPAD))) return rc;
  if ((rc = dalloc(ctx, &ctx->mlp_wb, (size_t)MW_N))) return rc;
  if ((rc = dalloc(ctx, &ctx->mlp_bias, (size_t)MB_N))) return rc;
  if ((rc = dalloc(ctx, &ctx->mlp_part, (size_t)MLP_PARTS * MLP_NPAD))) return rc;
  if ((rc = dalloc(ctx, &ctx->mlp_lpart, (size_t)MLP_PARTS))) return rc;
  if ((rc = dalloc(ctx, &ctx->mlp_loss, (size_t)1))) return rc;
  HIPCHK(hipMemsetAsync(ctx->mlp_wb, 0, MW_N * sizeof(short), ctx->st));
  HIPCHK(hipMemsetAsync(ctx->mlp_bias, 0, MB_N * sizeof(float), ctx->st));
  return GP_OK;
}

// compact parameter vector <-> padded layout
static void mlp_pad(const float* c, std::vector<float>& pd) {
  pd.assign(MLP_NPAD, 0.f);
  int o = 0;
  for (int h = 0; h < 64; ++h)
    for (int f = 0; f < 12; ++f) pd[MO_W1 + h * 16 + f] = c[o++];
  for (int h = 0; h < 64; ++h) pd[MO_B1 + h] = c[o++];
  for (int k = 0; k < 64 * 64; ++k) pd[MO_W2 + k] = c[o++];
  for (int h = 0; h < 64; ++h) pd[MO_B2 + h] = c[o++];
  for (int k = 0; k < 4 * 64; ++k) pd[MO_W3 + k] = c[o++];
  for (int k = 0; k < 4; ++k) pd[MO_B3 + k] = c[o++];
}
static void mlp_unpad(const float* pd, float* c) {
  int o = 0;
  for (int h = 0; h < 64; ++h)
    for (int f = 0; f < 12; ++f) c[o++] = pd[MO_W1 + h * 16 + f];
  for (int h = 0; h < 64; ++h) c[o++] = pd[MO_B1 + h];
  for (int k = 0; k < 64 * 64; ++k) c[o++] = pd[MO_W2 + k];
  for (int h = 0; h < 64; ++h) c[o++] = pd[MO_B2 + h];
  for (int k = 0; k < 4 * 64; ++k) c[o++] = pd[MO_W3 + k];
  for (int k = 0; k < 4; ++k) c[o++] = pd[MO_B3 + k];
}

static void mlp_launch_adam(gp_ctx* ctx, int update, float lr) {
  float c1 = 1.f, c2 = 1.f;
  if (update) {
    ++ctx->mlp_step;
    c1 = 1.f - powf(0.9f, (float)ctx->mlp_step);
    c2 = 1.f - powf(0.999f, (float)ctx->mlp_step);
  }
  short* wb = ctx->mlp_wb;
  float* bb = ctx->mlp_bias;
  hipLaunchKernelGGL(k_mlp_adam, dim3(nblk(MLP_NPAD)), dim3(BLOCK), 0, ctx->st, ctx->mlp_th,
                     ctx->mlp_m, ctx->mlp_v, ctx->mlp_grad, lr, c1, c2, wb + MW_W1, wb + MW_W2,
                     wb + MW_W3, wb + MW_W2T, wb + MW_W3T, bb + MB_B1, bb + MB_B2, bb + MB_B3,
                     update);
}

// one gradient over n rows already on the device (x [n][16], t [n][4])
static void mlp_launch_grad(gp_ctx* ctx, const float* dx, const float* dt, long long n) {
  const long long ntile = (n + 15) / 16;
  const int parts = (int)std::min<long long>(ntile, MLP_PARTS);
  hipLaunchKernelGGL(k_mlp_train_part, dim3(parts), dim3(64), 0, ctx->st, mlp_dev(ctx), dx, dt,
                     n, 1.0f / (float)n, ctx->mlp_part, ctx->mlp_lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3(nblk(MLP_NPAD)), dim3(BLOCK), 0, ctx->st, ctx->mlp_part,
                     parts, ctx->mlp_grad, ctx->mlp_lpart, ctx->mlp_loss);
}

// host rows (12 features; 4 targets: normal, flag) -> device rows (16, 4)
static int mlp_upload(gp_ctx* ctx, long long n, const float* x, const float* t, float** dx,
                      float** dt) {
  std::vector<float> hx((size_t)n * 16, 0.f), ht;
  for (long long k = 0; k < n; ++k)
    for (int c = 0; c < 12; ++c) hx[(size_t)k * 16 + c] = x[k * 12 + c];
  HIPCHK(hipMalloc((void**)dx, (size_t)n * 16 * sizeof(float)));
  HIPCHK(hipMemcpy(*dx, hx.data(), hx.size() * sizeof(float), hipMemcpyHostToDevice));
  if (t) {
    ht.assign((size_t)n * 4, 0.f);
    for (long long k = 0; k < n; ++k)
      for (int c = 0; c < 4; ++c) ht[(size_t)k * 4 + c] = t[k * 4 + c];
    HIPCHK(hipMalloc((void**)dt, (size_t)n * 4 * sizeof(float)));
    HIPCHK(hipMemcpy(*dt, ht.data(), ht.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  return GP_OK;
}

// epochs of mini-batch Adam over n device rows (batches in order)
static int mlp_train_device(gp_ctx* ctx, const float* dx, const float* dt, long long n,
                            int epochs, int batch, float lr, double* loss) {
  float h_loss = 0.f;
  double lsum = 0.0;
  long long lrows = 0;
  for (int e = 0; e < epochs; ++e) {
    lsum = 0.0;
    lrows = 0;
    for (long long b0 = 0; b0 < n; b0 += batch) {
      const long long nb = std::min<long long>(batch, n - b0);
      mlp_launch_grad(ctx, dx + b0 * 16, dt + b0 * 4, nb);
      mlp_launch_adam(ctx, 1, lr);
      if (e == epochs - 1) {
        HIPCHK(hipMemcpyAsync(&h_loss, ctx->mlp_loss, sizeof(float), hipMemcpyDeviceToHost,
                              ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        lsum += (double)h_loss;
        lrows += nb;
      }
    }
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (loss) *loss = lrows ? lsum / (double)lrows : 0.0;
  return GP_OK;
}

extern "C" {

int gp_seed_net_set(gp_ctx* ctx, const float* theta, int nparam) {
  if (!ctx || !theta || nparam != GP_SEED_NET_NPARAM) return GP_ERR_ARG;
  int rc;
  if ((rc = mlp_alloc(ctx))) return rc;
  std::vector<float> pd;
  mlp_pad(theta, pd);
  HIPCHK(hipMemcpyAsync(ctx->mlp_th, pd.data(), MLP_NPAD * sizeof(float), hipMemcpyHostToDevice,
                        ctx->st));
  HIPCHK(hipMemsetAsync(ctx->mlp_m, 0, MLP_NPAD * sizeof(float), ctx->st));
  HIPCHK(hipMemsetAsync(ctx->mlp_v, 0, MLP_NPAD * sizeof(float), ctx->st));
  mlp_launch_adam(ctx, 0, 0.f);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->mlp_step = 0;
  ctx->mlp_ready = true;
  return GP_OK;
}

int gp_seed_net_get(gp_ctx* ctx, float* theta, int nparam) {
  if (!ctx || !theta || nparam != GP_SEED_NET_NPARAM) return GP_ERR_ARG;
  if (!ctx->mlp_ready) {
    ctx->err = "the seed network has no parameters (gp_seed_net_set)";
    return GP_ERR_STATE;
  }
  std::vector<float> pd(MLP_NPAD);
  HIPCHK(hipMemcpyAsync(pd.data(), ctx->mlp_th, MLP_NPAD * sizeof(float), hipMemcpyDeviceToHost,
                        ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  mlp_unpad(pd.data(), theta);
  return GP_OK;
}

int gp_seed_net_forward(gp_ctx* ctx, int64_t n, const float* x, float* y) {
  if (!ctx || n < 1 || !x || !y) return GP_ERR_ARG;
  if (!ctx->mlp_ready) {
    ctx->err = "the seed network has no parameters (gp_seed_net_set)";
    return GP_ERR_STATE;
  }
  float *dx = nullptr, *dy = nullptr;
  int rc = mlp_upload(ctx, n, x, nullptr, &dx, nullptr);
  if (!rc) {
    HIPCHK(hipMalloc((void**)&dy, (size_t)n * 4 * sizeof(float)));
    const long long ntile = (n + 15) / 16;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)std::min<long long>(ntile, 4096)), dim3(64),
                       0, ctx->st, mlp_dev(ctx), dx, (long long)n, dy);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, dy, (size_t)n * 4 * sizeof(float), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
  }
  if (dx) hipFree(dx);
  if (dy) hipFree(dy);
  return rc;
}

int gp_seed_net_loss_grad(gp_ctx* ctx, int64_t n, const float* x, const float* t, float* grad,
                          double* loss) {
  if (!ctx || n < 1 || !x || !t || !grad) return GP_ERR_ARG;
  if (!ctx->mlp_ready) {
    ctx->err = "the seed network has no parameters (gp_seed_net_set)";
    return GP_ERR_STATE;
  }
  float *dx = nullptr, *dt = nullptr;
  int rc = mlp_upload(ctx, n, x, t, &dx, &dt);
  if (!rc) {
    mlp_launch_grad(ctx, dx, dt, n);
    HIPCHK(hipGetLastError());
    std::vector<float> pd(MLP_NPAD);
    float l = 0.f;
    HIPCHK(hipMemcpyAsync(pd.data(), ctx->mlp_grad, MLP_NPAD * sizeof(float),
                          hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipMemcpyAsync(&l, ctx->mlp_loss, sizeof(float), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    mlp_unpad(pd.data(), grad);
    if (loss) *loss = (double)l / (double)n;
  }
  if (dx) hipFree(dx);
  if (dt) hipFree(dt);
  return rc;
}

int gp_seed_net_train(gp_ctx* ctx, int64_t n, const float* x, const float* t, int epochs,
                      int batch, float lr, double* loss) {
  if (!ctx || n < 1 || !x || !t || epochs < 1 || batch < 1 || !(lr > 0.f)) return GP_ERR_ARG;
  if (!ctx->mlp_ready) {
    ctx->err = "the seed network has no parameters (gp_seed_net_set)";
    return GP_ERR_STATE;
  }
  float *dx = nullptr, *dt = nullptr;
  int rc = mlp_upload(ctx, n, x, t, &dx, &dt);
  if (!rc) rc = mlp_train_device(ctx, dx, dt, n, epochs, batch, lr, loss);
  if (dx) hipFree(dx);
  if (dt) hipFree(dt);
  return rc;
}

int gp_seed_net_pairs(gp_ctx* ctx, int64_t n, uint64_t seed, float* x, float* t, int* nfev,
                      double* poses) {
  if (!ctx || n < 1) return GP_ERR_ARG;
  float *dx = nullptr, *dt = nullptr;
  int* dn = nullptr;
  double* dp = nullptr;
  int rc = GP_OK;
  if (hipMalloc((void**)&dx, (size_t)n * 16 * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&dt, (size_t)n * 4 * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&dn, (size_t)n * sizeof(int)) != hipSuccess ||
      (poses && hipMalloc((void**)&dp, (size_t)n * 20 * sizeof(double)) != hipSuccess)) {
    rc = GP_ERR_NOMEM;
  } else {
    hipLaunchKernelGGL(k_mlp_pairs, dim3((unsigned)std::min<long long>((n + 63) / 64, 8192)),
                       dim3(64), 0, ctx->st, (unsigned long long)seed, 0ll, (long long)n, dx, dt,
                       dn, dp);
    if (poses && hipMemcpyAsync(poses, dp, (size_t)n * 20 * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->st) != hipSuccess)
      rc = GP_ERR_HIP;
    std::vector<float> hx((size_t)n * 16), ht((size_t)n * 4);
    if (hipMemcpyAsync(hx.data(), dx, hx.size() * sizeof(float), hipMemcpyDeviceToHost,
                       ctx->st) != hipSuccess ||
        hipMemcpyAsync(ht.data(), dt, ht.size() * sizeof(float), hipMemcpyDeviceToHost,
                       ctx->st) != hipSuccess ||
        (nfev && hipMemcpyAsync(nfev, dn, (size_t)n * sizeof(int), hipMemcpyDeviceToHost,
                                ctx->st) != hipSuccess) ||
        hipStreamSynchronize(ctx->st) != hipSuccess) {
      ctx->err = "seed-net pair generation failed";
      rc = GP_ERR_HIP;
    } else if (!rc) {
      for (long long k = 0; k < n; ++k) {
        if (x)
          for (int c = 0; c < 12; ++c) x[k * 12 + c] = hx[(size_t)k * 16 + c];
        if (t)
          for (int c = 0; c < 4; ++c) t[k * 4 + c] = ht[(size_t)k * 4 + c];
      }
    }
  }
  if (dx) hipFree(dx);
  if (dt) hipFree(dt);
  if (dn) hipFree(dn);
  if (dp) hipFree(dp);
  return rc;
}

int gp_seed_net_fit(gp_ctx* ctx, int64_t npairs, uint64_t seed, int epochs, int batch,
                    float lr, gp_seed_report* rep) {
  if (!ctx || npairs < 16 || epochs < 1 || batch < 16 || !(lr > 0.f)) return GP_ERR_ARG;
  if (!ctx->mlp_ready) {
    ctx->err = "the seed network needs initial parameters (gp_seed_net_set)";
    return GP_ERR_STATE;
  }
  float *dx = nullptr, *dt = nullptr;
  int* dn = nullptr;
  int rc = GP_OK;
  hipEvent_t e0, e1, e2;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventCreate(&e2);
  if (hipMalloc((void**)&dx, (size_t)npairs * 16 * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&dt, (size_t)npairs * 4 * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&dn, (size_t)npairs * sizeof(int)) != hipSuccess) {
    ctx->err = "hipMalloc failed for the training pairs";
    rc = GP_ERR_NOMEM;
  } else {
    hipEventRecord(e0, ctx->st);
    hipLaunchKernelGGL(k_mlp_pairs,
                       dim3((unsigned)std::min<long long>((npairs + 63) / 64, 16384)), dim3(64),
                       0, ctx->st, (unsigned long long)seed, 0ll, (long long)npairs, dx, dt, dn,
                       nullptr);
    hipEventRecord(e1, ctx->st);
    double loss = 0.0;
    rc = mlp_train_device(ctx, dx, dt, npairs, epochs, batch, lr, &loss);
    hipEventRecord(e2, ctx->st);
    if (!rc && hipEventSynchronize(e2) == hipSuccess && rep) {
      float ms1 = 0.f, ms2 = 0.f;
      hipEventElapsedTime(&ms1, e0, e1);
      hipEventElapsedTime(&ms2, e1, e2);
      std::vector<int> hn((size_t)npairs);
      hipMemcpy(hn.data(), dn, hn.size() * sizeof(int), hipMemcpyDeviceToHost);
      double s = 0.0;
      long long mx = 0;
      for (int v : hn) { s += v; mx = std::max<long long>(mx, v); }
      std::memset(rep, 0, sizeof(*rep));
      rep->npairs = npairs;
      rep->loss = loss;
      rep->label_ms = ms1;
      rep->train_ms = ms2;
      rep->nfev_ref_mean = s / (double)npairs;
      rep->nfev_ref_max = mx;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipEventDestroy(e2);
  if (dx) hipFree(dx);
  if (dt) hipFree(dt);
  if (dn) hipFree(dn);
  return rc;
}

int gp_seed_net_eval(gp_ctx* ctx, int64_t npairs, uint64_t seed, gp_seed_report* rep) {
  if (!ctx || npairs < 1 || !rep) return GP_ERR_ARG;
  if (!ctx->mlp_ready) {
    ctx->err = "the seed network has no parameters (gp_seed_net_set)";
    return GP_ERR_STATE;
  }
  const size_t n = (size_t)npairs;
  int *nr = nullptr, *nn = nullptr;
  unsigned char *fr = nullptr, *fn = nullptr, *fb = nullptr;
  float* cd = nullptr;
  int rc = GP_OK;
  if (hipMalloc((void**)&nr, n * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&nn, n * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&fr, n) != hipSuccess || hipMalloc((void**)&fn, n) != hipSuccess ||
      hipMalloc((void**)&fb, n) != hipSuccess ||
      hipMalloc((void**)&cd, n * sizeof(float)) != hipSuccess) {
    ctx->err = "hipMalloc failed for the evaluation pairs";
    rc = GP_ERR_NOMEM;
  } else {
    SeedEvalOut o{nr, nn, fr, fn, fb, cd};
    hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)std::min<long long>((npairs + 63) / 64, 16384)),
                       dim3(64), 0, ctx->st, (unsigned long long)seed, (long long)npairs,
                       mlp_dev(ctx), ctx->gate, o);
    // the copies below run on the null stream: wait for the context's stream
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    std::vector<int> hr(n), hn(n);
    std::vector<unsigned char> hfr(n), hfn(n), hfb(n);
    std::vector<float> hcd(n);
    if (hipMemcpy(hr.data(), nr, n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hn.data(), nn, n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hfr.data(), fr, n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hfn.data(), fn, n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hfb.data(), fb, n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hcd.data(), cd, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
      ctx->err = "seed-net evaluation failed";
      rc = GP_ERR_HIP;
    } else {
      std::memset(rep, 0, sizeof(*rep));
      rep->npairs = npairs;
      double sr = 0.0, sn = 0.0, mind = 1.0;
      for (size_t k = 0; k < n; ++k) {
        sr += hr[k];
        sn += hn[k];
        rep->nfev_ref_max = std::max<int64_t>(rep->nfev_ref_max, hr[k]);
        rep->nfev_net_max = std::max<int64_t>(rep->nfev_net_max, hn[k]);
        rep->flags_ref += hfr[k];
        rep->flags_net += hfn[k];
        rep->flag_agree += (hfr[k] == hfn[k]) ? 1 : 0;
        rep->fallbacks += (hfb[k] & 1) ? 1 : 0;
        rep->no_seed += (hfb[k] & 2) ? 1 : 0;
        if (hfr[k] && hfn[k]) mind = std::min<double>(mind, hcd[k]);
      }
      rep->nfev_ref_mean = sr / (double)n;
      rep->nfev_net_mean = sn / (double)n;
      rep->cn_min_dot = mind;
    }
  }
  for (void* b : {(void*)nr, (void*)nn, (void*)fr, (void*)fn, (void*)fb, (void*)cd})
    if (b) hipFree(b);
  return rc;
}

// ---------------------------------------------------------------------------
// on-disk formats (host only; no context, no GPU)
// ---------------------------------------------------------------------------
int gp_part_file_write(const char* path, int64_t npart, int npoly, const int* npclass,
                       const double* rhop, const double* emaj, double apr, const double* cols) {
  if (!path || npart < 0 || npart > INT_MAX || npoly < 1 || !npclass || !rhop || !emaj ||
      (npart > 0 && !cols) || npart * 8 > gpio::kMaxRecord)
    return GP_ERR_ARG;
  gpio::File f(path, "wb");
  if (!f.f) return GP_ERR_IO;
  const int32_t n32 = (int32_t)npart;
  bool ok = gpio::rec_write(f.f, &n32, 4) &&
            gpio::rec_write(f.f, npclass, 4 * (int64_t)npoly) &&
            gpio::rec_write(f.f, rhop, 8 * (int64_t)npoly) &&
            gpio::rec_write(f.f, emaj, 8 * (int64_t)npoly) && gpio::rec_write(f.f, &apr, 8);
  for (int c = 0; ok && c < 14; ++c) ok = gpio::rec_write(f.f, cols + (size_t)c * npart, 8 * npart);
  ok = f.close() && ok;
  return ok ? GP_OK : GP_ERR_IO;
}

int gp_part_file_info(const char* path, int64_t* npart, int* npoly) {
  if (!path || !npart || !npoly) return GP_ERR_ARG;
  gpio::File f(path, "rb");
  if (!f.f) return GP_ERR_IO;
  int32_t n32 = 0;
  int64_t len = 0;
  if (!gpio::rec_read(f.f, &n32, 4) || n32 < 0 || !gpio::rec_peek(f.f, &len) || len < 4 ||
      len % 4)
    return GP_ERR_IO;
  *npart = n32;
  *npoly = (int)(len / 4);
  return GP_OK;
}

int gp_part_file_read(const char* path, int64_t npart, int npoly, int* npclass, double* rhop,
                      double* emaj, double* apr, double* cols) {
  if (!path || npart < 0 || npoly < 1 || !npclass || !rhop || !emaj || !apr ||
      (npart > 0 && !cols))
    return GP_ERR_ARG;
  gpio::File f(path, "rb");
  if (!f.f) return GP_ERR_IO;
  int32_t n32 = 0;
  bool ok = gpio::rec_read(f.f, &n32, 4) && n32 == npart &&
            gpio::rec_read(f.f, npclass, 4 * (int64_t)npoly) &&
            gpio::rec_read(f.f, rhop, 8 * (int64_t)npoly) &&
            gpio::rec_read(f.f, emaj, 8 * (int64_t)npoly) && gpio::rec_read(f.f, apr, 8);
  for (int c = 0; ok && c < 14; ++c) ok = gpio::rec_read(f.f, cols + (size_t)c * npart, 8 * npart);
  return ok ? GP_OK : GP_ERR_IO;
}

int gp_field_file_write(const char* path, int n, const double* f3) {
  if (!path || n < 1 || !f3) return GP_ERR_ARG;
  gpio::File f(path, "wb");
  if (!f.f) return GP_ERR_IO;
  const int32_t hdr[3] = {n, n, n};
  const size_t cnt = (size_t)n * n * n;
  bool ok = std::fwrite(hdr, 4, 3, f.f) == 3 && std::fwrite(f3, 8, cnt, f.f) == cnt;
  ok = f.close() && ok;
  return ok ? GP_OK : GP_ERR_IO;
}

int gp_field_file_read(const char* path, int n, double* f3) {
  if (!path || n < 1 || !f3) return GP_ERR_ARG;
  gpio::File f(path, "rb");
  if (!f.f) return GP_ERR_IO;
  int32_t hdr[3] = {0, 0, 0};
  const size_t cnt = (size_t)n * n * n;
  if (std::fread(hdr, 4, 3, f.f) != 3 || hdr[0] != n || hdr[1] != n || hdr[2] != n)
    return GP_ERR_IO;
  if (std::fread(f3, 8, cnt, f.f) != cnt) return GP_ERR_IO;
  char extra;
  return std::fread(&extra, 1, 1, f.f) == 0 ? GP_OK : GP_ERR_IO;   // nothing after the data
}

}  // extern "C"

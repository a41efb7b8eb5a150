// gp_dns.hip -- the spectral DNS field supplier on the GPU: the reference's
// pseudo-spectral Navier-Stokes step (ADV_FLUID, adv_fluid.f90:78-234) with
// its rotational-form nonlinear term (NONLINEAR_TERMS, nonlinear_terms.f90:
// 96-206), divergence projection (PROJ_DIVFREE, proj_divfree.f90:46-60),
// smooth dealiasing filter (MESHING, meshing.f90:151-180), STEADY forcing
// (ADV_FORCING / ADD_FORCING3) and CHECK_CFL, in double precision.
//
// Layout: physical arrays [3][n][n][n] (x fastest, index (k n + j) n + i);
// spectral arrays [3][n][n][n/2+1] (the halved x axis fastest, as the
// reference's FSTART(1):FEND(1)), complex interleaved.  Transforms are
// rocFFT real-to-complex / complex-to-real plans batched over the three
// components, unnormalised both ways as P3DFFT's; the forward result is
// multiplied by FACTOR = 1/n^3 inside the kernel that consumes it.
//
// One ADV_FLUID is seven batched transforms and five element-wise kernels:
//   u -> z (copy), R2C z -> U                       (entry transform)
//   k_scale_curl: U *= FACTOR, A = i k x U
//   C2R A -> z (vorticity), k_cross: z = u x z, EPS_FLU partials
//   R2C z -> A, k_stage1: RHS = F + P(A FACTOR); Un = (U + dt/2 RHS) E,
//       U = (U + dt RHS) E, B = U, A = i k x U
//   C2R B -> u (stage velocity), C2R A -> z, k_cross, R2C z -> A
//   k_stage2: RHS = F + P(A FACTOR); U = Un + dt/2 RHS, projection, FILTER,
//       B = U
//   C2R B -> u
// (E = exp(-VISC |k|^2 DTIME), P the projection of NONLINEAR_TERMS's RHS
// assembly; the complex-to-real inputs are copies because the transform may
// overwrite its input.)
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/gazpart_gpu.h"

namespace gpdns {

constexpr int BLOCK = 256;
constexpr double ZERO_R = 1.0e-33;      // mod_dns_dim.f90:136
constexpr double INF_R = 1.0e33;        // mod_dns_dim.f90:137 INFINITY
constexpr double RTRUNC = 8.0 / 9.0;    // meshing.f90:38
constexpr int RED_BLOCKS = 1024;

struct Grid {
  int n, nh;            // n, n/2 + 1
  long long nr, nc;     // n^3, n^2 (n/2 + 1)
  double kf;            // KFIRST = 2 pi / L (init_run.f90:128)
  double klast;         // KLAST = pi / (L / n) (init_run.f90:129)
};

// wavenumbers of spectral element e (meshing.f90:96-121)
__device__ __forceinline__ void wavenum(const Grid& g, long long e, double& kx, double& ky,
                                        double& kz) {
  const int i = (int)(e % g.nh);
  const long long r = e / g.nh;
  const int j = (int)(r % g.n);
  const int k = (int)(r / g.n);
  kx = (double)i * g.kf;
  ky = (double)(j <= g.n / 2 ? j : j - g.n) * g.kf;
  kz = (double)(k <= g.n / 2 ? k : k - g.n) * g.kf;
}

// i k x U as nonlinear_terms.f90:104-112 forms it, term by term
__device__ __forceinline__ void curl(double kx, double ky, double kz, double2 u, double2 v,
                                     double2 w, double2& ox, double2& oy, double2& oz) {
  // (i a) k = (-a.y k, a.x k)
  ox.x = (-w.y * ky) - (-v.y * kz); ox.y = (w.x * ky) - (v.x * kz);
  oy.x = (-u.y * kz) - (-w.y * kx); oy.y = (u.x * kz) - (w.x * kx);
  oz.x = (-v.y * kx) - (-u.y * ky); oz.y = (v.x * kx) - (u.x * ky);
}

// RHS = F + P(T) as nonlinear_terms.f90:187-200 accumulates it
__device__ __forceinline__ void project_rhs(double kx, double ky, double kz, double2 t0,
                                            double2 t1, double2 t2, double2 f0, double2 f1,
                                            double2 f2, double2& r0, double2& r1, double2& r2) {
  double k2 = kx * kx + ky * ky + kz * kz;
  if (k2 <= ZERO_R) k2 = INF_R;
  const double a00 = (ky * ky + kz * kz) / k2, a01 = kx * ky / k2, a02 = kx * kz / k2;
  const double a10 = ky * kx / k2, a11 = (kx * kx + kz * kz) / k2, a12 = ky * kz / k2;
  const double a20 = kz * kx / k2, a21 = kz * ky / k2, a22 = (kx * kx + ky * ky) / k2;
  r0.x = f0.x + a00 * t0.x - a01 * t1.x - a02 * t2.x;
  r0.y = f0.y + a00 * t0.y - a01 * t1.y - a02 * t2.y;
  r1.x = f1.x - a10 * t0.x + a11 * t1.x - a12 * t2.x;
  r1.y = f1.y - a10 * t0.y + a11 * t1.y - a12 * t2.y;
  r2.x = f2.x - a20 * t0.x - a21 * t1.x + a22 * t2.x;
  r2.y = f2.y - a20 * t0.y - a21 * t1.y + a22 * t2.y;
}

__device__ __forceinline__ double2 cmul(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}

// the forcing of mode e (fmap: -1 or its index among nf modes)
__device__ __forceinline__ void forcing_at(const int* fmap, const double2* fc, int nf, long long e,
                                           double2& f0, double2& f1, double2& f2) {
  f0 = f1 = f2 = make_double2(0.0, 0.0);
  if (!fmap) return;
  const int m = fmap[e];
  if (m < 0) return;
  f0 = fc[m]; f1 = fc[nf + m]; f2 = fc[2 * nf + m];
}

__global__ __launch_bounds__(BLOCK) void k_scale_curl(Grid g, double2* __restrict__ U,
                                                      double2* __restrict__ A, double factor) {
  for (long long e = (long long)blockIdx.x * BLOCK + threadIdx.x; e < g.nc;
       e += (long long)gridDim.x * BLOCK) {
    double kx, ky, kz;
    wavenum(g, e, kx, ky, kz);
    const double2 u = cmul(U[e], factor), v = cmul(U[g.nc + e], factor),
                  w = cmul(U[2 * g.nc + e], factor);
    U[e] = u; U[g.nc + e] = v; U[2 * g.nc + e] = w;
    double2 ox, oy, oz;
    curl(kx, ky, kz, u, v, w, ox, oy, oz);
    A[e] = ox; A[g.nc + e] = oy; A[2 * g.nc + e] = oz;
  }
}

// z = u x z (nonlinear_terms.f90:131-136); with part != nullptr also the
// sum of z^2 per workgroup (EPS_FLU, :124-128)
__global__ __launch_bounds__(BLOCK) void k_cross(Grid g, const double* __restrict__ u,
                                                 double* __restrict__ z, double* part) {
  __shared__ double s[BLOCK];
  double acc = 0.0;
  for (long long p = (long long)blockIdx.x * BLOCK + threadIdx.x; p < g.nr;
       p += (long long)gridDim.x * BLOCK) {
    const double ux = u[p], uy = u[g.nr + p], uz = u[2 * g.nr + p];
    const double zx = z[p], zy = z[g.nr + p], zz = z[2 * g.nr + p];
    if (part) acc += zx * zx + zy * zy + zz * zz;
    z[p] = uy * zz - uz * zy;
    z[g.nr + p] = uz * zx - ux * zz;
    z[2 * g.nr + p] = ux * zy - uy * zx;
  }
  if (!part) return;
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// first stage (adv_fluid.f90:113-131) and the curl of its result for the
// second stage's nonlinear term
__global__ __launch_bounds__(BLOCK) void k_stage1(Grid g, double2* __restrict__ A,
                                                  double2* __restrict__ U, double2* __restrict__ Un,
                                                  double2* __restrict__ B, double factor,
                                                  double visc, double dt, const int* fmap,
                                                  const double2* fc, int nf) {
  for (long long e = (long long)blockIdx.x * BLOCK + threadIdx.x; e < g.nc;
       e += (long long)gridDim.x * BLOCK) {
    double kx, ky, kz;
    wavenum(g, e, kx, ky, kz);
    double2 f0, f1, f2, r0, r1, r2;
    forcing_at(fmap, fc, nf, e, f0, f1, f2);
    project_rhs(kx, ky, kz, cmul(A[e], factor), cmul(A[g.nc + e], factor),
                cmul(A[2 * g.nc + e], factor), f0, f1, f2, r0, r1, r2);
    const double ef = exp(-visc * (kx * kx + ky * ky + kz * kz) * dt);
    const double h = dt / 2;
    const double2 rr[3] = {r0, r1, r2};
    double2 un[3];
    for (int c = 0; c < 3; ++c) {
      const double2 u0 = U[c * g.nc + e];
      Un[c * g.nc + e] = cmul(cadd(u0, cmul(rr[c], h)), ef);
      un[c] = cmul(cadd(u0, cmul(rr[c], dt)), ef);
      U[c * g.nc + e] = un[c];
      B[c * g.nc + e] = un[c];
    }
    double2 ox, oy, oz;
    curl(kx, ky, kz, un[0], un[1], un[2], ox, oy, oz);
    A[e] = ox; A[g.nc + e] = oy; A[2 * g.nc + e] = oz;
  }
}

// second stage (adv_fluid.f90:166-190): U = Un + dt/2 RHS, PROJ_DIVFREE,
// FILTER; B = U for the backward transform
__global__ __launch_bounds__(BLOCK) void k_stage2(Grid g, const double2* __restrict__ A,
                                                  double2* __restrict__ U,
                                                  const double2* __restrict__ Un,
                                                  double2* __restrict__ B, double factor,
                                                  double dt, const int* fmap, const double2* fc,
                                                  int nf) {
  for (long long e = (long long)blockIdx.x * BLOCK + threadIdx.x; e < g.nc;
       e += (long long)gridDim.x * BLOCK) {
    double kx, ky, kz;
    wavenum(g, e, kx, ky, kz);
    double2 f0, f1, f2, r0, r1, r2;
    forcing_at(fmap, fc, nf, e, f0, f1, f2);
    project_rhs(kx, ky, kz, cmul(A[e], factor), cmul(A[g.nc + e], factor),
                cmul(A[2 * g.nc + e], factor), f0, f1, f2, r0, r1, r2);
    const double h = dt / 2;
    double2 u = cadd(Un[e], cmul(r0, h)), v = cadd(Un[g.nc + e], cmul(r1, h)),
            w = cadd(Un[2 * g.nc + e], cmul(r2, h));
    // PROJ_DIVFREE
    double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 <= ZERO_R) k2 = INF_R;
    const double2 dv = make_double2(kx * u.x + ky * v.x + kz * w.x, kx * u.y + ky * v.y + kz * w.y);
    u.x -= kx / k2 * dv.x; u.y -= kx / k2 * dv.y;
    v.x -= ky / k2 * dv.x; v.y -= ky / k2 * dv.y;
    w.x -= kz / k2 * dv.x; w.y -= kz / k2 * dv.y;
    // FILTER, ALIASCONTROL = 2
    const double k2s = (kx / g.klast) * (kx / g.klast) + (ky / g.klast) * (ky / g.klast) +
                       (kz / g.klast) * (kz / g.klast);
    double fl = 1.0;
    if (k2s > RTRUNC * RTRUNC && k2s <= 1.0) {
      const double c = cos((M_PI / 2.0) * (sqrt(k2s) - RTRUNC) / (1.0 - RTRUNC));
      fl = (c * c) * (c * c);
    } else if (k2s > 1.0) {
      fl = 0.0;
    }
    u = cmul(u, fl); v = cmul(v, fl); w = cmul(w, fl);
    U[e] = u; U[g.nc + e] = v; U[2 * g.nc + e] = w;
    B[e] = u; B[g.nc + e] = v; B[2 * g.nc + e] = w;
  }
}

__global__ __launch_bounds__(BLOCK) void k_curl(Grid g, const double2* __restrict__ U,
                                                double2* __restrict__ A) {
  for (long long e = (long long)blockIdx.x * BLOCK + threadIdx.x; e < g.nc;
       e += (long long)gridDim.x * BLOCK) {
    double kx, ky, kz;
    wavenum(g, e, kx, ky, kz);
    double2 ox, oy, oz;
    curl(kx, ky, kz, U[e], U[g.nc + e], U[2 * g.nc + e], ox, oy, oz);
    A[e] = ox; A[g.nc + e] = oy; A[2 * g.nc + e] = oz;
  }
}

// per-workgroup max |a| of the three components
__global__ __launch_bounds__(BLOCK) void k_absmax3(Grid g, const double* __restrict__ u,
                                                   double* part) {
  __shared__ double s[3][BLOCK];
  double m[3] = {0.0, 0.0, 0.0};
  for (long long p = (long long)blockIdx.x * BLOCK + threadIdx.x; p < g.nr;
       p += (long long)gridDim.x * BLOCK)
    for (int c = 0; c < 3; ++c) m[c] = fmax(m[c], fabs(u[c * g.nr + p]));
  for (int c = 0; c < 3; ++c) s[c][threadIdx.x] = m[c];
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < 3; ++c) s[c][threadIdx.x] = fmax(s[c][threadIdx.x], s[c][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) part[3 * blockIdx.x + c] = s[c][0];
}

// ---------------------------------------------------------------------------
// the reference's RAN1 (Park-Miller minimal standard, Schrage's method,
// Bays-Durham shuffle of 32) and GAUSS (polar Box-Muller, one value kept),
// adv_forcing.f90:130-215, as a host-side generator
// ---------------------------------------------------------------------------
struct Ran1 {
  static constexpr int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, NTAB = 32;
  int iv[NTAB] = {};
  int iy = 0;
  int iset = 0;
  double gset = 0.0;
  int idum = 0;
  static int lcg(int x) {
    const int k = x / IQ;
    x = IA * (x - k * IQ) - IR * k;
    return x < 0 ? x + IM : x;
  }
  double ran1() {
    const int ndiv = 1 + (IM - 1) / NTAB;
    const double am = 1.0 / IM, rnmx = 1.0 - 1.2e-7;
    if (idum <= 0 || iy == 0) {
      idum = std::max(-idum, 1);
      for (int j = NTAB + 8; j >= 1; --j) {
        idum = lcg(idum);
        if (j <= NTAB) iv[j - 1] = idum;
      }
      iy = iv[0];
    }
    idum = lcg(idum);
    const int j = iy / ndiv;
    iy = iv[j];
    iv[j] = idum;
    return std::min(am * iy, rnmx);
  }
  double gauss() {
    if (iset != 0) {
      iset = 0;
      return gset;
    }
    double v1, v2, rsq;
    do {
      v1 = 2.0 * ran1() - 1.0;
      v2 = 2.0 * ran1() - 1.0;
      rsq = v1 * v1 + v2 * v2;
    } while (!(rsq < 1.0 && rsq != 0.0));
    const double fac = std::sqrt(-2.0 * std::log(rsq) / rsq);
    gset = v1 * fac;
    iset = 1;
    return v2 * fac;
  }
};

}  // namespace gpdns

using namespace gpdns;

struct gp_dns {
  gp_dns_params prm{};
  Grid g{};
  hipStream_t st = nullptr;
  double* u = nullptr;     // [3][nr] velocity
  double* z = nullptr;     // [3][nr] vorticity / u x omega / copies
  double2* U = nullptr;    // [3][nc]
  double2* Un = nullptr;   // [3][nc] first-stage U_NP1
  double2* A = nullptr;    // [3][nc] curl / transformed products
  double2* B = nullptr;    // [3][nc] backward-transform inputs
  double* part = nullptr;  // reduction partials
  rocfft_plan fwd = nullptr, bwd = nullptr;
  rocfft_execution_info info = nullptr;
  void* work = nullptr;
  bool have_u = false;
  double eps = 0.0;
  // forcing: modes in the reference's loop order, amplitudes (host), the
  // dense mode map and the amplitudes on the device
  std::vector<long long> fmode;
  std::vector<double> fkf;
  std::vector<double> famp;    // [3][nf][2]
  int* fmap = nullptr;
  double2* fdev = nullptr;
  Ran1 rng;
  std::string err;
};

namespace {

#define DCHK(call)                                                        \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      d->err = std::string(#call " failed: ") + hipGetErrorString(e_);    \
      return GP_ERR_HIP;                                                  \
    }                                                                     \
  } while (0)

#define FCHK(call)                                                        \
  do {                                                                    \
    rocfft_status s_ = (call);                                            \
    if (s_ != rocfft_status_success) {                                    \
      d->err = std::string(#call " failed with rocfft status ") +        \
               std::to_string((int)s_);                                   \
      return GP_ERR_HIP;                                                  \
    }                                                                     \
  } while (0)

unsigned grid_of(long long n) {
  const long long b = (n + BLOCK - 1) / BLOCK;
  return (unsigned)std::min<long long>(b, 65536);
}

int r2c(gp_dns* d, double* in, double2* out) {
  void* ib[1] = {in};
  void* ob[1] = {out};
  FCHK(rocfft_execute(d->fwd, ib, ob, d->info));
  return GP_OK;
}

int c2r(gp_dns* d, double2* in, double* out) {
  void* ib[1] = {in};
  void* ob[1] = {out};
  FCHK(rocfft_execute(d->bwd, ib, ob, d->info));
  return GP_OK;
}

// ADV_FORCING (adv_forcing.f90:33-125) on the host, amplitudes to the device
int advance_forcing(gp_dns* d) {
  const int nf = (int)d->fmode.size();
  if (!d->prm.forcing || nf == 0) return GP_OK;
  const double dt = d->prm.dtime;
  const double alpha = dt / d->prm.time_force;
  const double beta = std::pow(2.0 * d->prm.sigma_force * d->prm.sigma_force * dt /
                                   d->prm.time_force, 0.5);
  const double kmax = d->prm.kforce_max;
  for (int m = 0; m < nf; ++m) {
    const double xampf = (kmax - d->fkf[m]) / 0.20 / kmax;
    const double xampl = (std::exp(2.0 * xampf) - 1.0) / (std::exp(2.0 * xampf) + 1.0);
    for (int c = 0; c < 3; ++c) {
      double* a = d->famp.data() + ((size_t)c * nf + m) * 2;
      double xe = d->rng.gauss() * xampl;
      const double re = a[0] * (1.0 - alpha) + xe * beta;
      xe = d->rng.gauss() * xampl;
      const double im = a[1] * (1.0 - alpha) + xe * beta;
      a[0] = re;
      a[1] = im;
    }
  }
  DCHK(hipMemcpyAsync(d->fdev, d->famp.data(), d->famp.size() * sizeof(double),
                      hipMemcpyHostToDevice, d->st));
  return GP_OK;
}

int one_step(gp_dns* d) {
  const Grid& g = d->g;
  const double factor = 1.0 / (double)g.nr;
  const unsigned gc = grid_of(g.nc), gr = grid_of(g.nr);
  const int nf = (int)d->fmode.size();
  const int* fmap = (d->prm.forcing && nf) ? d->fmap : nullptr;
  int rc;
  if (d->prm.reuse_spectral) {
    // U is the spectral state u came from (set_velocity or the last step)
    if ((rc = advance_forcing(d))) return rc;
    hipLaunchKernelGGL(k_curl, dim3(gc), dim3(BLOCK), 0, d->st, g, (const double2*)d->U, d->A);
  } else {
    // entry transform of the velocity (adv_fluid.f90:86-94); u is kept
    DCHK(hipMemcpyAsync(d->z, d->u, 3 * g.nr * sizeof(double), hipMemcpyDeviceToDevice, d->st));
    if ((rc = r2c(d, d->z, d->U))) return rc;
    if ((rc = advance_forcing(d))) return rc;
    hipLaunchKernelGGL(k_scale_curl, dim3(gc), dim3(BLOCK), 0, d->st, g, d->U, d->A, factor);
  }
  // stage 1
  if ((rc = c2r(d, d->A, d->z))) return rc;
  hipLaunchKernelGGL(k_cross, dim3(RED_BLOCKS), dim3(BLOCK), 0, d->st, g, (const double*)d->u,
                     d->z, d->part);
  if ((rc = r2c(d, d->z, d->A))) return rc;
  hipLaunchKernelGGL(k_stage1, dim3(gc), dim3(BLOCK), 0, d->st, g, d->A, d->U, d->Un, d->B,
                     factor, d->prm.visc, d->prm.dtime, fmap, (const double2*)d->fdev, nf);
  // stage 2
  if ((rc = c2r(d, d->B, d->u))) return rc;
  if ((rc = c2r(d, d->A, d->z))) return rc;
  hipLaunchKernelGGL(k_cross, dim3(gr), dim3(BLOCK), 0, d->st, g, (const double*)d->u, d->z,
                     (double*)nullptr);
  if ((rc = r2c(d, d->z, d->A))) return rc;
  hipLaunchKernelGGL(k_stage2, dim3(gc), dim3(BLOCK), 0, d->st, g, (const double2*)d->A, d->U,
                     (const double2*)d->Un, d->B, factor, d->prm.dtime, fmap,
                     (const double2*)d->fdev, nf);
  if ((rc = c2r(d, d->B, d->u))) return rc;
  DCHK(hipGetLastError());
  // EPS_FLU = VISC * sum(zeta^2) of the first stage
  std::vector<double> h(RED_BLOCKS);
  DCHK(hipMemcpyAsync(h.data(), d->part, RED_BLOCKS * sizeof(double), hipMemcpyDeviceToHost,
                      d->st));
  DCHK(hipStreamSynchronize(d->st));
  double s = 0.0;
  for (double v : h) s += v;
  d->eps = d->prm.visc * s;
  return GP_OK;
}

}  // namespace

extern "C" {

int gp_dns_create(const gp_dns_params* p, gp_dns** out) {
  if (!p || !out) return GP_ERR_ARG;
  *out = nullptr;
  if (p->n < 4 || p->n % 2 || p->n > 2048 || !(p->L > 0) || !(p->visc >= 0) ||
      !(p->dtime > 0))
    return GP_ERR_ARG;
  if (p->forcing && !(p->time_force > 0 && p->kforce_max > 0)) return GP_ERR_ARG;
  gp_dns* d = new gp_dns();
  d->prm = *p;
  Grid& g = d->g;
  g.n = p->n;
  g.nh = p->n / 2 + 1;
  g.nr = (long long)p->n * p->n * p->n;
  g.nc = (long long)p->n * p->n * g.nh;
  g.kf = 2.0 * M_PI / p->L;
  g.klast = M_PI / (p->L / p->n);
  auto fail = [&](int rc) {
    gp_dns_destroy(d);
    return rc;
  };
  if (hipStreamCreate(&d->st) != hipSuccess) return fail(GP_ERR_HIP);
  const size_t rb = 3 * (size_t)g.nr * sizeof(double), cb = 3 * (size_t)g.nc * sizeof(double2);
  if (hipMalloc((void**)&d->u, rb) != hipSuccess || hipMalloc((void**)&d->z, rb) != hipSuccess ||
      hipMalloc((void**)&d->U, cb) != hipSuccess || hipMalloc((void**)&d->Un, cb) != hipSuccess ||
      hipMalloc((void**)&d->A, cb) != hipSuccess || hipMalloc((void**)&d->B, cb) != hipSuccess ||
      hipMalloc((void**)&d->part, 3 * RED_BLOCKS * sizeof(double)) != hipSuccess)
    return fail(GP_ERR_NOMEM);
  if (rocfft_setup() != rocfft_status_success) return fail(GP_ERR_HIP);
  const size_t len[3] = {(size_t)g.n, (size_t)g.n, (size_t)g.n};
  if (rocfft_plan_create(&d->fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                         rocfft_precision_double, 3, len, 3, nullptr) != rocfft_status_success ||
      rocfft_plan_create(&d->bwd, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                         rocfft_precision_double, 3, len, 3, nullptr) != rocfft_status_success)
    return fail(GP_ERR_HIP);
  size_t w1 = 0, w2 = 0;
  rocfft_plan_get_work_buffer_size(d->fwd, &w1);
  rocfft_plan_get_work_buffer_size(d->bwd, &w2);
  const size_t wb = std::max(w1, w2);
  if (rocfft_execution_info_create(&d->info) != rocfft_status_success) return fail(GP_ERR_HIP);
  if (wb) {
    if (hipMalloc(&d->work, wb) != hipSuccess) return fail(GP_ERR_NOMEM);
    if (rocfft_execution_info_set_work_buffer(d->info, d->work, wb) != rocfft_status_success)
      return fail(GP_ERR_HIP);
  }
  if (rocfft_execution_info_set_stream(d->info, d->st) != rocfft_status_success)
    return fail(GP_ERR_HIP);
  // forcing modes in the order of adv_forcing.f90:45-62 (K, J, I)
  if (p->forcing) {
    std::vector<int> map((size_t)g.nc, -1);
    for (int k = 0; k < g.n; ++k)
      for (int j = 0; j < g.n; ++j)
        for (int i = 0; i < g.nh; ++i) {
          const double kx = i * g.kf, ky = (j <= g.n / 2 ? j : j - g.n) * g.kf,
                       kz = (k <= g.n / 2 ? k : k - g.n) * g.kf;
          const double kfm = std::pow(kx * kx + ky * ky + kz * kz, 0.5);
          if (kfm >= p->kforce_min && kfm <= p->kforce_max) {
            const long long e = ((long long)k * g.n + j) * g.nh + i;
            map[(size_t)e] = (int)d->fmode.size();
            d->fmode.push_back(e);
            d->fkf.push_back(kfm);
          }
        }
    const size_t nf = d->fmode.size();
    d->famp.assign(3 * nf * 2, 0.0);   // INITIATION_FORCING_NEW: F = 0
    d->rng.idum = p->seed;
    if (hipMalloc((void**)&d->fmap, (size_t)g.nc * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&d->fdev, std::max<size_t>(3 * nf, 1) * sizeof(double2)) != hipSuccess)
      return fail(GP_ERR_NOMEM);
    if (hipMemcpy(d->fmap, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemset(d->fdev, 0, std::max<size_t>(3 * nf, 1) * sizeof(double2)) != hipSuccess)
      return fail(GP_ERR_HIP);
  }
  *out = d;
  return GP_OK;
}

int gp_dns_destroy(gp_dns* d) {
  if (!d) return GP_ERR_ARG;
  if (d->st) hipStreamSynchronize(d->st);
  if (d->fwd) rocfft_plan_destroy(d->fwd);
  if (d->bwd) rocfft_plan_destroy(d->bwd);
  if (d->info) rocfft_execution_info_destroy(d->info);
  void* bufs[] = {d->u, d->z, d->U, d->Un, d->A, d->B, d->part, d->work, d->fmap, d->fdev};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (d->st) hipStreamDestroy(d->st);
  delete d;
  return GP_OK;
}

const char* gp_dns_error_string(const gp_dns* d) { return d ? d->err.c_str() : "null handle"; }

int gp_dns_set_velocity(gp_dns* d, const double* uvw) {
  if (!d || !uvw) return GP_ERR_ARG;
  DCHK(hipMemcpyAsync(d->u, uvw, 3 * d->g.nr * sizeof(double), hipMemcpyHostToDevice, d->st));
  // the spectral state of this velocity (for gp_dns_get_vorticity / bind)
  DCHK(hipMemcpyAsync(d->z, d->u, 3 * d->g.nr * sizeof(double), hipMemcpyDeviceToDevice, d->st));
  int rc;
  if ((rc = r2c(d, d->z, d->U))) return rc;
  hipLaunchKernelGGL(k_scale_curl, dim3(grid_of(d->g.nc)), dim3(BLOCK), 0, d->st, d->g, d->U,
                     d->A, 1.0 / (double)d->g.nr);
  DCHK(hipGetLastError());
  DCHK(hipStreamSynchronize(d->st));
  d->have_u = true;
  return GP_OK;
}

int gp_dns_get_velocity(gp_dns* d, double* uvw) {
  if (!d || !uvw) return GP_ERR_ARG;
  if (!d->have_u) {
    d->err = "no velocity (gp_dns_set_velocity)";
    return GP_ERR_STATE;
  }
  DCHK(hipMemcpyAsync(uvw, d->u, 3 * d->g.nr * sizeof(double), hipMemcpyDeviceToHost, d->st));
  DCHK(hipStreamSynchronize(d->st));
  return GP_OK;
}

static int vorticity_dev(gp_dns* d) {
  // i k x U -> z (z is free between steps)
  hipLaunchKernelGGL(k_curl, dim3(grid_of(d->g.nc)), dim3(BLOCK), 0, d->st, d->g,
                     (const double2*)d->U, d->A);
  DCHK(hipGetLastError());
  return c2r(d, d->A, d->z);
}

int gp_dns_get_vorticity(gp_dns* d, double* w) {
  if (!d || !w) return GP_ERR_ARG;
  if (!d->have_u) {
    d->err = "no velocity (gp_dns_set_velocity)";
    return GP_ERR_STATE;
  }
  int rc;
  if ((rc = vorticity_dev(d))) return rc;
  DCHK(hipMemcpyAsync(w, d->z, 3 * d->g.nr * sizeof(double), hipMemcpyDeviceToHost, d->st));
  DCHK(hipStreamSynchronize(d->st));
  return GP_OK;
}

int gp_dns_step(gp_dns* d, int nsteps) {
  if (!d || nsteps < 0) return GP_ERR_ARG;
  if (!d->have_u) {
    d->err = "no velocity (gp_dns_set_velocity)";
    return GP_ERR_STATE;
  }
  for (int s = 0; s < nsteps; ++s) {
    const int rc = one_step(d);
    if (rc) return rc;
  }
  return GP_OK;
}

static int absmax(gp_dns* d, double m[3]) {
  hipLaunchKernelGGL(k_absmax3, dim3(RED_BLOCKS), dim3(BLOCK), 0, d->st, d->g,
                     (const double*)d->u, d->part);
  DCHK(hipGetLastError());
  std::vector<double> h(3 * RED_BLOCKS);
  DCHK(hipMemcpyAsync(h.data(), d->part, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                      d->st));
  DCHK(hipStreamSynchronize(d->st));
  m[0] = m[1] = m[2] = 0.0;
  for (int b = 0; b < RED_BLOCKS; ++b)
    for (int c = 0; c < 3; ++c) m[c] = std::max(m[c], h[3 * b + c]);
  return GP_OK;
}

int gp_dns_cfl(gp_dns* d, double dtime_user, double* dtime) {
  if (!d) return GP_ERR_ARG;
  if (!d->have_u) {
    d->err = "no velocity (gp_dns_set_velocity)";
    return GP_ERR_STATE;
  }
  double m[3];
  int rc;
  if ((rc = absmax(d, m))) return rc;
  // check_cfl.f90:61-92 with DX = DY = DZ = L / (n - 1) (meshing.f90:45-47)
  const double dx = d->prm.L / (double)(d->g.n - 1);
  double dtc = INF_R;
  for (int c = 0; c < 3; ++c) dtc = std::min(dtc, m[c] > ZERO_R ? 0.05 * dx / m[c] : INF_R);
  const double dtf = std::min(INF_R, 1.0 * dx * dx / d->prm.visc);
  d->prm.dtime = std::min(std::min(dtc, dtf), dtime_user);
  if (dtime) *dtime = d->prm.dtime;
  return GP_OK;
}

int gp_dns_info(gp_dns* d, double* eps, double* umax3, int* nforce) {
  if (!d) return GP_ERR_ARG;
  if (eps) *eps = d->eps;
  if (nforce) *nforce = (int)d->fmode.size();
  if (umax3) {
    if (!d->have_u) {
      d->err = "no velocity (gp_dns_set_velocity)";
      return GP_ERR_STATE;
    }
    return absmax(d, umax3);
  }
  return GP_OK;
}

int gp_dns_bind(gp_dns* d, gp_ctx* ctx, int mesh_vorticity) {
  if (!d || !ctx) return GP_ERR_ARG;
  if (!d->have_u) {
    d->err = "no velocity (gp_dns_set_velocity)";
    return GP_ERR_STATE;
  }
  int rc;
  if ((rc = vorticity_dev(d))) return rc;
  const double vs = mesh_vorticity ? (double)(d->g.n - 1) / (double)d->g.n : 1.0;
  DCHK(hipStreamSynchronize(d->st));
  const int rc2 = gp_bind_field_planar(ctx, d->g.n, d->u, d->z, vs);
  if (rc2) d->err = gp_last_error_string(ctx);
  return rc2;
}

}  // extern "C"

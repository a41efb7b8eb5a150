// gp_contact.h -- FP64 ellipsoid contact detection for one candidate pair
// per thread: the common-normal Levenberg-Marquardt solve of
// ellipsoid_contact_detection.f90 (two unknowns, the azimuth and elevation
// of the normal; three residuals, the squared components of the gap between
// the two support points), with MINPACK's lmder/lmpar/qrfac/qrsolv/enorm
// (minpack_reduc.f90:180-1289) specialised to m = 3, n = 2 and held in
// registers.  The iteration logic (trust region, step acceptance, the seven
// termination tests, maxfev = 300) is the reference's; only the algebra of
// the support function is rearranged for the GPU (E symmetric, so
// E c c^T E = t t^T with t = E c).
#pragma once

#include "gp_device.h"

namespace gp {

struct Sym3 { double xx, xy, xz, yy, yz, zz; };

// E = R diag(a^2, b^2, c^2) R^T (calculate_ellipsoid_matrix.f90:33-71)
__device__ __forceinline__ Sym3 ell_matrix(const double q[4], double a, double b, double c) {
  Rot<double> R(q[0], q[1], q[2], q[3]);
  double a2 = a * a, b2 = b * b, c2 = c * c;
  Sym3 e;
  e.xx = R.r00 * R.r00 * a2 + R.r01 * R.r01 * b2 + R.r02 * R.r02 * c2;
  e.xy = R.r00 * R.r10 * a2 + R.r01 * R.r11 * b2 + R.r02 * R.r12 * c2;
  e.xz = R.r00 * R.r20 * a2 + R.r01 * R.r21 * b2 + R.r02 * R.r22 * c2;
  e.yy = R.r10 * R.r10 * a2 + R.r11 * R.r11 * b2 + R.r12 * R.r12 * c2;
  e.yz = R.r10 * R.r20 * a2 + R.r11 * R.r21 * b2 + R.r12 * R.r22 * c2;
  e.zz = R.r20 * R.r20 * a2 + R.r21 * R.r21 * b2 + R.r22 * R.r22 * c2;
  return e;
}

__device__ __forceinline__ void symv(const Sym3& e, const double v[3], double o[3]) {
  o[0] = e.xx * v[0] + e.xy * v[1] + e.xz * v[2];
  o[1] = e.xy * v[0] + e.yy * v[1] + e.yz * v[2];
  o[2] = e.xz * v[0] + e.yz * v[1] + e.zz * v[2];
}

// Perram-Wertheim separation test.  For ellipsoids x^T E^-1 x <= 1 about
// centres b1, b2 and r = b2 - b1,
//   F(l) = l (1 - l) r^T [(1 - l) E1 + l E2]^-1 r,   0 < l < 1,
// is concave with max F = s^2, s the common scale factor at which the two
// just touch: max F > 1 <=> disjoint.  Any l with F(l) > 1 + 1e-6 therefore
// proves every point of ellipsoid 2 to lie outside ellipsoid 1 by that
// margin, so the contact flag of the common-normal solve (a support point
// of one ellipsoid strictly inside the other) is false -- the solve can be
// skipped without changing any result.  Golden-section search on l with an
// early exit; false means "not proven disjoint" (the solve then decides).
__device__ __forceinline__ double pw_quad(const Sym3& A, const Sym3& B, double l,
                                          const double r[3], double& det) {
  const double m = 1.0 - l;
  const double xx = m * A.xx + l * B.xx, xy = m * A.xy + l * B.xy, xz = m * A.xz + l * B.xz;
  const double yy = m * A.yy + l * B.yy, yz = m * A.yz + l * B.yz, zz = m * A.zz + l * B.zz;
  const double c00 = yy * zz - yz * yz, c01 = xz * yz - xy * zz, c02 = xy * yz - xz * yy;
  const double c11 = xx * zz - xz * xz, c12 = xy * xz - xx * yz, c22 = xx * yy - xy * xy;
  det = xx * c00 + xy * c01 + xz * c02;
  // r^T adj(M) r
  return r[0] * r[0] * c00 + r[1] * r[1] * c11 + r[2] * r[2] * c22 +
         2.0 * (r[0] * r[1] * c01 + r[0] * r[2] * c02 + r[1] * r[2] * c12);
}

__device__ __forceinline__ bool pw_disjoint(const Sym3& A, const Sym3& B, const double r[3]) {
  constexpr double kMargin = 1.0 + 1.0e-6;
  constexpr double kInvPhi = 0.6180339887498949;
  // F(l) > margin  <=>  l (1-l) q > margin det   (det > 0: M is SPD)
  auto excess = [&](double l) {
    double det;
    const double q = pw_quad(A, B, l, r, det);
    return l * (1.0 - l) * q / det;
  };
  double lo = 0.0, hi = 1.0;
  double l1 = hi - kInvPhi * (hi - lo), l2 = lo + kInvPhi * (hi - lo);
  double f1 = excess(l1), f2 = excess(l2);
  for (int it = 0; it < 16; ++it) {
    if (f1 > kMargin || f2 > kMargin) return true;
    if (f1 < f2) {
      lo = l1; l1 = l2; f1 = f2;
      l2 = lo + kInvPhi * (hi - lo);
      f2 = excess(l2);
    } else {
      hi = l2; l2 = l1; f2 = f1;
      l1 = hi - kInvPhi * (hi - lo);
      f1 = excess(l1);
    }
  }
  return f1 > kMargin || f2 > kMargin;
}

// margin function (mod_particle_parallel.f90:145-201): (p-c)^T Q (p-c),
// Q = R diag(1/a^2,1/b^2,1/c^2) R^T, evaluated in the body frame.
__device__ __forceinline__ double fmargin(const double q[4], double a, double b, double c,
                                          const double p[3], const double ctr[3]) {
  Rot<double> R(q[0], q[1], q[2], q[3]);
  double v0, v1, v2;
  R.mulT(p[0] - ctr[0], p[1] - ctr[1], p[2] - ctr[2], v0, v1, v2);
  return v0 * v0 / (a * a) + v1 * v1 / (b * b) + v2 * v2 / (c * c);
}

struct ContactPair {
  Sym3 E1, E2;
  double b1[3], b2[3];
};

// The support points along the normal of angles x (ellipsoid_pair_depth.f90)
// and what the Jacobian at the same x needs of them.  lmder evaluates the
// Jacobian only at accepted points, whose residuals it has already
// evaluated, so the solve keeps the Jet of its current point and the
// Jacobian starts from it instead of recomputing the support (the same
// arithmetic on the same x, hence the same values).
struct Jet {
  double s0, c0, s1, c1;   // sin and cos of x[0], x[1]
  double t1[3], t2[3];     // E1 cv, E2 cv
  double l1, l2;           // sqrt(cv^T E cv / 4)
  double h1, h2;           // 1 / (2 l)
  double d[3];             // p1 - p2
};

// support of the normal cv(x) = (cos x0 cos x1, sin x0 cos x1, sin x1):
// p1 = b1 + E1 cv / (2 l1), p2 = b2 - E2 cv / (2 l2); residuals
// f_i = (p1 - p2)_i^2
__device__ __forceinline__ void support(const ContactPair& cp, const double x[2], Jet& s,
                                        double f[3]) {
  sincos(x[0], &s.s0, &s.c0);
  sincos(x[1], &s.s1, &s.c1);
  const double cv[3] = {s.c0 * s.c1, s.s0 * s.c1, s.s1};
  symv(cp.E1, cv, s.t1);
  symv(cp.E2, cv, s.t2);   // E2 (-cv) = -t2
  s.l1 = sqrt(0.25 * (cv[0] * s.t1[0] + cv[1] * s.t1[1] + cv[2] * s.t1[2]));
  s.l2 = sqrt(0.25 * (cv[0] * s.t2[0] + cv[1] * s.t2[1] + cv[2] * s.t2[2]));
  s.h1 = 1.0 / (2.0 * s.l1);
  s.h2 = 1.0 / (2.0 * s.l2);
  for (int i = 0; i < 3; ++i) {
    const double p1 = s.h1 * s.t1[i] + cp.b1[i];
    const double p2 = cp.b2[i] - s.h2 * s.t2[i];
    s.d[i] = p1 - p2;
    f[i] = s.d[i] * s.d[i];
  }
}

// normal and support points of a Jet
__device__ __forceinline__ void support_points(const ContactPair& cp, const Jet& s, double cv[3],
                                               double p1[3], double p2[3]) {
  cv[0] = s.c0 * s.c1; cv[1] = s.s0 * s.c1; cv[2] = s.s1;
  for (int i = 0; i < 3; ++i) {
    p1[i] = s.h1 * s.t1[i] + cp.b1[i];
    p2[i] = cp.b2[i] - s.h2 * s.t2[i];
  }
}

// Jacobian (ellipsoid_contact_detection.f90:290-330), column-major 3x2:
// d(depth)/dx = (C1 + C2) dcv/dx with Ck = Ek/(2 lk) - tk tk^T / (8 lk^3)
__device__ __forceinline__ void fcn_j(const ContactPair& cp, const Jet& s, double J[6]) {
  const double g0[3] = {-s.s0 * s.c1, s.c0 * s.c1, 0.0};
  const double g1[3] = {-s.c0 * s.s1, -s.s0 * s.s1, s.c1};
  const double k1 = 1.0 / (8.0 * s.l1 * s.l1 * s.l1), k2 = 1.0 / (8.0 * s.l2 * s.l2 * s.l2);
  double e1g0[3], e1g1[3], e2g0[3], e2g1[3];
  symv(cp.E1, g0, e1g0); symv(cp.E1, g1, e1g1);
  symv(cp.E2, g0, e2g0); symv(cp.E2, g1, e2g1);
  const double t1g0 = s.t1[0] * g0[0] + s.t1[1] * g0[1] + s.t1[2] * g0[2];
  const double t1g1 = s.t1[0] * g1[0] + s.t1[1] * g1[1] + s.t1[2] * g1[2];
  const double t2g0 = s.t2[0] * g0[0] + s.t2[1] * g0[1] + s.t2[2] * g0[2];
  const double t2g1 = s.t2[0] * g1[0] + s.t2[1] * g1[1] + s.t2[2] * g1[2];
  for (int i = 0; i < 3; ++i) {
    const double gd0 = s.h1 * e1g0[i] - k1 * s.t1[i] * t1g0 + s.h2 * e2g0[i] - k2 * s.t2[i] * t2g0;
    const double gd1 = s.h1 * e1g1[i] - k1 * s.t1[i] * t1g1 + s.h2 * e2g1[i] - k2 * s.t2[i] * t2g1;
    J[i] = 2.0 * (s.d[i] * gd0);
    J[3 + i] = 2.0 * (s.d[i] * gd1);
  }
}

// a[l] for l in {0, 1} (MINPACK's permutation indices), as a select: a
// dynamic index into a register array would go through scratch memory
__device__ __forceinline__ double pick(const double a[2], int l) { return l ? a[1] : a[0]; }

// enorm (minpack_reduc.f90:180-232), MINPACK's scaled sums for
// components that are tiny (<= rdwarf) or huge (>= rgiant / n).  Out of
// line: the solve reaches it only for such vectors, and keeping its code
// out of the solve's loop keeps that loop compact in the instruction cache.
__device__ __noinline__ double enorm_scaled(double x0, double x1, double x2, int n) {
  const double rdwarf = 3.834e-20, rgiant = 1.304e19;
  const double agiant = rgiant / double(n);
  const double xs[3] = {x0, x1, x2};
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, x1max = 0.0, x3max = 0.0;
  for (int i = 0; i < n; ++i) {
    double xabs = fabs(xs[i]);
    if (xabs > rdwarf && xabs < agiant) {
      s2 += xabs * xabs;
    } else if (xabs <= rdwarf) {
      if (xabs <= x3max) {
        if (xabs != 0.0) { double t = xabs / x3max; s3 += t * t; }
      } else {
        double t = x3max / xabs; s3 = 1.0 + s3 * t * t; x3max = xabs;
      }
    } else if (xabs <= x1max) {
      double t = xabs / x1max; s1 += t * t;
    } else {
      double t = x1max / xabs; s1 = 1.0 + s1 * t * t; x1max = xabs;
    }
  }
  if (s1 != 0.0) return x1max * sqrt(s1 + (s2 / x1max) / x1max);
  if (s2 == 0.0) return x3max * sqrt(s3);
  if (s2 >= x3max) return sqrt(s2 * (1.0 + (x3max / s2) * (x3max * s3)));
  return sqrt(x3max * ((s2 / x3max) + (x3max * s3)));
}

// enorm.  When every component is zero or of intermediate magnitude
// (rdwarf < |x| < rgiant / N) MINPACK's sums reduce to s1 = s3 = x3max = 0
// and the result to sqrt(s2), s2 the plain sum of squares in index order;
// that case is computed here, the others by enorm_scaled.
template <int N>
__device__ __forceinline__ double enorm(const double* x) {
  const double rdwarf = 3.834e-20, rgiant = 1.304e19;
  const double agiant = rgiant / double(N);
  // non-short-circuit operators: the test is a few compares, not a branch
  // per component
  bool mid = true;
  double sq = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double xabs = fabs(x[i]);
    mid = mid & ((xabs == 0.0) | ((xabs > rdwarf) & (xabs < agiant)));
    sq += xabs * xabs;
  }
  if (mid) return sqrt(sq);
  return enorm_scaled(x[0], x[1], N > 2 ? x[2] : 0.0, N);
}

constexpr double EPSMCH = 2.220446049250313e-16;
constexpr double DWARF = 2.2250738585072014e-308;

// qrfac with column pivoting (minpack_reduc.f90:1007-1120), m = 3, n = 2
__device__ __forceinline__ void qrfac32(double a[6], int ipvt[2], double rdiag[2],
                                        double acnorm[2], double wa[2]) {
  for (int j = 0; j < 2; ++j) {
    acnorm[j] = enorm<3>(a + 3 * j);
    rdiag[j] = acnorm[j];
    wa[j] = rdiag[j];
    ipvt[j] = j;
  }
  // j = 0
  if (rdiag[1] > rdiag[0]) {
    for (int i = 0; i < 3; ++i) { double t = a[i]; a[i] = a[3 + i]; a[3 + i] = t; }
    rdiag[1] = rdiag[0];
    wa[1] = wa[0];
    ipvt[0] = 1; ipvt[1] = 0;
  }
  {
    double ajnorm = enorm<3>(a);
    if (ajnorm != 0.0) {
      if (a[0] < 0.0) ajnorm = -ajnorm;
      for (int i = 0; i < 3; ++i) a[i] = a[i] / ajnorm;
      a[0] = a[0] + 1.0;
      double sum = a[0] * a[3] + a[1] * a[4] + a[2] * a[5];
      double temp = sum / a[0];
      for (int i = 0; i < 3; ++i) a[3 + i] = a[3 + i] - temp * a[i];
      if (rdiag[1] != 0.0) {
        temp = a[3] / rdiag[1];
        double t2 = 1.0 - temp * temp;
        rdiag[1] = rdiag[1] * sqrt(t2 > 0.0 ? t2 : 0.0);
        double t3 = rdiag[1] / wa[1];
        if (0.05 * (t3 * t3) <= EPSMCH) {
          rdiag[1] = enorm<2>(a + 4);
          wa[1] = rdiag[1];
        }
      }
    }
    rdiag[0] = -ajnorm;
  }
  // j = 1 (no pivot candidates left)
  {
    double ajnorm = enorm<2>(a + 4);
    if (ajnorm != 0.0) {
      if (a[4] < 0.0) ajnorm = -ajnorm;
      a[4] = a[4] / ajnorm;
      a[5] = a[5] / ajnorm;
      a[4] = a[4] + 1.0;
    }
    rdiag[1] = -ajnorm;
  }
}

// qrsolv (minpack_reduc.f90:1155-1289), n = 2, r column-major with ldr = 3
__device__ __forceinline__ void qrsolv2(double r[6], const int ipvt[2], const double diag[2],
                                        const double qtb[2], double x[2], double sdiag[2]) {
  double wa[2];
  // copy r and qtb; save the diagonal of r in x
  r[1] = r[3];             // r(1,0) = r(0,1)
  x[0] = r[0]; x[1] = r[4];
  wa[0] = qtb[0]; wa[1] = qtb[1];
  for (int j = 0; j < 2; ++j) {
    const double dl = pick(diag, ipvt[j]);
    if (dl != 0.0) {
      for (int k = j; k < 2; ++k) sdiag[k] = 0.0;
      sdiag[j] = dl;
      double qtbpj = 0.0;
      for (int k = j; k < 2; ++k) {
        if (sdiag[k] != 0.0) {
          double rkk = r[k * 4];   // r(k,k) at k + 3k
          double cs, sn;
          if (fabs(rkk) >= fabs(sdiag[k])) {
            double tn = sdiag[k] / rkk;
            cs = 0.5 / sqrt(0.25 + 0.25 * (tn * tn));
            sn = cs * tn;
          } else {
            double ct = rkk / sdiag[k];
            sn = 0.5 / sqrt(0.25 + 0.25 * (ct * ct));
            cs = sn * ct;
          }
          r[k * 4] = cs * rkk + sn * sdiag[k];
          double temp = cs * wa[k] + sn * qtbpj;
          qtbpj = -(sn * wa[k]) + cs * qtbpj;
          wa[k] = temp;
          if (k == 0) {   // i = 1: r(1,0)
            temp = cs * r[1] + sn * sdiag[1];
            sdiag[1] = -(sn * r[1]) + cs * sdiag[1];
            r[1] = temp;
          }
        }
      }
    }
    sdiag[j] = r[j * 4];
    r[j * 4] = x[j];
  }
  int nsing = 2;
  for (int j = 0; j < 2; ++j) {
    if (sdiag[j] == 0.0 && nsing == 2) nsing = j;
    if (nsing < 2) wa[j] = 0.0;
  }
  if (nsing >= 2) wa[1] = wa[1] / sdiag[1];
  if (nsing >= 1) {
    double sum = (nsing >= 2) ? r[1] * wa[1] : 0.0;
    wa[0] = (wa[0] - sum) / sdiag[0];
  }
  x[0] = ipvt[0] ? wa[1] : wa[0];   // x[ipvt[j]] = wa[j]
  x[1] = ipvt[0] ? wa[0] : wa[1];
}

// lmpar (minpack_reduc.f90:798-987), n = 2
__device__ __forceinline__ void lmpar2(double r[6], const int ipvt[2], const double diag[2],
                                       const double qtb[2], double delta, double& par,
                                       double x[2], double sdiag[2]) {
  const double p1 = 0.1, p001 = 1.0e-3;
  double wa1[2], wa2[2];
  int nsing = 2;
  for (int j = 0; j < 2; ++j) {
    wa1[j] = qtb[j];
    if (r[j * 4] == 0.0 && nsing == 2) nsing = j;
    if (nsing < 2) wa1[j] = 0.0;
  }
  if (nsing >= 2) { wa1[1] = wa1[1] / r[4]; wa1[0] = wa1[0] - r[3] * wa1[1]; }
  if (nsing >= 1) wa1[0] = wa1[0] / r[0];
  x[0] = ipvt[0] ? wa1[1] : wa1[0];   // x[ipvt[j]] = wa1[j]
  x[1] = ipvt[0] ? wa1[0] : wa1[1];
  wa2[0] = diag[0] * x[0]; wa2[1] = diag[1] * x[1];
  double dxnorm = enorm<2>(wa2);
  double fp = dxnorm - delta;
  if (fp <= p1 * delta) { par = 0.0; return; }
  double parl = 0.0;
  if (nsing >= 2) {
    for (int j = 0; j < 2; ++j) wa1[j] = pick(diag, ipvt[j]) * (pick(wa2, ipvt[j]) / dxnorm);
    wa1[0] = wa1[0] / r[0];
    wa1[1] = (wa1[1] - r[3] * wa1[0]) / r[4];
    double temp = enorm<2>(wa1);
    parl = ((fp / delta) / temp) / temp;
  }
  wa1[0] = (r[0] * qtb[0]) / pick(diag, ipvt[0]);
  wa1[1] = (r[3] * qtb[0] + r[4] * qtb[1]) / pick(diag, ipvt[1]);
  double gnorm = enorm<2>(wa1);
  double paru = gnorm / delta;
  if (paru == 0.0) paru = DWARF / fmin(delta, p1);
  par = fmax(par, parl);
  par = fmin(par, paru);
  if (par == 0.0) par = gnorm / dxnorm;
  for (int iter = 1;; ++iter) {
    if (par == 0.0) par = fmax(DWARF, p001 * paru);
    double temp = sqrt(par);
    wa1[0] = temp * diag[0]; wa1[1] = temp * diag[1];
    qrsolv2(r, ipvt, wa1, qtb, x, sdiag);
    wa2[0] = diag[0] * x[0]; wa2[1] = diag[1] * x[1];
    dxnorm = enorm<2>(wa2);
    temp = fp;
    fp = dxnorm - delta;
    if (fabs(fp) <= p1 * delta || (parl == 0.0 && fp <= temp && temp < 0.0) || iter == 10) break;
    for (int j = 0; j < 2; ++j) wa1[j] = pick(diag, ipvt[j]) * (pick(wa2, ipvt[j]) / dxnorm);
    wa1[0] = wa1[0] / sdiag[0];
    wa1[1] = wa1[1] - r[1] * wa1[0];
    wa1[1] = wa1[1] / sdiag[1];
    temp = enorm<2>(wa1);
    double parc = ((fp / delta) / temp) / temp;
    if (fp > 0.0) parl = fmax(parl, par);
    if (fp < 0.0) paru = fmin(paru, par);
    par = fmax(parl, par + parc);
  }
}

// lmder with lmder1's settings (minpack_reduc.f90:243-640, :732-745):
// mode 1, factor 100, maxfev 300, ftol = xtol = sqrt(eps), gtol = 0.
// Returns the number of function evaluations; *info_out = MINPACK's info
// (1 converged, 4 zero gradient, 5 maxfev, 6 tolerances too small for a
// further reduction: MINPACK's 6-8); cur = the Jet of the returned x.
// maxfev below 300 only for a network-seeded first attempt.
__device__ int lm_solve(const ContactPair& cp, double x[2], int* info_out, Jet& cur,
                        int maxfev = 300) {
  const double tol = 1.4901161193847656e-08;
  const double p1 = 0.1, p5 = 0.5, p25 = 0.25, p75 = 0.75, p0001 = 1.0e-4;
  double fvec[3], fjac[6], diag[2], qtf[2], wa1[2], wa2[2], wa3[2], wa4[3];
  int ipvt[2];
  int info = 0, nfev = 1, iter = 1;
  double delta = 0.0, xnorm = 0.0, par = 0.0;
  Jet trial;
  support(cp, x, cur, fvec);
  double fnorm = enorm<3>(fvec);
  for (;;) {
    fcn_j(cp, cur, fjac);
    qrfac32(fjac, ipvt, wa1, wa2, wa3);
    if (iter == 1) {
      for (int j = 0; j < 2; ++j) diag[j] = (wa2[j] == 0.0) ? 1.0 : wa2[j];
      wa3[0] = diag[0] * x[0]; wa3[1] = diag[1] * x[1];
      xnorm = enorm<2>(wa3);
      delta = 100.0 * xnorm;
      if (delta == 0.0) delta = 100.0;
    }
    wa4[0] = fvec[0]; wa4[1] = fvec[1]; wa4[2] = fvec[2];
    // Q^T fvec
    if (fjac[0] != 0.0) {
      double sum = fjac[0] * wa4[0] + fjac[1] * wa4[1] + fjac[2] * wa4[2];
      double temp = -sum / fjac[0];
      for (int i = 0; i < 3; ++i) wa4[i] += fjac[i] * temp;
    }
    fjac[0] = wa1[0];
    qtf[0] = wa4[0];
    if (fjac[4] != 0.0) {
      double sum = fjac[4] * wa4[1] + fjac[5] * wa4[2];
      double temp = -sum / fjac[4];
      wa4[1] += fjac[4] * temp;
      wa4[2] += fjac[5] * temp;
    }
    fjac[4] = wa1[1];
    qtf[1] = wa4[1];
    double gnorm = 0.0;
    if (fnorm != 0.0) {
      for (int j = 0; j < 2; ++j) {
        const double wl = pick(wa2, ipvt[j]);
        if (wl != 0.0) {
          double sum = (j == 0) ? fjac[0] * (qtf[0] / fnorm)
                                : fjac[3] * (qtf[0] / fnorm) + fjac[4] * (qtf[1] / fnorm);
          gnorm = fmax(gnorm, fabs(sum / wl));
        }
      }
    }
    if (gnorm <= 0.0) { info = 4; break; }   // gtol = 0
    diag[0] = fmax(diag[0], wa2[0]);
    diag[1] = fmax(diag[1], wa2[1]);
    for (;;) {
      lmpar2(fjac, ipvt, diag, qtf, delta, par, wa1, wa2);
      wa1[0] = -wa1[0]; wa1[1] = -wa1[1];
      wa2[0] = x[0] + wa1[0]; wa2[1] = x[1] + wa1[1];
      wa3[0] = diag[0] * wa1[0]; wa3[1] = diag[1] * wa1[1];
      double pnorm = enorm<2>(wa3);
      if (iter == 1) delta = fmin(delta, pnorm);
      support(cp, wa2, trial, wa4);
      nfev++;
      double fnorm1 = enorm<3>(wa4);
      double actred = -1.0;
      if (p1 * fnorm1 < fnorm) { double t = fnorm1 / fnorm; actred = 1.0 - t * t; }
      {
        double t0 = pick(wa1, ipvt[0]), t1 = pick(wa1, ipvt[1]);
        wa3[0] = fjac[0] * t0 + fjac[3] * t1;
        wa3[1] = fjac[4] * t1;
      }
      double temp1 = enorm<2>(wa3) / fnorm;
      double temp2 = (sqrt(par) * pnorm) / fnorm;
      double prered = temp1 * temp1 + (temp2 * temp2) / p5;
      double dirder = -(temp1 * temp1 + temp2 * temp2);
      double ratio = 0.0;
      if (prered != 0.0) ratio = actred / prered;
      if (ratio <= p25) {
        double temp = (actred >= 0.0) ? p5 : p5 * dirder / (dirder + p5 * actred);
        if (p1 * fnorm1 >= fnorm || temp < p1) temp = p1;
        delta = temp * fmin(delta, pnorm / p1);
        par = par / temp;
      } else if (par == 0.0 || ratio >= p75) {
        delta = pnorm / p5;
        par = p5 * par;
      }
      if (ratio >= p0001) {
        x[0] = wa2[0]; x[1] = wa2[1];
        wa2[0] = diag[0] * x[0]; wa2[1] = diag[1] * x[1];
        fvec[0] = wa4[0]; fvec[1] = wa4[1]; fvec[2] = wa4[2];
        cur = trial;
        xnorm = enorm<2>(wa2);
        fnorm = fnorm1;
        iter++;
      }
      bool small = fabs(actred) <= tol && prered <= tol && p5 * ratio <= 1.0;
      if (small || delta <= tol * xnorm) { info = 1; break; }
      if (nfev >= maxfev) { info = 5; break; }
      if ((fabs(actred) <= EPSMCH && prered <= EPSMCH && p5 * ratio <= 1.0) ||
          delta <= EPSMCH * xnorm || gnorm <= EPSMCH) {
        info = 6;
        break;
      }
      if (ratio >= p0001) break;
    }
    if (info != 0) break;
  }
  *info_out = info;
  return nfev;
}

struct ContactOut {
  double e1[3], e2[3], cn[3];
  int flag, nfev, info;
};

// ellipsoid_contact_detection (ellipsoid_contact_detection.f90:1-155) for
// ellipsoid i at bi and j at bj (bj already the minimum image).  seed, when
// given, replaces the centre-line initial guess (:77-87).
__device__ __forceinline__ ContactOut contact_detect(const double bi[3], const double qi[4],
                                                     const Species& si, const double bj[3],
                                                     const double qj[4], const Species& sj,
                                                     const double* seed, int maxfev = 300) {
  ContactPair cp;
  cp.E1 = ell_matrix(qi, si.a, si.b, si.c);
  cp.E2 = ell_matrix(qj, sj.a, sj.b, sj.c);
  for (int i = 0; i < 3; ++i) { cp.b1[i] = bi[i]; cp.b2[i] = bj[i]; }
  double x[2];
  if (seed) {
    x[0] = seed[0]; x[1] = seed[1];
  } else {
    double c0 = bi[0] - bj[0], c1 = bi[1] - bj[1], c2 = bi[2] - bj[2];
    double cn = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    if (cn != 0.0) { c0 /= cn; c1 /= cn; c2 /= cn; }
    x[1] = asin(c2);
    x[0] = acos(c0 / cos(x[1]));
  }
  ContactOut o;
  Jet s;
  o.nfev = lm_solve(cp, x, &o.info, s, maxfev);
  double cv[3];
  support_points(cp, s, cv, o.e1, o.e2);
  double d0 = o.e1[0] - o.e2[0], d1 = o.e1[1] - o.e2[1], d2 = o.e1[2] - o.e2[2];
  double contact = d0 * cv[0] + d1 * cv[1] + d2 * cv[2];
  o.flag = (contact > 0.0) && (fmargin(qi, si.a, si.b, si.c, o.e2, bi) < 1.0) &&
           (fmargin(qj, sj.a, sj.b, sj.c, o.e1, bj) < 1.0);
  o.cn[0] = cv[0]; o.cn[1] = cv[1]; o.cn[2] = cv[2];
  return o;
}

// ---------------------------------------------------------------------------
// COLLISION_RESPONSE (collision_response.f90:1-437): impulse exchange with
// restitution ecp / friction mup, unit masses, the inertia of the first
// particle's species used for both (:140-145).  u in the lab frame, o in the
// body frame, updated in place.  Returns false on the KE `stop` (:408-436).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool collision_response(const Species& s1, const double q1[4],
                                                   const double q2[4], double u1[3],
                                                   double o1[3], double u2[3], double o2[3],
                                                   const double r1[3], const double r2[3],
                                                   const double n[3]) {
  Rot<double> R1(q1[0], q1[1], q1[2], q1[3]), R2(q2[0], q2[1], q2[2], q2[3]);
  double w1[3], w2[3];
  R1.mul(o1[0], o1[1], o1[2], w1[0], w1[1], w1[2]);
  R2.mul(o2[0], o2[1], o2[2], w2[0], w2[1], w2[2]);
  double ur[3] = {(u2[0] + (w2[1] * r2[2] - w2[2] * r2[1])) - (u1[0] + (w1[1] * r1[2] - w1[2] * r1[1])),
                  (u2[1] + (w2[2] * r2[0] - w2[0] * r2[2])) - (u1[1] + (w1[2] * r1[0] - w1[0] * r1[2])),
                  (u2[2] + (w2[0] * r2[1] - w2[1] * r2[0])) - (u1[2] + (w1[0] * r1[1] - w1[1] * r1[0]))};
  double urn = ur[0] * n[0] + ur[1] * n[1] + ur[2] * n[2];
  double du[3] = {ur[0] + s1.ecp * urn * n[0], ur[1] + s1.ecp * urn * n[1],
                  ur[2] + s1.ecp * urn * n[2]};
  double ip1[9], ip2[9], iip1[9], iip2[9], K1[9], K2[9], K[9], iK[9];
  lab_tensor(R1, s1.ixx, s1.iyy, s1.izz, ip1);
  lab_tensor(R2, s1.ixx, s1.iyy, s1.izz, ip2);
  inv3(ip1, iip1);
  inv3(ip2, iip2);
  k_matrix(iip1, r1, K1);
  k_matrix(iip2, r2, K2);
  for (int i = 0; i < 9; ++i) K[i] = K1[i] + K2[i];
  inv3(K, iK);
  double pc[3];
  mat3v(iK, du, pc);
  pc[0] = -pc[0]; pc[1] = -pc[1]; pc[2] = -pc[2];
  double pcn = pc[0] * n[0] + pc[1] * n[1] + pc[2] * n[2];
  double pt0 = pc[0] - pcn * n[0], pt1 = pc[1] - pcn * n[1], pt2 = pc[2] - pcn * n[2];
  double pct = sqrt(pt0 * pt0 + pt1 * pt1 + pt2 * pt2);
  if (pct > s1.mup * pcn) {
    double tw[3] = {ur[0] - urn * n[0], ur[1] - urn * n[1], ur[2] - urn * n[2]};
    double ntw = sqrt(tw[0] * tw[0] + tw[1] * tw[1] + tw[2] * tw[2]);
    if (ntw != 0.0) { tw[0] = -tw[0] / ntw; tw[1] = -tw[1] / ntw; tw[2] = -tw[2] / ntw; }
    double vi[3] = {n[0] + s1.mup * tw[0], n[1] + s1.mup * tw[1], n[2] + s1.mup * tw[2]};
    double kn[3];
    mat3v(K, vi, kn);
    double kndn = kn[0] * n[0] + kn[1] * n[1] + kn[2] * n[2];
    double p = -(du[0] * n[0] + du[1] * n[1] + du[2] * n[2]);
    double s = p / kndn;
    pc[0] = s * vi[0]; pc[1] = s * vi[1]; pc[2] = s * vi[2];
  }
  double am[3];
  mat3v(ip1, w1, am);
  double ei = 0.5 * (u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]) +
              0.5 * (w1[0] * am[0] + w1[1] * am[1] + w1[2] * am[2]);
  mat3v(ip2, w2, am);
  ei += 0.5 * (u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]) +
        0.5 * (w2[0] * am[0] + w2[1] * am[1] + w2[2] * am[2]);
  for (int i = 0; i < 3; ++i) { u1[i] -= pc[i]; u2[i] += pc[i]; }
  // [r]x (-pc) and [r]x pc
  double a1[3] = {-(-r1[2] * pc[1] + r1[1] * pc[2]), -(r1[2] * pc[0] - r1[0] * pc[2]),
                  -(-r1[1] * pc[0] + r1[0] * pc[1])};
  double a2[3] = {-r2[2] * pc[1] + r2[1] * pc[2], r2[2] * pc[0] - r2[0] * pc[2],
                  -r2[1] * pc[0] + r2[0] * pc[1]};
  double d1[3], d2[3];
  mat3v(iip1, a1, d1);
  mat3v(iip2, a2, d2);
  for (int i = 0; i < 3; ++i) { w1[i] += d1[i]; w2[i] += d2[i]; }
  mat3v(ip1, w1, am);
  double ef = 0.5 * (u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]) +
              0.5 * (w1[0] * am[0] + w1[1] * am[1] + w1[2] * am[2]);
  mat3v(ip2, w2, am);
  ef += 0.5 * (u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]) +
        0.5 * (w2[0] * am[0] + w2[1] * am[1] + w2[2] * am[2]);
  R1.mulT(w1[0], w1[1], w1[2], o1[0], o1[1], o1[2]);
  R2.mulT(w2[0], w2[1], w2[2], o2[0], o2[1], o2[2]);
  return !(fabs((ei - ef) / ei) > 1.0e-9);
}

}  // namespace gp

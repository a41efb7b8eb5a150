/*
 * gazpart_oracle.c -- CPU FP64 restatement of GAZPART's Lagrangian ellipsoid
 * step.  TEST INFRASTRUCTURE: this is the checker the HIP library is compared
 * against, and the CPU baseline the benchmark times.  It is never linked
 * into the product library.
 *
 * Each function cites the reference routine (source/<file>:<line>) whose
 * arithmetic it restates.  Expressions keep the reference's evaluation order
 * (Fortran evaluates a*b*c as (a*b)*c and sums left to right) so that the
 * result equals the reference compiled with -fdefault-real-8 on the same
 * inputs; compile this file without FMA contraction (-ffp-contract=off).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   interpolation, force, torque, quaternion map, wall test/rebound,
 *   contact detection (LM), collision response, cell list and the full
 *   N-step trajectory are each compared with fixtures produced by the
 *   reference Fortran itself (tests/golden/) and with closed forms.
 */
#define _GNU_SOURCE
#include "gazpart_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PPI 3.141592653589793           /* 4*atan(1d0), init_run.f90:123 */

/* Where the reference build evaluates sin(t) and cos(t) of one argument
 * with a single sincos call (the support function and Jacobian of the
 * contact problem, the quaternion exponential map), the restatement does
 * the same: glibc's sincos can differ from cos in the last bit.  The force
 * and torque correlations keep separate sin and cos calls, as there.  Each
 * choice is checked bit for bit against the reference build. */
static inline void sc(double t, double *s, double *c) { sincos(t, s, c); }
#define ZERO_R 1.0e-33                  /* mod_dns_dim.f90:136 */

/* ------------------------------------------------------------------------ */
/* small 3x3 helpers, Fortran column-major semantics kept explicit          */
/* ------------------------------------------------------------------------ */
typedef struct { double m[3][3]; } m33;   /* m[row][col] */

/* matmul(A, B): C(i,j) = sum_k A(i,k) B(k,j), k ascending from 0 */
static m33 mm(const m33 *a, const m33 *b) {
  m33 c;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s = s + a->m[i][k] * b->m[k][j];
      c.m[i][j] = s;
    }
  return c;
}
static void mv(const m33 *a, const double *v, double *out) {
  double t[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + a->m[i][k] * v[k];
    t[i] = s;
  }
  out[0] = t[0]; out[1] = t[1]; out[2] = t[2];
}
static m33 transpose(const m33 *a) {
  m33 t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t.m[i][j] = a->m[j][i];
  return t;
}
static double dot3(const double *a, const double *b) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s = s + a[k] * b[k];
  return s;
}

/* rotation matrix of a quaternion, transform_basis.f90:44-54 */
static m33 rotmat(double q0, double q1, double q2, double q3) {
  m33 r;
  r.m[0][0] = 1.0 - 2.0 * (q2 * q2 + q3 * q3);
  r.m[0][1] = 2.0 * (q1 * q2 - q0 * q3);
  r.m[0][2] = 2.0 * (q1 * q3 + q0 * q2);
  r.m[1][0] = 2.0 * (q1 * q2 + q0 * q3);
  r.m[1][1] = 1.0 - 2.0 * (q1 * q1 + q3 * q3);
  r.m[1][2] = 2.0 * (q2 * q3 - q0 * q1);
  r.m[2][0] = 2.0 * (q1 * q3 - q0 * q2);
  r.m[2][1] = 2.0 * (q2 * q3 + q0 * q1);
  r.m[2][2] = 1.0 - 2.0 * (q1 * q1 + q2 * q2);
  return r;
}

/* transform_basis (vector case), transform_basis.f90:56-77 */
void oc_transform_basis_vec(const double q[4], double v[3]) {
  m33 r = rotmat(q[0], q[1], q[2], q[3]);
  mv(&r, v, v);
}
static void rot_conj(const double q[4], double v[3]) {
  m33 r = rotmat(q[0], -q[1], -q[2], -q[3]);
  mv(&r, v, v);
}
/* transform_basis (tensor case) R A R^T, transform_basis.f90:80-83 */
static m33 rot_tensor(const double q[4], const m33 *a) {
  m33 r = rotmat(q[0], q[1], q[2], q[3]);
  m33 rt = transpose(&r);
  m33 t = mm(a, &rt);
  return mm(&r, &t);
}

/* invert_ndim3_matrix.f90:1-39 */
static m33 inv3(const m33 *M) {
  const double (*m)[3] = M->m;
  double detinv = 1.0 / (m[0][0] * m[1][1] * m[2][2] - m[0][0] * m[1][2] * m[2][1]
                        - m[0][1] * m[1][0] * m[2][2] + m[0][1] * m[1][2] * m[2][0]
                        + m[0][2] * m[1][0] * m[2][1] - m[0][2] * m[1][1] * m[2][0]);
  m33 v;
  v.m[0][0] = +detinv * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
  v.m[1][0] = -detinv * (m[1][0] * m[2][2] - m[1][2] * m[2][0]);
  v.m[2][0] = +detinv * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  v.m[0][1] = -detinv * (m[0][1] * m[2][2] - m[0][2] * m[2][1]);
  v.m[1][1] = +detinv * (m[0][0] * m[2][2] - m[0][2] * m[2][0]);
  v.m[2][1] = -detinv * (m[0][0] * m[2][1] - m[0][1] * m[2][0]);
  v.m[0][2] = +detinv * (m[0][1] * m[1][2] - m[0][2] * m[1][1]);
  v.m[1][2] = -detinv * (m[0][0] * m[1][2] - m[0][2] * m[1][0]);
  v.m[2][2] = +detinv * (m[0][0] * m[1][1] - m[0][1] * m[1][0]);
  return v;
}

/* integer powers x**n for constant n, multiplied left to right as the
 * reference build evaluates them (checked against it bit for bit) */
static double ipow3(double x) { return (x * x) * x; }
static double ipow5(double x) { return (((x * x) * x) * x) * x; }

/* ------------------------------------------------------------------------ */
/* species constants                                                        */
/* ------------------------------------------------------------------------ */
/* PMASS fluid_force_ellipsoid.f90:59, dequiv :70, inertia
 * initiation_particle_position.f90:620-622, semi-axes
 * calculate_ellipsoid_matrix.f90:33-35 */
void oc_species_spheroid(double emaj, double apr, double rhop, double grav,
                         double ecp, double mup, double *row) {
  double a = emaj, b = a / apr, c = b;
  row[SP_A] = a; row[SP_B] = b; row[SP_C] = c;
  /* EMAJ**3.0 and APR**(2.0) are evaluated as products by the reference build */
  row[SP_PMASS] = rhop * (4.0 / 3.0) * PPI * ipow3(emaj) / (apr * apr);
  row[SP_DEQ] = (2.0 * emaj) / pow(apr, 2.0 / 3.0);
  row[SP_IXX] = (b * b + c * c) / 5.0;
  row[SP_IYY] = (a * a + c * c) / 5.0;
  row[SP_IZZ] = (a * a + b * b) / 5.0;
  row[SP_GRAV] = grav; row[SP_ECP] = ecp; row[SP_MUP] = mup;
  row[SP_RHOP] = rhop; row[SP_EMAJ] = emaj; row[SP_APR] = apr;
}
/* tri-axial extension (no reference counterpart): mass and volume-equivalent
 * diameter of a general ellipsoid; reduces to the spheroid values for b=c. */
void oc_species_triaxial(double a, double b, double c, double rhop,
                         double grav, double ecp, double mup, double *row) {
  row[SP_A] = a; row[SP_B] = b; row[SP_C] = c;
  row[SP_PMASS] = rhop * (4.0 / 3.0) * PPI * (a * b * c);
  row[SP_DEQ] = 2.0 * cbrt(a * b * c);
  row[SP_IXX] = (b * b + c * c) / 5.0;
  row[SP_IYY] = (a * a + c * c) / 5.0;
  row[SP_IZZ] = (a * a + b * b) / 5.0;
  row[SP_GRAV] = grav; row[SP_ECP] = ecp; row[SP_MUP] = mup;
  row[SP_RHOP] = rhop; row[SP_EMAJ] = a; row[SP_APR] = a / b;
}

/* ------------------------------------------------------------------------ */
/* field + interpolation                                                    */
/* ------------------------------------------------------------------------ */
int64_t oc_field_size(int n) { return (int64_t)n * (n + 4) * (n + 4); }

/* storage offset of UAIN(i, j, k), i in 1..n, j,k in -1..n+2 */
static inline int64_t fidx(int n, int i, int j, int k) {
  return (int64_t)(i - 1) + (int64_t)(j + 1) * n + (int64_t)(k + 1) * n * (n + 4);
}

/* ghost rule of FLUIDCOMM at NPROC = 1 (MPI_structures.f90:27-132):
 * index period n in y and z, applied to y first then z (edges included). */
void oc_fill_ghosts(int n, double *u) {
  for (int k = 1; k <= n; ++k)
    for (int g = 0; g < 4; ++g) {
      int j = (g == 0) ? -1 : (g == 1) ? 0 : (g == 2) ? n + 1 : n + 2;
      int js = (g == 0) ? n - 1 : (g == 1) ? n : (g == 2) ? 1 : 2;
      for (int i = 1; i <= n; ++i) u[fidx(n, i, j, k)] = u[fidx(n, i, js, k)];
    }
  for (int g = 0; g < 4; ++g) {
    int k = (g == 0) ? -1 : (g == 1) ? 0 : (g == 2) ? n + 1 : n + 2;
    int ks = (g == 0) ? n - 1 : (g == 1) ? n : (g == 2) ? 1 : 2;
    for (int j = -1; j <= n + 2; ++j)
      for (int i = 1; i <= n; ++i) u[fidx(n, i, j, k)] = u[fidx(n, i, j, ks)];
  }
}

/* INTERP_LAG3 interp_lag3.f90:119-290 (periodic x branch, WALL_BOUNDARY
 * false) and INTERP_LAG4 interp_lag4.f90:84-312 including the FYP2 row that
 * reads JP1 (interp_lag4.f90:196,222,248,274,300). */
static int interp_one(int scheme, int n, double dx, const double *U,
                      double xp, double yp, double zp, double *out) {
  int i0 = (int)(xp / dx) + 1, j0 = (int)(yp / dx) + 1, k0 = (int)(zp / dx) + 1;
  if (i0 < 1 || i0 > n || j0 < 1 || j0 > n || k0 < 1 || k0 > n ||
      !(xp == xp) || !(yp == yp) || !(zp == zp))
    return OC_ERR_ARG;
  double a = (xp - (double)(i0 - 1) * dx) / dx;
  double b = (yp - (double)(j0 - 1) * dx) / dx;
  double g = (zp - (double)(k0 - 1) * dx) / dx;
  int ii[5], jj[5], kk[5];
  double fx[5], fy[5], fz[5];
  int nt;
  if (scheme == 3) {
    const double c6 = 1.0 / 6.0, c2 = 1.0 / 2.0;
    nt = 4;
    int ip1 = (i0 == n) ? 1 : i0 + 1;
    int ip2 = (ip1 == n) ? 1 : ip1 + 1;
    int im1 = (i0 == 1) ? n : i0 - 1;
    ii[0] = im1; ii[1] = i0; ii[2] = ip1; ii[3] = ip2;
    jj[0] = j0 - 1; jj[1] = j0; jj[2] = j0 + 1; jj[3] = j0 + 2;
    kk[0] = k0 - 1; kk[1] = k0; kk[2] = k0 + 1; kk[3] = k0 + 2;
#define W3(F, t)                                                    \
  F[0] = -(c6 * (t - 2.0) * (t - 1.0) * t);                         \
  F[1] = c2 * (t - 2.0) * (t - 1.0) * (t + 1.0);                    \
  F[2] = -(c2 * (t - 2.0) * t * (t + 1.0));                         \
  F[3] = c6 * (t - 1.0) * t * (t + 1.0);
    W3(fx, a) W3(fy, b) W3(fz, g)
#undef W3
  } else if (scheme == 4) {
    const double c24 = 1.0 / 24.0, c6 = 1.0 / 6.0, c4 = 1.0 / 4.0;
    nt = 5;
    int ip1 = (i0 == n) ? 1 : i0 + 1;
    int ip2 = (ip1 == n) ? 1 : ip1 + 1;
    int im1 = (i0 == 1) ? n : i0 - 1;
    int im2 = (im1 == 1) ? n : im1 - 1;
    ii[0] = im2; ii[1] = im1; ii[2] = i0; ii[3] = ip1; ii[4] = ip2;
    jj[0] = j0 - 2; jj[1] = j0 - 1; jj[2] = j0; jj[3] = j0 + 1;
    jj[4] = j0 + 1;   /* reference reads JP1 for the FYP2 row */
    kk[0] = k0 - 2; kk[1] = k0 - 1; kk[2] = k0; kk[3] = k0 + 1; kk[4] = k0 + 2;
#define W4(F, t)                                                        \
  F[0] = c24 * (t - 2.0) * (t - 1.0) * t * (t + 1.0);                   \
  F[1] = -(c6 * (t - 2.0) * (t - 1.0) * t * (t + 2.0));                 \
  F[2] = c4 * (t - 2.0) * (t - 1.0) * (t + 1.0) * (t + 2.0);            \
  F[3] = -(c6 * (t - 2.0) * t * (t + 1.0) * (t + 2.0));                 \
  F[4] = c24 * (t - 1.0) * t * (t + 1.0) * (t + 2.0);
    W4(fx, a) W4(fy, b) W4(fz, g)
#undef W4
  } else {
    return OC_ERR_ARG;
  }
  double sz = 0.0;
  for (int kz = 0; kz < nt; ++kz) {
    double sy = 0.0;
    for (int jy = 0; jy < nt; ++jy) {
      double sx = 0.0;
      for (int ix = 0; ix < nt; ++ix)
        sx = sx + fx[ix] * U[fidx(n, ii[ix], jj[jy], kk[kz])];
      sy = sy + fy[jy] * sx;
    }
    sz = sz + fz[kz] * sy;
  }
  *out = sz;
  return OC_OK;
}

/* INTERPH interpH.f90:26-145 */
int oc_interph(int scheme, int n, double dx, const double *uain, int nint,
               const double *xint, const double *yint, const double *zint,
               double *uint_out) {
  int err = OC_OK;
  for (int p = 0; p < nint; ++p) {
    int e = interp_one(scheme, n, dx, uain, xint[p], yint[p], zint[p], &uint_out[p]);
    if (e && !err) err = e;
  }
  return err;
}

/* ------------------------------------------------------------------------ */
/* QUATERNION_INTEGRATION adv_quaternion.f90:1-88 (METHOD = 1) with
 * mult_quat = q_mult_q mod_quaternion.f90:136-145                          */
/* ------------------------------------------------------------------------ */
void oc_quaternion_integration(double q[4], double ox, double oy, double oz,
                               double dt) {
  double om[3] = {ox, oy, oz};
  oc_transform_basis_vec(q, om);
  double nrm = sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
  double m0, m1, m2, m3;
  if (nrm != 0.0) {
    double sh, ch;
    sc((nrm * dt) / 2.0, &sh, &ch);
    m0 = ch;
    m1 = sh * (om[0] / nrm);
    m2 = sh * (om[1] / nrm);
    m3 = sh * (om[2] / nrm);
  } else {
    m0 = 1.0; m1 = 0.0; m2 = 0.0; m3 = 0.0;
  }
  double a = m0 * q[0] - m1 * q[1] - m2 * q[2] - m3 * q[3];
  double b = m0 * q[1] + m1 * q[0] + m2 * q[3] - m3 * q[2];
  double c = m0 * q[2] - m1 * q[3] + m2 * q[0] + m3 * q[1];
  double d = m0 * q[3] + m1 * q[2] - m2 * q[1] + m3 * q[0];
  q[0] = a; q[1] = b; q[2] = c; q[3] = d;
}

/* ------------------------------------------------------------------------ */
/* FLUID_FORCE_ELLIPSOID fluid_force_ellipsoid.f90:1-161                    */
/* ------------------------------------------------------------------------ */
int oc_fluid_force(const oc_params *p, const double *sp, double up, double vp,
                   double wp, double uf, double vf, double wf,
                   const double q[4], double fdrag[3], double flift[3]) {
  const double a0 = 1.95, a1 = 18.12, a2 = 1.023, a3 = 4.26, a4 = 0.384,
               a5 = 21.52, a6 = 0.990, a7 = 2.86, a8 = 0.260;
  const double b1 = 0.083, b2 = -0.21, b3 = 1.582, b4 = 0.851, b5 = 1.842,
               b6 = -0.802, b7 = -0.006, b8 = 0.874, b9 = 0.009, b10 = 0.570;
  double pmass = sp[SP_PMASS];
  double ur[3] = {uf - up, vf - vp, wf - wp};
  double urel_mag = sqrt(ur[0] * ur[0] + ur[1] * ur[1] + ur[2] * ur[2]);
  double deq = sp[SP_DEQ];
  double rep = (urel_mag * deq) / p->visc;
  double u[3] = {uf - up, vf - vp, wf - wp};
  rot_conj(q, u);
  double phi = fabs(atan((sqrt(u[1] * u[1] + u[2] * u[2])) / (u[0])));
  double umag = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  double sphi = sin(phi), cphi = cos(phi);
  double cd0 = (a1 / pow(rep, a2)) + (a3 / pow(rep, a4));
  double cd90 = (a5 / pow(rep, a6)) + (a7 / pow(rep, a8));
  double cd = cd0 + (cd90 - cd0) * pow(sphi, a0);
  double cl = ((b1 / pow(rep, b2)) + (b3 / pow(rep, b4))) *
              pow(sphi, (b5 + b6 * pow(rep, b7))) *
              pow(cphi, (b8 + b9 * pow(rep, b10)));
  double pre = (1.0 / 2.0) * (p->rhof / pmass) * (1.0 / 4.0) * (PPI * (deq * deq));
  for (int i = 0; i < 3; ++i) fdrag[i] = pre * cd * umag * u[i];
  double syz = sqrt(u[1] * u[1] + u[2] * u[2]);
  flift[0] = pre * cl * (umag * umag) * sphi * copysign(1.0, -u[0]);
  flift[1] = pre * cl * (umag * umag) * cphi * (u[1] / syz);
  flift[2] = pre * cl * (umag * umag) * cphi * (u[2] / syz);
  oc_transform_basis_vec(q, fdrag);
  oc_transform_basis_vec(q, flift);
  if (cd != cd) return OC_ERR_NAN_DRAG;
  return OC_OK;
}

/* ------------------------------------------------------------------------ */
/* FLUID_TORQUE_ELLIPSOID fluid_torque_ellipsoid.f90:1-190.  The relative
 * velocity uses the stage-n particle velocity PART%UP (upn..) in every RK
 * stage (:88-105); the result stays in the body frame.                     */
/* ------------------------------------------------------------------------ */
int oc_fluid_torque(const oc_params *p, const double *sp, double ufap,
                    double vfap, double wfap, double upn, double vpn,
                    double wpn, double vxap, double vyap, double vzap,
                    double ox, double oy, double oz, const double q[4],
                    double tp[3], double tr[3]) {
  const double c1 = 0.935, c2 = 0.146, c3 = -0.469, c4 = 0.145, c5 = 0.116,
               c6 = 0.748, c7 = 0.041, c8 = 0.221, c9 = 0.657, c10 = 0.044;
  const double r1 = 0.573, r2 = -0.154, r3 = 116.61, r4 = 1.0;
  const double r1_ = 1.244, r2_ = 0.239, r3_ = 378.12, r4_ = 0.789;
  double pmass = sp[SP_PMASS];
  double ur[3] = {ufap - upn, vfap - vpn, wfap - wpn};
  double urel_mag = sqrt(ur[0] * ur[0] + ur[1] * ur[1] + ur[2] * ur[2]);
  double deq = sp[SP_DEQ];
  double rep = (urel_mag * deq) / p->visc;
  double u[3] = {ufap - upn, vfap - vpn, wfap - wpn};
  rot_conj(q, u);
  double phi = fabs(atan((sqrt((u[1]) * (u[1]) + u[2] * u[2])) / (u[0])));
  double umag = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  double curl[3] = {vxap, vyap, vzap};
  rot_conj(q, curl);
  double orel[3] = {0.5 * curl[0] - ox, 0.5 * curl[1] - oy, 0.5 * curl[2] - oz};
  double orel_mag = sqrt(orel[0] * orel[0] + orel[1] * orel[1] + orel[2] * orel[2]);
  double orel_yz = sqrt(orel[2] * orel[2] + orel[1] * orel[1]);
  double re_r = (orel_mag * (deq * deq)) / (p->visc);
  /* separate sin and cos here: matches the reference build bit for bit */
  double ctp = (c1 / pow(rep, c2) + c3 / pow(rep, c4)) *
               (pow(sin(phi), (c5 + c6 * (pow(rep, c7))))) *
               (pow(cos(phi), (c8 + c9 * (pow(rep, c10)))));
  double ctr = r1 * (pow(re_r, r2)) + r3 / (pow(re_r, r4));
  double ctr2 = r1_ * (pow(re_r, r2_)) + r3_ / (pow(re_r, r4_));
  double syz = sqrt(u[1] * u[1] + u[2] * u[2]);
  double pre = (1.0 / 4.0) * (p->rhof / pmass) * (1.0 / 4.0) * PPI * ipow3(deq) * ctp *
               (umag * umag);
  tp[0] = 0.0;
  tp[1] = pre * (fabs(u[2]) / syz) * copysign(1.0, u[0] * u[2]);
  tp[2] = pre * (fabs(u[1]) / syz) * copysign(1.0, u[0] * u[1]);
  double h5 = ipow5(deq / 2.0);
  tr[0] = ctr * (0.5 * p->rhof / pmass) * h5 * fabs(orel[0]) * orel[0];
  tr[1] = ctr2 * (0.5 * p->rhof / pmass) * h5 * orel_yz * orel[1];
  tr[2] = ctr2 * (0.5 * p->rhof / pmass) * h5 * orel_yz * orel[2];
  if (ox != ox) return OC_ERR_NAN_OMEGA;
  return OC_OK;
}

/* ------------------------------------------------------------------------ */
/* wall path                                                                */
/* ------------------------------------------------------------------------ */
/* ELL = R diag(1/a^2,1/b^2,1/c^2) R^T, closest_point_plane.f90:74-87 */
static m33 ell_quadric(const double *sp, const double q[4]) {
  m33 e;
  memset(&e, 0, sizeof e);
  e.m[0][0] = 1.0 / (sp[SP_A] * sp[SP_A]);
  e.m[1][1] = 1.0 / (sp[SP_B] * sp[SP_B]);
  e.m[2][2] = 1.0 / (sp[SP_C] * sp[SP_C]);
  return rot_tensor(q, &e);
}

/* WALL_BOUNDARY_CONDITION_NEW ellipsoid_wall_boundary_new.f90:1-112 with
 * closest_point_plane.f90 and max_encroachment_ellipsoid.f90 */
void oc_wall_test(const oc_params *p, const double *sp, double x, double y,
                  double z, const double q[4], double nrm[3], double arm[3],
                  double *delta_x, int *intersect) {
  double wpos[3];
  arm[0] = arm[1] = arm[2] = 0.0;
  *intersect = 0;
  if (x < p->L / 2.0) {
    wpos[0] = 0.0; nrm[0] = 1.0;
  } else {
    wpos[0] = p->L; nrm[0] = -1.0;
  }
  wpos[1] = y; wpos[2] = z; nrm[1] = 0.0; nrm[2] = 0.0;
  double c[3] = {x, y, z};
  m33 ell = ell_quadric(sp, q);
  m33 iell = inv3(&ell);
  double in_[3];
  mv(&iell, nrm, in_);
  double tnin = dot3(nrm, in_);
  double d[3] = {wpos[0] - c[0], wpos[1] - c[1], wpos[2] - c[2]};
  double tnwc = dot3(nrm, d);
  double lag = tnwc / tnin;
  double pv[3], vec[3], ev[3];
  for (int i = 0; i < 3; ++i) pv[i] = c[i] + lag * in_[i];
  for (int i = 0; i < 3; ++i) vec[i] = pv[i] - c[i];
  mv(&ell, vec, ev);
  double margin = dot3(vec, ev);
  if (margin < 1.0) {
    *intersect = 1;
    /* max_encroachment_ellipsoid.f90:67-95 */
    double sq = sqrt(tnin);
    double pt[3];
    for (int i = 0; i < 3; ++i) pt[i] = c[i] - (in_[i] / sq);
    arm[0] = (x - pt[0]); arm[1] = (y - pt[1]); arm[2] = (z - pt[2]);
    *delta_x = wpos[0] - pt[0];
  }
}

/* WALL_ELLIPSOID_REBOUND wall_ellipsoid_rebound.f90:1-386 */
int oc_wall_rebound(const oc_params *p, const double *sp, double *u,
                    double *v, double *w, double *ox, double *oy, double *oz,
                    const double q[4], const double R[3], const double n[3]) {
  double om[3] = {*ox, *oy, *oz};
  oc_transform_basis_vec(q, om);
  double vrx = *u + (om[1] * R[2] - om[2] * R[1]);
  double vry = *v + (om[2] * R[0] - om[0] * R[2]);
  double vrz = *w + (om[0] * R[1] - om[1] * R[0]);
  double vrn = vrx * n[0] + vry * n[1] + vrz * n[2];
  double dv[3] = {vrx + p->ecw * vrn * n[0], vry + p->ecw * vrn * n[1],
                  vrz + p->ecw * vrn * n[2]};
  m33 ip; memset(&ip, 0, sizeof ip);
  ip.m[0][0] = sp[SP_IXX]; ip.m[1][1] = sp[SP_IYY]; ip.m[2][2] = sp[SP_IZZ];
  ip = rot_tensor(q, &ip);
  m33 iip = inv3(&ip);
  m33 rx;
  rx.m[0][0] = 0.0;   rx.m[0][1] = -R[2]; rx.m[0][2] = R[1];
  rx.m[1][0] = R[2];  rx.m[1][1] = 0.0;   rx.m[1][2] = -R[0];
  rx.m[2][0] = -R[1]; rx.m[2][1] = R[0];  rx.m[2][2] = 0.0;
  m33 rxt = transpose(&rx);
  m33 t1 = mm(&iip, &rx);
  m33 t2 = mm(&rxt, &t1);
  m33 K;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) K.m[i][j] = (i == j ? 1.0 : 0.0) + t2.m[i][j];
  m33 K1 = inv3(&K);
  double pc[3];
  mv(&K1, dv, pc);
  for (int i = 0; i < 3; ++i) pc[i] = -pc[i];
  double pcn = pc[0] * n[0] + pc[1] * n[1] + pc[2] * n[2];
  double ptx = pc[0] - pcn * n[0], pty = pc[1] - pcn * n[1], ptz = pc[2] - pcn * n[2];
  double pct = sqrt(ptx * ptx + pty * pty + ptz * ptz);
  if (pct > p->muw * pcn) {
    double tw[3] = {vrx - vrn * n[0], vry - vrn * n[1], vrz - vrn * n[2]};
    double ntw = sqrt(tw[0] * tw[0] + tw[1] * tw[1] + tw[2] * tw[2]);
    if (ntw != 0.0)
      for (int i = 0; i < 3; ++i) tw[i] = -tw[i] / ntw;
    double vimp[3], kn[3];
    for (int i = 0; i < 3; ++i) vimp[i] = n[i] + p->muw * tw[i];
    mv(&K, vimp, kn);
    double kndn = kn[0] * n[0] + kn[1] * n[1] + kn[2] * n[2];
    pcn = -1.0 * (dv[0] * n[0] + dv[1] * n[1] + dv[2] * n[2]);
    for (int i = 0; i < 3; ++i) pc[i] = (pcn / kndn) * vimp[i];
  }
  double am[3];
  double ei_t = 0.5 * (*u * *u + *v * *v + *w * *w);
  mv(&ip, om, am);
  double ei_r = 0.5 * (om[0] * am[0] + om[1] * am[1] + om[2] * am[2]);
  double ei = ei_t + ei_r;
  *u = *u + pc[0]; *v = *v + pc[1]; *w = *w + pc[2];
  double t3[3], t4[3];
  mv(&rx, pc, t3);
  mv(&iip, t3, t4);
  for (int i = 0; i < 3; ++i) om[i] = om[i] + t4[i];
  double ef_t = 0.5 * (*u * *u + *v * *v + *w * *w);
  mv(&ip, om, am);
  double ef_r = 0.5 * (om[0] * am[0] + om[1] * am[1] + om[2] * am[2]);
  double ef = ef_t + ef_r;
  rot_conj(q, om);
  *ox = om[0]; *oy = om[1]; *oz = om[2];
  if (fabs((ei - ef) / ei) * 100.0 > 1.0e-9) return OC_ERR_KE_WALL;
  return OC_OK;
}

/* ------------------------------------------------------------------------ */
/* ADV_PARTICLE_ELLP adv_particle_new.f90:264-645 for one particle.
 * fl = fluid velocity and vorticity at x^n (UFAP..VORTZAP).                */
/* ------------------------------------------------------------------------ */
static int advance_one(const oc_params *p, const double *sp, double *s,
                       const double fl[6]) {
  const double dt = p->dtime;
  const double ixx = sp[SP_IXX], iyy = sp[SP_IYY], izz = sp[SP_IZZ];
  const double grav = sp[SP_GRAV];
  double x = s[0], y = s[1], z = s[2], up = s[3], vp = s[4], wp = s[5];
  double q[4] = {s[6], s[7], s[8], s[9]};
  double ox = s[10], oy = s[11], oz = s[12];
  double fd[3], flf[3], tp[3], tr[3];
  int err = OC_OK, e;

  double k1x[3] = {up, vp, wp};
  e = oc_fluid_force(p, sp, up, vp, wp, fl[0], fl[1], fl[2], q, fd, flf);
  if (e && !err) err = e;
  double ff[3] = {fd[0] + flf[0], fd[1] + flf[1], fd[2] + flf[2]};
  e = oc_fluid_torque(p, sp, fl[0], fl[1], fl[2], up, vp, wp, fl[3], fl[4], fl[5],
                      ox, oy, oz, q, tp, tr);
  if (e && !err) err = e;
  double ft[3] = {tp[0] + tr[0], tp[1] + tr[1], tp[2] + tr[2]};
  double k1u[3] = {ff[0], ff[1], ff[2] + grav};
  double k1o[3] = {(ft[0] + oy * oz * (iyy - izz)) / ixx,
                   (ft[1] + oz * ox * (izz - ixx)) / iyy,
                   (ft[2] + ox * oy * (ixx - iyy)) / izz};
  double xp1 = x + dt * k1x[0], yp1 = y + dt * k1x[1], zp1 = z + dt * k1x[2];
  double up1 = up + dt * k1u[0], vp1 = vp + dt * k1u[1], wp1 = wp + dt * k1u[2];
  double o1x = ox + dt * k1o[0], o1y = oy + dt * k1o[1], o1z = oz + dt * k1o[2];
  double q1[4] = {q[0], q[1], q[2], q[3]};
  oc_quaternion_integration(q1, o1x, o1y, o1z, dt);

  double wn[3], arm[3], dxw = 0.0;
  int hit;
  oc_wall_test(p, sp, xp1, yp1, zp1, q1, wn, arm, &dxw, &hit);
  double olab[3] = {o1x, o1y, o1z};
  oc_transform_basis_vec(q1, olab);
  double vpc[3] = {up1 + (olab[1] * arm[2] - olab[2] * arm[1]),
                   vp1 + (olab[2] * arm[0] - olab[0] * arm[2]),
                   wp1 + (olab[0] * arm[1] - olab[1] * arm[0])};

  if (hit && up1 * wn[0] < 0.0) {
    /* bounced particle, adv_particle_new.f90:358-511 */
    double xw = xp1 + dxw;
    xp1 = xw;
    double delt = fabs(dxw) / sqrt(vpc[0] * vpc[0] + vpc[1] * vpc[1] + vpc[2] * vpc[2]);
    yp1 = y + (dt - delt) * k1x[1];
    zp1 = z + (dt - delt) * k1x[2];
    up1 = up + (dt - delt) * k1u[0];
    vp1 = vp + (dt - delt) * k1u[1];
    wp1 = wp + (dt - delt) * k1u[2];
    o1x = ox + (dt - delt) * k1o[0];
    o1y = oy + (dt - delt) * k1o[1];
    o1z = oz + (dt - delt) * k1o[2];
    q1[0] = q[0]; q1[1] = q[1]; q1[2] = q[2]; q1[3] = q[3];
    oc_quaternion_integration(q1, o1x, o1y, o1z, dt - delt);
    e = oc_wall_rebound(p, sp, &up1, &vp1, &wp1, &o1x, &o1y, &o1z, q1, arm, wn);
    if (e && !err) err = e;
    double k2x[3] = {up1, vp1, wp1};
    e = oc_fluid_force(p, sp, up1, vp1, wp1, fl[0], fl[1], fl[2], q1, fd, flf);
    if (e && !err) err = e;
    double ff2[3] = {fd[0] + flf[0], fd[1] + flf[1], fd[2] + flf[2]};
    e = oc_fluid_torque(p, sp, fl[0], fl[1], fl[2], up, vp, wp, fl[3], fl[4], fl[5],
                        o1x, o1y, o1z, q1, tp, tr);
    if (e && !err) err = e;
    double ft2[3] = {tp[0] + tr[0], tp[1] + tr[1], tp[2] + tr[2]};
    double k2u[3] = {ff2[0], ff2[1], ff2[2] + grav};
    double k2o[3] = {(ft2[0] + o1y * o1z * (iyy - izz)) / ixx,
                     (ft2[1] + o1z * o1x * (izz - ixx)) / iyy,
                     (ft2[2] + o1x * o1y * (ixx - iyy)) / izz};
    s[0] = xp1 + (delt) * k2x[0];
    s[1] = yp1 + (delt) * k2x[1];
    s[2] = zp1 + (delt) * k2x[2];
    s[3] = up1 + (delt) * k2u[0];
    s[4] = vp1 + (delt) * k2u[1];
    s[5] = wp1 + (delt) * k2u[2];
    s[10] = o1x + (delt) * k2o[0];
    s[11] = o1y + (delt) * k2o[1];
    s[12] = o1z + (delt) * k2o[2];
    oc_quaternion_integration(q1, s[10], s[11], s[12], delt);
    s[6] = q1[0]; s[7] = q1[1]; s[8] = q1[2]; s[9] = q1[3];
  } else {
    /* unbounced particle, adv_particle_new.f90:514-640 */
    double k2x[3] = {up1, vp1, wp1};
    e = oc_fluid_force(p, sp, up1, vp1, wp1, fl[0], fl[1], fl[2], q1, fd, flf);
    if (e && !err) err = e;
    double ff2[3] = {fd[0] + flf[0], fd[1] + flf[1], fd[2] + flf[2]};
    e = oc_fluid_torque(p, sp, fl[0], fl[1], fl[2], up, vp, wp, fl[3], fl[4], fl[5],
                        o1x, o1y, o1z, q1, tp, tr);
    if (e && !err) err = e;
    double ft2[3] = {tp[0] + tr[0], tp[1] + tr[1], tp[2] + tr[2]};
    double k2u[3] = {ff2[0], ff2[1], ff2[2] + grav};
    double k2o[3] = {(ft2[0] + o1y * o1z * (iyy - izz)) / ixx,
                     (ft2[1] + o1z * o1x * (izz - ixx)) / iyy,
                     (ft2[2] + o1x * o1y * (ixx - iyy)) / izz};
    double xp2 = x + 0.5 * dt * (k1x[0] + k2x[0]);
    double yp2 = y + 0.5 * dt * (k1x[1] + k2x[1]);
    double zp2 = z + 0.5 * dt * (k1x[2] + k2x[2]);
    oc_wall_test(p, sp, xp2, yp2, zp2, q1, wn, arm, &dxw, &hit);
    if (hit) {
      s[0] = x + dt * k1x[0]; s[1] = y + dt * k1x[1]; s[2] = z + dt * k1x[2];
      s[3] = up + dt * k1u[0]; s[4] = vp + dt * k1u[1]; s[5] = wp + dt * k1u[2];
      s[10] = ox + dt * k1o[0]; s[11] = oy + dt * k1o[1]; s[12] = oz + dt * k1o[2];
    } else {
      s[0] = x + 0.5 * dt * (k1x[0] + k2x[0]);
      s[1] = y + 0.5 * dt * (k1x[1] + k2x[1]);
      s[2] = z + 0.5 * dt * (k1x[2] + k2x[2]);
      s[3] = up + 0.5 * dt * (k1u[0] + k2u[0]);
      s[4] = vp + 0.5 * dt * (k1u[1] + k2u[1]);
      s[5] = wp + 0.5 * dt * (k1u[2] + k2u[2]);
      s[10] = ox + 0.5 * dt * (k1o[0] + k2o[0]);
      s[11] = oy + 0.5 * dt * (k1o[1] + k2o[1]);
      s[12] = oz + 0.5 * dt * (k1o[2] + k2o[2]);
    }
    oc_quaternion_integration(q1, s[10], s[11], s[12], dt);
    s[6] = q1[0]; s[7] = q1[1]; s[8] = q1[2]; s[9] = q1[3];
  }
  if (p->renorm_quat) {
    double nq = sqrt(s[6] * s[6] + s[7] * s[7] + s[8] * s[8] + s[9] * s[9]);
    for (int i = 6; i < 10; ++i) s[i] = s[i] / nq;
  }
  return err;
}

/* BOUNDARY_PARTICLE boundary_particle.f90:45-64 (NPROC = 1) */
void oc_boundary_particle(const oc_params *p, int np, double *st) {
  double L = p->L;
  for (int c = 0; c < 3; ++c) {
    double *x = st + (int64_t)c * np;
    for (int i = 0; i < np; ++i) {
      if (x[i] > L) x[i] = x[i] - L;
      if (x[i] < 0.) x[i] = x[i] + L;
    }
  }
}

/* ADV_PARTICLE_ELLP adv_particle_new.f90:139-686: gather the six fields at
 * x^n (interp gating: FROZEN_FLOW > 1), advance, then wrap.
 * fluid_at_p (optional, double[6][np]) receives UFAP..VORTZAP. */
int oc_adv_particle_ellp(const oc_params *p, const double *field,
                         const double *spec, int np, double *state,
                         const int *species, double *fluid_at_p, int nthreads) {
  int err = OC_OK;
  const int64_t fs = oc_field_size(p->n);
  (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (int i = 0; i < np; ++i) {
    double s[OC_NSTATE], fl[6];
    for (int c = 0; c < OC_NSTATE; ++c) s[c] = state[(int64_t)c * np + i];
    int e = OC_OK;
    for (int c = 0; c < 6 && !e; ++c)
      e = interp_one(p->interp_scheme, p->n, p->dx, field + c * fs, s[0], s[1], s[2], &fl[c]);
    if (fluid_at_p)
      for (int c = 0; c < 6; ++c) fluid_at_p[(int64_t)c * np + i] = fl[c];
    if (!e) e = advance_one(p, spec + (int64_t)species[i] * OC_SP_STRIDE, s, fl);
    for (int c = 0; c < OC_NSTATE; ++c) state[(int64_t)c * np + i] = s[c];
    if (e) {
#pragma omp critical
      if (!err || e < err) err = e;
    }
  }
  oc_boundary_particle(p, np, state);
  return err;
}

/* ------------------------------------------------------------------------ */
/* collision cell list                                                      */
/* ------------------------------------------------------------------------ */
/* COLLISION_INITIATION collision_initiation.f90:50-102, INIT_GRID = 2,
 * NPROC = 1 (DELTA_COLL = ZERO); dpmax = 2 * max(EMAJ_PART). */
void oc_collision_grid_init(const oc_params *p, double dpmax, oc_cgrid *g) {
  double L = p->L, dy = p->dx;
  double dxbox = 2.0 * dpmax;
  int nxbox = (int)(L / dxbox);
  dxbox = L / (double)nxbox;
  double delta = ZERO_R;
  g->xminbox = 0.0;                               /* XMESH(1) */
  double ymesh_end = dy * (double)(p->n - 1);     /* YMESH(NY) */
  g->yminbox = 0.0 - delta;
  double ymaxbox = ymesh_end + dy + delta;
  g->zminbox = 0.0 - delta;
  double zmaxbox = ymesh_end + dy + delta;
  g->nxbox = nxbox;
  g->nybox = (int)((ymaxbox - g->yminbox) / dxbox);
  g->nzbox = (int)((zmaxbox - g->zminbox) / dxbox);
  g->dxbox = dxbox;
  g->dybox = (ymaxbox - g->yminbox) / (double)g->nybox;
  g->dzbox = (zmaxbox - g->zminbox) / (double)g->nzbox;
  g->nbox = g->nxbox * g->nybox * g->nzbox;
}

/* COLLISION_BOX_CONNECTIVITY collision_box_connectivity.f90:59-316:
 * 13-neighbour half shell, 1-based box numbers, con[(ijk-1)*13 + n]. */
void oc_box_connectivity(const oc_cgrid *g, int periodic, int *con, int *ncon) {
  int NX = g->nxbox, NY = g->nybox, NZ = g->nzbox;
  for (int i0 = 1; i0 <= NX; ++i0)
    for (int j0 = 1; j0 <= NY; ++j0)
      for (int k0 = 1; k0 <= NZ; ++k0) {
        int ijk = i0 + (j0 - 1) * NX + (k0 - 1) * NX * NY;
        int im1 = i0 - 1, ip1 = i0 + 1, jm1 = j0 - 1, jp1 = j0 + 1, km1 = k0 - 1, kp1 = k0 + 1;
        if (im1 < 1 && periodic) im1 = NX;
        if (ip1 > NX && periodic) ip1 = 1;
        if (jm1 < 1) jm1 = NY;
        if (jp1 > NY) jp1 = 1;
        if (km1 < 1) km1 = NZ;
        if (kp1 > NZ) kp1 = 1;
        int n = 0;
        int *c = con + (int64_t)(ijk - 1) * 13;
#define ID(I, J, K) ((I) + ((J)-1) * NX + ((K)-1) * NX * NY)
        if (km1 >= 1 && jp1 <= NY && ip1 <= NX) c[n++] = ID(ip1, jp1, km1);
        if (jm1 >= 1 && ip1 <= NX) c[n++] = ID(ip1, jm1, k0);
        if (ip1 <= NX) c[n++] = ID(ip1, j0, k0);
        if (jp1 <= NY) {
          c[n++] = ID(i0, jp1, k0);
          if (ip1 <= NX) c[n++] = ID(ip1, jp1, k0);
        }
        if (kp1 <= NZ) {
          if (jm1 >= 1) {
            c[n++] = ID(i0, jm1, kp1);
            if (ip1 <= NX) c[n++] = ID(ip1, jm1, kp1);
          }
          if (im1 >= 1) c[n++] = ID(im1, j0, kp1);
          c[n++] = ID(i0, j0, kp1);
          if (ip1 <= NX) c[n++] = ID(ip1, j0, kp1);
          if (jp1 <= NY) {
            if (im1 >= 1) c[n++] = ID(im1, jp1, kp1);
            c[n++] = ID(i0, jp1, kp1);
            if (ip1 <= NX) c[n++] = ID(ip1, jp1, kp1);
          }
        }
#undef ID
        ncon[ijk - 1] = n;
      }
}

/* COLLISION_DETECTION_CONNECTIVITY collision_detection_connectivity.f90:60-
 * 207: returns 1-based HOME/NEAR in the reference order. */
int oc_collision_detection(const oc_cgrid *g, const int *con, const int *ncon,
                           int np, const double *x, const double *y,
                           const double *z, int64_t nclosemax, int *home,
                           int *near_, int64_t *nclose) {
  int nbox = g->nbox, err = OC_OK;
  int NX = g->nxbox, NY = g->nybox;
  int *cell = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
  int *cnt = (int *)calloc((size_t)nbox + 1, sizeof(int));
  int *off = (int *)calloc((size_t)nbox + 2, sizeof(int));
  int *mem = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
  for (int i = 0; i < np; ++i) {
    int i0 = (int)((x[i] - g->xminbox) / g->dxbox) + 1;
    int j0 = (int)((y[i] - g->yminbox) / g->dybox) + 1;
    int k0 = (int)((z[i] - g->zminbox) / g->dzbox) + 1;
    int ijk = i0 + (j0 - 1) * NX + (k0 - 1) * NX * NY;
    if (i0 < 1 || i0 > NX || j0 < 1 || j0 > NY || k0 < 1 || k0 > g->nzbox) {
      if (!err) err = OC_ERR_OUT_OF_GRID;
      if (ijk < 1 || ijk > nbox) { free(cell); free(cnt); free(off); free(mem); *nclose = 0; return OC_ERR_OUT_OF_GRID; }
    }
    cell[i] = ijk;
    cnt[ijk]++;
  }
  for (int c = 1; c <= nbox; ++c) off[c + 1] = off[c] + cnt[c];
  {
    int *fill = (int *)calloc((size_t)nbox + 1, sizeof(int));
    for (int i = 0; i < np; ++i) mem[off[cell[i]] + fill[cell[i]]++] = i + 1;
    free(fill);
  }
  int64_t np0 = 0;
  for (int ijk = 1; ijk <= nbox; ++ijk) {
    int nh = cnt[ijk];
    if (!nh) continue;
    const int *hm = mem + off[ijk];
    for (int a = 0; a < nh; ++a) {
      for (int b = a + 1; b < nh; ++b) {
        if (np0 < nclosemax) { home[np0] = hm[a]; near_[np0] = hm[b]; }
        np0++;
      }
      for (int nv = 0; nv < ncon[ijk - 1]; ++nv) {
        int nb = con[(int64_t)(ijk - 1) * 13 + nv];
        const int *nm = mem + off[nb];
        for (int b = 0; b < cnt[nb]; ++b) {
          if (np0 < nclosemax) { home[np0] = hm[a]; near_[np0] = nm[b]; }
          np0++;
        }
      }
    }
  }
  *nclose = np0;
  if (np0 > nclosemax) err = OC_ERR_NCLOSE_OVERFLOW;
  free(cell); free(cnt); free(off); free(mem);
  return err;
}

/* ------------------------------------------------------------------------ */
/* MINPACK lmder1 / lmder / lmpar / qrfac / qrsolv / enorm                  */
/* (minpack_reduc.f90:180-1289), specialised to the contact problem's       */
/* callback but otherwise general in m and n.                               */
/* ------------------------------------------------------------------------ */
#define EPSMCH 2.220446049250313e-16
#define DWARF 2.2250738585072014e-308

static double enorm(int n, const double *x) {   /* minpack_reduc.f90:180-232 */
  const double rdwarf = 3.834e-20, rgiant = 1.304e19;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, x1max = 0.0, x3max = 0.0;
  double agiant = rgiant / (double)n, r;
  for (int i = 0; i < n; ++i) {
    double xabs = fabs(x[i]);
    if (xabs > rdwarf && xabs < agiant) {
      s2 = s2 + xabs * xabs;
    } else if (xabs <= rdwarf) {
      if (xabs <= x3max) {
        if (xabs != 0.0) { double t = xabs / x3max; s3 = s3 + t * t; }
      } else {
        double t = x3max / xabs; s3 = 1.0 + s3 * (t * t); x3max = xabs;
      }
    } else if (xabs <= x1max) {
      double t = xabs / x1max; s1 = s1 + t * t;
    } else {
      double t = x1max / xabs; s1 = 1.0 + s1 * (t * t); x1max = xabs;
    }
  }
  if (s1 != 0.0) {
    r = x1max * sqrt(s1 + (s2 / x1max) / x1max);
  } else if (s2 == 0.0) {
    r = x3max * sqrt(s3);
  } else {
    r = 0.0;
    if (s2 >= x3max) r = sqrt(s2 * (1.0 + (x3max / s2) * (x3max * s3)));
    if (s2 < x3max) r = sqrt(x3max * ((s2 / x3max) + (x3max * s3)));
  }
  return r;
}

/* qrfac with pivoting, minpack_reduc.f90:1007-1120; a is column-major lda */
static void qrfac(int m, int n, double *a, int lda, int *ipvt, double *rdiag,
                  double *acnorm, double *wa) {
  for (int j = 0; j < n; ++j) {
    acnorm[j] = enorm(m, a + j * lda);
    rdiag[j] = acnorm[j];
    wa[j] = rdiag[j];
    ipvt[j] = j;
  }
  int minmn = m < n ? m : n;
  for (int j = 0; j < minmn; ++j) {
    int kmax = j;
    for (int k = j; k < n; ++k)
      if (rdiag[k] > rdiag[kmax]) kmax = k;
    if (kmax != j) {
      for (int i = 0; i < m; ++i) {
        double t = a[i + j * lda]; a[i + j * lda] = a[i + kmax * lda]; a[i + kmax * lda] = t;
      }
      rdiag[kmax] = rdiag[j];
      wa[kmax] = wa[j];
      int k = ipvt[j]; ipvt[j] = ipvt[kmax]; ipvt[kmax] = k;
    }
    double ajnorm = enorm(m - j, a + j + j * lda);
    if (ajnorm != 0.0) {
      if (a[j + j * lda] < 0.0) ajnorm = -ajnorm;
      for (int i = j; i < m; ++i) a[i + j * lda] = a[i + j * lda] / ajnorm;
      a[j + j * lda] = a[j + j * lda] + 1.0;
      for (int k = j + 1; k < n; ++k) {
        double sum = 0.0;
        for (int i = j; i < m; ++i) sum = sum + a[i + j * lda] * a[i + k * lda];
        double temp = sum / a[j + j * lda];
        for (int i = j; i < m; ++i) a[i + k * lda] = a[i + k * lda] - temp * a[i + j * lda];
        if (rdiag[k] != 0.0) {
          temp = a[j + k * lda] / rdiag[k];
          double t2 = 1.0 - temp * temp;
          rdiag[k] = rdiag[k] * sqrt(t2 > 0.0 ? t2 : 0.0);
          double t3 = rdiag[k] / wa[k];
          if (5.0e-2 * (t3 * t3) <= EPSMCH) {
            rdiag[k] = enorm(m - j - 1, a + (j + 1) + k * lda);
            wa[k] = rdiag[k];
          }
        }
      }
    }
    rdiag[j] = -ajnorm;
  }
}

/* qrsolv minpack_reduc.f90:1155-1289 */
static void qrsolv(int n, double *r, int ldr, const int *ipvt,
                   const double *diag, const double *qtb, double *x,
                   double *sdiag, double *wa) {
  for (int j = 0; j < n; ++j) {
    for (int i = j; i < n; ++i) r[i + j * ldr] = r[j + i * ldr];
    x[j] = r[j + j * ldr];
    wa[j] = qtb[j];
  }
  for (int j = 0; j < n; ++j) {
    int l = ipvt[j];
    if (diag[l] != 0.0) {
      for (int k = j; k < n; ++k) sdiag[k] = 0.0;
      sdiag[j] = diag[l];
      double qtbpj = 0.0;
      for (int k = j; k < n; ++k) {
        if (sdiag[k] != 0.0) {
          double cs, sn;
          if (fabs(r[k + k * ldr]) >= fabs(sdiag[k])) {
            double tn = sdiag[k] / r[k + k * ldr];
            cs = 0.5 / sqrt(0.25 + 0.25 * (tn * tn));
            sn = cs * tn;
          } else {
            double ct = r[k + k * ldr] / sdiag[k];
            sn = 0.5 / sqrt(0.25 + 0.25 * (ct * ct));
            cs = sn * ct;
          }
          r[k + k * ldr] = cs * r[k + k * ldr] + sn * sdiag[k];
          double temp = cs * wa[k] + sn * qtbpj;
          qtbpj = -(sn * wa[k]) + cs * qtbpj;
          wa[k] = temp;
          for (int i = k + 1; i < n; ++i) {
            temp = cs * r[i + k * ldr] + sn * sdiag[i];
            sdiag[i] = -(sn * r[i + k * ldr]) + cs * sdiag[i];
            r[i + k * ldr] = temp;
          }
        }
      }
    }
    sdiag[j] = r[j + j * ldr];
    r[j + j * ldr] = x[j];
  }
  int nsing = n;
  for (int j = 0; j < n; ++j) {
    if (sdiag[j] == 0.0 && nsing == n) nsing = j;
    if (nsing < n) wa[j] = 0.0;
  }
  for (int kk = 0; kk < nsing; ++kk) {
    int j = nsing - 1 - kk;
    double sum = 0.0;
    for (int i = j + 1; i < nsing; ++i) sum = sum + r[i + j * ldr] * wa[i];
    wa[j] = (wa[j] - sum) / sdiag[j];
  }
  for (int j = 0; j < n; ++j) x[ipvt[j]] = wa[j];
}

static double fmaxf_(double a, double b) { return (a > b) ? a : b; }
static double fminf_(double a, double b) { return (a < b) ? a : b; }

/* lmpar minpack_reduc.f90:798-987 */
static void lmpar(int n, double *r, int ldr, const int *ipvt,
                  const double *diag, const double *qtb, double delta,
                  double *par, double *x, double *sdiag, double *wa1,
                  double *wa2) {
  const double p1 = 1.0e-1, p001 = 1.0e-3;
  int nsing = n;
  for (int j = 0; j < n; ++j) {
    wa1[j] = qtb[j];
    if (r[j + j * ldr] == 0.0 && nsing == n) nsing = j;
    if (nsing < n) wa1[j] = 0.0;
  }
  for (int kk = 0; kk < nsing; ++kk) {
    int j = nsing - 1 - kk;
    wa1[j] = wa1[j] / r[j + j * ldr];
    double temp = wa1[j];
    for (int i = 0; i < j; ++i) wa1[i] = wa1[i] - r[i + j * ldr] * temp;
  }
  for (int j = 0; j < n; ++j) x[ipvt[j]] = wa1[j];
  int iter = 0;
  for (int j = 0; j < n; ++j) wa2[j] = diag[j] * x[j];
  double dxnorm = enorm(n, wa2);
  double fp = dxnorm - delta;
  if (fp <= p1 * delta) {
    if (iter == 0) *par = 0.0;
    return;
  }
  double parl = 0.0;
  if (nsing >= n) {
    for (int j = 0; j < n; ++j) { int l = ipvt[j]; wa1[j] = diag[l] * (wa2[l] / dxnorm); }
    for (int j = 0; j < n; ++j) {
      double sum = 0.0;
      for (int i = 0; i < j; ++i) sum = sum + r[i + j * ldr] * wa1[i];
      wa1[j] = (wa1[j] - sum) / r[j + j * ldr];
    }
    double temp = enorm(n, wa1);
    parl = ((fp / delta) / temp) / temp;
  }
  for (int j = 0; j < n; ++j) {
    double sum = 0.0;
    for (int i = 0; i <= j; ++i) sum = sum + r[i + j * ldr] * qtb[i];
    int l = ipvt[j];
    wa1[j] = sum / diag[l];
  }
  double gnorm = enorm(n, wa1);
  double paru = gnorm / delta;
  if (paru == 0.0) paru = DWARF / fminf_(delta, p1);
  *par = fmaxf_(*par, parl);
  *par = fminf_(*par, paru);
  if (*par == 0.0) *par = gnorm / dxnorm;
  for (;;) {
    iter = iter + 1;
    if (*par == 0.0) *par = fmaxf_(DWARF, p001 * paru);
    double temp = sqrt(*par);
    for (int j = 0; j < n; ++j) wa1[j] = temp * diag[j];
    qrsolv(n, r, ldr, ipvt, wa1, qtb, x, sdiag, wa2);
    for (int j = 0; j < n; ++j) wa2[j] = diag[j] * x[j];
    dxnorm = enorm(n, wa2);
    temp = fp;
    fp = dxnorm - delta;
    if (fabs(fp) <= p1 * delta || (parl == 0.0 && fp <= temp && temp < 0.0) || iter == 10) {
      if (iter == 0) *par = 0.0;
      break;
    }
    for (int j = 0; j < n; ++j) { int l = ipvt[j]; wa1[j] = diag[l] * (wa2[l] / dxnorm); }
    for (int j = 0; j < n; ++j) {
      wa1[j] = wa1[j] / sdiag[j];
      double t = wa1[j];
      for (int i = j + 1; i < n; ++i) wa1[i] = wa1[i] - r[i + j * ldr] * t;
    }
    temp = enorm(n, wa1);
    double parc = ((fp / delta) / temp) / temp;
    if (fp > 0.0) parl = fmaxf_(parl, *par);
    if (fp < 0.0) paru = fminf_(paru, *par);
    *par = fmaxf_(parl, *par + parc);
  }
}

/* ---- contact problem: ellipsoid_contact_detection.f90:158-330 ---------- */
typedef struct {
  m33 E1, E2;        /* calculate_ELLIPSOID_matrix for each particle */
  double b1[3], b2[3];
} contact_ctx;

/* calculate_ELLIPSOID_matrix calculate_ellipsoid_matrix.f90:33-71 */
static m33 ell_matrix(const double *sp, const double q[4]) {
  m33 a; memset(&a, 0, sizeof a);
  a.m[0][0] = sp[SP_A] * sp[SP_A];
  a.m[1][1] = sp[SP_B] * sp[SP_B];
  a.m[2][2] = sp[SP_C] * sp[SP_C];
  return rot_tensor(q, &a);
}

/* support points of both ellipsoids along x (pair_depth,
 * ellipsoid_pair_depth.f90:45-73) */
static void support(const contact_ctx *c, const double x[2], double cv1[3],
                    double pt1[3], double pt2[3], double *l1, double *l2) {
  double s0, c0, s1, c1;
  sc(x[0], &s0, &c0);
  sc(x[1], &s1, &c1);
  cv1[0] = c0 * c1;
  cv1[1] = s0 * c1;
  cv1[2] = s1;
  double cv2[3] = {-cv1[0], -cv1[1], -cv1[2]};
  double t1[3], t2[3];
  mv(&c->E1, cv1, t1);
  mv(&c->E2, cv2, t2);
  double lam1 = 0.25 * (cv1[0] * t1[0] + cv1[1] * t1[1] + cv1[2] * t1[2]);
  lam1 = sqrt(lam1);
  double lam2 = 0.25 * (cv2[0] * t2[0] + cv2[1] * t2[1] + cv2[2] * t2[2]);
  lam2 = sqrt(lam2);
  double h1 = 1.0 / (2.0 * lam1), h2 = 1.0 / (2.0 * lam2);
  double m1[3], m2[3];
  mv(&c->E1, cv1, m1);
  mv(&c->E2, cv2, m2);
  for (int i = 0; i < 3; ++i) pt1[i] = h1 * m1[i] + c->b1[i];
  for (int i = 0; i < 3; ++i) pt2[i] = h2 * m2[i] + c->b2[i];
  *l1 = lam1; *l2 = lam2;
}

static m33 coeff_grad(const m33 *E, const double cv[3], double lam) {
  /* (1/(2 lam)) E - (1/(8 lam^3)) E c c^T E */
  double row[3];
  for (int j = 0; j < 3; ++j) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + cv[k] * E->m[k][j];
    row[j] = s;
  }
  m33 o;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o.m[i][j] = cv[i] * row[j];
  m33 pm = mm(E, &o);
  /* the reference build evaluates lambda**3.0 as lambda*lambda*lambda */
  double s1 = 1.0 / (2.0 * lam), s3 = 1.0 / (8.0 * ipow3(lam));
  m33 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = s1 * E->m[i][j] - s3 * pm.m[i][j];
  return r;
}

/* fcn of ellipsoid_contact_detection.f90:158-330 (m = 3, n = 2) */
static void contact_fcn(const contact_ctx *c, const double x[2], double *fvec,
                        double *fjac, int iflag) {
  double cv1[3], pt1[3], pt2[3], l1, l2;
  support(c, x, cv1, pt1, pt2, &l1, &l2);
  double d[3] = {pt1[0] - pt2[0], pt1[1] - pt2[1], pt1[2] - pt2[2]};
  if (iflag == 1) {
    fvec[0] = d[0] * d[0]; fvec[1] = d[1] * d[1]; fvec[2] = d[2] * d[2];
    return;
  }
  double g[3][2], s0, c0, s1, c1;
  sc(x[0], &s0, &c0);
  sc(x[1], &s1, &c1);
  g[0][0] = -s0 * c1;
  g[1][0] = c0 * c1;
  g[2][0] = 0.0;
  g[0][1] = -c0 * s1;
  g[1][1] = -s0 * s1;
  g[2][1] = c1;
  double cv2[3] = {-cv1[0], -cv1[1], -cv1[2]};
  m33 C1 = coeff_grad(&c->E1, cv1, l1);
  m33 C2 = coeff_grad(&c->E2, cv2, l2);
  double gp1[3][2], gp2[3][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) {
      double s = 0.0, t = 0.0;
      for (int k = 0; k < 3; ++k) s = s + C1.m[i][k] * g[k][j];
      for (int k = 0; k < 3; ++k) t = t + C2.m[i][k] * (-g[k][j]);
      gp1[i][j] = s; gp2[i][j] = t;
    }
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < 3; ++i)
      fjac[i + 3 * j] = 2.0 * (d[i] * (gp1[i][j] - gp2[i][j]));
}

/* lmder minpack_reduc.f90:243-640 with lmder1's settings (:732-745):
 * mode 1, factor 100, maxfev 300, ftol = xtol = tol, gtol = 0 */
static int lmder_contact(const contact_ctx *c, double x[2], double tol,
                         int *nfev_out) {
  const int m = 3, n = 2, ldfjac = 3, maxfev = 100 * (n + 1);
  const double ftol = tol, xtol = tol, gtol = 0.0, factor = 100.0;
  const double p1 = 1.0e-1, p5 = 5.0e-1, p25 = 2.5e-1, p75 = 7.5e-1, p0001 = 1.0e-4;
  double fvec[3], fjac[6], diag[2], qtf[2], wa1[2], wa2[2], wa3[2], wa4[3];
  int ipvt[2];
  int info = 0, nfev = 0, iter;
  double actred, delta = 0.0, dirder, fnorm, fnorm1, gnorm, par, pnorm, prered,
         ratio, sum, temp, temp1, temp2, xnorm = 0.0;

  contact_fcn(c, x, fvec, fjac, 1);
  nfev = 1;
  fnorm = enorm(m, fvec);
  par = 0.0;
  iter = 1;
  for (;;) {                       /* outer */
    contact_fcn(c, x, fvec, fjac, 2);
    qrfac(m, n, fjac, ldfjac, ipvt, wa1, wa2, wa3);
    if (iter == 1) {
      for (int j = 0; j < n; ++j) {
        diag[j] = wa2[j];
        if (wa2[j] == 0.0) diag[j] = 1.0;
      }
      for (int j = 0; j < n; ++j) wa3[j] = diag[j] * x[j];
      xnorm = enorm(n, wa3);
      delta = factor * xnorm;
      if (delta == 0.0) delta = factor;
    }
    for (int i = 0; i < m; ++i) wa4[i] = fvec[i];
    for (int j = 0; j < n; ++j) {
      if (fjac[j + j * ldfjac] != 0.0) {
        sum = 0.0;
        for (int i = j; i < m; ++i) sum = sum + fjac[i + j * ldfjac] * wa4[i];
        temp = -sum / fjac[j + j * ldfjac];
        for (int i = j; i < m; ++i) wa4[i] = wa4[i] + fjac[i + j * ldfjac] * temp;
      }
      fjac[j + j * ldfjac] = wa1[j];
      qtf[j] = wa4[j];
    }
    gnorm = 0.0;
    if (fnorm != 0.0) {
      for (int j = 0; j < n; ++j) {
        int l = ipvt[j];
        if (wa2[l] != 0.0) {
          sum = 0.0;
          for (int i = 0; i <= j; ++i) sum = sum + fjac[i + j * ldfjac] * (qtf[i] / fnorm);
          gnorm = fmaxf_(gnorm, fabs(sum / wa2[l]));
        }
      }
    }
    if (gnorm <= gtol) info = 4;
    if (info != 0) break;
    for (int j = 0; j < n; ++j) diag[j] = fmaxf_(diag[j], wa2[j]);
    int done = 0;
    for (;;) {                     /* inner */
      lmpar(n, fjac, ldfjac, ipvt, diag, qtf, delta, &par, wa1, wa2, wa3, wa4);
      for (int j = 0; j < n; ++j) {
        wa1[j] = -wa1[j];
        wa2[j] = x[j] + wa1[j];
        wa3[j] = diag[j] * wa1[j];
      }
      pnorm = enorm(n, wa3);
      if (iter == 1) delta = fminf_(delta, pnorm);
      contact_fcn(c, wa2, wa4, fjac, 1);
      nfev = nfev + 1;
      fnorm1 = enorm(m, wa4);
      actred = -1.0;
      if (p1 * fnorm1 < fnorm) { double t = fnorm1 / fnorm; actred = 1.0 - t * t; }
      for (int j = 0; j < n; ++j) {
        wa3[j] = 0.0;
        int l = ipvt[j];
        temp = wa1[l];
        for (int i = 0; i <= j; ++i) wa3[i] = wa3[i] + fjac[i + j * ldfjac] * temp;
      }
      temp1 = enorm(n, wa3) / fnorm;
      temp2 = (sqrt(par) * pnorm) / fnorm;
      prered = temp1 * temp1 + (temp2 * temp2) / p5;
      dirder = -(temp1 * temp1 + temp2 * temp2);
      ratio = 0.0;
      if (prered != 0.0) ratio = actred / prered;
      if (ratio <= p25) {
        if (actred >= 0.0) temp = p5;
        if (actred < 0.0) temp = p5 * dirder / (dirder + p5 * actred);
        if (p1 * fnorm1 >= fnorm || temp < p1) temp = p1;
        delta = temp * fminf_(delta, pnorm / p1);
        par = par / temp;
      } else if (par == 0.0 || ratio >= p75) {
        delta = pnorm / p5;
        par = p5 * par;
      }
      if (ratio >= p0001) {
        for (int j = 0; j < n; ++j) {
          x[j] = wa2[j];
          wa2[j] = diag[j] * x[j];
        }
        for (int i = 0; i < m; ++i) fvec[i] = wa4[i];
        xnorm = enorm(n, wa2);
        fnorm = fnorm1;
        iter = iter + 1;
      }
      if (fabs(actred) <= ftol && prered <= ftol && p5 * ratio <= 1.0) info = 1;
      if (delta <= xtol * xnorm) info = 2;
      if (fabs(actred) <= ftol && prered <= ftol && p5 * ratio <= 1.0 && info == 2) info = 3;
      if (info != 0) { done = 1; break; }
      if (nfev >= maxfev) info = 5;
      if (fabs(actred) <= EPSMCH && prered <= EPSMCH && p5 * ratio <= 1.0) info = 6;
      if (delta <= EPSMCH * xnorm) info = 7;
      if (gnorm <= EPSMCH) info = 8;
      if (info != 0) { done = 1; break; }
      if (ratio >= p0001) break;
    }
    if (done) break;
  }
  if (info == 8) info = 4;
  *nfev_out = nfev;
  return info;
}

/* fmargin = margin_function mod_particle_parallel.f90:145-201 */
static double fmargin(const double *sp, const double pt[3], const double c[3],
                      const double q[4]) {
  m33 e = ell_quadric(sp, q);
  double v[3] = {pt[0] - c[0], pt[1] - c[1], pt[2] - c[2]};
  double ev[3];
  mv(&e, v, ev);
  return dot3(v, ev);
}

/* ellipsoid_contact_detection ellipsoid_contact_detection.f90:1-155.
 * seed (optional) replaces the centre-line initial guess (:77-87). */
int oc_contact_detection(const double *spi, const double *spj,
                         const double bi[3], const double qi[4],
                         const double bj[3], const double qj[4],
                         const double *seed, int *flag, double ept1[3],
                         double ept2[3], double cnormal[3], int *nfev,
                         double xsol[2]) {
  contact_ctx c;
  c.E1 = ell_matrix(spi, qi);
  c.E2 = ell_matrix(spj, qj);
  for (int i = 0; i < 3; ++i) { c.b1[i] = bi[i]; c.b2[i] = bj[i]; }
  const double tol = 1.4901161193847656e-08;   /* sqrt(epsilon(1d0)) */
  double x[2];
  if (seed) {
    x[0] = seed[0]; x[1] = seed[1];
  } else {
    double cn[3] = {bi[0] - bj[0], bi[1] - bj[1], bi[2] - bj[2]};
    double cnorm = sqrt(cn[0] * cn[0] + cn[1] * cn[1] + cn[2] * cn[2]);
    if (cnorm != 0.0)
      for (int i = 0; i < 3; ++i) cn[i] = cn[i] / cnorm;
    x[1] = asin(cn[2]);
    x[0] = acos(cn[0] / cos(x[1]));
  }
  int nf = 0;
  lmder_contact(&c, x, tol, &nf);
  double cv1[3], pt1[3], pt2[3], l1, l2;
  support(&c, x, cv1, pt1, pt2, &l1, &l2);
  double d[3] = {pt1[0] - pt2[0], pt1[1] - pt2[1], pt1[2] - pt2[2]};
  double contact = d[0] * cv1[0] + d[1] * cv1[1] + d[2] * cv1[2];
  int f = (contact > 0) && (fmargin(spi, pt2, bi, qi) < 1.0) &&
          (fmargin(spj, pt1, bj, qj) < 1.0);
  for (int i = 0; i < 3; ++i) { ept1[i] = pt1[i]; ept2[i] = pt2[i]; cnormal[i] = cv1[i]; }
  *flag = f;
  if (nfev) *nfev = nf;
  if (xsol) { xsol[0] = x[0]; xsol[1] = x[1]; }
  return OC_OK;
}

/* ------------------------------------------------------------------------ */
/* COLLISION_RESPONSE collision_response.f90:1-437.  s1/s2 point at 13-
 * component particle records (x..oz); sp1 is the species of the first
 * particle, whose inertia is used for both (:140-145).                     */
/* ------------------------------------------------------------------------ */
int oc_collision_response(const double *sp1, double *s1, double *s2,
                          const double q1[4], const double q2[4],
                          const double r1[3], const double r2[3],
                          const double n[3]) {
  double u1[3] = {s1[3], s1[4], s1[5]};
  double o1[3] = {s1[10], s1[11], s1[12]};
  oc_transform_basis_vec(q1, o1);
  double u2[3] = {s2[3], s2[4], s2[5]};
  double o2[3] = {s2[10], s2[11], s2[12]};
  oc_transform_basis_vec(q2, o2);
  double upc1[3] = {s1[3] + (o1[1] * r1[2] - o1[2] * r1[1]),
                    s1[4] + (o1[2] * r1[0] - o1[0] * r1[2]),
                    s1[5] + (o1[0] * r1[1] - o1[1] * r1[0])};
  double upc2[3] = {s2[3] + (o2[1] * r2[2] - o2[2] * r2[1]),
                    s2[4] + (o2[2] * r2[0] - o2[0] * r2[2]),
                    s2[5] + (o2[0] * r2[1] - o2[1] * r2[0])};
  double ur[3] = {upc2[0] - upc1[0], upc2[1] - upc1[1], upc2[2] - upc1[2]};
  double urn = ur[0] * n[0] + ur[1] * n[1] + ur[2] * n[2];
  double ecp = sp1[SP_ECP], mup = sp1[SP_MUP];
  double du[3];
  for (int i = 0; i < 3; ++i) du[i] = ur[i] + ecp * urn * n[i];
  m33 ip1; memset(&ip1, 0, sizeof ip1);
  ip1.m[0][0] = sp1[SP_IXX]; ip1.m[1][1] = sp1[SP_IYY]; ip1.m[2][2] = sp1[SP_IZZ];
  m33 ip2 = ip1;
  ip1 = rot_tensor(q1, &ip1);
  ip2 = rot_tensor(q2, &ip2);
  m33 iip1 = inv3(&ip1), iip2 = inv3(&ip2);
  m33 rx1, rx2;
  rx1.m[0][0] = 0.0;    rx1.m[0][1] = -r1[2]; rx1.m[0][2] = r1[1];
  rx1.m[1][0] = r1[2];  rx1.m[1][1] = 0.0;    rx1.m[1][2] = -r1[0];
  rx1.m[2][0] = -r1[1]; rx1.m[2][1] = r1[0];  rx1.m[2][2] = 0.0;
  rx2.m[0][0] = 0.0;    rx2.m[0][1] = -r2[2]; rx2.m[0][2] = r2[1];
  rx2.m[1][0] = r2[2];  rx2.m[1][1] = 0.0;    rx2.m[1][2] = -r2[0];
  rx2.m[2][0] = -r2[1]; rx2.m[2][1] = r2[0];  rx2.m[2][2] = 0.0;
  m33 rxt1 = transpose(&rx1), rxt2 = transpose(&rx2);
  m33 a1 = mm(&iip1, &rx1), a2 = mm(&iip2, &rx2);
  m33 b1 = mm(&rxt1, &a1), b2 = mm(&rxt2, &a2);
  m33 K1, K2, K;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      K1.m[i][j] = (i == j ? 1.0 : 0.0) + b1.m[i][j];
      K2.m[i][j] = (i == j ? 1.0 : 0.0) + b2.m[i][j];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) K.m[i][j] = K1.m[i][j] + K2.m[i][j];
  m33 iK = inv3(&K);
  double pc[3];
  mv(&iK, du, pc);
  for (int i = 0; i < 3; ++i) pc[i] = -pc[i];
  double pcn = pc[0] * n[0] + pc[1] * n[1] + pc[2] * n[2];
  double ptx = pc[0] - pcn * n[0], pty = pc[1] - pcn * n[1], ptz = pc[2] - pcn * n[2];
  double pct = sqrt(ptx * ptx + pty * pty + ptz * ptz);
  if (pct > mup * pcn) {
    double tw[3] = {ur[0] - urn * n[0], ur[1] - urn * n[1], ur[2] - urn * n[2]};
    double ntw = sqrt(tw[0] * tw[0] + tw[1] * tw[1] + tw[2] * tw[2]);
    if (ntw != 0.0)
      for (int i = 0; i < 3; ++i) tw[i] = -tw[i] / ntw;
    double vimp[3], kn[3];
    for (int i = 0; i < 3; ++i) vimp[i] = n[i] + mup * tw[i];
    mv(&K, vimp, kn);
    double kndn = kn[0] * n[0] + kn[1] * n[1] + kn[2] * n[2];
    pcn = -1.0 * (du[0] * n[0] + du[1] * n[1] + du[2] * n[2]);
    for (int i = 0; i < 3; ++i) pc[i] = (pcn / kndn) * vimp[i];
  }
  double am[3];
  double ei_t1 = 0.5 * (u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  mv(&ip1, o1, am);
  double ei_r1 = 0.5 * (o1[0] * am[0] + o1[1] * am[1] + o1[2] * am[2]);
  double ei1 = ei_t1 + ei_r1;
  double ei_t2 = 0.5 * (u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]);
  mv(&ip2, o2, am);
  double ei_r2 = 0.5 * (o2[0] * am[0] + o2[1] * am[1] + o2[2] * am[2]);
  double ei2 = ei_t2 + ei_r2;
  double ei = ei1 + ei2;
  s1[3] = s1[3] - pc[0]; s1[4] = s1[4] - pc[1]; s1[5] = s1[5] - pc[2];
  double npc[3] = {-pc[0], -pc[1], -pc[2]}, t1[3], t2[3];
  mv(&rx1, npc, t1);
  mv(&iip1, t1, t2);
  for (int i = 0; i < 3; ++i) o1[i] = o1[i] + t2[i];
  s2[3] = s2[3] + pc[0]; s2[4] = s2[4] + pc[1]; s2[5] = s2[5] + pc[2];
  mv(&rx2, pc, t1);
  mv(&iip2, t1, t2);
  for (int i = 0; i < 3; ++i) o2[i] = o2[i] + t2[i];
  double ef_t1 = 0.5 * (s1[3] * s1[3] + s1[4] * s1[4] + s1[5] * s1[5]);
  mv(&ip1, o1, am);
  double ef_r1 = 0.5 * (o1[0] * am[0] + o1[1] * am[1] + o1[2] * am[2]);
  double ef1 = ef_t1 + ef_r1;
  double ef_t2 = 0.5 * (s2[3] * s2[3] + s2[4] * s2[4] + s2[5] * s2[5]);
  mv(&ip2, o2, am);
  double ef_r2 = 0.5 * (o2[0] * am[0] + o2[1] * am[1] + o2[2] * am[2]);
  double ef2 = ef_t2 + ef_r2;
  double ef = ef1 + ef2;
  rot_conj(q1, o1);
  rot_conj(q2, o2);
  s1[10] = o1[0]; s1[11] = o1[1]; s1[12] = o1[2];
  s2[10] = o2[0]; s2[11] = o2[1]; s2[12] = o2[2];
  if (fabs((ei - ef) / ei) > 1.0e-9) return OC_ERR_KE_COLLISION;
  return OC_OK;
}

/* ------------------------------------------------------------------------ */
/* COLLISION_INTERACTION collision_interaction.f90:190-419: contact tests
 * are independent of the collisions of earlier pairs (positions and
 * orientations are not changed, and a used particle is skipped), so they
 * run in parallel; the ordered one-collision-per-particle pass is serial.  */
/* ------------------------------------------------------------------------ */
typedef struct { int flag; double e1[3], e2[3], cn[3], vk[3]; } pair_res;

int oc_collision_interaction(const oc_params *p, const double *spec, int np,
                             double *st, const int *species, int64_t nclose,
                             const int *home, const int *near_, int *flag_out,
                             int *collided_out, int *ncoll, int nthreads) {
  return oc_collision_interaction_stats(p, spec, np, st, species, nclose, home, near_,
                                        flag_out, collided_out, ncoll, nthreads, 1, 0, NULL);
}

int oc_collision_interaction_stats(const oc_params *p, const double *spec, int np,
                                   double *st, const int *species, int64_t nclose,
                                   const int *home, const int *near_, int *flag_out,
                                   int *collided_out, int *ncoll, int nthreads,
                                   int npoly, int nbin, double *col) {
  const double L = p->L;
  /* MEAN_PART_COL(:,:,:) = ZERO (collision_interaction.f90:171) */
  if (col)
    for (int64_t b = 0; b < (int64_t)nbin * 8; ++b) col[b] = ZERO_R;
  pair_res *res = (pair_res *)malloc(sizeof(pair_res) * (size_t)(nclose > 0 ? nclose : 1));
  (void)nthreads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t nn = 0; nn < nclose; ++nn) {
    int i = home[nn] - 1, j = near_[nn] - 1;
    double xi = st[(int64_t)OC_X * np + i], yi = st[(int64_t)OC_Y * np + i], zi = st[(int64_t)OC_Z * np + i];
    double vkx = st[(int64_t)OC_X * np + j] - xi;
    double vky = st[(int64_t)OC_Y * np + j] - yi;
    double vkz = st[(int64_t)OC_Z * np + j] - zi;
    if (fabs(vkx) > L / 2) vkx = vkx - copysign(L, vkx);
    if (fabs(vky) > L / 2.) vky = vky - copysign(L, vky);
    if (fabs(vkz) > L / 2.) vkz = vkz - copysign(L, vkz);
    double bi[3] = {xi, yi, zi}, bj[3] = {xi + vkx, yi + vky, zi + vkz};
    double qi[4], qj[4];
    for (int c = 0; c < 4; ++c) {
      qi[c] = st[(int64_t)(OC_QA + c) * np + i];
      qj[c] = st[(int64_t)(OC_QA + c) * np + j];
    }
    pair_res *r = &res[nn];
    oc_contact_detection(spec + (int64_t)species[i] * OC_SP_STRIDE,
                         spec + (int64_t)species[j] * OC_SP_STRIDE, bi, qi, bj, qj,
                         NULL, &r->flag, r->e1, r->e2, r->cn, NULL, NULL);
    r->vk[0] = vkx; r->vk[1] = vky; r->vk[2] = vkz;
    if (flag_out) flag_out[nn] = r->flag;
  }
  char *used = (char *)calloc((size_t)(np > 0 ? np : 1), 1);
  int err = OC_OK, nc = 0;
  for (int64_t nn = 0; nn < nclose; ++nn) {
    int i = home[nn] - 1, j = near_[nn] - 1;
    if (collided_out) collided_out[nn] = 0;
    if (used[i] || used[j]) continue;
    pair_res *r = &res[nn];
    if (!r->flag) continue;
    double s1[OC_NSTATE], s2[OC_NSTATE];
    for (int c = 0; c < OC_NSTATE; ++c) {
      s1[c] = st[(int64_t)c * np + i];
      s2[c] = st[(int64_t)c * np + j];
    }
    double q1[4] = {s1[6], s1[7], s1[8], s1[9]}, q2[4] = {s2[6], s2[7], s2[8], s2[9]};
    double rx1[3] = {(r->e1[0] - s1[0]), (r->e1[1] - s1[1]), (r->e1[2] - s1[2])};
    double o1[3] = {s1[10], s1[11], s1[12]};
    oc_transform_basis_vec(q1, o1);
    double up1 = s1[3] + (o1[1] * rx1[2] - o1[2] * rx1[1]);
    double vp1 = s1[4] + (o1[2] * rx1[0] - o1[0] * rx1[2]);
    double wp1 = s1[5] + (o1[0] * rx1[1] - o1[1] * rx1[0]);
    double rx2[3] = {(r->e2[0] - (s1[0] + r->vk[0])), (r->e2[1] - (s1[1] + r->vk[1])),
                     (r->e2[2] - (s1[2] + r->vk[2]))};
    double o2[3] = {s2[10], s2[11], s2[12]};
    oc_transform_basis_vec(q2, o2);
    double up2 = s2[3] + (o2[1] * rx2[2] - o2[2] * rx2[1]);
    double vp2 = s2[4] + (o2[2] * rx2[0] - o2[0] * rx2[2]);
    double wp2 = s2[5] + (o2[0] * rx2[1] - o2[1] * rx2[0]);
    double vrx = up2 - up1, vry = vp2 - vp1, vrz = wp2 - wp1;
    double nrm_vrel = sqrt(vrx * vrx + vry * vry + vrz * vrz);
    double nvrel = vrx * r->cn[0] + vry * r->cn[1] + vrz * r->cn[2];
    /* collision_interaction.f90:326-345 */
    int64_t bin = (int64_t)species[i] * npoly + species[j] % npoly;
    double *cb = (col && bin < nbin) ? col + bin * 8 : NULL;
    if (cb) {
      double theta = acos(nvrel / nrm_vrel);
      double vpc[3] = {r->e1[0] - r->e2[0], r->e1[1] - r->e2[1], r->e1[2] - r->e2[2]};
      cb[0] = cb[0] + 1.0;
      cb[3] = cb[3] + sqrt(vpc[0] * vpc[0] + vpc[1] * vpc[1] + vpc[2] * vpc[2]);
      cb[5] = cb[5] + nvrel;
      cb[6] = cb[6] + nrm_vrel;
      cb[7] = cb[7] + theta;
    }
    if (nvrel < ZERO_R) {
      if (cb) cb[1] = cb[1] + 1.0;   /* :409 */
      int e = oc_collision_response(spec + (int64_t)species[i] * OC_SP_STRIDE, s1, s2, q1, q2,
                                    rx1, rx2, r->cn);
      if (e && !err) err = e;
      for (int c = 0; c < OC_NSTATE; ++c) {
        st[(int64_t)c * np + i] = s1[c];
        st[(int64_t)c * np + j] = s2[c];
      }
      used[i] = 1; used[j] = 1;
      if (collided_out) collided_out[nn] = 1;
      nc++;
    }
  }
  free(used);
  free(res);
  /* :423-441 */
  if (col)
    for (int64_t b = 0; b < nbin; ++b) {
      col[b * 8 + 2] = 2.0 * col[b * 8 + 1] / p->dtime;
      col[b * 8 + 4] = col[b * 8 + 3] / col[b * 8 + 0];
    }
  if (ncoll) *ncoll = nc;
  return err;
}

/* one full step: ADV_PARTICLE_ELLP (main.f90:466) then COLLISION
 * (main.f90:483) with COLLISION_DETECTION_CONNECTIVITY + INTERACTION */
int oc_step(const oc_params *p, const double *field, const double *spec,
            int np, double *state, const int *species, const oc_cgrid *g,
            const int *con, const int *ncon, int64_t nclosemax, int *home,
            int *near_, int64_t *nclose, int *ncoll, int nthreads) {
  int err = oc_adv_particle_ellp(p, field, spec, np, state, species, NULL, nthreads);
  if (err) return err;
  err = oc_collision_detection(g, con, ncon, np, state + (int64_t)OC_X * np,
                               state + (int64_t)OC_Y * np, state + (int64_t)OC_Z * np,
                               nclosemax, home, near_, nclose);
  if (err) return err;
  return oc_collision_interaction(p, spec, np, state, species, *nclose, home, near_,
                                  NULL, NULL, ncoll, nthreads);
}

// gp_device.h -- device-side physics of the Lagrangian ellipsoid step for
// gfx950.  One thread advances one particle; every routine is a register-
// resident restatement of the reference subroutine it cites, templated on
// the compute type CT (double for the FP64 form, float for the gather and
// drag of the mixed form).  Positions, quaternions, wall geometry,
// contact solve and impulses always run in double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gazpart_gpu.h"

namespace gp {

constexpr double PI = 3.141592653589793;   // 4*atan(1d0), init_run.f90:123
constexpr double ZERO_R = 1.0e-33;         // mod_dns_dim.f90:136

// error bits raised by kernels (mapped to GP_ERR_* on the host)
enum : unsigned {
  EB_NAN_DRAG = 1u << 0,
  EB_NAN_OMEGA = 1u << 1,
  EB_KE_COLLISION = 1u << 2,
  EB_KE_WALL = 1u << 3,
  EB_OUT_OF_GRID = 1u << 4,
  EB_OVERFLOW = 1u << 5,
  EB_INTERP_RANGE = 1u << 6,
  EB_ACTIVE_OVERFLOW = 1u << 7
};

struct Species {   // one row of the species table, device copy
  double a, b, c, pmass, deq, ixx, iyy, izz, grav, ecp, mup, rmax;
};

struct StepParams {
  int n, scheme, renorm, prefilter;
  double L, dx, visc, rhof, dtime, ecw, muw;
};

// ---------------------------------------------------------------------------
// rotation matrix of a quaternion (transform_basis.f90:44-54); the matrix of
// the conjugate quaternion is its exact transpose, so conj rotations use Rt.
// ---------------------------------------------------------------------------
template <typename T>
struct Rot {
  T r00, r01, r02, r10, r11, r12, r20, r21, r22;
  __device__ __forceinline__ Rot(T q0, T q1, T q2, T q3) {
    r00 = T(1) - T(2) * (q2 * q2 + q3 * q3);
    r01 = T(2) * (q1 * q2 - q0 * q3);
    r02 = T(2) * (q1 * q3 + q0 * q2);
    r10 = T(2) * (q1 * q2 + q0 * q3);
    r11 = T(1) - T(2) * (q1 * q1 + q3 * q3);
    r12 = T(2) * (q2 * q3 - q0 * q1);
    r20 = T(2) * (q1 * q3 - q0 * q2);
    r21 = T(2) * (q2 * q3 + q0 * q1);
    r22 = T(1) - T(2) * (q1 * q1 + q2 * q2);
  }
  __device__ __forceinline__ void mul(T x, T y, T z, T& ox, T& oy, T& oz) const {
    ox = r00 * x + r01 * y + r02 * z;
    oy = r10 * x + r11 * y + r12 * z;
    oz = r20 * x + r21 * y + r22 * z;
  }
  __device__ __forceinline__ void mulT(T x, T y, T z, T& ox, T& oy, T& oz) const {
    ox = r00 * x + r10 * y + r20 * z;
    oy = r01 * x + r11 * y + r21 * z;
    oz = r02 * x + r12 * y + r22 * z;
  }
};

template <typename T> __device__ __forceinline__ T dsqrt(T x);
template <> __device__ __forceinline__ double dsqrt<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return sqrtf(x); }

template <typename T> __device__ __forceinline__ T dlog(T x);
template <> __device__ __forceinline__ double dlog<double>(double x) { return log(x); }
template <> __device__ __forceinline__ float dlog<float>(float x) { return logf(x); }
template <typename T> __device__ __forceinline__ T dexp(T x);
template <> __device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float dexp<float>(float x) { return expf(x); }

// ---------------------------------------------------------------------------
// Zastawny correlations (fluid_force_ellipsoid.f90:48-56, 92-99 and
// fluid_torque_ellipsoid.f90:64-82, 137-142).  Every power of a Reynolds
// number is exp(a log Re) with log Re taken once; the incidence angle
// phi = |atan(sqrt(v'^2 + w'^2) / u')| in [0, pi/2] enters only through
// sin(phi) = sqrt(v'^2 + w'^2) / |u'| and cos(phi) = |u'_x| / |u'|.
// ---------------------------------------------------------------------------
template <typename CT>
struct DragCoef {          // functions of the particle Reynolds number only
  CT cd0, cd90;            // C_D at phi = 0 and 90 degrees
  CT cl0, esl, ecl;        // lift: prefactor and the sin / cos exponents
  __device__ __forceinline__ explicit DragCoef(CT rep) {
    const CT l = dlog(rep);
    cd0 = CT(18.12) * dexp(CT(-1.023) * l) + CT(4.26) * dexp(CT(-0.384) * l);
    cd90 = CT(21.52) * dexp(CT(-0.990) * l) + CT(2.86) * dexp(CT(-0.260) * l);
    cl0 = CT(0.083) * dexp(CT(0.21) * l) + CT(1.582) * dexp(CT(-0.851) * l);
    esl = CT(1.842) + CT(-0.802) * dexp(CT(-0.006) * l);
    ecl = CT(0.874) + CT(0.009) * dexp(CT(0.570) * l);
  }
};

template <typename CT>
struct PitchCoef {         // pitching-torque coefficient pieces, f(Re)
  CT ct0, esp, ecp;
  __device__ __forceinline__ explicit PitchCoef(CT rep) {
    const CT l = dlog(rep);
    ct0 = CT(0.935) * dexp(CT(-0.146) * l) + CT(-0.469) * dexp(CT(-0.145) * l);
    esp = CT(0.116) + CT(0.748) * dexp(CT(0.041) * l);
    ecp = CT(0.221) + CT(0.657) * dexp(CT(0.044) * l);
  }
};

// relative velocity in the body frame and sin / cos of the incidence angle
template <typename CT>
struct Incidence {
  CT u0, u1, u2, mag, syz, sphi, cphi;
  __device__ __forceinline__ Incidence(const Rot<CT>& R, CT urx, CT ury, CT urz) {
    R.mulT(urx, ury, urz, u0, u1, u2);
    syz = dsqrt(u1 * u1 + u2 * u2);
    mag = dsqrt(u0 * u0 + u1 * u1 + u2 * u2);
    sphi = syz / mag;
    cphi = fabs(u0) / mag;
  }
};

// x^e for x in [0, 1], e > 0 (0^e = 0)
template <typename CT>
__device__ __forceinline__ CT upow(CT lx, CT e) { return dexp(e * lx); }

// ---------------------------------------------------------------------------
// FLUID_FORCE_ELLIPSOID (fluid_force_ellipsoid.f90:1-161): drag + lift per
// unit mass in the lab frame for relative velocity (urx, ury, urz), whose
// Reynolds-number terms are in dc.  Returns false when C_D is NaN (the
// reference's `stop`, :150-158).
// ---------------------------------------------------------------------------
template <typename CT>
__device__ __forceinline__ bool fluid_force(const Rot<CT>& R, CT pmass, CT deq, CT rhof,
                                            const DragCoef<CT>& dc, CT urx, CT ury, CT urz,
                                            CT f[3]) {
  const Incidence<CT> in(R, urx, ury, urz);
  const CT ls = dlog(in.sphi), lc = dlog(in.cphi);
  const CT cd = dc.cd0 + (dc.cd90 - dc.cd0) * upow(ls, CT(1.95));
  const CT cl = dc.cl0 * upow(ls, dc.esl) * upow(lc, dc.ecl);
  const CT pre = CT(0.125) * (rhof / pmass) * (CT(PI) * (deq * deq));
  const CT kd = pre * cd * in.mag;
  const CT kl = pre * cl * (in.mag * in.mag);
  const CT sg = copysign(CT(1), -in.u0);
  const CT fx = kd * in.u0 + kl * in.sphi * sg;
  const CT fy = kd * in.u1 + kl * in.cphi * (in.u1 / in.syz);
  const CT fz = kd * in.u2 + kl * in.cphi * (in.u2 / in.syz);
  R.mul(fx, fy, fz, f[0], f[1], f[2]);
  return cd == cd;
}

// ---------------------------------------------------------------------------
// FLUID_TORQUE_ELLIPSOID (fluid_torque_ellipsoid.f90:1-190): pitching +
// rotational torque per unit mass, body frame.  The relative velocity is the
// stage-n one (PART%UFAP - PART%UP, :88-105) in every RK stage, so its
// Reynolds-number terms pc are shared by both stages; the curl is the
// interpolated vorticity.
// ---------------------------------------------------------------------------
template <typename CT>
__device__ __forceinline__ void fluid_torque(const Rot<CT>& R, CT pmass, CT deq, CT visc,
                                             CT rhof, const PitchCoef<CT>& pc, CT urx, CT ury,
                                             CT urz, CT wx, CT wy, CT wz, CT ox, CT oy, CT oz,
                                             CT t[3]) {
  const Incidence<CT> in(R, urx, ury, urz);
  const CT ctp = pc.ct0 * upow(dlog(in.sphi), pc.esp) * upow(dlog(in.cphi), pc.ecp);
  CT cx, cy, cz;
  R.mulT(wx, wy, wz, cx, cy, cz);
  const CT o0 = CT(0.5) * cx - ox, o1 = CT(0.5) * cy - oy, o2 = CT(0.5) * cz - oz;
  const CT omag = dsqrt(o0 * o0 + o1 * o1 + o2 * o2);
  const CT oyz = dsqrt(o2 * o2 + o1 * o1);
  const CT rer = (omag * (deq * deq)) / visc;
  const CT lr = dlog(rer);
  const CT ctr = CT(0.573) * dexp(CT(-0.154) * lr) + CT(116.61) / rer;      // r4 = 1
  const CT ctr2 = CT(1.244) * dexp(CT(0.239) * lr) + CT(378.12) * dexp(CT(-0.789) * lr);
  const CT pre = CT(0.0625) * (rhof / pmass) * CT(PI) * (deq * deq * deq) * ctp *
                 (in.mag * in.mag);
  const CT h = deq * CT(0.5);
  const CT h5 = h * h * h * h * h;
  const CT rr = (CT(0.5) * rhof / pmass) * h5;
  t[0] = ctr * rr * fabs(o0) * o0;
  t[1] = pre * (fabs(in.u2) / in.syz) * copysign(CT(1), in.u0 * in.u2) + ctr2 * rr * oyz * o1;
  t[2] = pre * (fabs(in.u1) / in.syz) * copysign(CT(1), in.u0 * in.u1) + ctr2 * rr * oyz * o2;
}

// ---------------------------------------------------------------------------
// QUATERNION_INTEGRATION (adv_quaternion.f90:39-83, METHOD = 1): exponential
// map of the lab-frame spin, left-multiplied (mod_quaternion.f90:136-145).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void quat_integrate(double q[4], double ox, double oy, double oz,
                                               double dt) {
  Rot<double> R(q[0], q[1], q[2], q[3]);
  double wx, wy, wz;
  R.mul(ox, oy, oz, wx, wy, wz);
  double nrm = sqrt(wx * wx + wy * wy + wz * wz);
  double m0 = 1.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
  if (nrm != 0.0) {
    double s, c;
    sincos((nrm * dt) * 0.5, &s, &c);
    double k = s / nrm;
    m0 = c; m1 = k * wx; m2 = k * wy; m3 = k * wz;
  }
  double a = m0 * q[0] - m1 * q[1] - m2 * q[2] - m3 * q[3];
  double b = m0 * q[1] + m1 * q[0] + m2 * q[3] - m3 * q[2];
  double c = m0 * q[2] - m1 * q[3] + m2 * q[0] + m3 * q[1];
  double d = m0 * q[3] + m1 * q[2] - m2 * q[1] + m3 * q[0];
  q[0] = a; q[1] = b; q[2] = c; q[3] = d;
}

// ---------------------------------------------------------------------------
// WALL_BOUNDARY_CONDITION_NEW (ellipsoid_wall_boundary_new.f90:1-112) with
// closest_point_plane / max_encroachment_ellipsoid in closed form: for the
// plane x = xw with normal +-e_x, the ellipsoid (x-c)^T E (x-c) = 1 with
// E^-1 = R diag(a^2,b^2,c^2) R^T reaches it iff (xw - x)^2 < (E^-1)_xx, and
// its deepest point is c - E^-1 n / sqrt(n^T E^-1 n).
// ---------------------------------------------------------------------------
struct WallHit {
  double nx;              // +-1
  double arm[3];          // impact arm (centre - deepest point)
  double dx;              // wall_pos - deepest point (x)
  bool hit;
};

__device__ __forceinline__ WallHit wall_test(double L, double a, double b, double c, double x,
                                             const double q[4]) {
  WallHit w;
  double xw = (x < L / 2.0) ? 0.0 : L;
  w.nx = (x < L / 2.0) ? 1.0 : -1.0;
  Rot<double> R(q[0], q[1], q[2], q[3]);
  double a2 = a * a, b2 = b * b, c2 = c * c;
  // column 0 of E^-1 = R diag(a2,b2,c2) R^T
  double e00 = R.r00 * R.r00 * a2 + R.r01 * R.r01 * b2 + R.r02 * R.r02 * c2;
  double e10 = R.r10 * R.r00 * a2 + R.r11 * R.r01 * b2 + R.r12 * R.r02 * c2;
  double e20 = R.r20 * R.r00 * a2 + R.r21 * R.r01 * b2 + R.r22 * R.r02 * c2;
  double d = xw - x;
  w.hit = (d * d) < e00;
  w.arm[0] = w.arm[1] = w.arm[2] = 0.0;
  w.dx = 0.0;
  if (w.hit) {
    double s = 1.0 / sqrt(e00);
    // deepest point p = c - n_x * E^-1 e_x / sqrt(e00); arm = c - p
    w.arm[0] = w.nx * e00 * s;
    w.arm[1] = w.nx * e10 * s;
    w.arm[2] = w.nx * e20 * s;
    w.dx = xw - (x - w.arm[0]);
  }
  return w;
}

// symmetric 3x3 inverse by cofactors (invert_ndim3_matrix.f90)
__device__ __forceinline__ void inv3(const double m[9], double o[9]) {
  double c00 = m[4] * m[8] - m[5] * m[7];
  double c01 = m[5] * m[6] - m[3] * m[8];
  double c02 = m[3] * m[7] - m[4] * m[6];
  double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  double id = 1.0 / det;
  o[0] = c00 * id;
  o[1] = (m[2] * m[7] - m[1] * m[8]) * id;
  o[2] = (m[1] * m[5] - m[2] * m[4]) * id;
  o[3] = c01 * id;
  o[4] = (m[0] * m[8] - m[2] * m[6]) * id;
  o[5] = (m[2] * m[3] - m[0] * m[5]) * id;
  o[6] = c02 * id;
  o[7] = (m[1] * m[6] - m[0] * m[7]) * id;
  o[8] = (m[0] * m[4] - m[1] * m[3]) * id;
}

// lab inertia R diag(I) R^T (row-major 9)
__device__ __forceinline__ void lab_tensor(const Rot<double>& R, double d0, double d1, double d2,
                                           double o[9]) {
  const double r[9] = {R.r00, R.r01, R.r02, R.r10, R.r11, R.r12, R.r20, R.r21, R.r22};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      o[i * 3 + j] = r[i * 3 + 0] * d0 * r[j * 3 + 0] + r[i * 3 + 1] * d1 * r[j * 3 + 1] +
                     r[i * 3 + 2] * d2 * r[j * 3 + 2];
}

__device__ __forceinline__ void mat3v(const double m[9], const double v[3], double o[3]) {
  o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  o[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  o[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

// K = I + [r]x^T Iinv [r]x for a skew matrix [r]x (wall_ellipsoid_rebound.f90:
// 171-200, collision_response.f90:150-212)
__device__ __forceinline__ void k_matrix(const double iinv[9], const double r[3], double K[9]) {
  const double S[9] = {0.0, -r[2], r[1], r[2], 0.0, -r[0], -r[1], r[0], 0.0};
  double T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      T[i * 3 + j] = iinv[i * 3 + 0] * S[0 * 3 + j] + iinv[i * 3 + 1] * S[1 * 3 + j] +
                     iinv[i * 3 + 2] * S[2 * 3 + j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      K[i * 3 + j] = (i == j ? 1.0 : 0.0) + S[0 * 3 + i] * T[0 * 3 + j] +
                     S[1 * 3 + i] * T[1 * 3 + j] + S[2 * 3 + i] * T[2 * 3 + j];
}

// ---------------------------------------------------------------------------
// WALL_ELLIPSOID_REBOUND (wall_ellipsoid_rebound.f90:1-386), unit mass.
// u (lab) and o (body) updated in place; returns false on the reference's
// kinetic-energy `stop` (:372-384).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool wall_rebound(const double q[4], double ixx, double iyy,
                                             double izz, double ecw, double muw, double u[3],
                                             double o[3], const double arm[3], double nx) {
  Rot<double> R(q[0], q[1], q[2], q[3]);
  double om[3];
  R.mul(o[0], o[1], o[2], om[0], om[1], om[2]);
  const double n[3] = {nx, 0.0, 0.0};
  double vr[3] = {u[0] + (om[1] * arm[2] - om[2] * arm[1]),
                  u[1] + (om[2] * arm[0] - om[0] * arm[2]),
                  u[2] + (om[0] * arm[1] - om[1] * arm[0])};
  double vrn = vr[0] * n[0];
  double dv[3] = {vr[0] + ecw * vrn * n[0], vr[1], vr[2]};
  double ip[9], iip[9], K[9], K1[9];
  lab_tensor(R, ixx, iyy, izz, ip);
  inv3(ip, iip);
  k_matrix(iip, arm, K);
  inv3(K, K1);
  double pc[3];
  mat3v(K1, dv, pc);
  pc[0] = -pc[0]; pc[1] = -pc[1]; pc[2] = -pc[2];
  double pcn = pc[0] * n[0];
  double pt0 = pc[0] - pcn * n[0], pt1 = pc[1], pt2 = pc[2];
  double pct = sqrt(pt0 * pt0 + pt1 * pt1 + pt2 * pt2);
  if (pct > muw * pcn) {
    double tw[3] = {vr[0] - vrn * n[0], vr[1], vr[2]};
    double ntw = sqrt(tw[0] * tw[0] + tw[1] * tw[1] + tw[2] * tw[2]);
    if (ntw != 0.0) { tw[0] = -tw[0] / ntw; tw[1] = -tw[1] / ntw; tw[2] = -tw[2] / ntw; }
    double vi[3] = {n[0] + muw * tw[0], muw * tw[1], muw * tw[2]};
    double kn[3];
    mat3v(K, vi, kn);
    double kndn = kn[0] * n[0];
    double p = -(dv[0] * n[0]);
    double s = p / kndn;
    pc[0] = s * vi[0]; pc[1] = s * vi[1]; pc[2] = s * vi[2];
  }
  double am[3];
  mat3v(ip, om, am);
  double ei = 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) +
              0.5 * (om[0] * am[0] + om[1] * am[1] + om[2] * am[2]);
  u[0] += pc[0]; u[1] += pc[1]; u[2] += pc[2];
  double t3[3] = {-arm[2] * pc[1] + arm[1] * pc[2], arm[2] * pc[0] - arm[0] * pc[2],
                  -arm[1] * pc[0] + arm[0] * pc[1]};
  double t4[3];
  mat3v(iip, t3, t4);
  om[0] += t4[0]; om[1] += t4[1]; om[2] += t4[2];
  mat3v(ip, om, am);
  double ef = 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) +
              0.5 * (om[0] * am[0] + om[1] * am[1] + om[2] * am[2]);
  R.mulT(om[0], om[1], om[2], o[0], o[1], o[2]);
  return !(fabs((ei - ef) / ei) * 100.0 > 1.0e-9);
}

// ---------------------------------------------------------------------------
// one particle through ADV_PARTICLE_ELLP (adv_particle_new.f90:264-645).
// fl = fluid velocity and vorticity at x^n in the compute type.
// Returns error bits; *wall = 1 when the particle bounced.
// ---------------------------------------------------------------------------
template <typename CT>
__device__ __forceinline__ unsigned advance_particle(const StepParams& P, const Species& sp,
                                                     double s[13], const CT fl[6], int* wall) {
  const double dt = P.dtime;
  const double ixx = sp.ixx, iyy = sp.iyy, izz = sp.izz;
  const CT pm = CT(sp.pmass), dq = CT(sp.deq), vs = CT(P.visc), rf = CT(P.rhof);
  unsigned err = 0;
  const double x = s[0], y = s[1], z = s[2], up = s[3], vp = s[4], wp = s[5];
  const double q[4] = {s[6], s[7], s[8], s[9]};
  const double ox = s[10], oy = s[11], oz = s[12];
  const CT ufx = fl[0], ufy = fl[1], ufz = fl[2];

  // stage 1 at (x^n, u^n, omega^n, q^n); the stage-n relative velocity and
  // its Reynolds number also serve the stage-2 torque
  CT f[3], t[3];
  const CT unx = ufx - CT(up), uny = ufy - CT(vp), unz = ufz - CT(wp);
  const CT ren = (dsqrt(unx * unx + uny * uny + unz * unz) * dq) / vs;
  const PitchCoef<CT> pcn(ren);
  {
    Rot<CT> R{CT(q[0]), CT(q[1]), CT(q[2]), CT(q[3])};
    const DragCoef<CT> dcn(ren);
    if (!fluid_force(R, pm, dq, rf, dcn, unx, uny, unz, f)) err |= EB_NAN_DRAG;
    fluid_torque(R, pm, dq, vs, rf, pcn, unx, uny, unz, fl[3], fl[4], fl[5], CT(ox), CT(oy),
                 CT(oz), t);
  }
  if (!(ox == ox)) err |= EB_NAN_OMEGA;
  const double k1u[3] = {double(f[0]), double(f[1]), double(f[2]) + sp.grav};
  const double k1o[3] = {(double(t[0]) + oy * oz * (iyy - izz)) / ixx,
                         (double(t[1]) + oz * ox * (izz - ixx)) / iyy,
                         (double(t[2]) + ox * oy * (ixx - iyy)) / izz};
  double xp1 = x + dt * up, yp1 = y + dt * vp, zp1 = z + dt * wp;
  double u1[3] = {up + dt * k1u[0], vp + dt * k1u[1], wp + dt * k1u[2]};
  double o1[3] = {ox + dt * k1o[0], oy + dt * k1o[1], oz + dt * k1o[2]};
  double q1[4] = {q[0], q[1], q[2], q[3]};
  quat_integrate(q1, o1[0], o1[1], o1[2], dt);

  // stage-2 force (at u1, q1) and torque (stage-n velocity, omega1, q1)
  auto stage2 = [&](const double* qq, const double* uu, const double* oo, CT* ff, CT* tt) {
    Rot<CT> R{CT(qq[0]), CT(qq[1]), CT(qq[2]), CT(qq[3])};
    const CT urx = ufx - CT(uu[0]), ury = ufy - CT(uu[1]), urz = ufz - CT(uu[2]);
    const DragCoef<CT> dc((dsqrt(urx * urx + ury * ury + urz * urz) * dq) / vs);
    if (!fluid_force(R, pm, dq, rf, dc, urx, ury, urz, ff)) err |= EB_NAN_DRAG;
    fluid_torque(R, pm, dq, vs, rf, pcn, unx, uny, unz, fl[3], fl[4], fl[5], CT(oo[0]),
                 CT(oo[1]), CT(oo[2]), tt);
  };

  WallHit w = wall_test(P.L, sp.a, sp.b, sp.c, xp1, q1);
  *wall = 0;
  if (w.hit && u1[0] * w.nx < 0.0) {
    // bounced particle (adv_particle_new.f90:358-511)
    *wall = 1;
    double olab[3];
    {
      Rot<double> R1(q1[0], q1[1], q1[2], q1[3]);
      R1.mul(o1[0], o1[1], o1[2], olab[0], olab[1], olab[2]);
    }
    double vpc0 = u1[0] + (olab[1] * w.arm[2] - olab[2] * w.arm[1]);
    double vpc1 = u1[1] + (olab[2] * w.arm[0] - olab[0] * w.arm[2]);
    double vpc2 = u1[2] + (olab[0] * w.arm[1] - olab[1] * w.arm[0]);
    xp1 = xp1 + w.dx;
    double delt = fabs(w.dx) / sqrt(vpc0 * vpc0 + vpc1 * vpc1 + vpc2 * vpc2);
    double tpart = dt - delt;
    yp1 = y + tpart * vp;
    zp1 = z + tpart * wp;
    u1[0] = up + tpart * k1u[0]; u1[1] = vp + tpart * k1u[1]; u1[2] = wp + tpart * k1u[2];
    o1[0] = ox + tpart * k1o[0]; o1[1] = oy + tpart * k1o[1]; o1[2] = oz + tpart * k1o[2];
    q1[0] = q[0]; q1[1] = q[1]; q1[2] = q[2]; q1[3] = q[3];
    quat_integrate(q1, o1[0], o1[1], o1[2], tpart);
    if (!wall_rebound(q1, ixx, iyy, izz, P.ecw, P.muw, u1, o1, w.arm, w.nx)) err |= EB_KE_WALL;
    stage2(q1, u1, o1, f, t);
    const double k2u[3] = {double(f[0]), double(f[1]), double(f[2]) + sp.grav};
    const double k2o[3] = {(double(t[0]) + o1[1] * o1[2] * (iyy - izz)) / ixx,
                           (double(t[1]) + o1[2] * o1[0] * (izz - ixx)) / iyy,
                           (double(t[2]) + o1[0] * o1[1] * (ixx - iyy)) / izz};
    s[0] = xp1 + delt * u1[0];
    s[1] = yp1 + delt * u1[1];
    s[2] = zp1 + delt * u1[2];
    s[3] = u1[0] + delt * k2u[0];
    s[4] = u1[1] + delt * k2u[1];
    s[5] = u1[2] + delt * k2u[2];
    s[10] = o1[0] + delt * k2o[0];
    s[11] = o1[1] + delt * k2o[1];
    s[12] = o1[2] + delt * k2o[2];
    quat_integrate(q1, s[10], s[11], s[12], delt);
  } else {
    // unbounced particle (adv_particle_new.f90:514-640)
    stage2(q1, u1, o1, f, t);
    const double k2u[3] = {double(f[0]), double(f[1]), double(f[2]) + sp.grav};
    const double k2o[3] = {(double(t[0]) + o1[1] * o1[2] * (iyy - izz)) / ixx,
                           (double(t[1]) + o1[2] * o1[0] * (izz - ixx)) / iyy,
                           (double(t[2]) + o1[0] * o1[1] * (ixx - iyy)) / izz};
    const double h = 0.5 * dt;
    double xp2 = x + h * (up + u1[0]);
    WallHit w2 = wall_test(P.L, sp.a, sp.b, sp.c, xp2, q1);
    if (w2.hit) {
      s[0] = xp1; s[1] = yp1; s[2] = zp1;
      s[3] = up + dt * k1u[0]; s[4] = vp + dt * k1u[1]; s[5] = wp + dt * k1u[2];
      s[10] = o1[0]; s[11] = o1[1]; s[12] = o1[2];
    } else {
      s[0] = xp2;
      s[1] = y + h * (vp + u1[1]);
      s[2] = z + h * (wp + u1[2]);
      s[3] = up + h * (k1u[0] + k2u[0]);
      s[4] = vp + h * (k1u[1] + k2u[1]);
      s[5] = wp + h * (k1u[2] + k2u[2]);
      s[10] = ox + h * (k1o[0] + k2o[0]);
      s[11] = oy + h * (k1o[1] + k2o[1]);
      s[12] = oz + h * (k1o[2] + k2o[2]);
    }
    quat_integrate(q1, s[10], s[11], s[12], dt);
  }
  if (P.renorm) {
    double nq = sqrt(q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2] + q1[3] * q1[3]);
    q1[0] /= nq; q1[1] /= nq; q1[2] /= nq; q1[3] /= nq;
  }
  s[6] = q1[0]; s[7] = q1[1]; s[8] = q1[2]; s[9] = q1[3];
  // BOUNDARY_PARTICLE (boundary_particle.f90:45-64)
  for (int c = 0; c < 3; ++c) {
    if (s[c] > P.L) s[c] = s[c] - P.L;
    if (s[c] < 0.) s[c] = s[c] + P.L;
  }
  return err;
}

}  // namespace gp

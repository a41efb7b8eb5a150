// gp_mlp.h -- the contact-seed network: 12 -> 64 -> 64 -> 4 with ReLU,
// bf16 weights and activations, fp32 accumulate on the CDNA matrix cores
// (v_mfma_f32_16x16x16_bf16).  From the relative pose and semi-axes of a
// candidate pair (body frame of the first particle) it predicts the common
// normal the FP64 contact solve converges to when it finds a contact
// (outputs 0..2) and the logit of the solve finding one (output 3).  When
// that probability exceeds MLP_SEED_P the solve starts from the predicted
// normal instead of the centre line; it still decides every contact.  No
// reference counterpart (the reference seeds from the centre line,
// ellipsoid_contact_detection.f90:53-62).
//
// Layout: every layer is computed transposed, activations as
// (units x pairs) 16 x 16 tiles.  For the 16x16x16 MFMA on a 64-lane
// wavefront, lane l holds
//   A (16 x 16, M x K): row l % 16, k = 4 (l / 16) + i      (i = 0..3)
//   B (16 x 16, K x N): col l % 16, k = 4 (l / 16) + i
//   D (16 x 16, M x N): col l % 16, rows 4 (l / 16) + i
// so the D registers of one layer (units x pairs) are, after bias, ReLU and
// rounding, exactly the B operand of the next.  Weight gradients (K =
// pairs) read operands staged in LDS as [unit][pair].
#pragma once

#include "gp_contact.h"

namespace gp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MLP_NIN = 12, MLP_NH = 64, MLP_NOUT = 4;
// seed only when the predicted contact probability exceeds 0.9
constexpr float MLP_SEED_LOGIT = 2.1972245773362196f;   // ln(0.9 / 0.1)
// padded device layout of the parameters (and of their gradients):
//   W1 [64][16] (k >= 12 zero), b1 [64], W2 [64][64], b2 [64],
//   W3 [16][64] (rows >= 4 zero), b3 [16] (>= 4 zero)
constexpr int MO_W1 = 0, MO_B1 = MO_W1 + 64 * 16, MO_W2 = MO_B1 + 64, MO_B2 = MO_W2 + 64 * 64,
              MO_W3 = MO_B2 + 64, MO_B3 = MO_W3 + 16 * 64, MLP_NPAD = MO_B3 + 16;
constexpr int MLP_NPARAM = 64 * 12 + 64 + 64 * 64 + 64 + 4 * 64 + 4;   // compact vector

__device__ __forceinline__ short f2bf(float x) {
  unsigned u = __float_as_uint(x);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;   // nearest even (finite x)
  return (short)u;
}
__device__ __forceinline__ float bf2f(short b) {
  return __uint_as_float(((unsigned)(unsigned short)b) << 16);
}
__device__ __forceinline__ bf16x4 to_bf(f32x4 v) {
  bf16x4 r;
  r[0] = f2bf(v[0]); r[1] = f2bf(v[1]); r[2] = f2bf(v[2]); r[3] = f2bf(v[3]);
  return r;
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 zero4() {
  f32x4 z = {0.f, 0.f, 0.f, 0.f};
  return z;
}

// bf16 weights for the products (padded, row-major) and their transposes
struct MlpDev {
  const short* w1;    // [64][16]
  const short* w2;    // [64][64]
  const short* w3;    // [16][64]
  const short* w2t;   // [64][64]  w2t[i][o] = w2[o][i]
  const short* w3t;   // [64][16]  w3t[h][o] = w3[o][h]
  const float* b1;    // [64]
  const float* b2;    // [64]
  const float* b3;    // [16]
};

// A operand from a row-major bf16 matrix with leading dimension ld:
// rows r0 + l % 16, columns k0 + 4 (l / 16) .. + 3
__device__ __forceinline__ bf16x4 lda(const short* W, int ld, int r0, int k0, int lane) {
  return *reinterpret_cast<const bf16x4*>(W + (r0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4));
}

struct MlpAct {
  f32x4 z1[4], z2[4];    // pre-activations (units x pairs tiles)
  bf16x4 h1[4], h2[4];   // rounded activations (next B operands)
};

// forward of one tile of 16 pairs; xb = B operand of the inputs (features
// 4 (l / 16) + i of pair l % 16).  Returns the output tile (outputs
// 4 (l / 16) + i of pair l % 16: lanes 0..15 hold outputs 0..3).
__device__ __forceinline__ f32x4 mlp_forward_tile(const MlpDev& W, bf16x4 xb, MlpAct& a,
                                                  int lane) {
  const int rq = 4 * (lane >> 4);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    f32x4 z = mfma(lda(W.w1, 16, 16 * u, 0, lane), xb, zero4());
#pragma unroll
    for (int i = 0; i < 4; ++i) z[i] += W.b1[16 * u + rq + i];
    a.z1[u] = z;
    f32x4 h;
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = fmaxf(z[i], 0.f);
    a.h1[u] = to_bf(h);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    f32x4 z = zero4();
#pragma unroll
    for (int t = 0; t < 4; ++t) z = mfma(lda(W.w2, 64, 16 * u, 16 * t, lane), a.h1[t], z);
#pragma unroll
    for (int i = 0; i < 4; ++i) z[i] += W.b2[16 * u + rq + i];
    a.z2[u] = z;
    f32x4 h;
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = fmaxf(z[i], 0.f);
    a.h2[u] = to_bf(h);
  }
  f32x4 y = zero4();
#pragma unroll
  for (int t = 0; t < 4; ++t) y = mfma(lda(W.w3, 64, 0, 16 * t, lane), a.h2[t], y);
#pragma unroll
  for (int i = 0; i < 4; ++i) y[i] += W.b3[rq + i];
  return y;
}

// ---------------------------------------------------------------------------
// features of a pair (FP64 in, float out): R1^T (b2 - b1) / s, a1 / s,
// a2 / s, rotation vector of q1* q2; s = max(a1) + max(a2)
__device__ __forceinline__ void pair_features(const double b1[3], const double q1[4],
                                              const double a1[3], const double b2[3],
                                              const double q2[4], const double a2[3],
                                              float f[12]) {
  const double s = fmax(a1[0], fmax(a1[1], a1[2])) + fmax(a2[0], fmax(a2[1], a2[2]));
  Rot<double> R(q1[0], q1[1], q1[2], q1[3]);
  double d0, d1, d2;
  R.mulT(b2[0] - b1[0], b2[1] - b1[1], b2[2] - b1[2], d0, d1, d2);
  f[0] = (float)(d0 / s); f[1] = (float)(d1 / s); f[2] = (float)(d2 / s);
  for (int k = 0; k < 3; ++k) {
    f[3 + k] = (float)(a1[k] / s);
    f[6 + k] = (float)(a2[k] / s);
  }
  // r = conj(q1) * q2
  const double w1 = q1[0], x1 = -q1[1], y1 = -q1[2], z1 = -q1[3];
  double rw = w1 * q2[0] - x1 * q2[1] - y1 * q2[2] - z1 * q2[3];
  double rx = w1 * q2[1] + x1 * q2[0] + y1 * q2[3] - z1 * q2[2];
  double ry = w1 * q2[2] - x1 * q2[3] + y1 * q2[0] + z1 * q2[1];
  double rz = w1 * q2[3] + x1 * q2[2] - y1 * q2[1] + z1 * q2[0];
  if (rw < 0.0) { rw = -rw; rx = -rx; ry = -ry; rz = -rz; }
  const double sv = sqrt(rx * rx + ry * ry + rz * rz);
  const double fac = sv > 1e-12 ? 2.0 * atan2(sv, rw) / sv : 2.0;
  f[9] = (float)(rx * fac); f[10] = (float)(ry * fac); f[11] = (float)(rz * fac);
}

// solve angles (azimuth, elevation) of the lab-frame normal R1 y / |y|,
// when the contact logit clears the seeding threshold
__device__ __forceinline__ bool seed_from_output(const double q1[4], float y0, float y1,
                                                 float y2, float logit, double x[2]) {
  if (!(logit > MLP_SEED_LOGIT)) return false;
  const double n = sqrt((double)y0 * y0 + (double)y1 * y1 + (double)y2 * y2);
  if (!(n > 1e-20) || !(n < 1e20)) return false;
  Rot<double> R(q1[0], q1[1], q1[2], q1[3]);
  double c0, c1, c2;
  R.mul(y0 / n, y1 / n, y2 / n, c0, c1, c2);
  x[0] = atan2(c1, c0);
  x[1] = asin(fmin(1.0, fmax(-1.0, c2)));
  return true;
}

// The network-seeded solve and its acceptance gate.  Started from the
// network's normal, the solve can end in another local minimum of its
// residual than from the reference's centre-line guess, and with it take
// another contact decision.  So the seeded attempt gets at most g.maxfev
// evaluations, and its result is kept only when it converged (MINPACK info
// other than 5) with its final normal within the angle acos(g.cos_keep) of
// the seed's: the network had then predicted the minimum the solve found.
// Otherwise the pair is solved again from the centre line with the full
// budget of 300 evaluations, as the reference solves it.  nfev counts both
// attempts; fallback reports a redone solve.
struct SeedGate {
  int maxfev;
  double cos_keep;
};
constexpr int SEED_GATE_MAXFEV = 6;
constexpr double SEED_GATE_KEEP = 0.02;   // radians

__device__ __forceinline__ ContactOut contact_detect_gated(
    const double bi[3], const double qi[4], const Species& si, const double bj[3],
    const double qj[4], const Species& sj, const double* seed, const SeedGate& g, int& nfev,
    bool& fallback) {
  ContactOut o;
  const double* sd = seed;
  fallback = false;
  for (int pass = 0; pass < 2; ++pass) {
    o = contact_detect(bi, qi, si, bj, qj, sj, sd, sd ? g.maxfev : 300);
    nfev += o.nfev;
    if (!sd) break;
    double s0, c0, s1, c1;
    sincos(sd[0], &s0, &c0);
    sincos(sd[1], &s1, &c1);
    const double dot = c0 * c1 * o.cn[0] + s0 * c1 * o.cn[1] + s1 * o.cn[2];
    if (o.info != 5 && dot >= g.cos_keep) break;
    sd = nullptr;
    fallback = true;
  }
  return o;
}

// ---------------------------------------------------------------------------
// synthetic solver-labelled pairs: tri-axial shapes (semi-axis ratios
// uniform in [1, 4], unit volume scale), uniform random orientations,
// centre offsets uniform in direction with distance uniform in
// [0.2, 1] x (sum of the largest semi-axes), kept only when not proven
// disjoint (the pairs that reach the solve)
struct PairPose {
  double b1[3], q1[4], a1[3], b2[3], q2[4], a2[3];
};

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
struct PairRng {
  unsigned long long s;
  __device__ double u() {   // uniform in (0, 1)
    s = splitmix64(s);
    return ((double)(s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  }
  __device__ double n() {   // standard normal (Box-Muller)
    const double a = u(), b = u();
    return sqrt(-2.0 * log(a)) * cospi(2.0 * b);
  }
};

__device__ __forceinline__ void sample_shape(PairRng& g, double a[3]) {
  const double rb = 1.0 + 3.0 * g.u(), rc = 1.0 + 3.0 * g.u();
  const double aa = cbrt(rb * rc);
  a[0] = aa; a[1] = aa / rb; a[2] = aa / rc;
}
__device__ __forceinline__ void sample_quat(PairRng& g, double q[4]) {
  double s = 0.0;
  for (int k = 0; k < 4; ++k) { q[k] = g.n(); s += q[k] * q[k]; }
  s = sqrt(s);
  for (int k = 0; k < 4; ++k) q[k] /= s;
}

__device__ __forceinline__ PairPose sample_pair(unsigned long long seed, long long idx) {
  PairRng g{splitmix64(seed ^ splitmix64((unsigned long long)idx))};
  PairPose p;
  for (int attempt = 0; attempt < 32; ++attempt) {
    sample_shape(g, p.a1);
    sample_shape(g, p.a2);
    sample_quat(g, p.q1);
    sample_quat(g, p.q2);
    double u[3] = {g.n(), g.n(), g.n()};
    const double un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double rs = fmax(p.a1[0], fmax(p.a1[1], p.a1[2])) + fmax(p.a2[0], fmax(p.a2[1], p.a2[2]));
    const double dist = (0.2 + 0.8 * g.u()) * rs;
    for (int k = 0; k < 3; ++k) {
      p.b1[k] = 0.0;
      p.b2[k] = dist * u[k] / un;
    }
    const Sym3 E1 = ell_matrix(p.q1, p.a1[0], p.a1[1], p.a1[2]);
    const Sym3 E2 = ell_matrix(p.q2, p.a2[0], p.a2[1], p.a2[2]);
    if (!pw_disjoint(E1, E2, p.b2)) break;
  }
  return p;
}

__device__ __forceinline__ Species pose_species(const double a[3]) {
  Species s{};
  s.a = a[0]; s.b = a[1]; s.c = a[2];
  s.rmax = fmax(a[0], fmax(a[1], a[2]));
  return s;
}

// ---------------------------------------------------------------------------
// kernels

// forward over n feature rows (x [n][16] floats, features 12..15 zero)
__global__ __launch_bounds__(64) void k_mlp_forward(MlpDev W, const float* __restrict__ x,
                                                    long long n, float* __restrict__ y) {
  const int lane = threadIdx.x;
  const long long ntile = (n + 15) / 16;
  for (long long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const long long p = tile * 16 + (lane & 15);
    f32x4 xv = zero4();
    if (p < n) xv = *reinterpret_cast<const f32x4*>(x + p * 16 + 4 * (lane >> 4));
    MlpAct a;
    const f32x4 out = mlp_forward_tile(W, to_bf(xv), a, lane);
    if (lane < 16 && p < n)
      for (int i = 0; i < 4; ++i) y[p * 4 + i] = out[i];
  }
}

// forward + backward of the mean pair loss over n rows (targets t [n][4]:
// unit normal and contact flag; loss flag |y0..2 - t|^2 + softplus(y3) -
// flag y3), scaled by inv_n; each wavefront (workgroup) writes its partial
// gradient (padded layout) and partial loss.
__global__ __launch_bounds__(64) void k_mlp_train_part(MlpDev W, const float* __restrict__ x,
                                                       const float* __restrict__ t, long long n,
                                                       float inv_n, float* __restrict__ part,
                                                       float* __restrict__ loss_part) {
  const int lane = threadIdx.x;
  const int rq = 4 * (lane >> 4), col = lane & 15;
  __shared__ short sX[16][16], sH1[64][16], sH2[64][16], sdY[16][16], sdH2[64][16],
      sdH1[64][16];
  f32x4 g1[4], g2[4][4], g3[4];
  f32x4 gb1[4], gb2[4], gb3 = zero4();
  for (int u = 0; u < 4; ++u) {
    g1[u] = zero4(); g3[u] = zero4(); gb1[u] = zero4(); gb2[u] = zero4();
    for (int v = 0; v < 4; ++v) g2[u][v] = zero4();
  }
  float lsum = 0.f;
  const long long ntile = (n + 15) / 16;
  for (long long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const long long p = tile * 16 + col;
    const bool valid = p < n;
    f32x4 xv = zero4();
    if (valid) xv = *reinterpret_cast<const f32x4*>(x + p * 16 + rq);
    const bf16x4 xb = to_bf(xv);
    MlpAct a;
    const f32x4 y = mlp_forward_tile(W, xb, a, lane);
    f32x4 dy = zero4();
    if (lane < 16 && valid) {
      const float flag = t[p * 4 + 3];
      for (int i = 0; i < 3; ++i) {
        const float e = y[i] - t[p * 4 + i];
        lsum += flag * e * e;
        dy[i] = 2.f * flag * e * inv_n;
      }
      const float z = y[3];
      lsum += fmaxf(z, 0.f) + log1pf(expf(-fabsf(z))) - flag * z;
      dy[3] = (1.f / (1.f + expf(-z)) - flag) * inv_n;
    }
    gb3 += dy;
    const bf16x4 dyb = to_bf(dy);
    // hidden-2 gradient: W3^T dY, masked by the ReLU
    bf16x4 dh2b[4];
    for (int u = 0; u < 4; ++u) {
      f32x4 d = mfma(lda(W.w3t, 16, 16 * u, 0, lane), dyb, zero4());
      for (int i = 0; i < 4; ++i) d[i] = a.z2[u][i] > 0.f ? d[i] : 0.f;
      gb2[u] += d;
      dh2b[u] = to_bf(d);
    }
    // hidden-1 gradient: W2^T dH2, masked
    bf16x4 dh1b[4];
    for (int u = 0; u < 4; ++u) {
      f32x4 d = zero4();
      for (int v = 0; v < 4; ++v) d = mfma(lda(W.w2t, 64, 16 * u, 16 * v, lane), dh2b[v], d);
      for (int i = 0; i < 4; ++i) d[i] = a.z1[u][i] > 0.f ? d[i] : 0.f;
      gb1[u] += d;
      dh1b[u] = to_bf(d);
    }
    // stage [unit][pair] tiles for the weight gradients (K = pairs)
    for (int i = 0; i < 4; ++i) {
      sX[rq + i][col] = xb[i];
      sdY[rq + i][col] = dyb[i];
      for (int u = 0; u < 4; ++u) {
        sH1[16 * u + rq + i][col] = a.h1[u][i];
        sH2[16 * u + rq + i][col] = a.h2[u][i];
        sdH2[16 * u + rq + i][col] = dh2b[u][i];
        sdH1[16 * u + rq + i][col] = dh1b[u][i];
      }
    }
    __syncthreads();
    auto lds = [&](short (*S)[16], int r0) {
      return *reinterpret_cast<const bf16x4*>(&S[r0 + col][rq]);
    };
    for (int u = 0; u < 4; ++u) {
      g3[u] = mfma(lds(sdY, 0), lds(sH2, 16 * u), g3[u]);          // dW3 [o][h]
      g1[u] = mfma(lds(sdH1, 16 * u), lds(sX, 0), g1[u]);          // dW1 [h][f]
      for (int v = 0; v < 4; ++v)
        g2[u][v] = mfma(lds(sdH2, 16 * u), lds(sH1, 16 * v), g2[u][v]);   // dW2 [o][i]
    }
    __syncthreads();
  }
  // partial gradients, padded layout; D tiles: rows rq + i, column col
  float* g = part + (long long)blockIdx.x * MLP_NPAD;
  for (int i = 0; i < 4; ++i) {
    for (int u = 0; u < 4; ++u) {
      g[MO_W1 + (16 * u + rq + i) * 16 + col] = g1[u][i];
      g[MO_W3 + (rq + i) * 64 + 16 * u + col] = g3[u][i];
      for (int v = 0; v < 4; ++v) g[MO_W2 + (16 * u + rq + i) * 64 + 16 * v + col] = g2[u][v][i];
    }
  }
  // bias gradients: sum over the 16 pair columns (lanes with equal l / 16)
  for (int o = 1; o < 16; o <<= 1) {
    for (int i = 0; i < 4; ++i) {
      gb3[i] += __shfl_xor(gb3[i], o);
      for (int u = 0; u < 4; ++u) {
        gb1[u][i] += __shfl_xor(gb1[u][i], o);
        gb2[u][i] += __shfl_xor(gb2[u][i], o);
      }
    }
  }
  if (col == 0) {
    for (int i = 0; i < 4; ++i) {
      g[MO_B3 + rq + i] = gb3[i];
      for (int u = 0; u < 4; ++u) {
        g[MO_B1 + 16 * u + rq + i] = gb1[u][i];
        g[MO_B2 + 16 * u + rq + i] = gb2[u][i];
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
  if (lane == 0) loss_part[blockIdx.x] = lsum;
}

// grad[j] = sum over partials in a fixed order (deterministic)
__global__ void k_mlp_reduce(const float* __restrict__ part, int nparts, float* __restrict__ grad,
                             const float* __restrict__ loss_part, float* __restrict__ loss) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < MLP_NPAD) {
    float s = 0.f;
    for (int w = 0; w < nparts; ++w) s += part[(long long)w * MLP_NPAD + j];
    grad[j] = s;
  }
  if (j == 0 && loss) {
    float s = 0.f;
    for (int w = 0; w < nparts; ++w) s += loss_part[w];
    *loss = s;
  }
}

// Adam on the fp32 masters (padded layout; padding stays zero since its
// gradient is zero), then the bf16 copies and transposes for the products
__global__ void k_mlp_adam(float* __restrict__ th, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ grad, float lr, float c1, float c2,
                           short* __restrict__ w1, short* __restrict__ w2,
                           short* __restrict__ w3, short* __restrict__ w2t,
                           short* __restrict__ w3t, float* __restrict__ b1,
                           float* __restrict__ b2, float* __restrict__ b3, int update) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= MLP_NPAD) return;
  float w = th[j];
  if (update) {
    const float gj = grad[j];
    const float mj = 0.9f * m[j] + 0.1f * gj;
    const float vj = 0.999f * v[j] + 0.001f * gj * gj;
    m[j] = mj;
    v[j] = vj;
    w = w - lr * (mj / c1) / (sqrtf(vj / c2) + 1e-8f);
    th[j] = w;
  }
  if (j < MO_B1) {
    w1[j] = f2bf(w);
  } else if (j < MO_W2) {
    b1[j - MO_B1] = w;
  } else if (j < MO_B2) {
    const int k = j - MO_W2, o = k / 64, i = k % 64;
    w2[k] = f2bf(w);
    w2t[i * 64 + o] = f2bf(w);
  } else if (j < MO_W3) {
    b2[j - MO_B2] = w;
  } else if (j < MO_B3) {
    const int k = j - MO_W3, o = k / 64, h = k % 64;
    w3[k] = f2bf(w);
    w3t[h * 16 + o] = f2bf(w);
  } else {
    b3[j - MO_B3] = w;
  }
}

// solver-labelled synthetic pairs: features x [n][16], targets t [n][4]
// (converged normal in the first body frame), reference-seed evaluations
__global__ __launch_bounds__(64) void k_mlp_pairs(unsigned long long seed, long long i0,
                                                  long long n, float* __restrict__ x,
                                                  float* __restrict__ t,
                                                  int* __restrict__ nfev,
                                                  double* __restrict__ poses) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const PairPose p = sample_pair(seed, i0 + k);
    float f[12];
    pair_features(p.b1, p.q1, p.a1, p.b2, p.q2, p.a2, f);
    for (int c = 0; c < 12; ++c) x[k * 16 + c] = f[c];
    for (int c = 12; c < 16; ++c) x[k * 16 + c] = 0.f;
    const ContactOut o =
        contact_detect(p.b1, p.q1, pose_species(p.a1), p.b2, p.q2, pose_species(p.a2), nullptr);
    Rot<double> R(p.q1[0], p.q1[1], p.q1[2], p.q1[3]);
    double n0, n1, n2;
    R.mulT(o.cn[0], o.cn[1], o.cn[2], n0, n1, n2);
    t[k * 4 + 0] = (float)n0; t[k * 4 + 1] = (float)n1; t[k * 4 + 2] = (float)n2;
    t[k * 4 + 3] = (float)o.flag;
    if (nfev) nfev[k] = o.nfev;
    if (poses) {
      double* q = poses + k * 20;
      for (int c = 0; c < 3; ++c) { q[c] = p.b1[c]; q[7 + c] = p.a1[c]; q[10 + c] = p.b2[c]; q[17 + c] = p.a2[c]; }
      for (int c = 0; c < 4; ++c) { q[3 + c] = p.q1[c]; q[13 + c] = p.q2[c]; }
    }
  }
}

}  // namespace gp

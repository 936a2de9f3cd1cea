// The time-of-impact proposer's network and the CCD sub-test quantities it
// works with, shared by the narrow phase (scoring, labels, seeded Newton
// solve) and the training kernels.  No counterpart in the reference.
//
// Network: 24 -> 128 -> 128 -> 2, ReLU, BF16 weights and activations, FP32
// accumulation on the gfx950 matrix cores (v_mfma_f32_16x16x16_bf16).  One
// wavefront computes a 16 x 16 block D += A B over K = 16 per instruction:
// lane l supplies A[l % 16][4 (l / 16) + j] and B[4 (l / 16) + j][l % 16],
// j = 0..3, and receives D[4 (l / 16) + i][l % 16], i = 0..3.  B is read from
// "Bt" stored as Bt[n][k] (a weight matrix in its natural [out][in] layout),
// so both operands are four contiguous BF16 values.
#pragma once
#include "cs_device.h"

namespace cs {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int MF_IN = 24, MF_INP = 32, MF_H = 128, MF_OUT = 2;
// FP32 master parameters: W1[128][24], b1, W2[128][128], b2, W3[2][128], b3
constexpr int MO_W1 = 0, MO_B1 = MO_W1 + MF_H * MF_IN, MO_W2 = MO_B1 + MF_H;
constexpr int MO_B2 = MO_W2 + MF_H * MF_H, MO_W3 = MO_B2 + MF_H, MO_B3 = MO_W3 + MF_OUT * MF_H;
constexpr int MF_NPARAM = MO_B3 + MF_OUT;
static_assert(MF_NPARAM == 19970, "parameter count");
// BF16 weight copies for the matrix cores: W1 [128][32] (inputs 24..31 zero),
// W2 [128][128], W3 [16][128] (outputs 2..15 zero), and the transposes the
// backward pass multiplies by: W2t [128][128], W3t [128][16]
constexpr int MB_W1 = 0, MB_W2 = MB_W1 + MF_H * MF_INP, MB_W3 = MB_W2 + MF_H * MF_H;
constexpr int MB_W2T = MB_W3 + 16 * MF_H, MB_W3T = MB_W2T + MF_H * MF_H;
constexpr int MB_TOTAL = MB_W3T + MF_H * 16;
constexpr int MT = 64;  // samples per tile (4 wavefronts x 16 rows)
constexpr int MLP_THREADS = 256;

__device__ __host__ inline uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __host__ inline float bf2f(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__device__ inline bf16x4 ld4(const uint16_t* p) {  // four contiguous BF16 (8-byte aligned)
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  bf16x4 r;
  r[0] = (short)(v.x & 0xffffu);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffffu);
  r[3] = (short)(v.y >> 16);
  return r;
}
__device__ inline bf16x4 ld4_strided(const uint16_t* p, int stride) {  // p[0], p[s], p[2s], p[3s]
  bf16x4 r;
  r[0] = (short)p[0];
  r[1] = (short)p[stride];
  r[2] = (short)p[2 * stride];
  r[3] = (short)p[3 * stride];
  return r;
}
__device__ inline floatx4 mfma16(bf16x4 a, bf16x4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// The wavefront's 16 rows of A (row-major, lda) times Bt^T for NT column
// tiles of 16, K a multiple of 16.
template <int NT>
__device__ inline void wave_gemm(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, int K,
                                 floatx4 acc[NT]) {
  const int lane = threadIdx.x & 63, r = lane & 15, kq = (lane >> 4) * 4;
  for (int kk = 0; kk < K; kk += 16) {
    const bf16x4 a = ld4(A + r * lda + kk + kq);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = mfma16(a, ld4(Bt + (16 * nt + r) * ldb + kk + kq), acc[nt]);
  }
}

// One 64-sample tile through the network (256 threads).  x: [MT][32] BF16
// in LDS (rows of absent samples zero); h1, h2: [MT][128] BF16 in LDS
// (post-ReLU); out: [MT][2] FP32 in LDS (logit, raw seed).
__device__ inline void tile_forward(const uint16_t* x, uint16_t* h1, uint16_t* h2, float* out,
                                    const uint16_t* __restrict__ wb, const float* __restrict__ w) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = 16 * wave, col = lane & 15, rq = (lane >> 4) * 4;
  floatx4 acc[8];
#pragma unroll
  for (int nt = 0; nt < 8; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
  wave_gemm<8>(x + r0 * MF_INP, MF_INP, wb + MB_W1, MF_INP, MF_INP, acc);
#pragma unroll
  for (int nt = 0; nt < 8; ++nt) {
    const int n = 16 * nt + col;
    const float b = w[MO_B1 + n];
#pragma unroll
    for (int i = 0; i < 4; ++i) h1[(r0 + rq + i) * MF_H + n] = f2bf(fmaxf(acc[nt][i] + b, 0.f));
  }
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < 8; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
  wave_gemm<8>(h1 + r0 * MF_H, MF_H, wb + MB_W2, MF_H, MF_H, acc);
#pragma unroll
  for (int nt = 0; nt < 8; ++nt) {
    const int n = 16 * nt + col;
    const float b = w[MO_B2 + n];
#pragma unroll
    for (int i = 0; i < 4; ++i) h2[(r0 + rq + i) * MF_H + n] = f2bf(fmaxf(acc[nt][i] + b, 0.f));
  }
  __syncthreads();
  floatx4 o[1] = {floatx4{0.f, 0.f, 0.f, 0.f}};
  wave_gemm<1>(h2 + r0 * MF_H, MF_H, wb + MB_W3, MF_H, MF_H, o);
  if (col < MF_OUT) {
    const float b = w[MO_B3 + col];
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(r0 + rq + i) * MF_OUT + col] = o[0][i] + b;
  }
  __syncthreads();
}

__device__ inline float mlp_sigmoid(float z) { return 1.f / (1.f + __expf(-z)); }

// The 24 features of a CCD sub-test (four points in sub-test order): x_old
// relative to the points' centroid, then avgVel * dt, both over the
// thickness h
__device__ inline void subtest_features(const double xo[4][3], const double v[4][3], double dt,
                                        double h, float f[MF_IN]) {
  double c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) c[k] = 0.25 * (((xo[0][k] + xo[1][k]) + xo[2][k]) + xo[3][k]);
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      f[3 * q + k] = (float)((xo[q][k] - c[k]) / h);
      f[12 + 3 * q + k] = (float)(v[q][k] * dt / h);
    }
}

// Coefficients of det[x_i + t v_i] relative to point 0, formed as
// isCoplanar forms them (dcollid3d.cpp:389-427)
__device__ inline void cubic_coeffs(const double xo[4][3], const double vs[4][3], double* a,
                                    double* b, double* c, double* d) {
  double v[4][3], x[4][3];
#pragma unroll
  for (int i = 1; i < 4; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      v[i][k] = vs[i][k] - vs[0][k];
      x[i][k] = xo[i][k] - xo[0][k];
    }
  double vv[3], vx[3], xx[3];
  vv[0] = v[1][1] * v[2][2] - v[1][2] * v[2][1];
  vv[1] = v[1][0] * v[2][2] - v[1][2] * v[2][0];
  vv[2] = v[1][0] * v[2][1] - v[1][1] * v[2][0];
  vx[0] = v[1][1] * x[2][2] - v[1][2] * x[2][1] - v[2][1] * x[1][2] + v[2][2] * x[1][1];
  vx[1] = v[1][0] * x[2][2] - v[1][2] * x[2][0] - v[2][0] * x[1][2] + v[2][2] * x[1][0];
  vx[2] = v[1][0] * x[2][1] - v[1][1] * x[2][0] - v[2][0] * x[1][1] + v[2][1] * x[1][0];
  xx[0] = x[1][1] * x[2][2] - x[1][2] * x[2][1];
  xx[1] = x[1][0] * x[2][2] - x[1][2] * x[2][0];
  xx[2] = x[1][0] * x[2][1] - x[1][1] * x[2][0];
  *a = v[3][0] * vv[0] - v[3][1] * vv[1] + v[3][2] * vv[2];
  *b = x[3][0] * vv[0] - x[3][1] * vv[1] + x[3][2] * vv[2] + v[3][0] * vx[0] - v[3][1] * vx[1] +
       v[3][2] * vx[2];
  *c = x[3][0] * vx[0] - x[3][1] * vx[1] + x[3][2] * vx[2] + v[3][0] * xx[0] - v[3][1] * xx[1] +
       v[3][2] * xx[2];
  *d = x[3][0] * xx[0] - x[3][1] * xx[1] + x[3][2] * xx[2];
}

// First root of a t^3 + b t^2 + c t + d in [t0, t1].  The interval is cut at
// the critical points into monotone pieces; the first piece whose ends differ
// in sign (or hold a zero) is the bracket.  Newton steps start from `seed`
// (the bracket's midpoint when the seed lies outside it) and fall back to
// bisection whenever they leave the bracket, which shrinks with every step;
// the solve stops when a step moves less than 1e-15 (t1 - t0) or the value is
// zero.  Returns the iterations, or -1 when no piece changes sign.
__device__ inline int newton_first_root(double a, double b, double c, double d, double t0,
                                        double t1, double seed, double* root) {
  const double dt = t1 - t0;
  double cut[4];
  int nc = 0;
  cut[nc++] = t0;
  const double qa = 3.0 * a, qb = 2.0 * b, qc = c;
  if (qa != 0.0) {
    const double disc = qb * qb - 4.0 * qa * qc;
    if (disc > 0.0) {
      const double s = sqrt(disc);
      double r1 = (-qb - s) / (2.0 * qa), r2 = (-qb + s) / (2.0 * qa);
      if (r1 > r2) { const double t = r1; r1 = r2; r2 = t; }
      if (r1 > t0 && r1 < t1) cut[nc++] = r1;
      if (r2 > t0 && r2 < t1) cut[nc++] = r2;
    }
  } else if (qb != 0.0) {
    const double r1 = -qc / qb;
    if (r1 > t0 && r1 < t1) cut[nc++] = r1;
  }
  cut[nc++] = t1;
  for (int p = 0; p + 1 < nc; ++p) {
    double lo = cut[p], hi = cut[p + 1];
    const double flo = ((a * lo + b) * lo + c) * lo + d;
    const double fhi = ((a * hi + b) * hi + c) * hi + d;
    if (flo == 0.0) { *root = lo; return 0; }
    if ((flo < 0.0) == (fhi < 0.0) && fhi != 0.0) continue;
    const bool neg_lo = flo < 0.0;
    double t = (seed > lo && seed < hi) ? seed : 0.5 * (lo + hi);
    int it = 0;
    for (; it < 200; ++it) {
      const double ft = ((a * t + b) * t + c) * t + d;
      if (ft == 0.0) break;
      if ((ft < 0.0) == neg_lo) lo = t; else hi = t;
      const double d1 = (3.0 * a * t + 2.0 * b) * t + c;
      double tn = d1 != 0.0 ? t - ft / d1 : 0.5 * (lo + hi);
      if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
      const bool done = fabs(tn - t) <= 1e-15 * dt;
      t = tn;
      if (done) { ++it; break; }
    }
    *root = t;
    return it;
  }
  return -1;
}

}  // namespace cs

// Device-side building blocks shared by the collision kernels (gfx950).
//
// FP64 arithmetic here follows the reference operation by operation:
// left-to-right sums, no fused multiply-add (the library is compiled with
// -ffp-contract=off), IEEE division and square root.  The FronTier vector
// macros the reference relies on are defined as
//   Dot3d = a0*b0 + a1*b1 + a2*b2, Cross3d = (a1b2-a2b1, a2b0-a0b2, a0b1-a1b0),
//   Mag3d = sqrt(Dot3d), distance = sqrt(sum_i (p_i-q_i)^2), MACH_EPS = DBL_EPSILON.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CS_MACH_EPS 2.220446049250313e-16   // FronTier MACH_EPS (DBL_EPSILON)
#define CS_ROUND_EPS 1e-10                   // collid.h:17
#define CS_BOX_EPS 1e-6                      // AABB.cpp:10-33
#define CS_SUPERELAS_TOL 0.10                // dcollid.cpp:489
#define CS_PI 3.14159265358979323846

struct __align__(16) dvec4 {
  double x, y, z, w;
};

struct __align__(16) dbox {  // FP64 element box, inflated by 1e-6
  double lo[3];
  double hi[3];
};

// LBVH node, one 32-byte record.  Nodes are stored in depth-first pre-order
// (root first; an internal node's first child is the next record).
//   internal: FP32 box of the subtree (rounded outward), left = -1,
//             rope = next node after the subtree in depth-first order
//   leaf:     FP32 box of the element (rounded outward), left = element id,
//             rope as above
// rope = -1 ends a traversal.  Depth-first traversal needs no stack: on an
// overlap go to the next record (internal) or handle the leaf and go to
// `rope`; on a miss go to `rope`.  Every step moves forward in memory.
struct __align__(16) bvh_node {
  float lo[3];
  int32_t left;
  float hi[3];
  int32_t rope;
};

// A narrow-phase hit: one sub-test of one element pair that reported
// contact, with the per-point impulse and friction increments it adds.
struct __align__(16) hit_rec {
  uint64_t key;        // (lo << 36) | (hi << 8) | subtest  (canonical order)
  int32_t v[4];        // vertex ids; -1 for static points (no contribution)
  double imp[4][3];
  double fric[4][3];
};

__host__ __device__ inline uint64_t hit_key(uint32_t lo, uint32_t hi, uint32_t sub) {
  return ((uint64_t)lo << 36) | ((uint64_t)hi << 8) | (uint64_t)sub;
}

__device__ inline double smin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ inline double smax(double a, double b) { return (a < b) ? b : a; }  // std::max

__device__ inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double mag3(const double* a) { return sqrt(dot3(a, a)); }
__device__ inline void sub3(const double* a, const double* b, double* c) {
  c[0] = a[0] - b[0];
  c[1] = a[1] - b[1];
  c[2] = a[2] - b[2];
}
__device__ inline double dist3(const double* a, const double* b) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double d = a[i] - b[i];
    s += d * d;
  }
  return sqrt(s);
}
__device__ inline double det3(const double a[3][3]) {  // myDet3d, dcollid.cpp:977-981
  return a[0][0] * (a[1][1] * a[2][2] - a[2][1] * a[1][2]) -
         a[0][1] * (a[1][0] * a[2][2] - a[2][0] * a[1][2]) +
         a[0][2] * (a[1][0] * a[2][1] - a[2][0] * a[1][1]);
}

// ---------------------------------------------------------------------------
// Deterministic elementary functions.  The reference calls libm acos, cos,
// sin and pow(x, 1/3) (dcollid3d.cpp:163-169, 436-444).  These are fixed
// textbook algorithms built from IEEE +,-,*,/, sqrt, fma, frexp, ldexp, so
// their results do not depend on a vendor math library:
//   sin/cos: Cody-Waite reduction by pi/2 (two-part constant via fma),
//            Taylor polynomials on |r| <= pi/4.
//   acos:    pi/2 - asin(x) on |x| <= 1/2, half-angle identities outside;
//            asin by its Maclaurin series (29 terms).
//   cbrt:    exponent split + Newton iterations (stands in for pow(x,1/3)).
// ---------------------------------------------------------------------------
#define CS_PIO2_HI 1.5707963267948966
#define CS_PIO2_LO 6.123233995736766e-17
#define CS_PI_HI 3.141592653589793
#define CS_PI_LO 1.2246467991473532e-16

__device__ inline double cs_asin_series(double z) {
  const double c[29] = {
      1.0, 0.16666666666666666, 0.075, 0.044642857142857144, 0.030381944444444444,
      0.022372159090909092, 0.017352764423076924, 0.01396484375, 0.011551800896139705,
      0.009761609529194078, 0.008390335809616815, 0.0073125258735988454,
      0.006447210311889649, 0.005740037670841924, 0.005153309682319905,
      0.004660143486915096, 0.004240907093679363, 0.003880964558837669,
      0.0035692053938259347, 0.003297059503473485, 0.0030578216492580306,
      0.002846178401108942, 0.00265787063820729, 0.0024894486782468836,
      0.002338091892111975, 0.0022014739737101384, 0.0020776610325181676,
      0.0019650336162772837, 0.0018622264064031275};
  double z2 = z * z;
  double acc = c[28];
#pragma unroll
  for (int k = 27; k >= 0; --k) acc = acc * z2 + c[k];
  return acc * z;
}

__device__ inline double cs_sin_poly(double r) {
  const double c[12] = {1.0, -0.16666666666666666, 0.008333333333333333,
                        -0.0001984126984126984, 2.7557319223985893e-06, -2.505210838544172e-08,
                        1.6059043836821613e-10, -7.647163731819816e-13, 2.8114572543455206e-15,
                        -8.22063524662433e-18, 1.9572941063391263e-20, -3.868170170630684e-23};
  double r2 = r * r;
  double acc = c[11];
#pragma unroll
  for (int k = 10; k >= 0; --k) acc = acc * r2 + c[k];
  return acc * r;
}

__device__ inline double cs_cos_poly(double r) {
  const double c[12] = {1.0, -0.5, 0.041666666666666664, -0.001388888888888889,
                        2.48015873015873e-05, -2.755731922398589e-07, 2.08767569878681e-09,
                        -1.1470745597729725e-11, 4.779477332387385e-14, -1.5619206968586225e-16,
                        4.110317623312165e-19, -8.896791392450574e-22};
  double r2 = r * r;
  double acc = c[11];
#pragma unroll
  for (int k = 10; k >= 0; --k) acc = acc * r2 + c[k];
  return acc;
}

__device__ inline int cs_reduce_pio2(double x, double* r) {
  double k = floor(x * (1.0 / CS_PIO2_HI) + 0.5);
  double t = fma(-k, CS_PIO2_HI, x);
  t = fma(-k, CS_PIO2_LO, t);
  *r = t;
  long long q = (long long)k;
  return (int)(q & 3);
}

__device__ inline double cs_det_cos(double x) {
  if (!(fabs(x) < 1e9)) return x - x;
  double r;
  int q = cs_reduce_pio2(x, &r);
  if (q == 0) return cs_cos_poly(r);
  if (q == 1) return -cs_sin_poly(r);
  if (q == 2) return -cs_cos_poly(r);
  return cs_sin_poly(r);
}

__device__ inline double cs_det_sin(double x) {
  if (!(fabs(x) < 1e9)) return x - x;
  double r;
  int q = cs_reduce_pio2(x, &r);
  if (q == 0) return cs_sin_poly(r);
  if (q == 1) return cs_cos_poly(r);
  if (q == 2) return -cs_sin_poly(r);
  return -cs_cos_poly(r);
}

__device__ inline double cs_det_acos(double x) {
  if (x != x || x > 1.0 || x < -1.0) return (x - x) / (x - x);
  if (x <= 0.5 && x >= -0.5) return (CS_PIO2_HI - cs_asin_series(x)) + CS_PIO2_LO;
  if (x > 0.5) {
    double z = sqrt((1.0 - x) * 0.5);
    return 2.0 * cs_asin_series(z);
  }
  double z = sqrt((1.0 + x) * 0.5);
  return (CS_PI_HI - 2.0 * cs_asin_series(z)) + CS_PI_LO;
}

__device__ inline double cs_det_cbrt(double x) {
  if (x == 0.0 || x != x || x > 1.7976931348623157e308) return x;
  int e;
  double m = frexp(x, &e);
  int q = e / 3, rem = e - 3 * q;
  if (rem < 0) {
    rem += 3;
    q -= 1;
  }
  m = ldexp(m, rem);
  double t = 0.6 + 0.3 * m;
#pragma unroll
  for (int i = 0; i < 6; ++i) t = t - (t * t * t - m) / (3.0 * t * t);
  return ldexp(t, q);
}

// Round a double to float toward -inf / +inf (conservative FP32 boxes).
__device__ inline float f32_down(double v) { return __double2float_rd(v); }
__device__ inline float f32_up(double v) { return __double2float_ru(v); }

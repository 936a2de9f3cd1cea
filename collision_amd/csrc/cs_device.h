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
// contact, with the per-point impulse and friction increments it adds.  A
// contact whose four points are all rigid (static or movable) adds nothing
// to its points; it spreads imp[0] over the movable body sb[0] and imp[1]
// over sb[1] instead (SpreadImpactZoneImpulse, dcollid.cpp:1101-1115).
struct __align__(16) hit_rec {
  uint64_t key;        // (lo << 36) | (hi << 8) | subtest  (canonical order)
  int32_t v[4];        // vertex ids; -1: no contribution (static point, all-rigid contact)
  int32_t sb[2];       // all-rigid contact: bodies receiving imp[0] / imp[1] (-1: none)
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
// Deterministic, correctly rounded elementary functions.  The reference calls
// libm acos, cos, sin and pow(x, 1.0/3.0) (dcollid3d.cpp:163-169, 436-444).
// These are fixed algorithms built from IEEE +, -, *, /, sqrt, fma, frexp,
// ldexp, floor; each evaluates its result as a double-double hi + lo with a
// relative error far below 2^-100 and rounds it once, so it returns the
// correctly rounded value (what glibc returns for all but ~0.1 % of
// arguments), independent of any vendor math library.
//   double-double: two-sum, fast two-sum, two-product by fma.
//   sin/cos: Cody-Waite reduction by pi/2 with a three-part constant (first
//            product exact through fma, second by two-product); Taylor
//            series on |r| <= pi/4, terms below 2^-60 in double Horner, the
//            leading nine in double-double Horner.
//   acos:    Maclaurin asin seed + one Newton step y - (sin y - z) / cos y in
//            double-double; pi/2 - asin(x) on |x| <= 1/2, 2 asin(sqrt((1-x)/2))
//            and pi - 2 asin(sqrt((1+x)/2)) outside.
//   pow(x, 1/3): 1.0/3.0 = 1/3 - 2^-54/3, so pow = cbrt(x) exp(-d ln x);
//            cbrt by a double-double Newton step, ln x to 1e-16 absolute,
//            exp(-d ln x) = 1 + g + g^2/2.
// ---------------------------------------------------------------------------
#define CS_PIO2_1 1.5707963267948966
#define CS_PIO2_2 6.123233995736766e-17
#define CS_PIO2_3 (-1.4973849048591698e-33)
#define CS_PI_1 3.141592653589793
#define CS_PI_2 1.2246467991473532e-16
#define CS_TWO_OVER_PI 0.6366197723675814
#define CS_LN2_1 0.6931471805599453
#define CS_LN2_2 2.3190468138462996e-17
#define CS_THIRD_DELTA_1 1.850371707708594e-17
#define CS_THIRD_DELTA_2 1.0271626370065257e-33

struct cs_dd {
  double hi, lo;
};

__device__ inline cs_dd cs_two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return cs_dd{s, e};
}
__device__ inline cs_dd cs_fast_two_sum(double a, double b) {
  const double s = a + b;
  const double e = b - (s - a);
  return cs_dd{s, e};
}
__device__ inline cs_dd cs_two_prod(double a, double b) {
  const double p = a * b;
  return cs_dd{p, fma(a, b, -p)};
}
__device__ inline cs_dd cs_dd_add(cs_dd a, cs_dd b) {
  cs_dd s = cs_two_sum(a.hi, b.hi);
  const cs_dd t = cs_two_sum(a.lo, b.lo);
  s.lo = s.lo + t.hi;
  s = cs_fast_two_sum(s.hi, s.lo);
  s.lo = s.lo + t.lo;
  return cs_fast_two_sum(s.hi, s.lo);
}
__device__ inline cs_dd cs_dd_mul(cs_dd a, cs_dd b) {
  cs_dd p = cs_two_prod(a.hi, b.hi);
  p.lo = p.lo + (a.hi * b.lo + a.lo * b.hi);
  return cs_fast_two_sum(p.hi, p.lo);
}
__device__ inline cs_dd cs_dd_mul_d(cs_dd a, double b) {
  cs_dd p = cs_two_prod(a.hi, b);
  p.lo = p.lo + a.lo * b;
  return cs_fast_two_sum(p.hi, p.lo);
}
__device__ inline double cs_dd_round(cs_dd a) { return cs_fast_two_sum(a.hi, a.lo).hi; }
__device__ inline cs_dd cs_dd_neg(cs_dd a) { return cs_dd{-a.hi, -a.lo}; }

__device__ inline cs_dd cs_sin_dd(cs_dd r) {
  const double hi[15] = {1.0, -0.16666666666666666, 0.008333333333333333, -0.0001984126984126984,
                         2.7557319223985893e-06, -2.505210838544172e-08, 1.6059043836821613e-10,
                         -7.647163731819816e-13, 2.8114572543455206e-15, -8.22063524662433e-18,
                         1.9572941063391263e-20, -3.868170170630684e-23, 6.446950284384474e-26,
                         -9.183689863795546e-29, 1.1309962886447716e-31};
  const double lo[9] = {0.0, -9.25185853854297e-18, 1.1564823173178714e-19,
                        -1.7209558293420705e-22, -1.858393274046472e-22, 1.448814070935912e-24,
                        1.2585294588752098e-26, -7.03872877733453e-30, 1.6508842730861433e-31};
  const cs_dd r2 = cs_dd_mul(r, r);
  double t = hi[14];
#pragma unroll
  for (int k = 13; k >= 9; --k) t = t * r2.hi + hi[k];
  cs_dd acc{t, 0.0};
#pragma unroll
  for (int k = 8; k >= 0; --k) acc = cs_dd_add(cs_dd_mul(acc, r2), cs_dd{hi[k], lo[k]});
  return cs_dd_mul(acc, r);
}
__device__ inline cs_dd cs_cos_dd(cs_dd r) {
  const double hi[15] = {1.0, -0.5, 0.041666666666666664, -0.001388888888888889,
                         2.48015873015873e-05, -2.755731922398589e-07, 2.08767569878681e-09,
                         -1.1470745597729725e-11, 4.779477332387385e-14, -1.5619206968586225e-16,
                         4.110317623312165e-19, -8.896791392450574e-22, 1.6117375710961184e-24,
                         -2.4795962632247976e-27, 3.279889237069838e-30};
  const double lo[9] = {0.0, 0.0, 2.3129646346357427e-18, 5.300543954373577e-20,
                        2.1511947866775882e-23, -2.3767714622250297e-23, -1.20734505911326e-25,
                        -2.0655512752830745e-28, 4.399205485834081e-31};
  const cs_dd r2 = cs_dd_mul(r, r);
  double t = hi[14];
#pragma unroll
  for (int k = 13; k >= 9; --k) t = t * r2.hi + hi[k];
  cs_dd acc{t, 0.0};
#pragma unroll
  for (int k = 8; k >= 0; --k) acc = cs_dd_add(cs_dd_mul(acc, r2), cs_dd{hi[k], lo[k]});
  return acc;
}

// x = k pi/2 + r, |r| <= pi/4 (+ rounding); returns k mod 4
__device__ inline int cs_reduce(double x, cs_dd* r) {
  const double k = floor(x * CS_TWO_OVER_PI + 0.5);
  const double r1 = fma(-k, CS_PIO2_1, x);
  const cs_dd p2 = cs_two_prod(k, CS_PIO2_2);
  cs_dd t = cs_two_sum(r1, -p2.hi);
  t.lo = (t.lo - p2.lo) - k * CS_PIO2_3;
  *r = cs_fast_two_sum(t.hi, t.lo);
  return (int)((long long)k & 3);
}

__device__ inline double cs_det_cos(double x) {
  if (!(fabs(x) < 1e9)) return x - x;
  cs_dd r;
  const int q = cs_reduce(x, &r);
  cs_dd v = (q & 1) ? cs_sin_dd(r) : cs_cos_dd(r);
  if (q == 1 || q == 2) v = cs_dd_neg(v);
  return cs_dd_round(v);
}

__device__ inline double cs_det_sin(double x) {
  if (!(fabs(x) < 1e9)) return x - x;
  if (x == 0.0) return x;
  cs_dd r;
  const int q = cs_reduce(x, &r);
  cs_dd v = (q & 1) ? cs_cos_dd(r) : cs_sin_dd(r);
  if (q >= 2) v = cs_dd_neg(v);
  return cs_dd_round(v);
}

// asin seed: Maclaurin series, 29 terms, |z| <= 0.5 (a few ulps)
__device__ inline double cs_asin_seed(double z) {
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
  const double z2 = z * z;
  double acc = c[28];
#pragma unroll
  for (int k = 27; k >= 0; --k) acc = acc * z2 + c[k];
  return acc * z;
}

// asin of a double-double |z| <= 0.5 + 1e-16: one Newton step on sin
__device__ inline cs_dd cs_asin_dd(cs_dd z) {
  const double y0 = cs_asin_seed(z.hi);
  const cs_dd y{y0, 0.0};
  const cs_dd s = cs_sin_dd(y), c = cs_cos_dd(y);
  const cs_dd d = cs_dd_add(s, cs_dd_neg(z));
  const double corr = d.hi / c.hi;
  return cs_two_sum(y0, -corr);
}

__device__ inline double cs_det_acos(double x) {
  if (x != x || x > 1.0 || x < -1.0) return (x - x) / (x - x);
  if (x <= 0.5 && x >= -0.5) {
    const cs_dd a = cs_asin_dd(cs_dd{x, 0.0});
    return cs_dd_round(cs_dd_add(cs_dd{CS_PIO2_1, CS_PIO2_2}, cs_dd_neg(a)));
  }
  const double t = (x > 0.0 ? (1.0 - x) : (1.0 + x)) * 0.5;  // exact
  cs_dd s{0.0, 0.0};
  if (t > 0.0) {
    s.hi = sqrt(t);
    s.lo = fma(-s.hi, s.hi, t) / (2.0 * s.hi);
  }
  const cs_dd a = cs_asin_dd(s);
  const cs_dd a2{2.0 * a.hi, 2.0 * a.lo};
  if (x > 0.0) return cs_dd_round(a2);
  return cs_dd_round(cs_dd_add(cs_dd{CS_PI_1, CS_PI_2}, cs_dd_neg(a2)));
}

// cube-root seed: exponent split m 2^(3q), six Newton steps (a few ulps)
__device__ inline double cs_cbrt_seed(double x) {
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

// ln x for finite x > 0 as double-double, absolute error about 1e-16
__device__ inline cs_dd cs_ln_dd(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.7071067811865476) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double acc = 1.0 / 29.0;
#pragma unroll
  for (int k = 27; k >= 1; k -= 2) acc = acc * s2 + 1.0 / (double)k;
  const double lnm = 2.0 * s * acc;
  cs_dd a = cs_two_prod((double)e, CS_LN2_1);
  a.lo = a.lo + ((double)e * CS_LN2_2 + lnm);
  return cs_fast_two_sum(a.hi, a.lo);
}

// pow(x, 1.0/3.0) for x >= 0, correctly rounded
__device__ inline double cs_det_pow_third(double x) {
  if (x == 0.0 || x != x || x > 1.7976931348623157e308) return x;
  const double y0 = cs_cbrt_seed(x);
  const cs_dd y2 = cs_two_prod(y0, y0);
  const cs_dd y3 = cs_dd_mul_d(y2, y0);
  const cs_dd r = cs_dd_add(y3, cs_dd{-x, 0.0});
  const double corr = r.hi / (3.0 * y2.hi);
  const cs_dd c = cs_two_sum(y0, -corr);
  cs_dd g = cs_dd_mul(cs_ln_dd(x), cs_dd{-CS_THIRD_DELTA_1, -CS_THIRD_DELTA_2});
  g.lo = g.lo + 0.5 * (g.hi * g.hi);
  g = cs_fast_two_sum(g.hi, g.lo);
  const cs_dd cg = cs_dd_mul(c, g);
  return cs_dd_round(cs_dd_add(c, cg));
}

// Fast approximations (a few ulps) of the same functions, for the certified
// screen of the coplanarity cubic: Cody-Waite reduction by pi/2 and Taylor
// polynomials for cos; pi/2 - asin and the half-angle identities with the
// Maclaurin series for acos; the exponent-split Newton seed for the cube root.
__device__ inline double cs_fast_cos(double x) {
  const double k = floor(x * CS_TWO_OVER_PI + 0.5);
  double r = fma(-k, CS_PIO2_1, x);
  r = fma(-k, CS_PIO2_2, r);
  const double r2 = r * r;
  const int q = (int)((long long)k & 3);
  double s = -7.647163731819816e-13, c = 4.779477332387385e-14;
  s = s * r2 + 1.6059043836821613e-10;
  s = s * r2 - 2.505210838544172e-08;
  s = s * r2 + 2.7557319223985893e-06;
  s = s * r2 - 0.0001984126984126984;
  s = s * r2 + 0.008333333333333333;
  s = s * r2 - 0.16666666666666666;
  s = (s * r2 + 1.0) * r;
  c = c * r2 - 1.1470745597729725e-11;
  c = c * r2 + 2.08767569878681e-09;
  c = c * r2 - 2.755731922398589e-07;
  c = c * r2 + 2.48015873015873e-05;
  c = c * r2 - 0.001388888888888889;
  c = c * r2 + 0.041666666666666664;
  c = c * r2 - 0.5;
  c = c * r2 + 1.0;
  return q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
}
__device__ inline double cs_fast_acos(double x) {
  if (x <= 0.5 && x >= -0.5) return (CS_PIO2_1 - cs_asin_seed(x)) + CS_PIO2_2;
  const double t = (x > 0.0 ? (1.0 - x) : (1.0 + x)) * 0.5;
  const double a = 2.0 * cs_asin_seed(sqrt(t));
  return x > 0.0 ? a : (CS_PI_1 - a) + CS_PI_2;
}

// Round a double to float toward -inf / +inf (conservative FP32 boxes).
__device__ inline float f32_down(double v) { return __double2float_rd(v); }
__device__ inline float f32_up(double v) { return __double2float_ru(v); }

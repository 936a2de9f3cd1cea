/*
 * Deterministic, correctly rounded elementary functions for the oracle (test
 * infrastructure).
 *
 * The reference evaluates acos, cos, sin and pow(x, 1.0/3.0) through the host
 * libm (dcollid3d.cpp:436-444, 163-169).  The oracle's second math mode
 * computes these four functions with the fixed algorithms below, built only
 * from IEEE-754 +, -, *, /, sqrt, fma, frexp and ldexp.  The HIP kernels use
 * the same algorithms with the same constants and operation order, so oracle
 * and GPU are bitwise comparable.
 *
 * Every function returns the correctly rounded (round-to-nearest) value: it
 * evaluates the result as an unevaluated double-double sum hi + lo with a
 * relative error far below 2^-100 and rounds that sum once.  glibc's
 * acos/cos/sin/pow are correctly rounded in all but a tiny fraction of
 * arguments (its documented worst cases are 0.5x ulp), so this mode returns
 * libm's bits for nearly every argument; the tests count the exceptions and
 * check them against 60-digit decimal arithmetic.
 *
 * Methods (textbook):
 *   double-double: Knuth's two-sum, Dekker's fast two-sum, two-product by
 *             fma.
 *   sin/cos : Cody-Waite reduction by pi/2 with a three-part constant (the
 *             first product is exact through fma, the second by two-product),
 *             then the Taylor series on |r| <= pi/4: terms below 2^-60 in
 *             double Horner, the leading nine in double-double Horner.
 *   acos    : a seed from the Maclaurin series of asin, refined by one Newton
 *             step y - (sin y - z) / cos y with double-double sin/cos;
 *             acos(x) = pi/2 - asin(x) for |x| <= 1/2, 2 asin(sqrt((1-x)/2))
 *             for x > 1/2, pi - 2 asin(sqrt((1+x)/2)) for x < -1/2 (the
 *             square root carried in double-double).
 *   pow(x, 1/3) : the double 1.0/3.0 equals 1/3 - d with d = 2^-54/3 exactly,
 *             so pow(x, 1.0/3.0) = cbrt(x) * exp(-d ln x); cbrt by a
 *             double-double Newton step from an exponent-split seed, ln x to
 *             1e-16 absolute (e ln2 in double-double plus an atanh series),
 *             exp(-d ln x) = 1 + g + g^2/2 (|g| < 1.4e-14).
 */
#ifndef ORC_DETMATH_H
#define ORC_DETMATH_H

#include <math.h>

#define ORC_PIO2_1 1.5707963267948966
#define ORC_PIO2_2 6.123233995736766e-17
#define ORC_PIO2_3 (-1.4973849048591698e-33)
#define ORC_PI_1 3.141592653589793
#define ORC_PI_2 1.2246467991473532e-16
#define ORC_TWO_OVER_PI 0.6366197723675814
#define ORC_LN2_1 0.6931471805599453
#define ORC_LN2_2 2.3190468138462996e-17
#define ORC_THIRD_DELTA_1 1.850371707708594e-17    /* 1/3 - (double)(1.0/3.0) */
#define ORC_THIRD_DELTA_2 1.0271626370065257e-33

typedef struct { double hi, lo; } orc_dd;

static inline orc_dd orc_two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    orc_dd r = {s, e};
    return r;
}
static inline orc_dd orc_fast_two_sum(double a, double b) { /* |a| >= |b| or a == 0 */
    double s = a + b;
    double e = b - (s - a);
    orc_dd r = {s, e};
    return r;
}
static inline orc_dd orc_two_prod(double a, double b) {
    double p = a * b;
    orc_dd r = {p, fma(a, b, -p)};
    return r;
}
static inline orc_dd orc_dd_add(orc_dd a, orc_dd b) {
    orc_dd s = orc_two_sum(a.hi, b.hi);
    orc_dd t = orc_two_sum(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = orc_fast_two_sum(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return orc_fast_two_sum(s.hi, s.lo);
}
static inline orc_dd orc_dd_mul(orc_dd a, orc_dd b) {
    orc_dd p = orc_two_prod(a.hi, b.hi);
    p.lo = p.lo + (a.hi * b.lo + a.lo * b.hi);
    return orc_fast_two_sum(p.hi, p.lo);
}
static inline orc_dd orc_dd_mul_d(orc_dd a, double b) {
    orc_dd p = orc_two_prod(a.hi, b);
    p.lo = p.lo + a.lo * b;
    return orc_fast_two_sum(p.hi, p.lo);
}
/* the double nearest to hi + lo */
static inline double orc_dd_round(orc_dd a) { return orc_fast_two_sum(a.hi, a.lo).hi; }

/* (-1)^k / (2k+1)! and (-1)^k / (2k)! as double-double, k = 0..14 */
static const double orc_sin_hi[15] = {
    1.0, -0.16666666666666666, 0.008333333333333333, -0.0001984126984126984,
    2.7557319223985893e-06, -2.505210838544172e-08, 1.6059043836821613e-10,
    -7.647163731819816e-13, 2.8114572543455206e-15, -8.22063524662433e-18,
    1.9572941063391263e-20, -3.868170170630684e-23, 6.446950284384474e-26,
    -9.183689863795546e-29, 1.1309962886447716e-31};
static const double orc_sin_lo[9] = {
    0.0, -9.25185853854297e-18, 1.1564823173178714e-19, -1.7209558293420705e-22,
    -1.858393274046472e-22, 1.448814070935912e-24, 1.2585294588752098e-26,
    -7.03872877733453e-30, 1.6508842730861433e-31};
static const double orc_cos_hi[15] = {
    1.0, -0.5, 0.041666666666666664, -0.001388888888888889, 2.48015873015873e-05,
    -2.755731922398589e-07, 2.08767569878681e-09, -1.1470745597729725e-11,
    4.779477332387385e-14, -1.5619206968586225e-16, 4.110317623312165e-19,
    -8.896791392450574e-22, 1.6117375710961184e-24, -2.4795962632247976e-27,
    3.279889237069838e-30};
static const double orc_cos_lo[9] = {
    0.0, 0.0, 2.3129646346357427e-18, 5.300543954373577e-20, 2.1511947866775882e-23,
    -2.3767714622250297e-23, -1.20734505911326e-25, -2.0655512752830745e-28,
    4.399205485834081e-31};

/* sin(r) and cos(r) for a double-double |r| <= pi/4 + 1e-12 */
static inline orc_dd orc_sin_dd(orc_dd r) {
    orc_dd r2 = orc_dd_mul(r, r);
    double t = orc_sin_hi[14];
    for (int k = 13; k >= 9; --k) t = t * r2.hi + orc_sin_hi[k];
    orc_dd acc = {t, 0.0};
    for (int k = 8; k >= 0; --k) {
        orc_dd c = {orc_sin_hi[k], orc_sin_lo[k]};
        acc = orc_dd_add(orc_dd_mul(acc, r2), c);
    }
    return orc_dd_mul(acc, r);
}
static inline orc_dd orc_cos_dd(orc_dd r) {
    orc_dd r2 = orc_dd_mul(r, r);
    double t = orc_cos_hi[14];
    for (int k = 13; k >= 9; --k) t = t * r2.hi + orc_cos_hi[k];
    orc_dd acc = {t, 0.0};
    for (int k = 8; k >= 0; --k) {
        orc_dd c = {orc_cos_hi[k], orc_cos_lo[k]};
        acc = orc_dd_add(orc_dd_mul(acc, r2), c);
    }
    return acc;
}

/* x = k pi/2 + r, |r| <= pi/4 (+ rounding); returns k mod 4.  Exact first
 * step: x - k P1 is representable for |k| < 2^20, so the fma does not round. */
static inline int orc_reduce(double x, orc_dd *r) {
    double k = floor(x * ORC_TWO_OVER_PI + 0.5);
    double r1 = fma(-k, ORC_PIO2_1, x);
    orc_dd p2 = orc_two_prod(k, ORC_PIO2_2);
    orc_dd t = orc_two_sum(r1, -p2.hi);
    t.lo = (t.lo - p2.lo) - k * ORC_PIO2_3;
    *r = orc_fast_two_sum(t.hi, t.lo);
    return (int)((long long)k & 3);
}

static inline orc_dd orc_cos_full(double x) {
    orc_dd r;
    int q = orc_reduce(x, &r);
    orc_dd v = (q & 1) ? orc_sin_dd(r) : orc_cos_dd(r);
    if (q == 1 || q == 2) { v.hi = -v.hi; v.lo = -v.lo; }
    return v;
}
static inline orc_dd orc_sin_full(double x) {
    orc_dd r;
    int q = orc_reduce(x, &r);
    orc_dd v = (q & 1) ? orc_cos_dd(r) : orc_sin_dd(r);
    if (q >= 2) { v.hi = -v.hi; v.lo = -v.lo; }
    return v;
}

static inline double orc_det_cos(double x) {
    if (!(fabs(x) < 1e9)) return x - x;
    return orc_dd_round(orc_cos_full(x));
}
static inline double orc_det_sin(double x) {
    if (!(fabs(x) < 1e9)) return x - x;
    if (x == 0.0) return x;
    return orc_dd_round(orc_sin_full(x));
}

/* asin seed: Maclaurin series, 29 terms, |z| <= 0.5 (a few ulps) */
static const double orc_asin_c[29] = {
    1.0, 0.16666666666666666, 0.075, 0.044642857142857144, 0.030381944444444444,
    0.022372159090909092, 0.017352764423076924, 0.01396484375, 0.011551800896139705,
    0.009761609529194078, 0.008390335809616815, 0.0073125258735988454,
    0.006447210311889649, 0.005740037670841924, 0.005153309682319905,
    0.004660143486915096, 0.004240907093679363, 0.003880964558837669,
    0.0035692053938259347, 0.003297059503473485, 0.0030578216492580306,
    0.002846178401108942, 0.00265787063820729, 0.0024894486782468836,
    0.002338091892111975, 0.0022014739737101384, 0.0020776610325181676,
    0.0019650336162772837, 0.0018622264064031275};
static inline double orc_asin_seed(double z) {
    double z2 = z * z;
    double acc = orc_asin_c[28];
    for (int k = 27; k >= 0; --k) acc = acc * z2 + orc_asin_c[k];
    return acc * z;
}

/* asin(z) for a double-double |z| <= 0.5 + 1e-16: one Newton step on sin */
static inline orc_dd orc_asin_dd(orc_dd z) {
    double y0 = orc_asin_seed(z.hi);
    orc_dd y = {y0, 0.0};
    orc_dd s = orc_sin_dd(y), c = orc_cos_dd(y);
    orc_dd mz = {-z.hi, -z.lo};
    orc_dd d = orc_dd_add(s, mz);
    double corr = d.hi / c.hi;
    return orc_two_sum(y0, -corr);
}

static inline double orc_det_acos(double x) {
    if (x != x || x > 1.0 || x < -1.0) return (x - x) / (x - x);
    if (x <= 0.5 && x >= -0.5) {
        orc_dd z = {x, 0.0};
        orc_dd a = orc_asin_dd(z);
        orc_dd p = {ORC_PIO2_1, ORC_PIO2_2};
        orc_dd ma = {-a.hi, -a.lo};
        return orc_dd_round(orc_dd_add(p, ma));
    }
    double t = (x > 0.0 ? (1.0 - x) : (1.0 + x)) * 0.5;   /* exact */
    orc_dd s = {0.0, 0.0};
    if (t > 0.0) {
        s.hi = sqrt(t);
        s.lo = fma(-s.hi, s.hi, t) / (2.0 * s.hi);
    }
    orc_dd a = orc_asin_dd(s);
    orc_dd a2 = {2.0 * a.hi, 2.0 * a.lo};
    if (x > 0.0) return orc_dd_round(a2);
    orc_dd p = {ORC_PI_1, ORC_PI_2};
    orc_dd ma = {-a2.hi, -a2.lo};
    return orc_dd_round(orc_dd_add(p, ma));
}

/* cube-root seed: exponent split m 2^(3q), six Newton steps (a few ulps) */
static inline double orc_cbrt_seed(double x) {
    int e;
    double m = frexp(x, &e);
    int q = e / 3, rem = e - 3 * q;
    if (rem < 0) { rem += 3; q -= 1; }
    m = ldexp(m, rem);
    double t = 0.6 + 0.3 * m;
    for (int i = 0; i < 6; ++i) t = t - (t * t * t - m) / (3.0 * t * t);
    return ldexp(t, q);
}

/* ln x for finite x > 0 as double-double, absolute error about 1e-16 */
static inline orc_dd orc_ln_dd(double x) {
    int e;
    double m = frexp(x, &e);                 /* [0.5, 1) */
    if (m < 0.7071067811865476) { m = m * 2.0; e = e - 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double acc = 1.0 / 29.0;
    for (int k = 27; k >= 1; k -= 2) acc = acc * s2 + 1.0 / (double)k;
    double lnm = 2.0 * s * acc;
    orc_dd a = orc_two_prod((double)e, ORC_LN2_1);
    a.lo = a.lo + ((double)e * ORC_LN2_2 + lnm);
    return orc_fast_two_sum(a.hi, a.lo);
}

/* pow(x, 1.0/3.0) for x >= 0, correctly rounded */
static inline double orc_det_pow_third(double x) {
    if (x == 0.0 || x != x || x > 1.7976931348623157e308) return x;
    double y0 = orc_cbrt_seed(x);
    orc_dd y2 = orc_two_prod(y0, y0);
    orc_dd y3 = orc_dd_mul_d(y2, y0);
    orc_dd mx = {-x, 0.0};
    orc_dd r = orc_dd_add(y3, mx);
    double corr = r.hi / (3.0 * y2.hi);
    orc_dd c = orc_two_sum(y0, -corr);
    /* g = -d ln x + (d ln x)^2 / 2 */
    orc_dd md = {-ORC_THIRD_DELTA_1, -ORC_THIRD_DELTA_2};
    orc_dd g = orc_dd_mul(orc_ln_dd(x), md);
    g.lo = g.lo + 0.5 * (g.hi * g.hi);
    g = orc_fast_two_sum(g.hi, g.lo);
    orc_dd cg = orc_dd_mul(c, g);
    return orc_dd_round(orc_dd_add(c, cg));
}

#endif

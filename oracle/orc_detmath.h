/*
 * Deterministic elementary functions for the oracle (test infrastructure).
 *
 * The reference evaluates acos, cos, sin and pow(x, 1/3) through the host
 * libm (dcollid3d.cpp:436-444, 163-169).  A GPU's math library rounds those
 * differently in the last bits, so the oracle offers a second mode in which
 * these four functions are computed by the fixed algorithms below, built only
 * from IEEE-754 +, -, *, /, sqrt, fma, frexp and ldexp.  The same algorithms,
 * with the same constants and operation order, are what the HIP kernels use,
 * which makes oracle and GPU bitwise comparable.  Accuracy against libm is
 * a few ulps; the tests check it.
 *
 * Methods (textbook):
 *   sin/cos : Cody-Waite reduction by pi/2 with a two-part constant applied
 *             through fma, then Taylor polynomials on |r| <= pi/4 (Horner).
 *   acos    : acos(x) = pi/2 - asin(x) for |x| <= 1/2, and the half-angle
 *             identities acos(x) = 2 asin(sqrt((1-x)/2)) (x > 1/2),
 *             pi - 2 asin(sqrt((1+x)/2)) (x < -1/2); asin by its Maclaurin
 *             series sum_n (2n)!/(4^n (n!)^2 (2n+1)) z^(2n+1), 29 terms.
 *   cbrt    : exponent split m*2^(3q), 4 Newton steps t <- t - (t^3-m)/(3t^2)
 *             from a linear seed; stands in for pow(x, 1.0/3.0), x >= 0.
 */
#ifndef ORC_DETMATH_H
#define ORC_DETMATH_H

#include <math.h>

#define ORC_PIO2_HI 1.5707963267948966
#define ORC_PIO2_LO 6.123233995736766e-17
#define ORC_PI_HI 3.141592653589793
#define ORC_PI_LO 1.2246467991473532e-16

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

static const double orc_sin_c[12] = {
    1.0, -0.16666666666666666, 0.008333333333333333, -0.0001984126984126984,
    2.7557319223985893e-06, -2.505210838544172e-08, 1.6059043836821613e-10,
    -7.647163731819816e-13, 2.8114572543455206e-15, -8.22063524662433e-18,
    1.9572941063391263e-20, -3.868170170630684e-23};

static const double orc_cos_c[12] = {
    1.0, -0.5, 0.041666666666666664, -0.001388888888888889, 2.48015873015873e-05,
    -2.755731922398589e-07, 2.08767569878681e-09, -1.1470745597729725e-11,
    4.779477332387385e-14, -1.5619206968586225e-16, 4.110317623312165e-19,
    -8.896791392450574e-22};

static inline double orc_sin_poly(double r) {
    double r2 = r * r;
    double acc = orc_sin_c[11];
    for (int k = 10; k >= 0; --k) acc = acc * r2 + orc_sin_c[k];
    return acc * r;
}

static inline double orc_cos_poly(double r) {
    double r2 = r * r;
    double acc = orc_cos_c[11];
    for (int k = 10; k >= 0; --k) acc = acc * r2 + orc_cos_c[k];
    return acc;
}

/* reduce x = k*pi/2 + r, |r| <= ~pi/4; returns k mod 4 */
static inline int orc_reduce_pio2(double x, double *r) {
    double k = floor(x * (1.0 / ORC_PIO2_HI) + 0.5);
    double t = fma(-k, ORC_PIO2_HI, x);
    t = fma(-k, ORC_PIO2_LO, t);
    *r = t;
    long long q = (long long)k;
    return (int)(q & 3);
}

static inline double orc_det_cos(double x) {
    if (!(fabs(x) < 1e9)) return x - x; /* nan/inf -> nan; huge args unused */
    double r;
    int q = orc_reduce_pio2(x, &r);
    switch (q) {
        case 0: return orc_cos_poly(r);
        case 1: return -orc_sin_poly(r);
        case 2: return -orc_cos_poly(r);
        default: return orc_sin_poly(r);
    }
}

static inline double orc_det_sin(double x) {
    if (!(fabs(x) < 1e9)) return x - x;
    double r;
    int q = orc_reduce_pio2(x, &r);
    switch (q) {
        case 0: return orc_sin_poly(r);
        case 1: return orc_cos_poly(r);
        case 2: return -orc_sin_poly(r);
        default: return -orc_cos_poly(r);
    }
}

static inline double orc_asin_series(double z) { /* |z| <= 0.5 */
    double z2 = z * z;
    double acc = orc_asin_c[28];
    for (int k = 27; k >= 0; --k) acc = acc * z2 + orc_asin_c[k];
    return acc * z;
}

static inline double orc_det_acos(double x) {
    if (x != x || x > 1.0 || x < -1.0) return (x - x) / (x - x); /* nan */
    if (x <= 0.5 && x >= -0.5)
        return (ORC_PIO2_HI - orc_asin_series(x)) + ORC_PIO2_LO;
    if (x > 0.5) {
        double z = sqrt((1.0 - x) * 0.5);
        return 2.0 * orc_asin_series(z);
    }
    double z = sqrt((1.0 + x) * 0.5);
    return (ORC_PI_HI - 2.0 * orc_asin_series(z)) + ORC_PI_LO;
}

/* cube root of x >= 0 (stands in for pow(x, 1.0/3.0)) */
static inline double orc_det_cbrt(double x) {
    if (x == 0.0 || x != x || x > 1.7976931348623157e308) return x;
    int e;
    double m = frexp(x, &e);           /* x = m 2^e, m in [0.5, 1) */
    int q = e / 3, rem = e - 3 * q;    /* e = 3q + rem, rem in {-2..2} */
    if (rem < 0) { rem += 3; q -= 1; } /* rem in {0,1,2} */
    m = ldexp(m, rem);                 /* m in [0.5, 4) */
    double t = 0.6 + 0.3 * m;          /* seed within ~25% on [0.5, 4) */
    for (int i = 0; i < 6; ++i) t = t - (t * t * t - m) / (3.0 * t * t);
    return ldexp(t, q);
}

#endif

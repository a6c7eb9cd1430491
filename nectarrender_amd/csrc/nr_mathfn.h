// nr_mathfn.h — deterministic float transcendentals for the path tracer.
//
// The reference draws volume free-flight distances from a hash of a ray's
// float bits (hittable/include/hittable/volumes.h:35, random/include/hash.h:107-124),
// so a one-ulp difference in any earlier sin/cos/log turns into an unrelated
// path.  To keep host and device bit-identical, the device-side functions of
// the hot path (sin, cos, acos, atan2, log, pow) are evaluated here in IEEE
// double with a fixed operation order and rounded once to float; +,-,*,/,sqrt
// in double are correctly rounded on x86-64 and gfx950 alike.  Methods:
//   sin/cos  Cody-Waite reduction by pi/2 (two-part constant) + Taylor to x^21
//   log      x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)) series
//   exp      x = k ln2 + r, |r| <= ln2/2, Taylor to r^15, scaled by 2^k
//   atan     two half-angle reductions t -> t / (1 + sqrt(1 + t^2)), Taylor to t^27
//   acos     atan2(sqrt(1 - x^2), x) (x^2 and 1 - x^2 are exact in double for float x)
//   pow      exp(y log x) in double, with the IEEE special cases the path tracer meets
//   log10    log / ln 10 in double; 10^k for whole k by repeated multiplication
// Accuracy: < 1 ulp of float against float64 references (tests/test_oracle.py, test_mathfn_*).
// Compile with -ffp-contract=off so no step is fused.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define NRM_FN __host__ __device__ inline
#else
#define NRM_FN inline
#endif

namespace nrm {

NRM_FN uint64_t dbits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
NRM_FN double bitsd(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }
NRM_FN double dabs(double x) { return x < 0.0 ? -x : x; }
NRM_FN double dsqrt(double x);

// 2^k for integer k in the normal double range
NRM_FN double pow2i(int k) { return bitsd((uint64_t)(k + 1023) << 52); }

NRM_FN double sin_poly(double r) {
    double z = r * r;
    double p = 1.9572941063391263e-20;
    p = p * z + -8.22063524662433e-18;
    p = p * z + 2.8114572543455206e-15;
    p = p * z + -7.647163731819816e-13;
    p = p * z + 1.6059043836821613e-10;
    p = p * z + -2.505210838544172e-08;
    p = p * z + 2.7557319223985893e-06;
    p = p * z + -0.0001984126984126984;
    p = p * z + 0.008333333333333333;
    p = p * z + -0.16666666666666666;
    return r + (r * z) * p;
}
NRM_FN double cos_poly(double r) {
    double z = r * r;
    double p = 4.110317623312165e-19;
    p = p * z + -1.5619206968586225e-16;
    p = p * z + 4.779477332387385e-14;
    p = p * z + -1.1470745597729725e-11;
    p = p * z + 2.08767569878681e-09;
    p = p * z + -2.755731922398589e-07;
    p = p * z + 2.48015873015873e-05;
    p = p * z + -0.001388888888888889;
    p = p * z + 0.041666666666666664;
    p = p * z + -0.5;
    return 1.0 + z * p;
}
// reduce x by multiples of pi/2: returns r in about [-pi/4, pi/4] and the quadrant
NRM_FN double reduce_pio2(double x, int& q) {
    double k = x * 0.6366197723675814;
    k = k < 0.0 ? -(double)(int64_t)(0.5 - k) : (double)(int64_t)(k + 0.5);
    q = (int)((int64_t)k & 3);
    return (x - k * 1.5707963267341256) - k * 6.077100506506192e-11;
}

NRM_FN float sinf_(float xf) {
    if (xf != xf) return xf;
    double x = (double)xf;
    if (dabs(x) > 1.0e9) return 0.0f;   // outside the range the path tracer uses
    int q;
    double r = reduce_pio2(x, q);
    double v = (q == 0) ? sin_poly(r) : (q == 1) ? cos_poly(r) : (q == 2) ? -sin_poly(r) : -cos_poly(r);
    return (float)v;
}
NRM_FN float cosf_(float xf) {
    if (xf != xf) return xf;
    double x = (double)xf;
    if (dabs(x) > 1.0e9) return 1.0f;
    int q;
    double r = reduce_pio2(x, q);
    double v = (q == 0) ? cos_poly(r) : (q == 1) ? -sin_poly(r) : (q == 2) ? -cos_poly(r) : sin_poly(r);
    return (float)v;
}

// natural log of a positive finite double
NRM_FN double log_pos(double x) {
    uint64_t b = dbits(x);
    int e = (int)((b >> 52) & 0x7FF);
    if (e == 0) {                       // subnormal: scale up by 2^54
        x = x * 18014398509481984.0;
        b = dbits(x);
        e = (int)((b >> 52) & 0x7FF) - 54;
    }
    e -= 1023;
    double m = bitsd((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);   // [1, 2)
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = 0.047619047619047616;
    p = p * z + 0.05263157894736842;
    p = p * z + 0.058823529411764705;
    p = p * z + 0.06666666666666667;
    p = p * z + 0.07692307692307693;
    p = p * z + 0.09090909090909091;
    p = p * z + 0.1111111111111111;
    p = p * z + 0.14285714285714285;
    p = p * z + 0.2;
    p = p * z + 0.3333333333333333;
    double lm = 2.0 * s + (2.0 * s * z) * p;
    return (double)e * 0.6931471805599453 + lm;
}
NRM_FN float logf_(float xf) {
    if (xf != xf) return xf;
    if (xf < 0.0f) return bitsd(0x7FF8000000000000ull);
    if (xf == 0.0f) return -bitsd(0x7FF0000000000000ull);
    if (xf == bitsd(0x7FF0000000000000ull)) return xf;
    return (float)log_pos((double)xf);
}

NRM_FN double exp_d(double x) {
    if (x > 709.0) return bitsd(0x7FF0000000000000ull);
    if (x < -745.0) return 0.0;
    double kf = x * 1.4426950408889634;
    kf = kf < 0.0 ? -(double)(int64_t)(0.5 - kf) : (double)(int64_t)(kf + 0.5);
    int k = (int)kf;
    double r = (x - kf * 0.6931471803691238) - kf * 1.9082149292705877e-10;
    double p = 7.647163731819816e-13;
    p = p * r + 1.1470745597729725e-11;
    p = p * r + 1.6059043836821613e-10;
    p = p * r + 2.08767569878681e-09;
    p = p * r + 2.505210838544172e-08;
    p = p * r + 2.755731922398589e-07;
    p = p * r + 2.7557319223985893e-06;
    p = p * r + 2.48015873015873e-05;
    p = p * r + 0.0001984126984126984;
    p = p * r + 0.001388888888888889;
    p = p * r + 0.008333333333333333;
    p = p * r + 0.041666666666666664;
    p = p * r + 0.16666666666666666;
    p = p * r + 0.5;
    double er = 1.0 + (r + (r * r) * p);
    if (k > 1000) return er * pow2i(1000) * pow2i(k - 1000);
    if (k < -1000) return er * pow2i(-1000) * pow2i(k + 1000);
    return er * pow2i(k);
}

// atan of |t| <= 1 in double
NRM_FN double atan_unit(double t) {
    double t1 = t / (1.0 + dsqrt(1.0 + t * t));
    double t2 = t1 / (1.0 + dsqrt(1.0 + t1 * t1));
    double z = t2 * t2;
    double p = -0.037037037037037035;
    p = p * z + 0.04;
    p = p * z + -0.043478260869565216;
    p = p * z + 0.047619047619047616;
    p = p * z + -0.05263157894736842;
    p = p * z + 0.058823529411764705;
    p = p * z + -0.06666666666666667;
    p = p * z + 0.07692307692307693;
    p = p * z + -0.09090909090909091;
    p = p * z + 0.1111111111111111;
    p = p * z + -0.14285714285714285;
    p = p * z + 0.2;
    p = p * z + -0.3333333333333333;
    return 4.0 * (t2 + (t2 * z) * p);
}

NRM_FN double atan2_d(double y, double x) {
    const double PI = 3.141592653589793, PIO2 = 1.5707963267948966;
    if (x != x || y != y) return x + y;
    double ax = dabs(x), ay = dabs(y);
    bool yneg = (dbits(y) >> 63) != 0, xneg = (dbits(x) >> 63) != 0;
    double a;
    const double INF = bitsd(0x7FF0000000000000ull);
    if (ay == 0.0) a = xneg ? PI : 0.0;
    else if (ax == 0.0) a = PIO2;
    else if (ax == INF && ay == INF) a = xneg ? 3.0 * PIO2 * 0.5 : PIO2 * 0.5;
    else if (ax == INF) a = xneg ? PI : 0.0;
    else if (ay == INF) a = PIO2;
    else {
        a = ay <= ax ? atan_unit(ay / ax) : PIO2 - atan_unit(ax / ay);
        if (xneg) a = PI - a;
    }
    return yneg ? -a : a;
}
NRM_FN float atan2f_(float y, float x) { return (float)atan2_d((double)y, (double)x); }

NRM_FN float acosf_(float xf) {
    if (xf != xf) return xf;
    if (xf > 1.0f || xf < -1.0f) return bitsd(0x7FF8000000000000ull);
    double x = (double)xf;
    return (float)atan2_d(dsqrt(1.0 - x * x), x);
}

// powf for the cases the path tracer meets (x >= 0, or x < 0 with integral y)
NRM_FN float powf_(float xf, float yf) {
    if (yf == 0.0f) return 1.0f;
    if (xf != xf || yf != yf) return xf + yf;
    double x = (double)xf, y = (double)yf;
    if (xf == 0.0f) return yf > 0.0f ? 0.0f : bitsd(0x7FF0000000000000ull);
    bool neg = false;
    if (x < 0.0) {
        double yi = (double)(int64_t)y;
        if (yi != y) return bitsd(0x7FF8000000000000ull);
        neg = ((int64_t)y & 1) != 0;
        x = -x;
    }
    if (x == bitsd(0x7FF0000000000000ull)) return (float)(yf > 0.0f ? (neg ? -x : x) : 0.0);
    double v = exp_d(y * log_pos(x));
    return (float)(neg ? -v : v);
}

// log10 of a positive finite float: log(x) / ln 10 in double, rounded once
NRM_FN float log10f_pos(float xf) { return (float)(log_pos((double)xf) / 2.302585092994046); }

// 10^k for an integer k >= 0 (a float exponent that is a whole number): exact in
// double up to 10^22, so the float result is the correctly rounded power there
NRM_FN float pow10f_int(int k) {
    double v = 1.0;
    for (int i = 0; i < k; i++) v = v * 10.0;
    return (float)v;
}

NRM_FN double dsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_sqrt(x);
#else
    return __builtin_sqrt(x);
#endif
}

}  // namespace nrm

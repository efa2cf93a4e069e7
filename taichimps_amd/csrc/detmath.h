// sin and exp for the servo's strain styles (wiggle, trate), written out so
// that every build of them (device or host) rounds the same way: the GPU's
// libm and the host's glibc may differ in the last bit, and one bit of the
// box changes a trajectory. Within two ulps of libm (exp within one).
//
// Method: Cody-Waite range reduction (pi/2 in three parts, ln 2 in two; the
// leading parts have 33 and 32 significant bits, so k * part is exact for the
// |k| < 2^20 this path meets), then Taylor polynomials in Horner form on
// |r| <= pi/4 (sin, cos; truncation < 1e-19) and |r| <= ln(2)/2 (exp;
// truncation < 1e-20), then the quadrant or the power of two.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace tmps {

__host__ __device__ inline double det_sin_poly(double r) {
#pragma clang fp contract(off)
    const double z = r * r;
    double p = -2.8114572543455206e-15;         // -1/17!
    p = p * z + 7.647163731819816e-13;          //  1/15!
    p = p * z - 1.6059043836821613e-10;         // -1/13!
    p = p * z + 2.505210838544172e-08;          //  1/11!
    p = p * z - 2.7557319223985893e-06;         // -1/9!
    p = p * z + 0.0001984126984126984;          //  1/7!
    p = p * z - 0.008333333333333333;           // -1/5!
    p = p * z + 0.16666666666666666;            //  1/3!
    return r - (r * z) * p;
}

__host__ __device__ inline double det_cos_poly(double r) {
#pragma clang fp contract(off)
    const double z = r * r;
    double p = 1.5619206968586225e-16;          //  1/18!
    p = p * z - 4.779477332387385e-14;          // -1/16!
    p = p * z + 1.1470745597729725e-11;         //  1/14!
    p = p * z - 2.08767569878681e-09;           // -1/12!
    p = p * z + 2.755731922398589e-07;          //  1/10!
    p = p * z - 2.48015873015873e-05;           // -1/8!
    p = p * z + 0.001388888888888889;           //  1/6!
    p = p * z - 0.041666666666666664;           // -1/4!
    p = p * z + 0.5;                            //  1/2!
    return 1.0 - z * p;
}

__host__ __device__ inline double det_sin(double x) {
#pragma clang fp contract(off)
    const double k = rint(x * 0.6366197723675814);  // 2/pi
    const double r = ((x - k * 1.5707963267341256) - k * 6.077100506303966e-11) -
                     k * 2.0222662487959506e-21;
    const long long q = (long long)k & 3;
    if (q == 0) return det_sin_poly(r);
    if (q == 1) return det_cos_poly(r);
    if (q == 2) return -det_sin_poly(r);
    return -det_cos_poly(r);
}

__host__ __device__ inline double det_exp(double x) {
#pragma clang fp contract(off)
    const double k = rint(x * 1.4426950408889634);  // 1/ln 2
    const double r = (x - k * 0.6931471806019545) - k * -4.2009150726810846e-11;
    double p = 7.647163731819816e-13;           // 1/15!
    p = p * r + 1.1470745597729725e-11;         // 1/14!
    p = p * r + 1.6059043836821613e-10;         // 1/13!
    p = p * r + 2.08767569878681e-09;           // 1/12!
    p = p * r + 2.505210838544172e-08;          // 1/11!
    p = p * r + 2.755731922398589e-07;          // 1/10!
    p = p * r + 2.7557319223985893e-06;         // 1/9!
    p = p * r + 2.48015873015873e-05;           // 1/8!
    p = p * r + 0.0001984126984126984;          // 1/7!
    p = p * r + 0.001388888888888889;           // 1/6!
    p = p * r + 0.008333333333333333;           // 1/5!
    p = p * r + 0.041666666666666664;           // 1/4!
    p = p * r + 0.16666666666666666;            // 1/3!
    p = p * r + 0.5;                            // 1/2!
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)k);
}

}  // namespace tmps

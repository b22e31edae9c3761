// noise.h - counter-based white noise for the neuron updates.
//
// The reference draws curand_normal from a 48-byte XORWOW state per neuron
// (src/ephys.cu:4-10,82).  Here a draw is a pure function of (seed, step,
// neuron): no state is stored or read, graphs replay correctly, and any
// subset of neurons or steps can be regenerated.
//
//   words  = Philox4x64-10(counter = {step, neuron, kNoisePurpose, 0},
//                          key = {seed, 0})         (Salmon, Moraes, Dror,
//            Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11)
//   u1     = ((words[0] >> 11) + 1) 2^-53  in (0, 1]
//   u2     = (words[1] >> 11) 2^-53        in [0, 1)
//   z      = sqrt(-2 ln u1) cos(2 pi u2)   (Box & Muller, Ann. Math.
//            Statist. 29, 1958), sqrt correctly rounded
//
// ln and cos are this file's own, a fixed sequence of IEEE double
// operations (explicit fma), so a CPU restatement produces the same bits:
//   ln u  = e ln2 + ln m with u = m 2^e, m in [sqrt(1/2), sqrt(2)):
//           ln m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^23/23),
//           s = (m - 1) / (m + 1), |s| <= 0.1716; ln2 split into a 42-bit
//           head (e ln2_hi exact) and a tail;
//   cos 2 pi u: t = u - rint(u), a = |t| in [0, 1/2] reduced to an octant
//           (cos 2 pi a = C(a), S(1/4 - a) or -C(1/2 - a), each argument
//           exact and within [-1/8, 1/8]); C and S are the Taylor series of
//           cos 2 pi r (to r^18) and sin 2 pi r (to r^19) in r, whose
//           coefficients (2 pi)^k / k! are rounded to double.
// (Range reduction and polynomial evaluation as in J.-M. Muller,
// "Elementary Functions: Algorithms and Implementation".)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nnx {

constexpr uint64_t kNoisePurpose = 0x4e4f495345ull;  // "NOISE"

__device__ __forceinline__ void philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0,
                                              uint64_t k1, uint64_t out[4]) {
    constexpr uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    constexpr uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += W0;
            k1 += W1;
        }
        const uint64_t hi0 = __umul64hi(M0, c0), lo0 = M0 * c0;
        const uint64_t hi1 = __umul64hi(M1, c2), lo1 = M1 * c2;
        const uint64_t y0 = hi1 ^ c1 ^ k0, y2 = hi0 ^ c3 ^ k1;
        c0 = y0;
        c1 = lo1;
        c2 = y2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// ln u for u in (0, 1] (normal numbers)
__device__ __forceinline__ double noise_log(double u) {
    const uint64_t b = (uint64_t)__double_as_longlong(u);
    int e = (int)((b >> 52) & 0x7ff) - 1022;  // u = f 2^e, f in [1/2, 1)
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffull) | 0x3fe0000000000000ull));
    if (m < 0x1.6a09e667f3bcdp-1) {
        m = m + m;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 0x1.642c8590b2164p-4;
    p = __builtin_fma(p, z, 0x1.8618618618618p-4);
    p = __builtin_fma(p, z, 0x1.af286bca1af28p-4);
    p = __builtin_fma(p, z, 0x1.e1e1e1e1e1e1ep-4);
    p = __builtin_fma(p, z, 0x1.1111111111111p-3);
    p = __builtin_fma(p, z, 0x1.3b13b13b13b14p-3);
    p = __builtin_fma(p, z, 0x1.745d1745d1746p-3);
    p = __builtin_fma(p, z, 0x1.c71c71c71c71cp-3);
    p = __builtin_fma(p, z, 0x1.2492492492492p-2);
    p = __builtin_fma(p, z, 0x1.999999999999ap-2);
    p = __builtin_fma(p, z, 0x1.5555555555555p-1);
    const double lm = __builtin_fma(s * z, p, s + s);
    const double de = (double)e;
    return __builtin_fma(de, 0x1.62e42fefa3800p-1, __builtin_fma(de, 0x1.ef35793c76730p-45, lm));
}

// cos 2 pi r and sin 2 pi r for |r| <= 1/8
__device__ __forceinline__ double noise_cos2pi_r(double r) {
    const double y = r * r;
    double p = -0x1.2a0c591af8314p-5;
    p = __builtin_fma(p, y, 0x1.20c62c2f2d7f5p-2);
    p = __builtin_fma(p, y, -0x1.b6e24f44b128fp+0);
    p = __builtin_fma(p, y, 0x1.f9d38a3763cc3p+2);
    p = __builtin_fma(p, y, -0x1.a6d1f2a204a8cp+4);
    p = __builtin_fma(p, y, 0x1.e1f506891babbp+5);
    p = __builtin_fma(p, y, -0x1.55d3c7e3cbffap+6);
    p = __builtin_fma(p, y, 0x1.03c1f081b5ac4p+6);
    p = __builtin_fma(p, y, -0x1.3bd3cc9be45dep+4);
    return __builtin_fma(p, y, 1.0);
}
__device__ __forceinline__ double noise_sin2pi_r(double r) {
    const double y = r * r;
    double p = -0x1.8a404211f9547p-7;
    p = __builtin_fma(p, y, 0x1.aaec32af93359p-4);
    p = __builtin_fma(p, y, -0x1.6fadb9f155744p-1);
    p = __builtin_fma(p, y, 0x1.e8f434d018d63p+1);
    p = __builtin_fma(p, y, -0x1.e3074fde8871fp+3);
    p = __builtin_fma(p, y, 0x1.50783487ee782p+5);
    p = __builtin_fma(p, y, -0x1.32d2cce62bd86p+6);
    p = __builtin_fma(p, y, 0x1.466bc6775aae2p+6);
    p = __builtin_fma(p, y, -0x1.4abbce625be53p+5);
    p = __builtin_fma(p, y, 0x1.921fb54442d18p+2);
    return p * r;
}

// cos 2 pi u for u in [0, 1)
__device__ __forceinline__ double noise_cos2pi(double u) {
    const double t = u - __builtin_rint(u);
    const double a = t < 0.0 ? -t : t;
    if (a <= 0.125) return noise_cos2pi_r(a);
    if (a <= 0.375) return noise_sin2pi_r(0.25 - a);
    return -noise_cos2pi_r(0.5 - a);
}

// Correctly rounded sqrt of x >= 0: the hardware-seeded estimate y, then
// Tuckerman's test with exact fma residuals picks y or a neighbour:
// sqrt(x) rounds to y iff y (y - u-) < x <= y (y + u+), u-/u+ the spacings
// below and above y.
__device__ __forceinline__ double noise_sqrt(double x) {
    if (!(x > 0.0)) return x;
    double y = __builtin_sqrt(x);
    const double yu = __longlong_as_double(__double_as_longlong(y) + 1);
    const double yd = __longlong_as_double(__double_as_longlong(y) - 1);
    if (__builtin_fma(y, yu, -x) < 0.0) y = yu;
    else if (__builtin_fma(y, yd, -x) >= 0.0) y = yd;
    return y;
}

// standard normal draw of (seed, step, neuron)
__device__ __forceinline__ double noise_normal(uint64_t seed, uint64_t step, uint64_t neuron) {
    uint64_t w[4];
    philox4x64_10(step, neuron, kNoisePurpose, 0, seed, 0, w);
    const double u1 = (double)((w[0] >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(w[1] >> 11) * 0x1.0p-53;
    return noise_sqrt(-2.0 * noise_log(u1)) * noise_cos2pi(u2);
}

}  // namespace nnx

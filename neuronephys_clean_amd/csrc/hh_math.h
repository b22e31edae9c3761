// hh_math.h - device-side Hodgkin-Huxley arithmetic (gfx950).
//
// Every function is a fixed sequence of IEEE-754 operations (the library is
// compiled with -ffp-contract=off; fused multiply-adds are written as fma),
// so a run is reproducible bit for bit and can be checked against a CPU
// restatement performing the same sequence.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nnx {

// ---------------------------------------------------------------------------
// exp(x): x = k ln2 + r with |r| <= ln2/2 (Cody-Waite: ln2 = hi + lo, hi with
// 16 (float) / 42 (double) significant bits so k*hi is exact), e^r by its
// Taylor polynomial (degree 7 / 13, Horner with fma), scaled by 2^k with
// ldexp (exact).  Arguments are clamped so that the result stays normal.
// Max error 1 ulp against a correctly rounded exp over the clamped range.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float dexp(float x) {
    if (x != x) return x;
    x = __builtin_amdgcn_fmed3f(x, -80.0f, 88.0f);  // the clamp to [-80, 88] (x is not NaN here)
    const float k = __builtin_rintf(x * 0x1.715476p+0f);
    float r = __builtin_fmaf(-k, 0x1.62e4p-1f, x);
    r = __builtin_fmaf(-k, 0x1.7f7d1cp-20f, r);
    float p = 0x1.a01a02p-13f;
    p = __builtin_fmaf(p, r, 0x1.6c16c2p-10f);
    p = __builtin_fmaf(p, r, 0x1.111112p-7f);
    p = __builtin_fmaf(p, r, 0x1.555556p-5f);
    p = __builtin_fmaf(p, r, 0x1.555556p-3f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return __builtin_amdgcn_ldexpf(p, (int)k);  // exact: the result is normal
}

__device__ __forceinline__ double dexp(double x) {
    if (x != x) return x;
    x = x > 708.0 ? 708.0 : x;
    x = x < -700.0 ? -700.0 : x;
    const double k = __builtin_rint(x * 0x1.71547652b82fep+0);
    double r = __builtin_fma(-k, 0x1.62e42fefa38p-1, x);
    r = __builtin_fma(-k, 0x1.ef35793c7673p-45, r);
    double p = 0x1.6124613a86d09p-33;
    p = __builtin_fma(p, r, 0x1.1eed8eff8d898p-29);
    p = __builtin_fma(p, r, 0x1.ae64567f544e4p-26);
    p = __builtin_fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = __builtin_fma(p, r, 0x1.71de3a556c734p-19);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-16);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, r, 0x1.1111111111111p-7);
    p = __builtin_fma(p, r, 0x1.5555555555555p-5);
    p = __builtin_fma(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)k);
}

// exp for arguments computed from a neuron's V: the same sequence without
// the NaN test (a NaN V stays NaN through the update whatever these return)
__device__ __forceinline__ float dexp_v(float x) {
    x = __builtin_amdgcn_fmed3f(x, -80.0f, 88.0f);
    const float k = __builtin_rintf(x * 0x1.715476p+0f);
    float r = __builtin_fmaf(-k, 0x1.62e4p-1f, x);
    r = __builtin_fmaf(-k, 0x1.7f7d1cp-20f, r);
    float p = 0x1.a01a02p-13f;
    p = __builtin_fmaf(p, r, 0x1.6c16c2p-10f);
    p = __builtin_fmaf(p, r, 0x1.111112p-7f);
    p = __builtin_fmaf(p, r, 0x1.555556p-5f);
    p = __builtin_fmaf(p, r, 0x1.555556p-3f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return __builtin_amdgcn_ldexpf(p, (int)k);
}
__device__ __forceinline__ double dexp_v(double x) {
    x = x > 708.0 ? 708.0 : x;
    x = x < -700.0 ? -700.0 : x;
    const double k = __builtin_rint(x * 0x1.71547652b82fep+0);
    double r = __builtin_fma(-k, 0x1.62e42fefa38p-1, x);
    r = __builtin_fma(-k, 0x1.ef35793c7673p-45, r);
    double p = 0x1.6124613a86d09p-33;
    p = __builtin_fma(p, r, 0x1.1eed8eff8d898p-29);
    p = __builtin_fma(p, r, 0x1.ae64567f544e4p-26);
    p = __builtin_fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = __builtin_fma(p, r, 0x1.71de3a556c734p-19);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-16);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, r, 0x1.1111111111111p-7);
    p = __builtin_fma(p, r, 0x1.5555555555555p-5);
    p = __builtin_fma(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)k);
}

template <class T> __device__ __forceinline__ T dfma(T a, T b, T c);
template <> __device__ __forceinline__ float dfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <class T> __device__ __forceinline__ T dabs(T a) { return a < T(0) ? -a : a; }

// The Bernoulli series of a / (1 - e^-a): 1 + a/2 + a^2/12 - a^4/720 +
// a^6/30240 - a^8/1209600, used for |a| < 1/2 (it removes the 0/0 at a = 0).
template <class T> __device__ __forceinline__ T vtrap_ser(T a) {
    const T a2 = a * a;
    T q = dfma(a2, T(-1.0 / 1209600.0), T(1.0 / 30240.0));
    q = dfma(a2, q, T(-1.0 / 720.0));
    q = dfma(a2, q, T(1.0 / 12.0));
    return dfma(a2, q, dfma(a, T(0.5), T(1)));
}

template <class T> struct Rates { T am, bm, ah, bh, an, bn; };

// HH rates of the reference (src/ephys.cu:57-63, written for a -60 mV rest),
// evaluated at u = V + v_shift with three exponentials:
//   a = 3.5 + 0.1u, ea = e^-a:  am = a/(1-e^-a); b = a + 1.5, e^-b = ea e^-1.5,
//   an = 0.1 b/(1-e^-b); bh = 1/(e^(-3-0.1u) + 1) = 1/(ea e^0.5 + 1);
//   with w = u + 60:  bn = 0.125 e^(-w/80), ah = 0.07 e^(-u/20-3) =
//   0.07 (e^(-w/80))^4, bm = 4 e^(-w/18); the arguments -w/80 and -w/18 are
//   formed as fma(u, -1/80, -60/80) and fma(u, -1/18, -60/18).
// The three quotients share one division: with d1 = 1 - e^-a, d2 = 1 - e^-b,
// d3 = ea e^0.5 + 1 (a denominator replaced by 1 where its series is used)
// and inv = 1 / (d1 d2 d3): am = a (d2 d3) inv, an = 0.1 b (d1 d3) inv,
// bh = (d1 d2) inv.  Below a = -20 (V near -240 mV and lower) the product
// could overflow, and three divisions are used.
template <class T> __device__ __forceinline__ Rates<T> rates(T u) {
    Rates<T> R;
    const T a = dfma(T(0.1), u, T(3.5));
    const T ea = dexp_v(-a);
    const T b = a + T(1.5);
    const T d1 = T(1) - ea;
    const T d2 = dfma(-ea, T(0x1.c8f87724b5c1dp-3), T(1));
    const T d3 = dfma(ea, T(0x1.a61298e1e069cp+0), T(1));
    const bool sa = dabs(a) < T(0.5), sb = dabs(b) < T(0.5);
    if (a > T(-20)) {
        const T e1 = sa ? T(1) : d1, e2 = sb ? T(1) : d2;
        const T e12 = e1 * e2;
        const T inv = T(1) / (e12 * d3);
        const T t1 = e2 * d3, t2 = e1 * d3;
        R.am = (a * t1) * inv;
        R.an = T(0.1) * ((b * t2) * inv);
        R.bh = e12 * inv;
        // the series near a = 0 or b = 0 (V within 5 mV of -35 or -50 in the
        // reference's convention): evaluated only by wavefronts that need it
        if (__any(sa || sb)) {
            if (sa) R.am = vtrap_ser(a);
            if (sb) R.an = T(0.1) * vtrap_ser(b);
        }
    } else {
        R.am = a / d1;
        R.an = T(0.1) * (b / d2);
        R.bh = T(1) / d3;
    }
    const T e80 = dexp_v(dfma(u, T(-0.0125), T(-0.75)));
    const T e40 = e80 * e80;
    R.bn = T(0.125) * e80;
    R.ah = T(0.07) * (e40 * e40);
    R.bm = T(4) * dexp_v(dfma(u, T(-1.0 / 18.0), T(-60.0 / 18.0)));
    return R;
}

// Model constants in T (cast once on the host from the double parameters).
template <class T> struct HHConst {
    T C, invC, gna, gk, gl, ena, ek, el, vshift, esyn;
    T tmax, vt, kp, ar, ad, g_local, e_local, g_i, e_i, g_e, e_e, noise_mean, noise_std;
    T v_thresh;
    // the level rule compares the output, widened, with this double (the
    // reference's CheckFiringOperator tests a Float against the literal 0.1)
    double out_thresh;
    // step constants: dt, dt/2, dt/6, exp(-dt/(2 tau)), exp(-dt/tau)
    T dt, hdt, dt6, eh, ef;
};

// Right-hand side of the RK4 model: gates as a - (a + b) x, the current
// sum as a chain of fma: i = i_ext - g (V - E) for Na, K, leak, synapse.
template <class T>
__device__ __forceinline__ void deriv(const HHConst<T>& p, T V, T m, T h, T n, T s, T iext, T& dV,
                                      T& dm, T& dh, T& dn) {
    const Rates<T> R = rates(V + p.vshift);
    dm = dfma(-(R.am + R.bm), m, R.am);
    dh = dfma(-(R.ah + R.bh), h, R.ah);
    dn = dfma(-(R.an + R.bn), n, R.an);
    const T m3h = m * m * m * h;
    const T n2 = n * n;
    const T n4 = n2 * n2;
    T i = dfma(-(p.gna * m3h), V - p.ena, iext);
    i = dfma(-(p.gk * n4), V - p.ek, i);
    i = dfma(-p.gl, V - p.el, i);
    i = dfma(-s, V - p.esyn, i);
    dV = i * p.invC;
}

// x + c (k1 + 2 k2 + 2 k3 + k4) with the sum left to right and one fma at the end
template <class T> __device__ __forceinline__ T rk4_comb(T x, T c, T k1, T k2, T k3, T k4) {
    T t = dfma(T(2), k2, k1);
    t = dfma(T(2), k3, t);
    return dfma(c, t + k4, x);
}

// Classical RK4 over (V, m, h, n), stage arguments x + c k as fma.  The
// synaptic conductance jumps by the coupling input at the start of the step
// and decays exactly within it: s0 = s + g_in, s(dt/2) = s0 eh, s(dt) = s0 ef.
template <class T>
__device__ __forceinline__ void rk4(const HHConst<T>& p, T& V, T& m, T& h, T& n, T& s, T iext, T gin) {
    const T s0 = s + gin;
    const T sh = s0 * p.eh;
    const T sf = s0 * p.ef;
    T k1V, k1m, k1h, k1n, k2V, k2m, k2h, k2n, k3V, k3m, k3h, k3n, k4V, k4m, k4h, k4n;
    deriv(p, V, m, h, n, s0, iext, k1V, k1m, k1h, k1n);
    deriv(p, dfma(p.hdt, k1V, V), dfma(p.hdt, k1m, m), dfma(p.hdt, k1h, h), dfma(p.hdt, k1n, n), sh, iext, k2V,
          k2m, k2h, k2n);
    deriv(p, dfma(p.hdt, k2V, V), dfma(p.hdt, k2m, m), dfma(p.hdt, k2h, h), dfma(p.hdt, k2n, n), sh, iext, k3V,
          k3m, k3h, k3n);
    deriv(p, dfma(p.dt, k3V, V), dfma(p.dt, k3m, m), dfma(p.dt, k3h, h), dfma(p.dt, k3n, n), sf, iext, k4V, k4m,
          k4h, k4n);
    V = rk4_comb(V, p.dt6, k1V, k2V, k3V, k4V);
    m = rk4_comb(m, p.dt6, k1m, k2m, k3m, k4m);
    h = rk4_comb(h, p.dt6, k1h, k2h, k3h, k4h);
    n = rk4_comb(n, p.dt6, k1n, k2n, k3n, k4n);
    s = sf;
}

// Kinetic synapse of the reference (src/ephys.cu:69-76): trapezoid update
// of Y' = Ar T (1 - Y) - Ad Y with T = IN / COUNT (0 without inputs).
template <class T>
__device__ __forceinline__ T kinetic(const HHConst<T>& p, T in, T cnt, T Y) {
    const T tavg = cnt > T(0) ? in / cnt : T(0);
    return (p.ar * tavg * p.dt + (T(1) - p.hdt * (p.ar * tavg + p.ad)) * Y) /
           (p.hdt * (p.ar * tavg + p.ad) + T(1));
}

}  // namespace nnx

// oracle.hpp - CPU restatement of the neuronephys per-timestep algorithm.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (neuronephys_clean_amd/,
// include/) includes, links or calls this file; only tests/, the smoke check
// and bench.py's CPU-baseline leg load it (through liboracle.so).
//
// Templated on T in {float, double}.  Built with -ffp-contract=off and
// explicit std::fma wherever a fused multiply-add is meant, so that every
// result is a fixed sequence of IEEE-754 operations: the device kernels
// perform the same sequence and must agree bit for bit.  Accuracy of that
// sequence is pinned separately (libm, scipy DOP853, closed forms).
//
// Reference files restated (cited per function below):
//   src/ephys.cu:28-96          Leapfrog_HH
//   src/defines.cuh:87-102      hh_params
//   src/coupling.cu:34-48,103-178, src_wholebrain/coupling.cu:12-69
//   src/network_factory.cu:5-27 CSR + in-degree
//   src/paracrine.cu:19-134     27-point CN-CG
//   src/main.cu:166-181         step order coupling -> paracrine -> ephys
//   src/electrode_net.cu:3-47, src_wholebrain/electrode_net.cu:3-126
//                               electrode readout and synchrony
//   src_wholebrain/hebbian.cu:3-69   Hebbian learning
//   src/ephys.cu:250-347        Leapfrog_diekman (SCN neuron)
//   src/ephys.cu:105-141,203-245  Euler_an_ie_model (SAN neuron)
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace oracle {

// --------------------------------------------------------------------------
// exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, Taylor polynomial of
// e^r by Horner with fma, scaled by 2^k with ldexp (exact).
// ln2 is split so that k*ln2_hi is exact (16 bits for float, 42 for double).
// The argument is clamped to [xmin, xmax] so the result stays normal.
// --------------------------------------------------------------------------
template <class T> struct ExpK;
template <> struct ExpK<float> {
    static constexpr float log2e = 0x1.715476p+0f;
    static constexpr float ln2_hi = 0x1.62e4p-1f;
    static constexpr float ln2_lo = 0x1.7f7d1cp-20f;
    static constexpr float xmax = 88.0f;
    static constexpr float xmin = -80.0f;
    static constexpr int deg = 7;
    static constexpr float c[8] = {1.0f, 1.0f, 0.5f, 0x1.555556p-3f, 0x1.555556p-5f,
                                   0x1.111112p-7f, 0x1.6c16c2p-10f, 0x1.a01a02p-13f};
};
template <> struct ExpK<double> {
    static constexpr double log2e = 0x1.71547652b82fep+0;
    static constexpr double ln2_hi = 0x1.62e42fefa38p-1;
    static constexpr double ln2_lo = 0x1.ef35793c7673p-45;
    static constexpr double xmax = 708.0;
    static constexpr double xmin = -700.0;
    static constexpr int deg = 13;
    static constexpr double c[14] = {1.0, 1.0, 0.5, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
                                     0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10,
                                     0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16,
                                     0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22,
                                     0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29,
                                     0x1.6124613a86d09p-33};
};

template <class T> inline T oexp(T x) {
    using K = ExpK<T>;
    if (x != x) return x;
    x = x > K::xmax ? K::xmax : x;
    x = x < K::xmin ? K::xmin : x;
    T k = std::rint(x * K::log2e);
    T r = std::fma(-k, K::ln2_hi, x);
    r = std::fma(-k, K::ln2_lo, r);
    T p = K::c[K::deg];
    for (int i = K::deg - 1; i >= 0; --i) p = std::fma(p, r, K::c[i]);
    return std::ldexp(p, (int)k);  // exact: the result is normal
}

// The Bernoulli series of a / (1 - e^-a): 1 + a/2 + a^2/12 - a^4/720 +
// a^6/30240 - a^8/1209600, selected for |a| < 1/2; it removes the 0/0 of the
// reference's alpha_m and alpha_n at V = -35 and V = -50 (src/ephys.cu:57-58).
template <class T> inline T vtrap_ser(T a) {
    const T a2 = a * a;
    T q = std::fma(a2, T(-1.0 / 1209600.0), T(1.0 / 30240.0));
    q = std::fma(a2, q, T(-1.0 / 720.0));
    q = std::fma(a2, q, T(1.0 / 12.0));
    return std::fma(a2, q, std::fma(a, T(0.5), T(1)));
}

template <class T> struct Rates { T am, bm, ah, bh, an, bn; };

// The reference's six HH rates (src/ephys.cu:57-63), -60 mV-rest form,
// evaluated at u = V + v_shift with three exponentials:
//   a = 3.5 + 0.1u, ea = e^-a:  am = a/(1-e^-a), an = 0.1 b/(1-e^-b) with
//   b = a + 1.5 and e^-b = ea e^-1.5, bh = 1/(e^(-3-0.1u) + 1) = 1/(ea e^0.5 + 1);
//   w = u + 60:  bn = 0.125 e^(-w/80), ah = 0.07 e^(-u/20-3) = 0.07 (e^(-w/80))^4,
//   bm = 4 e^(-w/18).
// a = fma(0.1, u, 3.5), d2 = fma(-ea, e^-1.5, 1), d3 = fma(ea, e^0.5, 1); the
// arguments -w/80 and -w/18 are fma(u, -1/80, -60/80) and fma(u, -1/18, -60/18).
// The three quotients share one division: with d1 = 1 - e^-a, d2 = 1 - e^-b,
// d3 = ea e^0.5 + 1 (a denominator replaced by 1 where its series is used) and
// inv = 1 / (d1 d2 d3): am = a (d2 d3) inv, an = 0.1 b (d1 d3) inv,
// bh = (d1 d2) inv; below a = -20 three divisions (the product could overflow).
template <class T> inline Rates<T> rates(T u) {
    Rates<T> R;
    const T a = std::fma(T(0.1), u, T(3.5));
    const T ea = oexp(-a);
    const T b = a + T(1.5);
    const T d1 = T(1) - ea;
    const T d2 = std::fma(-ea, T(0x1.c8f87724b5c1dp-3), T(1));
    const T d3 = std::fma(ea, T(0x1.a61298e1e069cp+0), T(1));
    const bool sa = std::fabs(a) < T(0.5), sb = std::fabs(b) < T(0.5);
    if (a > T(-20)) {
        const T e1 = sa ? T(1) : d1, e2 = sb ? T(1) : d2;
        const T e12 = e1 * e2;
        const T inv = T(1) / (e12 * d3);
        const T t1 = e2 * d3, t2 = e1 * d3;
        R.am = sa ? vtrap_ser(a) : (a * t1) * inv;
        R.an = T(0.1) * (sb ? vtrap_ser(b) : (b * t2) * inv);
        R.bh = e12 * inv;
    } else {
        R.am = a / d1;
        R.an = T(0.1) * (b / d2);
        R.bh = T(1) / d3;
    }
    const T e80 = oexp(std::fma(u, T(-0.0125), T(-0.75)));
    const T e40 = e80 * e80;
    R.bn = T(0.125) * e80;
    R.ah = T(0.07) * (e40 * e40);
    R.bm = T(4) * oexp(std::fma(u, T(-1.0 / 18.0), T(-60.0 / 18.0)));
    return R;
}

// --------------------------------------------------------------------------
// white noise: a standard normal per (seed, step, neuron) from Philox4x64-10
// (Salmon et al., SC'11; counter {step, neuron, purpose, 0}, key {seed, 0})
// and the Box-Muller transform (Box & Muller 1958) with ln u from the atanh
// series after splitting off the exponent and cos 2 pi u from Taylor series
// after an octant reduction (J.-M. Muller, "Elementary Functions"); every
// operation an IEEE double one, fma where written, sqrt correctly rounded.
// Replaces curand_normal of Leapfrog_HH (src/ephys.cu:82), whose XORWOW
// stream is not reproduced.
// --------------------------------------------------------------------------
namespace noise {
constexpr uint64_t kPurpose = 0x4e4f495345ull;

inline void philox(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0, uint64_t k1, uint64_t o[4]) {
    const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B97F4A7C15ull;
            k1 += 0xBB67AE8584CAA73Bull;
        }
        const __uint128_t p0 = (__uint128_t)M0 * c0, p1 = (__uint128_t)M1 * c2;
        const uint64_t n0 = (uint64_t)(p1 >> 64) ^ c1 ^ k0, n2 = (uint64_t)(p0 >> 64) ^ c3 ^ k1;
        c1 = (uint64_t)p1;
        c3 = (uint64_t)p0;
        c0 = n0;
        c2 = n2;
    }
    o[0] = c0, o[1] = c1, o[2] = c2, o[3] = c3;
}

inline double bits_to_double(uint64_t b) {
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}
inline uint64_t double_to_bits(double d) {
    uint64_t b;
    std::memcpy(&b, &d, 8);
    return b;
}

inline double log_u(double u) {
    const uint64_t b = double_to_bits(u);
    int e = (int)((b >> 52) & 0x7ff) - 1022;
    double m = bits_to_double((b & 0x000fffffffffffffull) | 0x3fe0000000000000ull);
    if (m < 0x1.6a09e667f3bcdp-1) {
        m = m + m;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0), z = s * s;
    static const double c[11] = {0x1.642c8590b2164p-4, 0x1.8618618618618p-4, 0x1.af286bca1af28p-4,
                                 0x1.e1e1e1e1e1e1ep-4, 0x1.1111111111111p-3, 0x1.3b13b13b13b14p-3,
                                 0x1.745d1745d1746p-3, 0x1.c71c71c71c71cp-3, 0x1.2492492492492p-2,
                                 0x1.999999999999ap-2, 0x1.5555555555555p-1};
    double p = c[0];
    for (int k = 1; k < 11; ++k) p = std::fma(p, z, c[k]);
    const double lm = std::fma(s * z, p, s + s);
    const double de = (double)e;
    return std::fma(de, 0x1.62e42fefa3800p-1, std::fma(de, 0x1.ef35793c76730p-45, lm));
}

inline double cos_r(double r) {
    static const double c[9] = {0x1.20c62c2f2d7f5p-2, -0x1.b6e24f44b128fp+0, 0x1.f9d38a3763cc3p+2,
                                -0x1.a6d1f2a204a8cp+4, 0x1.e1f506891babbp+5, -0x1.55d3c7e3cbffap+6,
                                0x1.03c1f081b5ac4p+6, -0x1.3bd3cc9be45dep+4, 1.0};
    const double y = r * r;
    double p = -0x1.2a0c591af8314p-5;
    for (int k = 0; k < 9; ++k) p = std::fma(p, y, c[k]);
    return p;
}
inline double sin_r(double r) {
    static const double c[9] = {0x1.aaec32af93359p-4, -0x1.6fadb9f155744p-1, 0x1.e8f434d018d63p+1,
                                -0x1.e3074fde8871fp+3, 0x1.50783487ee782p+5, -0x1.32d2cce62bd86p+6,
                                0x1.466bc6775aae2p+6, -0x1.4abbce625be53p+5, 0x1.921fb54442d18p+2};
    const double y = r * r;
    double p = -0x1.8a404211f9547p-7;
    for (int k = 0; k < 9; ++k) p = std::fma(p, y, c[k]);
    return p * r;
}
inline double cos2pi(double u) {
    const double t = u - std::rint(u);
    const double a = t < 0.0 ? -t : t;
    if (a <= 0.125) return cos_r(a);
    if (a <= 0.375) return sin_r(0.25 - a);
    return -cos_r(0.5 - a);
}

inline double normal(uint64_t seed, uint64_t step, uint64_t neuron) {
    uint64_t w[4];
    philox(step, neuron, kPurpose, 0, seed, 0, w);
    const double u1 = (double)((w[0] >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(w[1] >> 11) * 0x1.0p-53;
    return std::sqrt(-2.0 * log_u(u1)) * cos2pi(u2);
}
}  // namespace noise

// Parameters cast to T once (the host passes doubles; both sides cast the
// same doubles, so both see the same T values).
template <class T> struct P {
    T C, invC, gna, gk, gl, ena, ek, el, vshift, esyn;
    T tmax, vt, kp, ar, ad, g_local, e_local, g_i, e_i, g_e, e_e, noise_mean;
    T v_thresh;
    double out_thresh;  // the level rule widens the output (CheckFiringOperator: Float > 0.1)
};

// dy/dt of the RK4 model (V, m, h, n) with synaptic conductance s.
template <class T>
inline void hh_deriv(const P<T>& p, T V, T m, T h, T n, T s, T iext, T& dV, T& dm, T& dh, T& dn) {
    Rates<T> R = rates(V + p.vshift);
    // gates as a - (a + b) x, currents as an fma chain i = i_ext - g (V - E)
    dm = std::fma(-(R.am + R.bm), m, R.am);
    dh = std::fma(-(R.ah + R.bh), h, R.ah);
    dn = std::fma(-(R.an + R.bn), n, R.an);
    T m3h = m * m * m * h;
    T n2 = n * n;
    T n4 = n2 * n2;
    T i = std::fma(-(p.gna * m3h), V - p.ena, iext);
    i = std::fma(-(p.gk * n4), V - p.ek, i);
    i = std::fma(-p.gl, V - p.el, i);
    i = std::fma(-s, V - p.esyn, i);
    dV = i * p.invC;
}

// x + c (k1 + 2 k2 + 2 k3 + k4), the sum left to right, one fma at the end
template <class T> inline T rk4_comb(T x, T c, T k1, T k2, T k3, T k4) {
    T t = std::fma(T(2), k2, k1);
    t = std::fma(T(2), k3, t);
    return std::fma(c, t + k4, x);
}

// Classical RK4 over (V, m, h, n); s is integrated exactly: s0 = s + g_in at
// the start of the step, s(t) = s0 e^{-t/tau}; eh = e^{-dt/(2 tau)},
// ef = e^{-dt/tau}, hdt = dt/2, dt6 = dt/6 (all cast from double by the host).
template <class T>
inline void rk4_step(const P<T>& p, T dt, T hdt, T dt6, T eh, T ef, T& V, T& m, T& h, T& n, T& s,
                     T iext, T gin) {
    T s0 = s + gin;
    T sh = s0 * eh;
    T sf = s0 * ef;
    T k1V, k1m, k1h, k1n, k2V, k2m, k2h, k2n, k3V, k3m, k3h, k3n, k4V, k4m, k4h, k4n;
    hh_deriv(p, V, m, h, n, s0, iext, k1V, k1m, k1h, k1n);
    hh_deriv(p, std::fma(hdt, k1V, V), std::fma(hdt, k1m, m), std::fma(hdt, k1h, h), std::fma(hdt, k1n, n), sh,
             iext, k2V, k2m, k2h, k2n);
    hh_deriv(p, std::fma(hdt, k2V, V), std::fma(hdt, k2m, m), std::fma(hdt, k2h, h), std::fma(hdt, k2n, n), sh,
             iext, k3V, k3m, k3h, k3n);
    hh_deriv(p, std::fma(dt, k3V, V), std::fma(dt, k3m, m), std::fma(dt, k3h, h), std::fma(dt, k3n, n), sf, iext,
             k4V, k4m, k4h, k4n);
    V = rk4_comb(V, dt6, k1V, k2V, k3V, k4V);
    m = rk4_comb(m, dt6, k1m, k2m, k3m, k4m);
    h = rk4_comb(h, dt6, k1h, k2h, k3h, k4h);
    n = rk4_comb(n, dt6, k1n, k2n, k3n, k4n);
    s = sf;
}

// Reference semi-implicit ("leapfrog") update, src/ephys.cu:57-94: gates by
// the trapezoid rule with rates frozen at V_n, three kinetic synapses with
// T = IN/COUNT, then V linear in the new G and E, then the sigmoid output.
// powf(N, 4) is written (N*N)*(N*N).  All arithmetic is in T.
template <class T> struct LeapfrogIO {
    T V, M, N, H, Yl, Yi, Ye;
    T in_l, cnt_l, in_i, cnt_i, in_e, cnt_e, appcur;
    T out;
};
template <class T> inline T kinetic(const P<T>& p, T dt, T hdt, T in, T cnt, T Y) {
    T tavg = cnt > T(0) ? in / cnt : T(0);
    return (p.ar * tavg * dt + (T(1) - hdt * (p.ar * tavg + p.ad)) * Y) /
           (hdt * (p.ar * tavg + p.ad) + T(1));
}
template <class T> inline void leapfrog_step(const P<T>& p, T dt, T hdt, T noise, LeapfrogIO<T>& s) {
    Rates<T> R = rates(s.V + p.vshift);
    T nM = (R.am * dt + (T(1) - hdt * (R.am + R.bm)) * s.M) / (hdt * (R.am + R.bm) + T(1));
    T nN = (R.an * dt + (T(1) - hdt * (R.an + R.bn)) * s.N) / (hdt * (R.an + R.bn) + T(1));
    T nH = (R.ah * dt + (T(1) - hdt * (R.ah + R.bh)) * s.H) / (hdt * (R.ah + R.bh) + T(1));
    T nYl = kinetic(p, dt, hdt, s.in_l, s.cnt_l, s.Yl);
    T nYi = kinetic(p, dt, hdt, s.in_i, s.cnt_i, s.Yi);
    T nYe = kinetic(p, dt, hdt, s.in_e, s.cnt_e, s.Ye);
    T m3h = p.gna * nM * nM * nM * nH;
    T n2 = nN * nN;
    T n4 = p.gk * (n2 * n2);
    T G = m3h + n4 + p.gl + p.g_local * nYl + p.g_i * nYi + p.g_e * nYe;
    T E = m3h * p.ena + n4 * p.ek + p.gl * p.el + p.g_local * nYl * p.e_local +
          p.g_i * nYi * p.e_i + p.g_e * nYe * p.e_e;
    T Vn = ((noise + s.appcur) * dt + E * dt + (T(1) - hdt * G) * s.V) / (T(1) + hdt * G);
    s.V = Vn;
    s.M = nM;
    s.N = nN;
    s.H = nH;
    s.Yl = nYl;
    s.Yi = nYi;
    s.Ye = nYe;
    s.out = p.tmax / (T(1) + oexp(-(Vn - p.vt) / p.kp));
}

// --------------------------------------------------------------------------
// coupling (src/coupling.cu:34-48,111-141; src_wholebrain/coupling.cu:12-69)
// --------------------------------------------------------------------------
inline void compact(int64_t n, const uint8_t* fire, std::vector<int32_t>& list) {
    list.clear();
    for (int64_t i = 0; i < n; ++i)
        if (fire[i]) list.push_back((int32_t)i);
}

// mode 0/1 weighted (w), 2 count (+1), 3 signal (+out[src]); the sum for each
// target runs over its spiking sources in ascending order (the CSR is walked
// source by source in ascending order).  g is overwritten.
template <class T>
inline void couple_push(int64_t n, const int32_t* off, const int32_t* tgt, const float* w,
                        const std::vector<int32_t>& list, int mode, const T* out, T* g) {
    for (int64_t i = 0; i < n; ++i) g[i] = T(0);
    for (int32_t j : list)
        for (int32_t e = off[j]; e < off[j + 1]; ++e) {
            T add = mode == 2 ? T(1) : mode == 3 ? out[j] : T(w[e]);
            g[tgt[e]] = g[tgt[e]] + add;
        }
}

// --------------------------------------------------------------------------
// grid operators (x fastest: i + nx*(j + ny*k), Dirichlet 0 outside)
// --------------------------------------------------------------------------
struct Grid {
    int nx, ny, nz, stencil;
    double c, decay_dt;
    const float *kx, *ky, *kz;  // face coefficients (nx+1)*ny*nz, nx*(ny+1)*nz, nx*ny*(nz+1)
};

template <class T> inline T at(const T* x, const Grid& g, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return T(0);
    return x[i + (int64_t)g.nx * (j + (int64_t)g.ny * k)];
}

// Reference 27-point stencil sum (weights corner 1, edge 3, face 14, centre
// -128; src/paracrine.cu:40): ((corners + 3*edges) + 14*faces) - 128*x, each
// group accumulated in (dk, dj, di) lexicographic order from -1 to 1.
template <class T> inline T s27(const T* x, const Grid& g, int i, int j, int k) {
    T corners = T(0), edges = T(0), faces = T(0);
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                int nzc = (di != 0) + (dj != 0) + (dk != 0);
                if (nzc == 0) continue;
                T v = at(x, g, i + di, j + dj, k + dk);
                if (nzc == 3) corners = corners + v;
                else if (nzc == 2) edges = edges + v;
                else faces = faces + v;
            }
    return ((corners + T(3) * edges) + T(14) * faces) - T(128) * at(x, g, i, j, k);
}

// y = A x.  STENCIL_7 constant: y = diag*x - c*(((((xm + xp) + ym) + yp) + zm) + zp).
// STENCIL_7 variable: y = (1 + sum faces)*x - (((((kxm xm + kxp xp) + kym ym) + kyp yp) + kzm zm) + kzp zp).
// STENCIL_27 (reference CN): y = a0*x - w*(((corners + 3*edges) + 14*faces) - 128*x),
// a0 = 1 + decay_dt/2, w = c/60 (= (dt D / 2) / (30 h^2), src/paracrine.cu:40-43,80-81).
template <class T> inline void apply(const Grid& g, const T* x, T* y) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int64_t id = i + (int64_t)nx * (j + (int64_t)ny * k);
                T xc = x[id];
                if (g.stencil == 7 && !g.kx) {
                    T c = T(g.c), diag = T(1.0 + 6.0 * g.c);
                    T s = at(x, g, i - 1, j, k) + at(x, g, i + 1, j, k);
                    s = s + at(x, g, i, j - 1, k);
                    s = s + at(x, g, i, j + 1, k);
                    s = s + at(x, g, i, j, k - 1);
                    s = s + at(x, g, i, j, k + 1);
                    y[id] = diag * xc - c * s;
                } else if (g.stencil == 7) {
                    int64_t fx = i + (int64_t)(nx + 1) * (j + (int64_t)ny * k);
                    int64_t fy = i + (int64_t)nx * (j + (int64_t)(ny + 1) * k);
                    int64_t fz = id;
                    int64_t sz = (int64_t)nx * ny;
                    T kxm = T(g.kx[fx]), kxp = T(g.kx[fx + 1]);
                    T kym = T(g.ky[fy]), kyp = T(g.ky[fy + nx]);
                    T kzm = T(g.kz[fz]), kzp = T(g.kz[fz + sz]);
                    T d = T(1) + kxm + kxp + kym + kyp + kzm + kzp;
                    T s = kxm * at(x, g, i - 1, j, k) + kxp * at(x, g, i + 1, j, k);
                    s = s + kym * at(x, g, i, j - 1, k);
                    s = s + kyp * at(x, g, i, j + 1, k);
                    s = s + kzm * at(x, g, i, j, k - 1);
                    s = s + kzp * at(x, g, i, j, k + 1);
                    y[id] = d * xc - s;
                } else {
                    T a0 = T(1.0 + 0.5 * g.decay_dt), w = T(g.c / 60.0);
                    y[id] = a0 * xc - w * s27(x, g, i, j, k);
                }
            }
}

template <class T> inline double diag_of(const Grid& g, int64_t id) {
    if (g.stencil == 27) return 1.0 + 0.5 * g.decay_dt + 128.0 * g.c / 60.0;
    if (!g.kx) return 1.0 + 6.0 * g.c;
    int64_t nx = g.nx, ny = g.ny;
    int64_t i = id % nx, j = (id / nx) % ny, k = id / (nx * ny);
    int64_t fx = i + (nx + 1) * (j + ny * k), fy = i + nx * (j + (ny + 1) * k);
    return 1.0 + (double)g.kx[fx] + g.kx[fx + 1] + g.ky[fy] + g.ky[fy + nx] + g.kz[id] +
           g.kz[id + nx * ny];
}

// FP64 dot in fixed blocks of 4096 summed in order: independent of threads.
template <class A, class B> inline double dot(const A* a, const B* b, int64_t n) {
    const int64_t bs = 4096, nb = (n + bs - 1) / bs;
    std::vector<double> part(nb);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nb; ++q) {
        double s = 0;
        for (int64_t i = q * bs; i < std::min(n, (q + 1) * bs); ++i) s += (double)a[i] * (double)b[i];
        part[q] = s;
    }
    double s = 0;
    for (double v : part) s += v;
    return s;
}

// Jacobi-preconditioned CG, vectors in T, dots and scalars in FP64.  Stops on
// the recursive residual ||r|| <= rtol ||b||.  Returns iterations.
template <class T>
inline int cg(const Grid& g, const T* b, T* x, double rtol, int maxit, bool jacobi, bool use_x0,
              double* relres) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    std::vector<T> r(n), z(n), p(n), q(n), dinv(n);
    for (int64_t i = 0; i < n; ++i) dinv[i] = jacobi ? T(1.0 / diag_of<T>(g, i)) : T(1);
    if (!use_x0)
        for (int64_t i = 0; i < n; ++i) x[i] = T(0);
    apply(g, x, q.data());
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
    for (int64_t i = 0; i < n; ++i) z[i] = r[i] * dinv[i];
    p = z;
    double bnorm = std::sqrt(dot(b, b, n));
    if (bnorm == 0) bnorm = 1;
    double rz = dot(r.data(), z.data(), n);
    double rr = dot(r.data(), r.data(), n);
    int it = 0;
    while (std::sqrt(rr) > rtol * bnorm && it < maxit) {
        apply(g, p.data(), q.data());
        double alpha = rz / dot(p.data(), q.data(), n);
        T a = T(alpha);
        for (int64_t i = 0; i < n; ++i) {
            x[i] = x[i] + a * p[i];
            r[i] = r[i] - a * q[i];
            z[i] = r[i] * dinv[i];
        }
        double rz2 = dot(r.data(), z.data(), n);
        rr = dot(r.data(), r.data(), n);
        T beta = T(rz2 / rz);
        rz = rz2;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    if (relres) *relres = std::sqrt(rr) / bnorm;
    return it;
}

// FP64 iterative refinement around the float CG: r = b - A x in double,
// inner float CG on A d = r to max(inner_rtol, 0.5 rtol / relres), x += d.
inline int refine(const Grid& g, const double* b, double* x, double rtol, double inner_rtol,
                  int maxit, int max_outer, double* relres, int* inner_total) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    std::vector<double> r(n);
    std::vector<float> rf(n), d(n);
    double bnorm = std::sqrt(dot(b, b, n));
    if (bnorm == 0) bnorm = 1;
    int outer = 0, tot = 0;
    double rel = 0;
    for (;;) {
        apply(g, x, r.data());
        for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
        rel = std::sqrt(dot(r.data(), r.data(), n)) / bnorm;
        if (rel <= rtol || outer >= max_outer) break;
        for (int64_t i = 0; i < n; ++i) rf[i] = (float)r[i];
        double ir;
        const double tol = std::max(inner_rtol, 0.5 * rtol / rel);
        tot += cg<float>(g, rf.data(), d.data(), tol, maxit, true, false, &ir);
        for (int64_t i = 0; i < n; ++i) x[i] += (double)d[i];
        ++outer;
    }
    if (relres) *relres = rel;
    if (inner_total) *inner_total = tot;
    return outer;
}

// Reference update_density (src/paracrine.cu:63-134), restated on the
// 27-point operator: explicit-Euler guess, b with h where dt is meant
// (src/paracrine.cu:105), plain CG with the absolute squared-residual stop and
// at most max_iter iterations, early return without writing the guess
// (src/paracrine.cu:114-117).  Vectors float, dots double.
inline int update_density_reference(int nx, int ny, int nz, double h, double D, double k,
                                    double dt, double eb, int max_iter, float* data,
                                    double* rr_out) {
    const int64_t n = (int64_t)nx * ny * nz;
    Grid L{nx, ny, nz, 27, 0, 0, nullptr, nullptr, nullptr};
    // L27 u / h^2 via the operator with a0 = 1, w = c/60: y = x - (c/60) S x.
    // Pick c = -60 / h^2 / 30 ... instead evaluate S x directly:
    std::vector<float> lap(n), xg(n), bb(n), r(n), p(n), q(n);
    {
        const float inv = (float)(1.0 / (h * h * 30.0));
#pragma omp parallel for schedule(static)
        for (int kk = 0; kk < nz; ++kk)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    lap[i + (int64_t)nx * (j + (int64_t)ny * kk)] = s27(data, L, i, j, kk) * inv;
    }
    for (int64_t i = 0; i < n; ++i)
        xg[i] = (float)(1.0 - dt * k) * data[i] + (float)(dt * D) * lap[i];
    for (int64_t i = 0; i < n; ++i)
        bb[i] = (float)(1.0 - h * k / 2.0) * data[i] + (float)(h * D / 2.0) * lap[i];
    Grid A{nx, ny, nz, 27, dt * D / (h * h), dt * k, nullptr, nullptr, nullptr};
    apply(A, xg.data(), q.data());
    for (int64_t i = 0; i < n; ++i) r[i] = bb[i] - q[i];
    p = r;
    double rr = dot(r.data(), r.data(), n);
    if (rr_out) *rr_out = rr;
    if (rr < eb) return 0;
    int it = 0;
    for (; it < max_iter;) {
        apply(A, p.data(), q.data());
        double pq = dot(p.data(), q.data(), n);
        float alpha = (float)(rr / pq);
        for (int64_t i = 0; i < n; ++i) {
            xg[i] = xg[i] + alpha * p[i];
            r[i] = r[i] - alpha * q[i];
        }
        double rr2 = dot(r.data(), r.data(), n);
        ++it;
        if (rr2 < eb) {
            rr = rr2;
            break;
        }
        float beta = (float)(rr2 / rr);
        for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        rr = rr2;
    }
    if (rr_out) *rr_out = rr;
    std::memcpy(data, xg.data(), n * sizeof(float));
    (void)L;
    return it;
}

// --------------------------------------------------------------------------
// Electrode readout (src_wholebrain/electrode_net.cu:3-30, which adds the
// firing count to src/electrode_net.cu:3-21): one region at a time, its
// voltages added in index order in T, divided by the region's size (NaN for
// an empty region, 0/0 as there).  Outside the reference's domain, and as the
// device defines it: an index outside [0, n) contributes NaN and does not
// count as firing; offsets are clamped to [0, nnz].  fire: 0/1 per neuron
// (the reference's is_firing), NULL = no counts.
// --------------------------------------------------------------------------
template <class T>
void electrode_measure(int64_t nreg, int64_t nnz, const int32_t* off, const int32_t* idx, int64_t n,
                       const T* V, const int32_t* fire, T* avg, int32_t* cnt) {
    for (int64_t r = 0; r < nreg; ++r) {
        int64_t b = std::min<int64_t>(std::max<int64_t>(off[r], 0), nnz);
        int64_t e = std::min<int64_t>(std::max<int64_t>(off[r + 1], b), nnz);
        T s = 0;
        int32_t c = 0;
        for (int64_t j = b; j < e; ++j) {
            const int32_t i = idx[j];
            if (i >= 0 && i < n) {
                s += V[i];
                if (fire) c += fire[i];
            } else {
                s += std::numeric_limits<T>::quiet_NaN();
            }
        }
        avg[r] = s / T(e - b);
        if (fire) cnt[r] = c;
    }
}

// Mean of V (ElectrodeNet::MeasureAverageEntireNetwork,
// src/electrode_net.cu:44-47).  The reference's fp32 thrust::reduce has no
// specified order; this is the order the device uses: 2^18 lanes (1024
// blocks of 256), lane L adding elements L, L + 2^18, ... in FP64; a block
// total is the butterfly of each 64-lane wave (pairs (l, l+o) for o = 32 ..
// 1) followed by its 4 waves in order; the 1024 block totals are summed the
// same way by 256 lanes taking t, t+256, t+512, t+768; mean = total / n,
// rounded to T.
inline double wave_tree64(double* v) {
    for (int o = 32; o > 0; o >>= 1)
        for (int l = 0; l < o; ++l) v[l] = v[l] + v[l + o];
    return v[0];
}
inline double block_total(const double* lanes, int nt) {
    std::vector<double> w(64);
    double t = 0;
    for (int wv = 0; wv < nt / 64; ++wv) {
        std::copy(lanes + 64 * wv, lanes + 64 * wv + 64, w.begin());
        const double x = wave_tree64(w.data());
        t = wv == 0 ? x : t + x;
    }
    return t;
}
template <class T> T network_mean(int64_t n, const T* V) {
    const int NB = 1024, NT = 256;
    const int64_t stride = (int64_t)NB * NT;
    std::vector<double> lanes(stride, 0.0), part(NB), l2(NT, 0.0);
    for (int64_t i = 0; i < n; ++i) lanes[i % stride] += (double)V[i];
    for (int b = 0; b < NB; ++b) part[b] = block_total(&lanes[(size_t)b * NT], NT);
    for (int t = 0; t < NT; ++t)
        for (int k = t; k < NB; k += NT) l2[t] += part[k];
    return (T)(block_total(l2.data(), NT) / (double)n);
}

// Synchrony accumulation of one timestep (SynchronyNet::MeasureVariances,
// src_wholebrain/electrode_net.cu:54-97): per-neuron sums with the square
// formed in T; each sub-population's Vsum is an FP64 sum in the device's
// order (1024 lanes, lane L adding entries L, L+1024, ...; block total as
// in network_mean with 16 waves) rounded to T (the reference: an fp32
// thrust::reduce of unspecified order), Vavg = Vsum / count in T.
template <class T>
void synchrony_accumulate(int64_t n, const T* V, int64_t n_idx, const int32_t* idx, int n_sub,
                          const int32_t* counts, double* vi, double* vi2, double* vavg, double* vavg2) {
    auto get = [&](int32_t i) { return (i >= 0 && i < n) ? V[i] : std::numeric_limits<T>::quiet_NaN(); };
    for (int64_t j = 0; j < n_idx; ++j) {
        const T x = get(idx[j]);
        vi2[j] += (double)(x * x);
        vi[j] += (double)x;
    }
    for (int s = 0; s < n_sub; ++s) {
        const int32_t cnt = counts[s];
        double d = std::numeric_limits<double>::quiet_NaN();
        if (cnt >= 1 && cnt <= n_idx) {
            std::vector<double> lanes(1024, 0.0);
            for (int32_t j = 0; j < cnt; ++j) lanes[j % 1024] += (double)get(idx[j]);
            const T vs = (T)block_total(lanes.data(), 1024);
            d = (double)(vs / (T)cnt);
        }
        vavg[s] += d;
        vavg2[s] += d * d;
    }
}

// chi^2 per sub-population (SynchronyNet::GetSynchronyMeasures,
// src_wholebrain/electrode_net.cu:99-126), the same operations in order.
inline void synchrony_measures(int64_t n_idx, int n_sub, const int32_t* counts, const double* vi,
                               const double* vi2, const double* vavg, const double* vavg2, int64_t steps,
                               double* chi) {
    const double T_ = (double)steps;
    for (int s = 0; s < n_sub; ++s) {
        const int32_t cnt = counts[s];
        if (cnt < 1 || cnt > n_idx) {
            chi[s] = std::numeric_limits<double>::quiet_NaN();
            continue;
        }
        double denom = 0;
        for (int32_t j = 0; j < cnt; ++j) {
            double a2 = vi2[j], a1 = vi[j];
            if (steps > 0) {
                a2 /= T_;
                a1 /= T_;
            }
            denom += a2 - a1 * a1;
        }
        const double sig = (vavg2[s] - vavg[s] * vavg[s] / T_) / T_;
        chi[s] = (double)cnt * sig / denom;
    }
}

// --------------------------------------------------------------------------
// Hebbian::UpdateWeights (src_wholebrain/hebbian.cu:3-69), as written: the
// countdown of every out-synapse of a firing source restarts at `window`
// (set_timers), then every synapse with a running countdown decrements it;
// at zero its weight goes down by heb_inc * exp(-|w - 1|), otherwise, if its
// target fires, it goes up by as much and the countdown stops.  The exp is
// this file's oexp (the reference calls CUDA's expf).  fire: 0/1 per neuron.
// --------------------------------------------------------------------------
inline void hebbian_update(int64_t n, const int32_t* off, const int32_t* tgt, const uint8_t* fire, int32_t window,
                           float heb_inc, int32_t* timers, float* w) {
    for (int64_t p = 0; p < n; ++p)
        if (fire[p])
            for (int32_t e = off[p]; e < off[p + 1]; ++e) timers[e] = window;
    const int64_t nnz = off[n];
    for (int64_t e = 0; e < nnz; ++e) {
        if (timers[e] == 0) continue;
        timers[e] -= 1;
        const float alpha = oexp(-std::fabs(w[e] - 1.0f));
        if (timers[e] == 0)
            w[e] -= heb_inc * alpha;
        else if (fire[tgt[e]]) {
            w[e] += heb_inc * alpha;
            timers[e] = 0;
        }
    }
}

// --------------------------------------------------------------------------
// Diekman SCN neuron, one step (Leapfrog_diekman, src/ephys.cu:250-347).  In
// fp32 the reference's operations keep their C++ types: its double literals
// widen parts of the arithmetic to double (written here as W), expf takes a
// float argument, exp a double one; powf(x, 0.25), powf(x, 1.5) and
// powf(x, 4) are sqrt(sqrt(x)), x sqrt(x) and (x^2)^2.  fp64 is double
// throughout.  c: the constants (fp32: rounded to float as diekman_params
// stores them).  x: V, M, H, N, Rl, Rnl, Fnl, S, P, Y, Cas, Cac (updated);
// in: Tin, num_in, egaba, gkca, gkleak; returns IOUTPUT.
// --------------------------------------------------------------------------
template <class T> struct DkC {
    T gna, gnaleak, gk, gcal, gcanl, ena, ek, eca, k1, k2, kcas, kcac, taucas, taucac, bcas, bcac;
    T ar, ad, tmax, vt, kp, g_gaba, appcur, gnap, tau_rl, tau_rnl, tau_p;
};
inline float dk_exf(float, double a) { return oexp((float)a); }
inline double dk_exf(double, double a) { return oexp(a); }

template <class T> T diekman_step(const DkC<T>& c, T dt, T* x, const T* in) {
    using W = double;
    const T V = x[0], M = x[1], H = x[2], N = x[3], Rl = x[4], Rnl = x[5], Fnl = x[6], S = x[7], P = x[8], Y = x[9];
    const T Cas = x[10], Cac = x[11];
    const T Tin = in[0], num_in = in[1], egaba = in[2], gkca = in[3], gkleak = in[4];
    const T z = 0;
    const T M_inf = (T)(1.0 / (1.0 + (W)dk_exf(z, -((W)V + 35.2) / 8.1)));
    const T H_inf = (T)(1.0 / (1.0 + (W)dk_exf(z, ((W)V + 62.0) / 2.0)));
    const T nb = (T)(1.0 + (W)dk_exf(z, ((W)V - 14.0) / (-17.0)));
    const T N_inf = (T)(1.0 / (W)std::sqrt(std::sqrt(nb)));
    const T Rl_inf = (T)(1.0 / (1.0 + (W)dk_exf(z, -((W)V + 36.0) / 5.1)));
    const T Rnl_inf = (T)(1.0 / (1.0 + (W)dk_exf(z, -((W)V + 21.6) / 6.7)));
    const T Fnl_inf = (T)(1.0 / (1.0 + oexp(((W)V + 260.0) / 65.0)));
    const W cs = 1e7 * (W)Cas * (W)Cas;
    const T S_inf = (T)(cs / (cs + 5.6));
    const T pb = (T)(1.0 + (W)dk_exf(z, -((W)V + 25.0) / 7.4));
    const T P_inf = (T)(1.0 / (W)(pb * std::sqrt(pb)));
    const T tM = dk_exf(z, -((W)V + 286.0) / 160.0);
    const T tH = (T)(0.51 + (W)dk_exf(z, -((W)V + 26.6) / 7.1));
    const T tN = dk_exf(z, -((W)V - 67.0) / 68.0);
    const T tF = dk_exf(z, -((W)V - 444.0) / 220.0);
    const T tS = (T)(500.0 / (cs + 5.6));
    auto gate = [&](T tau, T xi, T xo) -> T {
        return (T)(2.0 * (W)dt / (2.0 * (W)tau + (W)dt) * (W)xi + (2.0 * (W)tau - (W)dt) / (2.0 * (W)tau + (W)dt) * (W)xo);
    };
    const T nM = gate(tM, M_inf, M), nH = gate(tH, H_inf, H), nN = gate(tN, N_inf, N);
    const T nRl = gate(c.tau_rl, Rl_inf, Rl), nRnl = gate(c.tau_rnl, Rnl_inf, Rnl), nFnl = gate(tF, Fnl_inf, Fnl);
    const T nP = gate(c.tau_p, P_inf, P);
    const T dv = V - c.eca, h = dt / T(2);
    const T ex = -c.kcas * (c.gcal * Rl * c.k1 / (c.k2 + Cas) + c.gcanl * Rnl * Fnl) * dv - Cas / c.taucas + c.bcas;
    const T dn = T(1) + h / c.taucas;
    const T B = (c.k2 - Cas - h * ex + h * c.kcas * c.gcanl * nRnl * nFnl * dv + h / c.taucas * c.k2 - c.bcas * dt / T(2)) / dn;
    const T C = (-c.k2 * Cas - c.k2 * dt / T(2) * ex + h * c.kcas * c.gcal * nRl * c.k1 * dv +
                 h * c.kcas * c.gcanl * nRnl * nFnl * dv * c.k2 - c.bcas * dt / T(2) * c.k2) / dn;
    const T nCas = (std::sqrt(B * B - T(4) * C) - B) / T(2);
    const W ncs = 1e7 * (W)nCas * (W)nCas;
    const T S_inf2 = (T)(ncs / (ncs + 5.6));
    const T tS2 = (T)(500.0 / (ncs + 5.6));
    const T nS = (T)(1.0 / (1.0 + (W)dt / (2.0 * (W)tS2)) *
                     ((W)S * (1.0 - (W)dt / (2.0 * (W)tS)) + (W)dt / 2.0 * (W)(S_inf / tS + S_inf2 / tS2)));
    const T sum = c.gcal * nRl * (c.k1 / (c.k2 + Cac)) + c.gcanl * nRl * nFnl + c.gcal * Rl * (c.k1 / (c.k2 + Cas)) +
                  c.gcanl * Rnl * Fnl;
    const T nCac = (T)(1.0 / (1.0 + (W)dt / (2.0 * (W)c.taucac)) *
                       ((W)Cac * (1.0 - (W)dt / (2.0 * (W)c.taucac)) + (W)(c.bcac * dt) -
                        (W)(dt * c.kcac) / 2.0 * (W)sum * (W)dv));
    const T Ta = num_in == T(0) ? T(0) : Tin / num_in;
    const T ya = c.ar * Ta + c.ad;
    const T nY = (c.ar * Ta * dt + (T(1) - h * ya) * Y) / (h * ya + T(1));
    const T n2 = nN * nN, n4 = n2 * n2, fL = c.k1 / (c.k2 + nCas), gn = c.gna * nM * nM * nM * nH;
    const T G = gn + c.gk * n4 + c.gcal * nRl * fL + c.gcanl * nRnl * nFnl + gkca * nS * nS + gkleak + c.gnaleak +
                c.gnap * nP + nY;
    const T E = c.ena * (gn + c.gnaleak + c.gnap * nP) + c.ek * (c.gk * n4 + gkca * nS * nS + gkleak) +
                c.eca * (c.gcal * nRl * fL + c.gcanl * nRnl * nFnl) + c.g_gaba * nY * egaba;
    const T nV = (c.appcur * dt + E * dt + (T(1) - h * G) * V) / (T(1) + h * G);
    const T nx[12] = {nV, nM, nH, nN, nRl, nRnl, nFnl, nS, nP, nY, nCas, nCac};
    for (int k = 0; k < 12; ++k) x[k] = nx[k];
    return (T)((W)c.tmax / (1.0 + (W)oexp(-(nV - c.vt) / c.kp)));
}

// --------------------------------------------------------------------------
// SAN ("AN" slow-wave) neuron, one forward-Euler step (Euler_an_ie_model
// with dxdt_an_ie_model, src/ephys.cu:105-141,203-245); the white noise is
// its mean.  fp32 keeps the reference's types: expf of float arguments, exp
// of double ones, double literals (W); powf(x, 3.5) = x^3 sqrt(x),
// powf(x, 4) = (x^2)^2.  c: 24 constants in nnx_san_params order.  x: V, nK,
// Ca, Ye, Yi (updated); in: Ein, Ecount, Iin, Icount; returns the output.
// --------------------------------------------------------------------------
template <class T> T san_step(const T* c, T dt, T* x, const T* in) {
    using W = double;
    const T C = c[0], A = c[1], vL = c[2], vNa = c[3], vK = c[4], vCa = c[5], kD = c[6], vAMPA = c[7], vGABA = c[8];
    const T alphaCa = c[9], tauAMPA = c[10], tauGABA = c[11], gL = c[12], gK = c[13], gNaP = c[14], gCa = c[15];
    const T gKCa = c[16], gAMPA = c[17], gGABA = c[18], tauCa = c[19], noise = c[20], Tmax = c[21], Vt = c[22];
    const T Kp = c[23];
    const T V = x[0], nK = x[1], Ca = x[2], Ye = x[3], Yi = x[4];
    const T eT = in[1] == T(0) ? T(0) : in[0] / in[1];
    const T iT = in[3] == T(0) ? T(0) : in[2] / in[3];
    const T mCa = T(1) / (T(1) + oexp(-(V + T(20)) / T(9)));
    const T r = kD / Ca;
    const T mKCa = T(1) / (T(1) + r * r * r * std::sqrt(r));
    const T mNaP = (T)(1.0 / (1.0 + (W)oexp(-((W)V + 55.7) / 7.7)));
    const T an = (T)(V == T(-34) ? 0.1 : 0.01 * (W)(V + T(34)) / (W)(T(1) - oexp(-(V + T(34)) / T(10))));
    const T bn = (T)(0.125 * (W)oexp(-(V + T(44)) / T(25)));
    const T dn = T(4) * (an * (T(1) - nK) - bn * nK);
    const T dye = (T)(3.48 * (W)eT - (W)(Ye / tauAMPA));
    const T dyi = iT - Yi / tauGABA;
    const T ICa = gCa * mCa * mCa * (V - vCa);
    const T dca = -alphaCa * (T(10) * A * ICa) - Ca / tauCa;
    const T n2 = nK * nK;
    T dv = noise - gK * (n2 * n2) * (V - vK) - ICa - gKCa * mKCa * (V - vK) - gNaP * mNaP * mNaP * mNaP * (V - vNa) -
           gL * (V - vL) - gAMPA * Ye * (V - vAMPA) - gGABA * Yi * (V - vGABA);
    dv = dv / C;
    x[0] = V + dt * dv;
    x[1] = nK + dt * dn;
    x[2] = Ca + dt * dca;
    x[3] = Ye + dt * dye;
    x[4] = Yi + dt * dyi;
    return (T)((W)Tmax / (1.0 + (W)oexp(-(x[0] - Vt) / Kp)));
}

}  // namespace oracle

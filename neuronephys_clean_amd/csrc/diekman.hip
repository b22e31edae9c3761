// diekman.hip - the Diekman suprachiasmatic-nucleus neuron step.
//
// Replaces DiekmanEphys::SimulateEphys and its kernel Leapfrog_diekman
// (src/ephys.cu:250-353).  One thread per neuron over the SoA state, the
// whole update in registers, the level spike rule on the new output folded
// into the same pass (one ballot word per wavefront).  In fp32 the reference
// mixes float and double arithmetic through its double literals; the same
// operations are performed here in the same types and order (so that the
// step is reproducible bit for bit by a CPU restatement), with this
// library's exponentials; fp64 states are computed in double throughout.
#include "internal.h"

namespace nnx {

namespace {

template <class T> struct DkConst {
    T gna, gnaleak, gk, gcal, gcanl, ena, ek, eca, k1, k2, kcas, kcac, taucas, taucac, bcas, bcac;
    T ar, ad, tmax, vt, kp, g_gaba, appcur, gnap, tau_rl, tau_rnl, tau_p;
    double out_thresh;  // compared with the widened output, as CheckFiringOperator does
};

// expf of an argument computed in double (rounded to float first), as the
// reference's expf(double expression); plain exp in fp64
__device__ __forceinline__ float ex(float, double a) { return dexp((float)a); }
__device__ __forceinline__ double ex(double, double a) { return dexp(a); }
__device__ __forceinline__ float rt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double rt(double x) { return __builtin_sqrt(x); }

template <class T> struct DkPtr {
    T *v, *m, *h, *ng, *rl, *rnl, *fnl, *s, *p, *y, *cas, *cac, *out;
    const T *iin, *icount, *egaba, *gkca, *gkleak;
};

template <class T>
__global__ __launch_bounds__(256) void k_diekman(int64_t n, DkConst<T> c, T dt, DkPtr<T> a, uint64_t* __restrict__ bits,
                                                 int64_t nwords) {
    using W = double;
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    bool spike = false;
    if (i < n) {
        const T V = a.v[i], M = a.m[i], H = a.h[i], N = a.ng[i], Rl = a.rl[i], Rnl = a.rnl[i], Fnl = a.fnl[i];
        const T S = a.s[i], P = a.p[i], Y = a.y[i], Cas = a.cas[i], Cac = a.cac[i];
        const T Tin = a.iin[i], num_in = a.icount[i], egaba = a.egaba[i], gkca = a.gkca[i], gkleak = a.gkleak[i];
        const T z = T(0);
        // steady states and time constants (src/ephys.cu:286-302)
        const T M_inf = (T)(1.0 / (1.0 + (W)ex(z, -((W)V + 35.2) / 8.1)));
        const T H_inf = (T)(1.0 / (1.0 + (W)ex(z, ((W)V + 62.0) / 2.0)));
        const T n_b = (T)(1.0 + (W)ex(z, ((W)V - 14.0) / (-17.0)));
        const T N_inf = (T)(1.0 / (W)rt(rt(n_b)));  // powf(x, 0.25)
        const T Rl_inf = (T)(1.0 / (1.0 + (W)ex(z, -((W)V + 36.0) / 5.1)));
        const T Rnl_inf = (T)(1.0 / (1.0 + (W)ex(z, -((W)V + 21.6) / 6.7)));
        const T Fnl_inf = (T)(1.0 / (1.0 + dexp(((W)V + 260.0) / 65.0)));  // exp in double
        const W cas2 = 1e7 * (W)Cas * (W)Cas;
        const T S_inf = (T)(cas2 / (cas2 + 5.6));
        const T p_b = (T)(1.0 + (W)ex(z, -((W)V + 25.0) / 7.4));
        const T P_inf = (T)(1.0 / (W)(p_b * rt(p_b)));  // powf(x, 1.5)
        const T tau_M = ex(z, -((W)V + 286.0) / 160.0);
        const T tau_H = (T)(0.51 + (W)ex(z, -((W)V + 26.6) / 7.1));
        const T tau_N = ex(z, -((W)V - 67.0) / 68.0);
        const T tau_Fnl = ex(z, -((W)V - 444.0) / 220.0);
        const T tau_S = (T)(500.0 / (cas2 + 5.6));
        // trapezoid gate updates (src/ephys.cu:305-311)
        auto gate = [&](T tau, T xinf, T x) -> T {
            return (T)(2.0 * (W)dt / (2.0 * (W)tau + (W)dt) * (W)xinf +
                       (2.0 * (W)tau - (W)dt) / (2.0 * (W)tau + (W)dt) * (W)x);
        };
        const T nM = gate(tau_M, M_inf, M), nH = gate(tau_H, H_inf, H), nN = gate(tau_N, N_inf, N);
        const T nRl = gate(c.tau_rl, Rl_inf, Rl), nRnl = gate(c.tau_rnl, Rnl_inf, Rnl);
        const T nFnl = gate(tau_Fnl, Fnl_inf, Fnl), nP = gate(c.tau_p, P_inf, P);
        // submembrane Ca: the positive root of the trapezoid step's quadratic (src/ephys.cu:312-316)
        const T dV = V - c.eca;
        const T hdt = dt / T(2);
        const T expr = -c.kcas * (c.gcal * Rl * c.k1 / (c.k2 + Cas) + c.gcanl * Rnl * Fnl) * dV - Cas / c.taucas + c.bcas;
        const T den = T(1) + hdt / c.taucas;
        const T B = (c.k2 - Cas - hdt * expr + hdt * c.kcas * c.gcanl * nRnl * nFnl * dV + hdt / c.taucas * c.k2 -
                     c.bcas * dt / T(2)) / den;
        const T C = (-c.k2 * Cas - c.k2 * dt / T(2) * expr + hdt * c.kcas * c.gcal * nRl * c.k1 * dV +
                     hdt * c.kcas * c.gcanl * nRnl * nFnl * dV * c.k2 - c.bcas * dt / T(2) * c.k2) / den;
        const T nCas = (rt(B * B - T(4) * C) - B) / T(2);
        // Ca-activated K gate (src/ephys.cu:318-321)
        const W ncas2 = 1e7 * (W)nCas * (W)nCas;
        const T S_inf2 = (T)(ncas2 / (ncas2 + 5.6));
        const T tau_S2 = (T)(500.0 / (ncas2 + 5.6));
        const T nS = (T)(1.0 / (1.0 + (W)dt / (2.0 * (W)tau_S2)) *
                         ((W)S * (1.0 - (W)dt / (2.0 * (W)tau_S)) + (W)dt / 2.0 * (W)(S_inf / tau_S + S_inf2 / tau_S2)));
        // cytosolic Ca (src/ephys.cu:323-324; the non-L current uses new_Rl
        // as there)
        const T ca_sum = c.gcal * nRl * (c.k1 / (c.k2 + Cac)) + c.gcanl * nRl * nFnl + c.gcal * Rl * (c.k1 / (c.k2 + Cas)) +
                         c.gcanl * Rnl * Fnl;
        const T nCac = (T)(1.0 / (1.0 + (W)dt / (2.0 * (W)c.taucac)) *
                           ((W)Cac * (1.0 - (W)dt / (2.0 * (W)c.taucac)) + (W)(c.bcac * dt) -
                            (W)(dt * c.kcac) / 2.0 * (W)ca_sum * (W)dV));
        // kinetic GABA synapse (src/ephys.cu:326-328)
        const T Tavg = num_in == T(0) ? T(0) : Tin / num_in;
        const T ya = c.ar * Tavg + c.ad;
        const T nY = (c.ar * Tavg * dt + (T(1) - hdt * ya) * Y) / (hdt * ya + T(1));
        // membrane potential, semi-implicit (src/ephys.cu:330-334)
        const T n2 = nN * nN;
        const T n4 = n2 * n2;  // powf(x, 4)
        const T fL = c.k1 / (c.k2 + nCas);
        const T gna_ = c.gna * nM * nM * nM * nH;
        const T G = gna_ + c.gk * n4 + c.gcal * nRl * fL + c.gcanl * nRnl * nFnl + gkca * nS * nS + gkleak + c.gnaleak +
                    c.gnap * nP + nY;
        const T E = c.ena * (gna_ + c.gnaleak + c.gnap * nP) + c.ek * (c.gk * n4 + gkca * nS * nS + gkleak) +
                    c.eca * (c.gcal * nRl * fL + c.gcanl * nRnl * nFnl) + c.g_gaba * nY * egaba;
        const T nV = (c.appcur * dt + E * dt + (T(1) - hdt * G) * V) / (T(1) + hdt * G);
        const T out = (T)((W)c.tmax / (1.0 + (W)dexp(-(nV - c.vt) / c.kp)));
        a.v[i] = nV;
        a.m[i] = nM;
        a.h[i] = nH;
        a.ng[i] = nN;
        a.rl[i] = nRl;
        a.rnl[i] = nRnl;
        a.fnl[i] = nFnl;
        a.s[i] = nS;
        a.p[i] = nP;
        a.y[i] = nY;
        a.cas[i] = nCas;
        a.cac[i] = nCac;
        a.out[i] = out;
        spike = (double)out > c.out_thresh;
    }
    const uint64_t word = __ballot(spike);
    if (bits && (threadIdx.x & 63) == 0) {
        const int64_t w = i >> 6;
        if (w < nwords) bits[w] = word;
    }
}

template <class T> DkConst<T> dk_const(const nnx_diekman_params* q) {
    DkConst<T> c;
    c.gna = (T)q->gna, c.gnaleak = (T)q->gnaleak, c.gk = (T)q->gk, c.gcal = (T)q->gcal, c.gcanl = (T)q->gcanl;
    c.ena = (T)q->ena, c.ek = (T)q->ek, c.eca = (T)q->eca, c.k1 = (T)q->k1, c.k2 = (T)q->k2, c.kcas = (T)q->kcas;
    c.kcac = (T)q->kcac, c.taucas = (T)q->taucas, c.taucac = (T)q->taucac, c.bcas = (T)q->bcas, c.bcac = (T)q->bcac;
    c.ar = (T)q->ar, c.ad = (T)q->ad, c.tmax = (T)q->tmax, c.vt = (T)q->vt, c.kp = (T)q->kp, c.g_gaba = (T)q->g_gaba;
    c.appcur = (T)q->appcur, c.gnap = (T)q->gnap, c.tau_rl = (T)q->tau_rl, c.tau_rnl = (T)q->tau_rnl;
    c.tau_p = (T)q->tau_p, c.out_thresh = q->out_thresh;
    return c;
}

template <class T>
nnx_status launch(nnx_ctx* c, const nnx_diekman_state* st, const nnx_diekman_params* q, double dt, uint64_t* bits) {
    DkPtr<T> a;
    a.v = (T*)st->v, a.m = (T*)st->m, a.h = (T*)st->h, a.ng = (T*)st->ng, a.rl = (T*)st->rl, a.rnl = (T*)st->rnl;
    a.fnl = (T*)st->fnl, a.s = (T*)st->s, a.p = (T*)st->p, a.y = (T*)st->y, a.cas = (T*)st->cas, a.cac = (T*)st->cac;
    a.out = (T*)st->ioutput;
    a.iin = (const T*)st->iin, a.icount = (const T*)st->icount, a.egaba = (const T*)st->egaba;
    a.gkca = (const T*)st->gkca, a.gkleak = (const T*)st->gkleak;
    const int64_t n = st->n;
    k_diekman<T><<<(unsigned)ceil_div(n, 256), 256, 0, c->stream>>>(n, dk_const<T>(q), (T)dt, a, bits, ceil_div(n, 64));
    NNX_LAUNCHED(c);
    return NNX_OK;
}

}  // namespace

}  // namespace nnx

using namespace nnx;

extern "C" {

void nnx_diekman_params_reference(nnx_diekman_params* p) {
    // diekman_params (src/defines.cuh) and the constants of Leapfrog_diekman
    *p = nnx_diekman_params{};
    p->gna = 229, p->gnaleak = 0.0576, p->gk = 3, p->gcal = 6, p->gcanl = 20;
    p->ena = 45, p->ek = -97, p->eca = 54, p->k1 = 3.93e-5, p->k2 = 6.55e-4;
    p->kcas = 1.65e-4, p->kcac = 8.59e-9, p->taucas = 0.1, p->taucac = 1.75e3, p->bcas = 5.425e-4, p->bcac = 3.1e-8;
    p->ar = 5.0, p->ad = 0.18, p->tmax = 1, p->vt = -20, p->kp = 3, p->g_gaba = 0.01, p->appcur = 3.5;
    p->gnap = 2.85, p->tau_rl = 3.1, p->tau_rnl = 3.1, p->tau_p = 100;
    p->out_thresh = 0.1;
}

nnx_status nnx_diekman_step(nnx_ctx* c, const nnx_diekman_state* st, const nnx_diekman_params* q, double dt,
                            uint64_t* bits) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, st && q, "null state or params");
    NNX_CHECK(c, st->dtype == NNX_F32 || st->dtype == NNX_F64, "bad dtype %d", st->dtype);
    NNX_CHECK(c, st->n >= 0 && st->n < ((int64_t)1 << 31), "n out of range: %lld", (long long)st->n);
    NNX_CHECK(c, dt > 0, "dt must be > 0");
    if (st->n == 0) return NNX_OK;
    NNX_CHECK(c, st->v && st->m && st->h && st->ng && st->rl && st->rnl && st->fnl && st->s && st->p && st->y &&
                     st->cas && st->cac && st->iin && st->icount && st->ioutput && st->egaba && st->gkca && st->gkleak,
              "every Diekman array is needed");
    if (st->dtype == NNX_F64) return launch<double>(c, st, q, dt, bits);
    return launch<float>(c, st, q, dt, bits);
}

}  // extern "C"

// san.hip - the slow-wave "AN" neuron with excitatory and inhibitory
// synapses, one forward-Euler step.
//
// Replaces SANEphys::SimulateEphys with Euler_an_ie_model and
// dxdt_an_ie_model (src/ephys.cu:105-141,203-245).  One thread per neuron,
// the level spike rule on the new output folded in.  fp32 keeps the
// reference's float/double mix (expf of float arguments, exp of double ones,
// double literals); the white-noise drive is its mean.
#include "internal.h"

namespace nnx {

namespace {

template <class T> struct SanC {
    T c, a, v_l, v_na, v_k, v_ca, kd, v_ampa, v_gaba, alpha_ca, tau_ampa, tau_gaba;
    T g_l, g_k, g_nap, g_ca, g_kca, g_ampa, g_gaba, tau_ca, noise, tmax, vt, kp;
    double out_thresh;
};

__device__ __forceinline__ float rt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double rt(double x) { return __builtin_sqrt(x); }

template <class T> struct SanPtr {
    T *v, *nk, *ca, *ye, *yi, *eout, *iout;
    const T *ein, *ecount, *iin, *icount;
};

template <class T>
__global__ __launch_bounds__(256) void k_san(int64_t n, SanC<T> c, T dt, SanPtr<T> a, uint64_t* __restrict__ bits,
                                             int64_t nwords) {
    using W = double;
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    bool spike = false;
    if (i < n) {
        const T V = a.v[i], nK = a.nk[i], Ca = a.ca[i], Ye = a.ye[i], Yi = a.yi[i];
        const T ec = a.ecount[i], ic = a.icount[i];
        const T eT = ec == T(0) ? T(0) : a.ein[i] / ec;
        const T iT = ic == T(0) ? T(0) : a.iin[i] / ic;
        // dxdt_an_ie_model (src/ephys.cu:105-141)
        const T mCa_inf = T(1) / (T(1) + dexp(-(V + T(20)) / T(9)));
        const T r = c.kd / Ca;
        const T mKCa_inf = T(1) / (T(1) + r * r * r * rt(r));  // powf(x, 3.5)
        const T mNaP_inf = (T)(1.0 / (1.0 + (W)dexp(-((W)V + 55.7) / 7.7)));
        const T alpha_n = (T)(V == T(-34) ? 0.1 : 0.01 * (W)(V + T(34)) / (W)(T(1) - dexp(-(V + T(34)) / T(10))));
        const T beta_n = (T)(0.125 * (W)dexp(-(V + T(44)) / T(25)));
        const T dn = T(4) * (alpha_n * (T(1) - nK) - beta_n * nK);
        const T dye = (T)(3.48 * (W)eT - (W)(Ye / c.tau_ampa));
        const T dyi = iT - Yi / c.tau_gaba;
        const T ICa = c.g_ca * mCa_inf * mCa_inf * (V - c.v_ca);
        const T dca = -c.alpha_ca * (T(10) * c.a * ICa) - Ca / c.tau_ca;
        const T n2 = nK * nK;
        T dv = c.noise - c.g_k * (n2 * n2) * (V - c.v_k) - ICa - c.g_kca * mKCa_inf * (V - c.v_k) -
               c.g_nap * mNaP_inf * mNaP_inf * mNaP_inf * (V - c.v_na) - c.g_l * (V - c.v_l) -
               c.g_ampa * Ye * (V - c.v_ampa) - c.g_gaba * Yi * (V - c.v_gaba);
        dv = dv / c.c;
        // Euler_an_ie_model (src/ephys.cu:226-238)
        const T nV = V + dt * dv;
        a.v[i] = nV;
        a.nk[i] = nK + dt * dn;
        a.ca[i] = Ca + dt * dca;
        a.ye[i] = Ye + dt * dye;
        a.yi[i] = Yi + dt * dyi;
        const T out = (T)((W)c.tmax / (1.0 + (W)dexp(-(nV - c.vt) / c.kp)));
        a.iout[i] = out;
        a.eout[i] = out;
        spike = (double)out > c.out_thresh;
    }
    const uint64_t word = __ballot(spike);
    if (bits && (threadIdx.x & 63) == 0) {
        const int64_t w = i >> 6;
        if (w < nwords) bits[w] = word;
    }
}

template <class T> SanC<T> san_const(const nnx_san_params* q) {
    SanC<T> c;
    c.c = (T)q->c, c.a = (T)q->a, c.v_l = (T)q->v_l, c.v_na = (T)q->v_na, c.v_k = (T)q->v_k, c.v_ca = (T)q->v_ca;
    c.kd = (T)q->kd, c.v_ampa = (T)q->v_ampa, c.v_gaba = (T)q->v_gaba, c.alpha_ca = (T)q->alpha_ca;
    c.tau_ampa = (T)q->tau_ampa, c.tau_gaba = (T)q->tau_gaba, c.g_l = (T)q->g_l, c.g_k = (T)q->g_k;
    c.g_nap = (T)q->g_nap, c.g_ca = (T)q->g_ca, c.g_kca = (T)q->g_kca, c.g_ampa = (T)q->g_ampa;
    c.g_gaba = (T)q->g_gaba, c.tau_ca = (T)q->tau_ca, c.noise = (T)q->noise_mean, c.tmax = (T)q->tmax;
    c.vt = (T)q->vt, c.kp = (T)q->kp, c.out_thresh = q->out_thresh;
    return c;
}

template <class T> nnx_status launch(nnx_ctx* c, const nnx_san_state* st, const nnx_san_params* q, double dt, uint64_t* bits) {
    SanPtr<T> a;
    a.v = (T*)st->v, a.nk = (T*)st->nk, a.ca = (T*)st->ca, a.ye = (T*)st->ye, a.yi = (T*)st->yi;
    a.eout = (T*)st->eoutput, a.iout = (T*)st->ioutput;
    a.ein = (const T*)st->ein, a.ecount = (const T*)st->ecount, a.iin = (const T*)st->iin;
    a.icount = (const T*)st->icount;
    const int64_t n = st->n;
    k_san<T><<<(unsigned)ceil_div(n, 256), 256, 0, c->stream>>>(n, san_const<T>(q), (T)dt, a, bits, ceil_div(n, 64));
    NNX_LAUNCHED(c);
    return NNX_OK;
}

}  // namespace

}  // namespace nnx

using namespace nnx;

extern "C" {

void nnx_san_params_reference(nnx_san_params* p) {
    // an_ie_params (src/defines.cuh:118-149)
    *p = nnx_san_params{};
    p->c = 1, p->a = 0.02, p->v_l = -60.95, p->v_na = 55, p->v_k = -100, p->v_ca = 120, p->kd = 30;
    p->v_ampa = 0, p->v_gaba = -70, p->alpha_ca = 0.5, p->tau_ampa = 2, p->tau_gaba = 10;
    p->g_l = 0.016307, p->g_k = 19.20436, p->g_nap = 0.63314, p->g_ca = 0.1624, p->g_kca = 0.7506;
    p->g_ampa = 0.01, p->g_gaba = 0.5, p->tau_ca = 0.5 * 739.09, p->noise_mean = 0.1, p->tmax = 1.0;
    p->vt = -45, p->kp = 1, p->out_thresh = 0.1;
}

nnx_status nnx_san_step(nnx_ctx* c, const nnx_san_state* st, const nnx_san_params* q, double dt, uint64_t* bits) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, st && q, "null state or params");
    NNX_CHECK(c, st->dtype == NNX_F32 || st->dtype == NNX_F64, "bad dtype %d", st->dtype);
    NNX_CHECK(c, st->n >= 0 && st->n < ((int64_t)1 << 31), "n out of range: %lld", (long long)st->n);
    NNX_CHECK(c, dt > 0, "dt must be > 0");
    if (st->n == 0) return NNX_OK;
    NNX_CHECK(c, st->v && st->nk && st->ca && st->ye && st->yi && st->ein && st->ecount && st->iin && st->icount &&
                     st->eoutput && st->ioutput,
              "every SAN array is needed");
    if (st->dtype == NNX_F64) return launch<double>(c, st, q, dt, bits);
    return launch<float>(c, st, q, dt, bits);
}

}  // extern "C"

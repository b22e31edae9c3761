// ephys.hip - one Hodgkin-Huxley timestep for every neuron.
//
// Replaces HHEphys::SimulateEphys / Leapfrog_HH (src/ephys.cu:28-103,
// src_wholebrain/ephys.cu:149-224).  One thread per neuron over the SoA
// state, the whole update in registers, one 64-bit spike word per wavefront
// written from a ballot (no atomics).  The reference launches 32 blocks of 64
// threads with a grid-stride loop and synchronises the device afterwards; here
// the grid covers N (4096 blocks of 256 at N = 1M) and nothing synchronises.
#include "coupling_dev.h"
#include "internal.h"
#include "noise.h"

namespace nnx {

constexpr int kBlock = 256;

// EVENT: the coupling input is the target's pending event-driven sum.  A
// target with at most kSortMax bucket entries sums them itself (gather_one);
// heavier targets - and every target of a burst step, where the push was
// skipped - are pulled by the whole wavefront cooperatively over their
// in-edges (wave_pull_one).  Both give the ascending-source-order sum.
// Otherwise the input is read from gin (NULL = none).
template <class T, bool EVENT>
__global__ __launch_bounds__(kBlock) void k_hh_rk4(int64_t n, HHConst<T> p, T* __restrict__ V,
                                                   T* __restrict__ M, T* __restrict__ H,
                                                   T* __restrict__ NG, T* __restrict__ S,
                                                   const T* __restrict__ iext,
                                                   const T* __restrict__ gin, EventIn ev,
                                                   uint64_t* __restrict__ bits, int64_t nwords) {
    const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    T g = T(0);
    if (EVENT) {
        const bool dense = *ev.dense != 0;
        const int32_t c = (i < n && !dense) ? ev.cnt[i] : 0;
        bool heavy = i < n && (dense || c > kSortMax);
        if (!heavy && i < n && c > 0) {
            const int32_t b = ev.in_off[i];
            if (c <= ev.in_off[i + 1] - b) g = gather_one<T>((int32_t)i, c, ev.in_off, ev.in_src, ev.in_w, ev.prev_bits, ev.slot);
            else heavy = true;
        }
        uint64_t hm = __ballot(heavy);
        const int64_t wave0 = i - (threadIdx.x & 63);
        while (hm) {
            const int l = __ffsll((unsigned long long)hm) - 1;
            hm &= hm - 1;
            const T acc = wave_pull_one<T>(wave0 + l, ev.in_off, ev.in_src, ev.in_w, ev.prev_bits);
            if ((threadIdx.x & 63) == l) g = acc;
        }
        if (c) ev.cnt[i] = 0;
    } else if (gin && i < n) {
        g = gin[i];
    }
    bool spike = false;
    if (i < n) {
        T v = V[i], m = M[i], h = H[i], ng = NG[i], s = S[i];
        const T v0 = v;
        rk4(p, v, m, h, ng, s, iext[i], g);
        V[i] = v;
        M[i] = m;
        H[i] = h;
        NG[i] = ng;
        S[i] = s;
        spike = (v0 < p.v_thresh) && (v >= p.v_thresh);
    }
    const uint64_t word = __ballot(spike);
    if (bits && (threadIdx.x & 63) == 0) {
        const int64_t w = i >> 6;
        if (w < nwords) bits[w] = word;
    }
}

template <class T> struct LeapPtr {
    T *V, *M, *H, *NG, *Yl, *Yi, *Ye, *out_l, *out_i, *out_e;
    const T *in_l, *cnt_l, *in_i, *cnt_i, *in_e, *cnt_e, *appcur;
};

template <class T> __device__ __forceinline__ T ld0(const T* p, int64_t i) { return p ? p[i] : T(0); }

// White noise of the step: noise_std z + noise_mean with z the standard
// normal of (seed, step, neuron) (noise.h); the mean alone when noise_std
// is 0 (the deterministic mode).  The step number comes from *stepp when
// given (the device step counter of a captured step sequence), else `step`.
struct NoiseArg {
    uint64_t seed;
    uint64_t step;
    const int64_t* stepp;
    int on;
};

// src/ephys.cu:57-94
template <class T>
__global__ __launch_bounds__(kBlock) void k_hh_leapfrog(int64_t n, HHConst<T> p, LeapPtr<T> a,
                                                        int level, uint64_t* __restrict__ bits,
                                                        int64_t nwords, NoiseArg nz) {
    const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    bool spike = false;
    if (i < n) {
        T wn = p.noise_mean;
        if (nz.on) {
            const uint64_t step = nz.stepp ? (uint64_t)*nz.stepp : nz.step;
            wn = p.noise_std * T(noise_normal(nz.seed, step, (uint64_t)i)) + p.noise_mean;
        }
        const T V = a.V[i], M = a.M[i], NG = a.NG[i], H = a.H[i];
        const Rates<T> R = rates(V + p.vshift);
        const T dt = p.dt, hdt = p.hdt;
        const T nM = (R.am * dt + (T(1) - hdt * (R.am + R.bm)) * M) / (hdt * (R.am + R.bm) + T(1));
        const T nN = (R.an * dt + (T(1) - hdt * (R.an + R.bn)) * NG) / (hdt * (R.an + R.bn) + T(1));
        const T nH = (R.ah * dt + (T(1) - hdt * (R.ah + R.bh)) * H) / (hdt * (R.ah + R.bh) + T(1));
        const T nYl = kinetic(p, ld0(a.in_l, i), ld0(a.cnt_l, i), ld0((const T*)a.Yl, i));
        const T nYi = kinetic(p, ld0(a.in_i, i), ld0(a.cnt_i, i), ld0((const T*)a.Yi, i));
        const T nYe = kinetic(p, ld0(a.in_e, i), ld0(a.cnt_e, i), ld0((const T*)a.Ye, i));
        const T m3h = p.gna * nM * nM * nM * nH;
        const T n2 = nN * nN;
        const T n4 = p.gk * (n2 * n2);
        const T G = m3h + n4 + p.gl + p.g_local * nYl + p.g_i * nYi + p.g_e * nYe;
        const T E = m3h * p.ena + n4 * p.ek + p.gl * p.el + p.g_local * nYl * p.e_local +
                    p.g_i * nYi * p.e_i + p.g_e * nYe * p.e_e;
        const T Vn = ((wn + ld0(a.appcur, i)) * dt + E * dt + (T(1) - hdt * G) * V) / (T(1) + hdt * G);
        const T out = p.tmax / (T(1) + dexp(-(Vn - p.vt) / p.kp));
        a.V[i] = Vn;
        a.M[i] = nM;
        a.NG[i] = nN;
        a.H[i] = nH;
        if (a.Yl) a.Yl[i] = nYl;
        if (a.Yi) a.Yi[i] = nYi;
        if (a.Ye) a.Ye[i] = nYe;
        if (a.out_l) a.out_l[i] = out;
        if (a.out_i) a.out_i[i] = out;
        if (a.out_e) a.out_e[i] = out;
        spike = level ? ((double)out > p.out_thresh) : (V < p.v_thresh && Vn >= p.v_thresh);
    }
    const uint64_t word = __ballot(spike);
    if (bits && (threadIdx.x & 63) == 0) {
        const int64_t w = i >> 6;
        if (w < nwords) bits[w] = word;
    }
}

template <class T>
static nnx_status hh_step_t(nnx_ctx* c, const nnx_hh_state* st, const nnx_hh_params* q,
                            int integrator, int rule, double dt, const void* g_in,
                            uint64_t* bits, uint64_t step, const int64_t* stepp) {
    const int64_t n = st->n;
    const int64_t nwords = ceil_div(n, 64);
    const unsigned blocks = (unsigned)ceil_div(n, kBlock);
    HHConst<T> p = make_const<T>(q, dt);
    if (integrator == NNX_INT_RK4) {
        NNX_CHECK(c, rule == NNX_SPIKE_CROSSING, "RK4 supports the CROSSING spike rule only");
        NNX_CHECK(c, st->v && st->m && st->h && st->ng && st->s && st->i_ext,
                  "RK4 needs v, m, h, ng, s, i_ext");
        NNX_CHECK(c, q->tau_syn > 0, "tau_syn must be > 0");
        EventIn none{};
        k_hh_rk4<T, false><<<blocks, kBlock, 0, c->stream>>>(n, p, (T*)st->v, (T*)st->m, (T*)st->h,
                                                             (T*)st->ng, (T*)st->s, (const T*)st->i_ext,
                                                             (const T*)g_in, none, bits, nwords);
    } else {
        NNX_CHECK(c, st->v && st->m && st->h && st->ng, "leapfrog needs v, m, h, ng");
        LeapPtr<T> a;
        a.V = (T*)st->v;
        a.M = (T*)st->m;
        a.H = (T*)st->h;
        a.NG = (T*)st->ng;
        a.Yl = (T*)st->y_local;
        a.Yi = (T*)st->y_i;
        a.Ye = (T*)st->y_e;
        a.out_l = (T*)st->out_local;
        a.out_i = (T*)st->out_i;
        a.out_e = (T*)st->out_e;
        a.in_l = (const T*)st->in_local;
        a.cnt_l = (const T*)st->count_local;
        a.in_i = (const T*)st->in_i;
        a.cnt_i = (const T*)st->count_i;
        a.in_e = (const T*)st->in_e;
        a.cnt_e = (const T*)st->count_e;
        a.appcur = (const T*)st->appcur;
        NoiseArg nz{q->noise_seed, step, stepp, q->noise_std != 0.0};
        k_hh_leapfrog<T><<<blocks, kBlock, 0, c->stream>>>(n, p, a, rule == NNX_SPIKE_LEVEL, bits,
                                                           nwords, nz);
    }
    NNX_LAUNCHED(c);
    return NNX_OK;
}

template <class T> __global__ void k_finite(int64_t n, const T* x, int32_t* flag) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const T v = x[i];
        if (!(v == v) || v - v != T(0)) flag[0] = 1;
    }
}

nnx_status hh_rk4_event_launch(nnx_ctx* c, const nnx_hh_state* st, const nnx_hh_params* q,
                               double dt, const EventIn& ev, uint64_t* bits) {
    const int64_t n = st->n;
    const int64_t nwords = ceil_div(n, 64);
    const unsigned blocks = (unsigned)ceil_div(n, kBlock);
    if (st->dtype == NNX_F32) {
        HHConst<float> p = make_const<float>(q, dt);
        k_hh_rk4<float, true><<<blocks, kBlock, 0, c->stream>>>(
            n, p, (float*)st->v, (float*)st->m, (float*)st->h, (float*)st->ng, (float*)st->s,
            (const float*)st->i_ext, nullptr, ev, bits, nwords);
    } else {
        HHConst<double> p = make_const<double>(q, dt);
        k_hh_rk4<double, true><<<blocks, kBlock, 0, c->stream>>>(
            n, p, (double*)st->v, (double*)st->m, (double*)st->h, (double*)st->ng, (double*)st->s,
            (const double*)st->i_ext, nullptr, ev, bits, nwords);
    }
    NNX_LAUNCHED(c);
    return NNX_OK;
}

}  // namespace nnx

using namespace nnx;

__global__ void k_noise_normal(uint64_t seed, uint64_t step, int64_t first, int64_t n, double* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = noise_normal(seed, step, (uint64_t)(first + i));
}

nnx_status nnx::hh_step_launch(nnx_ctx* c, const nnx_hh_state* st, const nnx_hh_params* q, int integrator, int rule,
                               double dt, uint64_t step, const int64_t* stepp, const void* g_in, uint64_t* bits) {
    NNX_CHECK(c, st && q, "null state or params");
    NNX_CHECK(c, st->n >= 0 && st->n < (int64_t)1 << 31, "n out of range: %lld", (long long)st->n);
    NNX_CHECK(c, dt > 0, "dt must be > 0");
    NNX_CHECK(c, integrator == NNX_INT_RK4 || integrator == NNX_INT_LEAPFROG, "bad integrator %d",
              integrator);
    NNX_CHECK(c, rule == NNX_SPIKE_LEVEL || rule == NNX_SPIKE_CROSSING, "bad spike rule %d", rule);
    if (q->noise_std != 0.0 && integrator != NNX_INT_LEAPFROG)
        return fail(c, NNX_ERR_UNSUPPORTED, "white noise (noise_std != 0) is drawn by the leapfrog step only");
    if (st->n == 0) return NNX_OK;
    if (st->dtype == NNX_F32) return hh_step_t<float>(c, st, q, integrator, rule, dt, g_in, bits, step, stepp);
    if (st->dtype == NNX_F64) return hh_step_t<double>(c, st, q, integrator, rule, dt, g_in, bits, step, stepp);
    return fail(c, NNX_ERR_ARG, "bad dtype %d", st->dtype);
}

extern "C" nnx_status nnx_hh_step(nnx_ctx* c, const nnx_hh_state* st, const nnx_hh_params* q,
                                  int integrator, int rule, double dt, uint64_t step,
                                  const void* g_in, uint64_t* bits) {
    if (!c) return NNX_ERR_ARG;
    return hh_step_launch(c, st, q, integrator, rule, dt, step, nullptr, g_in, bits);
}

extern "C" nnx_status nnx_noise_normal(nnx_ctx* c, uint64_t seed, uint64_t step, int64_t first, int64_t n,
                                       double* d_out) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, n >= 0 && first >= 0 && (d_out || n == 0), "bad range or null output");
    if (n == 0) return NNX_OK;
    const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 256), 8192);
    k_noise_normal<<<blocks, 256, 0, c->stream>>>(seed, step, first, n, d_out);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

extern "C" nnx_status nnx_check_finite(nnx_ctx* c, int dtype, int64_t n, const void* x) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, x || n == 0, "null array");
    if (n == 0) return NNX_OK;
    void* f;
    NNX_TRY(scratch(c, "finite_flag", 256, true, &f));
    NNX_HIP(c, hipMemsetAsync(f, 0, 4, c->stream));
    const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 256), 2048);
    if (dtype == NNX_F32)
        k_finite<float><<<blocks, 256, 0, c->stream>>>(n, (const float*)x, (int32_t*)f);
    else
        k_finite<double><<<blocks, 256, 0, c->stream>>>(n, (const double*)x, (int32_t*)f);
    NNX_LAUNCHED(c);
    NNX_HIP(c, hipMemcpyAsync(c->pinned, f, 4, hipMemcpyDeviceToHost, c->stream));
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    if (*(int32_t*)c->pinned) return fail(c, NNX_ERR_NONFINITE, "non-finite value found");
    return NNX_OK;
}

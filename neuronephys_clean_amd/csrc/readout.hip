// readout.hip - electrode readout: per-region mean voltage and firing
// counts, the network mean, and the synchrony statistics of a region.
//
// Replaces ElectrodeNet (src/electrode_net.cu:3-47,
// src_wholebrain/electrode_net.cu:3-52) and SynchronyNet
// (src_wholebrain/electrode_net.cu:54-126).  The reference gives each region
// to one thread that adds its voltages one after the other, synchronises the
// device and copies the result to the host at every print step.  Here the
// sum keeps that order (so the means are the reference's bits) but a region
// of more than a few neurons is given to a whole wavefront: the 64 lanes
// gather 64 voltages at once and the sequential sum runs over the lanes with
// readlane, so only the adds are serial, not the memory latency.  Nothing is
// copied back: the step graph writes one row per print step into a device
// buffer.
#include <algorithm>

#include "internal.h"

namespace nnx {

namespace {

constexpr int kMeanBlocks = 1024;  // fixed grid of the network mean (fixes its summation order)
constexpr int kMeanThreads = 256;
// regions with at most this many neurons on average go one per thread
constexpr int64_t kThreadRegionMax = 8;

template <class T> struct RegK {
    int64_t nreg, nnz, n;
    const int32_t* off;
    const int32_t* idx;
    const T* v;
    const uint64_t* bits;  // NULL: no counts
    T* avg;
    int32_t* firing;
    // row selection inside a step graph: row = (*stepp + add) / every (rows
    // >= cap dropped); stepp NULL = row 0 always
    const int64_t* stepp;
    int64_t add, every, cap;
};

template <class T> __device__ __forceinline__ T qnan();
template <> __device__ __forceinline__ float qnan<float>() { return __builtin_nanf(""); }
template <> __device__ __forceinline__ double qnan<double>() { return __builtin_nan(""); }

__device__ __forceinline__ float rdlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rdlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// the output row of this launch (-1: dropped)
template <class T> __device__ __forceinline__ int64_t out_row(const RegK<T>& a) {
    if (!a.stepp) return 0;
    const int64_t row = (*a.stepp + a.add) / a.every;
    return row < a.cap ? row : -1;
}

__device__ __forceinline__ void region_bounds(const int32_t* off, int64_t r, int64_t nnz, int64_t& b,
                                              int64_t& e) {
    b = off[r];
    e = off[r + 1];
    b = b < 0 ? 0 : (b > nnz ? nnz : b);
    e = e < b ? b : (e > nnz ? nnz : e);
}

// One wavefront per region.  Lanes past the region's end contribute +0,
// which leaves the running sum unchanged (it starts at +0 and can never be
// -0), so every chunk adds all 64 lanes without a data-dependent bound.
template <class T>
__global__ __launch_bounds__(256) void k_region_wave(RegK<T> a) {
    const int64_t r = (blockIdx.x * (int64_t)256 + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (r >= a.nreg) return;
    const int64_t row = out_row(a);
    if (row < 0) return;
    int64_t b, e;
    region_bounds(a.off, r, a.nnz, b, e);
    T s = T(0);
    int32_t cnt = 0;
    for (int64_t base = b; base < e; base += 64) {
        const int64_t j = base + lane;
        T x = T(0);
        bool f = false;
        if (j < e) {
            const int32_t i = a.idx[j];
            if (i >= 0 && (int64_t)i < a.n) {
                x = a.v[i];
                if (a.bits) f = (a.bits[i >> 6] >> (i & 63)) & 1;
            } else {
                x = qnan<T>();
            }
        }
        cnt += __popcll(__ballot(f));
#pragma unroll
        for (int l = 0; l < 64; ++l) s = s + rdlane(x, l);
    }
    if (lane == 0) {
        a.avg[row * a.nreg + r] = s / T(e - b);
        if (a.bits && a.firing) a.firing[row * a.nreg + r] = cnt;
    }
}

// One thread per region (small regions): the reference's loop.
template <class T>
__global__ __launch_bounds__(256) void k_region_thread(RegK<T> a) {
    const int64_t r = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (r >= a.nreg) return;
    const int64_t row = out_row(a);
    if (row < 0) return;
    int64_t b, e;
    region_bounds(a.off, r, a.nnz, b, e);
    T s = T(0);
    int32_t cnt = 0;
    for (int64_t j = b; j < e; ++j) {
        const int32_t i = a.idx[j];
        if (i >= 0 && (int64_t)i < a.n) {
            s = s + a.v[i];
            if (a.bits) cnt += (int32_t)((a.bits[i >> 6] >> (i & 63)) & 1);
        } else {
            s = s + qnan<T>();
        }
    }
    a.avg[row * a.nreg + r] = s / T(e - b);
    if (a.bits && a.firing) a.firing[row * a.nreg + r] = cnt;
}

// ---------------------------------------------------------------------------
// fixed-order FP64 block sums (the same tree for every launch shape used)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_tree(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block total into thread 0: each wave's butterfly, then the waves in order.
template <int NT> __device__ __forceinline__ double block_total(double v) {
    __shared__ double sh[NT / 64];
    v = wave_tree(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0) {
        t = sh[0];
        for (int w = 1; w < NT / 64; ++w) t += sh[w];
    }
    return t;
}

// network mean: thread t of block b adds elements b*256 + t + k*2^18 in
// order; block totals; then one block adds partials t, t+256, t+512, t+768
// per thread, and the block total divided by n is rounded to T once
template <class T>
__global__ __launch_bounds__(kMeanThreads) void k_mean_part(int64_t n, const T* __restrict__ v,
                                                            double* __restrict__ part) {
    double acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)kMeanThreads + threadIdx.x; i < n;
         i += (int64_t)kMeanBlocks * kMeanThreads)
        acc += (double)v[i];
    const double t = block_total<kMeanThreads>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <class T>
__global__ __launch_bounds__(kMeanThreads) void k_mean_final(int64_t n, const double* __restrict__ part,
                                                             T* out, const int64_t* stepp, int64_t add,
                                                             int64_t every, int64_t cap) {
    int64_t row = 0;
    if (stepp) {
        row = (*stepp + add) / every;
        if (row >= cap) return;
    }
    double acc = 0;
    for (int k = threadIdx.x; k < kMeanBlocks; k += kMeanThreads) acc += part[k];
    const double t = block_total<kMeanThreads>(acc);
    if (threadIdx.x == 0) out[row] = (T)(t / (double)n);
}

// ---------------------------------------------------------------------------
// synchrony (SynchronyNet)
// ---------------------------------------------------------------------------
template <class T> __device__ __forceinline__ T gather_v(const T* v, int64_t n, int32_t i) {
    return (i >= 0 && (int64_t)i < n) ? v[i] : qnan<T>();
}

// src_wholebrain/electrode_net.cu:54-68: per-neuron running sums; the
// square is formed in T and widened, as there
template <class T>
__global__ __launch_bounds__(256) void k_sync_neurons(int64_t n, const T* __restrict__ v, int64_t n_idx,
                                                      const int32_t* __restrict__ idx, double* vi,
                                                      double* vi2) {
    for (int64_t j = blockIdx.x * (int64_t)256 + threadIdx.x; j < n_idx; j += (int64_t)gridDim.x * 256) {
        const T x = gather_v(v, n, idx[j]);
        vi2[j] += (double)(x * x);
        vi[j] += (double)x;
    }
}

// src_wholebrain/electrode_net.cu:80-92: one block of 1024 per
// sub-population; thread t adds entries t, t+1024, ... in order, then the
// block tree; Vsum rounded to T, Vavg = Vsum / count in T
template <class T>
__global__ __launch_bounds__(1024) void k_sync_subs(int64_t n, const T* __restrict__ v, int64_t n_idx,
                                                    const int32_t* __restrict__ idx,
                                                    const int32_t* __restrict__ counts, double* vavg,
                                                    double* vavg2) {
    const int s = blockIdx.x;
    const int32_t cnt = counts[s];
    const bool ok = cnt >= 1 && (int64_t)cnt <= n_idx;
    double acc = 0;
    if (ok)
        for (int32_t j = threadIdx.x; j < cnt; j += 1024) acc += (double)gather_v(v, n, idx[j]);
    const double t = block_total<1024>(acc);
    if (threadIdx.x == 0) {
        double d = __builtin_nan("");
        if (ok) {
            const T vs = (T)t;
            d = (double)(vs / (T)cnt);
        }
        vavg[s] += d;
        vavg2[s] += d * d;
    }
}

// src_wholebrain/electrode_net.cu:99-126: one thread per sub-population
__global__ void k_sync_chi(int64_t n_idx, int32_t n_sub, const int32_t* counts, const double* vi,
                           const double* vi2, const double* vavg, const double* vavg2, int64_t steps,
                           double* chi) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_sub) return;
    const int32_t cnt = counts[s];
    if (cnt < 1 || (int64_t)cnt > n_idx) {
        chi[s] = __builtin_nan("");
        return;
    }
    const double T_ = (double)steps;
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

template <class T>
nnx_status region_launch(nnx_ctx* c, const nnx_regions* r, int64_t n, const void* v, const uint64_t* bits,
                         void* avg, int32_t* firing, const int64_t* stepp, int64_t add, int64_t every,
                         int64_t cap) {
    RegK<T> a;
    a.nreg = r->n_regions, a.nnz = r->nnz, a.n = n;
    a.off = r->offsets, a.idx = r->idx, a.v = (const T*)v, a.bits = bits;
    a.avg = (T*)avg, a.firing = firing;
    a.stepp = stepp, a.add = add, a.every = every, a.cap = cap;
    if (r->nnz <= kThreadRegionMax * r->n_regions) {
        k_region_thread<T><<<(unsigned)ceil_div(r->n_regions, 256), 256, 0, c->stream>>>(a);
    } else {
        k_region_wave<T><<<(unsigned)ceil_div(r->n_regions, 4), 256, 0, c->stream>>>(a);
    }
    NNX_LAUNCHED(c);
    return NNX_OK;
}

template <class T>
nnx_status mean_launch(nnx_ctx* c, int64_t n, const void* v, void* out, const int64_t* stepp, int64_t add,
                       int64_t every, int64_t cap) {
    void* part;
    NNX_TRY(scratch(c, "mean_part", kMeanBlocks * sizeof(double), false, &part));
    k_mean_part<T><<<kMeanBlocks, kMeanThreads, 0, c->stream>>>(n, (const T*)v, (double*)part);
    NNX_LAUNCHED(c);
    k_mean_final<T><<<1, kMeanThreads, 0, c->stream>>>(n, (const double*)part, (T*)out, stepp, add, every, cap);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

nnx_status check_regions(nnx_ctx* c, const nnx_regions* r) {
    NNX_CHECK(c, r, "null regions");
    NNX_CHECK(c, r->n_regions >= 0 && r->n_regions < ((int64_t)1 << 31), "n_regions out of range: %lld",
              (long long)r->n_regions);
    NNX_CHECK(c, r->nnz >= 0 && r->nnz < ((int64_t)1 << 31), "nnz out of range: %lld", (long long)r->nnz);
    NNX_CHECK(c, r->n_regions == 0 || r->offsets, "null offsets");
    NNX_CHECK(c, r->nnz == 0 || r->idx, "null idx");
    return NNX_OK;
}

}  // namespace

nnx_status check_readout(nnx_ctx* c, const nnx_step_plan* p) {
    if (p->readout_every <= 0) return NNX_OK;
    NNX_TRY(check_regions(c, &p->regions));
    NNX_CHECK(c, p->readout_cap >= 0, "readout_cap < 0");
    NNX_CHECK(c, p->regions.n_regions == 0 || p->readout_avg, "readout needs readout_avg");
    return NNX_OK;
}

nnx_status readout_prepare(nnx_ctx* c, const nnx_step_plan* p) {
    if (p->readout_every <= 0 || !p->readout_mean) return NNX_OK;
    void* part;
    return scratch(c, "mean_part", kMeanBlocks * sizeof(double), false, &part);
}

nnx_status readout_launch(nnx_ctx* c, const nnx_step_plan* p, const int64_t* stepp, int64_t add,
                          const uint64_t* bits) {
    const int64_t n = p->state.n;
    const nnx_regions& r = p->regions;
    const bool f64 = p->state.dtype == NNX_F64;
    if (r.n_regions > 0) {
        const uint64_t* b = p->readout_firing ? bits : nullptr;
        if (f64)
            NNX_TRY(region_launch<double>(c, &r, n, p->state.v, b, p->readout_avg, p->readout_firing, stepp, add,
                                          p->readout_every, p->readout_cap));
        else
            NNX_TRY(region_launch<float>(c, &r, n, p->state.v, b, p->readout_avg, p->readout_firing, stepp, add,
                                         p->readout_every, p->readout_cap));
    }
    if (p->readout_mean && n > 0) {
        if (f64)
            NNX_TRY(mean_launch<double>(c, n, p->state.v, p->readout_mean, stepp, add, p->readout_every,
                                        p->readout_cap));
        else
            NNX_TRY(mean_launch<float>(c, n, p->state.v, p->readout_mean, stepp, add, p->readout_every,
                                       p->readout_cap));
    }
    return NNX_OK;
}

}  // namespace nnx

using namespace nnx;

extern "C" {

nnx_status nnx_electrode_measure(nnx_ctx* c, const nnx_regions* r, int dtype, int64_t n, const void* v,
                                 const uint64_t* bits, void* avg, int32_t* firing) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(check_regions(c, r));
    NNX_CHECK(c, dtype == NNX_F32 || dtype == NNX_F64, "bad dtype %d", dtype);
    NNX_CHECK(c, n >= 0 && n < ((int64_t)1 << 31), "n out of range: %lld", (long long)n);
    if (r->n_regions == 0) return NNX_OK;
    NNX_CHECK(c, avg, "null avg");
    NNX_CHECK(c, v || n == 0, "null v");
    NNX_CHECK(c, !bits || firing, "bits given without firing");
    if (dtype == NNX_F64) return region_launch<double>(c, r, n, v, bits, avg, firing, nullptr, 0, 1, 1);
    return region_launch<float>(c, r, n, v, bits, avg, firing, nullptr, 0, 1, 1);
}

nnx_status nnx_network_mean(nnx_ctx* c, int dtype, int64_t n, const void* v, void* mean) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, dtype == NNX_F32 || dtype == NNX_F64, "bad dtype %d", dtype);
    NNX_CHECK(c, n >= 1 && n < ((int64_t)1 << 31), "n out of range: %lld", (long long)n);
    NNX_CHECK(c, v && mean, "null pointer");
    if (dtype == NNX_F64) return mean_launch<double>(c, n, v, mean, nullptr, 0, 1, 1);
    return mean_launch<float>(c, n, v, mean, nullptr, 0, 1, 1);
}

nnx_status nnx_synchrony_accumulate(nnx_ctx* c, int dtype, int64_t n, const void* v, int64_t n_idx,
                                    const int32_t* idx, int32_t n_sub, const int32_t* counts, double* vi,
                                    double* vi2, double* vavg, double* vavg2) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, dtype == NNX_F32 || dtype == NNX_F64, "bad dtype %d", dtype);
    NNX_CHECK(c, n >= 0 && n < ((int64_t)1 << 31), "n out of range: %lld", (long long)n);
    NNX_CHECK(c, n_idx >= 0 && n_idx < ((int64_t)1 << 31), "n_idx out of range: %lld", (long long)n_idx);
    NNX_CHECK(c, n_sub >= 0 && n_sub <= 65535, "n_sub out of range: %d", n_sub);
    NNX_CHECK(c, (v || n == 0) && (idx || n_idx == 0), "null v or idx");
    NNX_CHECK(c, n_idx == 0 || (vi && vi2), "null vi_sum / vi2_sum");
    NNX_CHECK(c, n_sub == 0 || (counts && vavg && vavg2), "null sub-population arrays");
    const bool f64 = dtype == NNX_F64;
    if (n_idx > 0) {
        const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n_idx, 256), 4096);
        if (f64)
            k_sync_neurons<double><<<blocks, 256, 0, c->stream>>>(n, (const double*)v, n_idx, idx, vi, vi2);
        else
            k_sync_neurons<float><<<blocks, 256, 0, c->stream>>>(n, (const float*)v, n_idx, idx, vi, vi2);
        NNX_LAUNCHED(c);
    }
    if (n_sub > 0) {
        if (f64)
            k_sync_subs<double><<<n_sub, 1024, 0, c->stream>>>(n, (const double*)v, n_idx, idx, counts, vavg,
                                                               vavg2);
        else
            k_sync_subs<float><<<n_sub, 1024, 0, c->stream>>>(n, (const float*)v, n_idx, idx, counts, vavg,
                                                              vavg2);
        NNX_LAUNCHED(c);
    }
    return NNX_OK;
}

nnx_status nnx_synchrony_measures(nnx_ctx* c, int64_t n_idx, int32_t n_sub, const int32_t* counts,
                                  const double* vi, const double* vi2, const double* vavg, const double* vavg2,
                                  int64_t n_steps, double* chi) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, n_idx >= 0 && n_idx < ((int64_t)1 << 31), "n_idx out of range: %lld", (long long)n_idx);
    NNX_CHECK(c, n_sub >= 0 && n_sub <= 65535, "n_sub out of range: %d", n_sub);
    NNX_CHECK(c, n_steps >= 0, "n_steps < 0");
    if (n_sub == 0) return NNX_OK;
    NNX_CHECK(c, counts && vavg && vavg2 && chi, "null sub-population arrays");
    NNX_CHECK(c, n_idx == 0 || (vi && vi2), "null vi_sum / vi2_sum");
    k_sync_chi<<<(unsigned)ceil_div(n_sub, 64), 64, 0, c->stream>>>(n_idx, n_sub, counts, vi, vi2, vavg, vavg2,
                                                                   n_steps, chi);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

}  // extern "C"

// hebbian.hip - Hebbian learning of synapse weights, event by event.
//
// Replaces Hebbian::UpdateWeights with its set_timers and hebbian_rule_kernel
// (src_wholebrain/hebbian.cu:3-69).  The reference restarts a countdown of
// `window` calls on every out-synapse of a firing source and then visits
// every synapse of the network on every call to decrement the running
// countdowns; at C2 that is 2.1e8 read-modify-writes per call.  Here the
// countdown is implicit: a synapse is "armed" from its source's last firing
// step, and a call only visits
//   A. the out-synapses of the sources firing now (countdown restarted; the
//      reference's same-call rule applied),
//   B. the in-synapses of the targets firing now whose source is not firing
//      now (an armed synapse whose window has not ended goes up),
//   C. the out-synapses of the sources whose window ends now (fired exactly
//      window-1 calls ago and not since; armed synapses go down),
// found from the firing bitmap and a ring of the last `window` bitmaps.  The
// three sets of synapses are disjoint, so one launch does all three without
// atomics: one wavefront per 64-neuron word of the bitmap, the edges of each
// neuron across the lanes.
#include "internal.h"

namespace nnx {

namespace {

struct HebK {
    int64_t n, nwords;
    int32_t W;
    float inc;
    const int32_t* out_off;
    const int32_t* out_tgt;
    const int32_t* in_off;
    const int32_t* in_src;
    const int32_t* in_to_out;
    float* w;
    uint8_t* armed;
    int32_t* last;
    uint64_t* ring;
    const uint64_t* fire;
    const int64_t* stepp;  // step = *stepp + add (NULL: step_host)
    int64_t add, step_host;
};

__device__ __forceinline__ bool bit(const uint64_t* b, int32_t i) { return (b[i >> 6] >> (i & 63)) & 1; }

// the reference's increment heb_inc * exp(-|w - 1|)
__device__ __forceinline__ float dw(float inc, float w) { return inc * dexp(-__builtin_fabsf(w - 1.0f)); }

__global__ __launch_bounds__(256) void k_hebbian(HebK k) {
    const int lane = threadIdx.x & 63;
    const int64_t wv = (blockIdx.x * (int64_t)256 + threadIdx.x) >> 6;
    if (wv >= k.nwords) return;
    const int64_t step = k.stepp ? *k.stepp + k.add : k.step_host;
    const int32_t t = (int32_t)step;  // differences are taken in 32 bits
    const int64_t slot_now = ((step % k.W) + k.W) % k.W;
    const int64_t slot_end = (((step - k.W + 1) % k.W) + k.W) % k.W;
    uint64_t f = k.fire[wv];
    if (wv == k.nwords - 1 && (k.n & 63)) f &= (1ull << (k.n & 63)) - 1;  // bits past n are not neurons
    // sources whose window ends now: fired window-1 calls ago (W = 1: none,
    // that case is settled within the firing call)
    uint64_t ending = k.W > 1 ? k.ring[slot_end * k.nwords + wv] & ~f : 0ull;
    if (lane == 0) k.ring[slot_now * k.nwords + wv] = f;
    const int32_t base = (int32_t)(wv << 6);
    for (uint64_t fm = f; fm; fm &= fm - 1) {
        const int32_t p = base + __ffsll((unsigned long long)fm) - 1;
        // A: restart the countdown of p's out-synapses; after this call's
        // decrement it reads window-1: zero -> down, else up if the target fires
        const int32_t e1 = k.out_off[p + 1];
        for (int32_t e = k.out_off[p] + lane; e < e1; e += 64) {
            const float w = k.w[e];
            if (k.W == 1) {
                k.w[e] = w - dw(k.inc, w);
                k.armed[e] = 0;
            } else if (bit(k.fire, k.out_tgt[e])) {
                k.w[e] = w + dw(k.inc, w);
                k.armed[e] = 0;
            } else {
                k.armed[e] = 1;
            }
        }
        // B: p as a target; synapses from sources firing now are A's
        const int32_t c1 = k.in_off[p + 1];
        for (int32_t c = k.in_off[p] + lane; c < c1; c += 64) {
            const int32_t src = k.in_src[c];
            if (bit(k.fire, src)) continue;
            const int32_t e = k.in_to_out[c];
            if (!k.armed[e]) continue;
            if ((int32_t)((uint32_t)t - (uint32_t)k.last[src]) < k.W - 1) {  // == W-1: the window ends now (C)
                const float w = k.w[e];
                k.w[e] = w + dw(k.inc, w);
                k.armed[e] = 0;
            }
        }
    }
    // every read of last[] by B and C concerns sources not firing now, so
    // the firing sources' steps can be written as soon as their A is done
    if (lane == 0)
        for (uint64_t fm = f; fm; fm &= fm - 1) k.last[base + __ffsll((unsigned long long)fm) - 1] = t;
    for (; ending; ending &= ending - 1) {
        const int32_t p = base + __ffsll((unsigned long long)ending) - 1;
        if (k.last[p] != (int32_t)((uint32_t)t - (uint32_t)(k.W - 1))) continue;  // fired again since: runs on
        const int32_t e1 = k.out_off[p + 1];
        for (int32_t e = k.out_off[p] + lane; e < e1; e += 64) {
            if (!k.armed[e]) continue;
            const float w = k.w[e];
            k.w[e] = w - dw(k.inc, w);
            k.armed[e] = 0;
        }
    }
}

__global__ void k_heb_clear(int64_t nnz, int64_t n, int64_t nring, uint8_t* armed, int32_t* last, uint64_t* ring) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += stride) armed[i] = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) last[i] = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nring; i += stride) ring[i] = 0;
}

__global__ void k_invert(int64_t nnz, const int32_t* __restrict__ out_to_in, int32_t* __restrict__ in_to_out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
        in_to_out[out_to_in[e]] = (int32_t)e;
}

nnx_status check_heb(nnx_ctx* c, const nnx_graph* g, const nnx_hebbian* h) {
    NNX_CHECK(c, g && h, "null graph or hebbian state");
    NNX_CHECK(c, g->n >= 0 && g->n < ((int64_t)1 << 31) && g->nnz >= 0 && g->nnz < ((int64_t)1 << 31),
              "graph size out of range");
    NNX_CHECK(c, h->window >= 1 && h->window <= (1 << 24), "window must be in [1, 2^24]: %d", h->window);
    NNX_CHECK(c, g->out_offsets && g->out_targets && g->in_offsets && g->in_sources, "graph needs CSR and CSC");
    NNX_CHECK(c, g->nnz == 0 || (h->weights && h->armed && h->in_to_out), "null weights / armed / in_to_out");
    NNX_CHECK(c, h->last_fire && h->ring, "null last_fire / ring");
    return NNX_OK;
}

}  // namespace

nnx_status hebbian_launch(nnx_ctx* c, const nnx_graph* g, const nnx_hebbian* h, const int64_t* stepp, int64_t add,
                          int64_t step, const uint64_t* fire) {
    if (g->n == 0) return NNX_OK;
    HebK k;
    k.n = g->n;
    k.nwords = ceil_div(g->n, 64);
    k.W = h->window;
    k.inc = h->heb_inc;
    k.out_off = g->out_offsets, k.out_tgt = g->out_targets, k.in_off = g->in_offsets, k.in_src = g->in_sources;
    k.in_to_out = h->in_to_out;
    k.w = h->weights, k.armed = h->armed, k.last = h->last_fire, k.ring = h->ring;
    k.fire = fire;
    k.stepp = stepp, k.add = add, k.step_host = step;
    k_hebbian<<<(unsigned)ceil_div(k.nwords, 4), 256, 0, c->stream>>>(k);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

nnx_status check_hebbian_plan(nnx_ctx* c, const nnx_step_plan* p) {
    if (p->hebbian.window == 0) return NNX_OK;
    return check_heb(c, &p->graph, &p->hebbian);
}

}  // namespace nnx

using namespace nnx;

extern "C" {

nnx_status nnx_hebbian_init(nnx_ctx* c, const nnx_graph* g, const nnx_hebbian* h) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(check_heb(c, g, h));
    NNX_CHECK(c, g->nnz == 0 || g->out_to_in, "graph needs out_to_in");
    const int64_t nring = (int64_t)h->window * ceil_div(g->n, 64);
    k_heb_clear<<<1024, 256, 0, c->stream>>>(g->nnz, g->n, nring, h->armed, h->last_fire, h->ring);
    NNX_LAUNCHED(c);
    if (g->nnz) {
        k_invert<<<2048, 256, 0, c->stream>>>(g->nnz, g->out_to_in, h->in_to_out);
        NNX_LAUNCHED(c);
    }
    return NNX_OK;
}

nnx_status nnx_hebbian_update(nnx_ctx* c, const nnx_graph* g, const nnx_hebbian* h, int64_t step,
                              const uint64_t* fire) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(check_heb(c, g, h));
    NNX_CHECK(c, fire || g->n == 0, "null firing bitmap");
    NNX_CHECK(c, step >= 0, "step < 0");
    return hebbian_launch(c, g, h, nullptr, 0, step, fire);
}

}  // extern "C"

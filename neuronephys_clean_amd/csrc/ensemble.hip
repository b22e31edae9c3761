// ensemble.hip - many independent HH networks of uneven size stepped by one
// persistent fused kernel per timestep.
//
// The networks share one index space (network i owns neurons
// [begin_i, begin_i + size_i), begin_i a multiple of 64, so that no spike
// word straddles two networks) and one block-diagonal CSR/CSC.  A timestep
// is one launch of a fixed grid of workgroups that walk a static list of
// tiles (a network's neurons cut into pieces of `tile`, largest first,
// dealt round-robin).  For each tile: the event-driven coupling input of
// the previous step's spikes, the RK4 update, the spike word, and the push
// of the new spikes for the next step (event storage and spike words
// ping-pong by step parity, so no workgroup waits for another within a
// step).
//
// Event storage is one bit per in-edge, at the edge's CSC position, plus a
// per-target flag: a push sets the bit (atomic OR, no return value) and the
// flag (plain store), so it never waits on memory.  A flagged target reads
// the bit words of its CSC range and adds the weights of the set bits in
// ascending position - ascending source - order, clearing the bits as it
// goes.  After a step that fired more than `bulk` neurons the next step's
// pushes leave the flags alone and every target reads its bit words (a word
// per parity records which form the pending events took): once a sizeable
// part of the targets receives events, the random flag stores cost more
// than scanning every target's words.  Because every sum has a fixed
// order, the trajectory is bit-identical for every (block, tile, grid,
// bulk) configuration; the autotuner uses this to time candidate
// configurations on real steps.
//
// No counterpart in the reference, which simulates one network
// (src/main.cu:112-182).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "coupling_dev.h"
#include "internal.h"

namespace nnx {

struct EnsK {
    float *V, *M, *H, *NG, *S;
    const float* iext;
    const int32_t *out_off, *out_tgt, *out_to_in, *in_off, *in_src;
    const float* in_w;
    int32_t *flag0, *flag1;  // per target: events pending
    uint32_t *eb0, *eb1;     // per in-edge (CSC position): event bit
    uint64_t *bits0, *bits1;
    const int32_t* voxel_of;
    int32_t* voxel_counts;
    int32_t* net_spikes;
    int32_t* raster;
    unsigned long long* raster_len;
    int64_t raster_cap;
    const int4* tiles;  // (first neuron, end, network, 0)
    int32_t ntiles;
    int64_t* stepctr;   // step number by parity
    int32_t* spk3;      // spikes of each step, by step mod 3
    int32_t* evmode;    // by step parity: 1 when that step pushed its events without target flags
    int32_t burst;      // above this many spikes the next step pulls densely (< 0: never)
    int32_t bulk;       // above this many spikes the next step pushes without flags (< 0: never)
};

// Sum of the pending events of the target whose CSC range is [e0, e1)
// (e1 > e0): weights of the set bits of eb in that range, in ascending
// position.  The bits are cleared: words that hold only this target's edges
// by a store, the first and last word (which may hold a neighbour's edges,
// read and cleared concurrently) by an atomic AND of the bits read.
__device__ __forceinline__ float gather_bits(int32_t e0, int32_t e1, uint32_t* __restrict__ eb,
                                             const float* __restrict__ in_w) {
    float acc = 0.f;
    const int32_t wf = e0 >> 5, wl = (e1 - 1) >> 5;
    const uint32_t mf = ~0u << (e0 & 31);
    const uint32_t ml = 0xffffffffu >> (31 - ((e1 - 1) & 31));
    for (int32_t w0 = wf; w0 <= wl; w0 += 8) {
        uint32_t m[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) m[k] = (w0 + k <= wl) ? eb[w0 + k] : 0u;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t w = w0 + k;
            uint32_t x = m[k];
            if (w == wf) x &= mf;
            if (w == wl) x &= ml;
            if (x) {
                if (w == wf || w == wl)
                    atomicAnd(&eb[w], ~x);
                else
                    eb[w] = 0u;
                do {
                    const int b = __ffs(x) - 1;
                    x &= x - 1;
                    acc = acc + in_w[w * 32 + b];
                } while (x);
            }
        }
    }
    return acc;
}

// Clears the event bits of the CSC range [e0, e1) (e1 > e0) without
// reading them (a dense step recomputes the sums from the spike bitmap).
__device__ __forceinline__ void clear_bits(int32_t e0, int32_t e1, uint32_t* __restrict__ eb) {
    const int32_t wf = e0 >> 5, wl = (e1 - 1) >> 5;
    const uint32_t mf = ~0u << (e0 & 31);
    const uint32_t ml = 0xffffffffu >> (31 - ((e1 - 1) & 31));
    if (wf == wl) {
        atomicAnd(&eb[wf], ~(mf & ml));
        return;
    }
    atomicAnd(&eb[wf], ~mf);
    for (int32_t w = wf + 1; w < wl; ++w) eb[w] = 0u;
    atomicAnd(&eb[wl], ~ml);
}

// par: parity of this step; push = 0 leaves the new spikes unpushed (the
// caller pushes them from the compacted list before the next step).
// Burst steps: every wavefront adds its spike count to spk3[step mod 3] and
// pushes only while the running total stays within `burst`; if the final
// total exceeds it, some pushes were skipped, and the next step pulls every
// target over its in-edges with the spike bitmap instead (same sums, same
// order), clearing the partly set event bits.
template <int B>
__global__ __launch_bounds__(B) void k_ens_step(EnsK d, HHConst<float> p, int par, int push) {
    const int q = par ^ 1;
    int32_t* __restrict__ flag_r = q ? d.flag1 : d.flag0;  // set by the previous step
    uint32_t* __restrict__ eb_r = q ? d.eb1 : d.eb0;
    int32_t* __restrict__ flag_w = par ? d.flag1 : d.flag0;
    uint32_t* __restrict__ eb_w = par ? d.eb1 : d.eb0;
    const uint64_t* __restrict__ bits_r = q ? d.bits1 : d.bits0;
    uint64_t* __restrict__ bits_w = par ? d.bits1 : d.bits0;
    const int64_t step = d.stepctr[q] + 1;
    const int cur3 = (int)(step % 3);
    const int prev_spk = d.spk3[(cur3 + 2) % 3];  // final: written by the previous launch
    const bool dense = d.burst >= 0 && prev_spk > d.burst;
    // flagless events: the previous step pushed without setting target
    // flags (bulk_r), so every target scans its event bits; this step does
    // the same when the previous one fired more than `bulk` (bulk_w)
    const bool bulk_r = d.evmode[q] != 0;
    const bool bulk_w = d.bulk >= 0 && prev_spk > d.bulk;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        d.stepctr[par] = step;
        d.spk3[(cur3 + 1) % 3] = 0;  // the next step's counter (last read by the previous step)
        d.evmode[par] = bulk_w;
    }
    const int lane = threadIdx.x & 63;
    for (int t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        const int4 tl = d.tiles[t];
        for (int32_t i0 = tl.x; i0 < tl.y; i0 += B) {
            const int32_t i = i0 + (int32_t)threadIdx.x;
            const bool live = i < tl.y;
            // the state is loaded first, so that its latency overlaps the
            // event gather's
            float v = 0.f, m = 0.f, h = 0.f, ng = 0.f, s = 0.f, ie = 0.f;
            if (live) v = d.V[i], m = d.M[i], h = d.H[i], ng = d.NG[i], s = d.S[i], ie = d.iext[i];
            // coupling input
            float g = 0.f;
            const int32_t wave0 = i - lane;
            if (!dense) {
                if (live && (bulk_r || flag_r[i])) {
                    const int32_t e0 = d.in_off[i], e1 = d.in_off[i + 1];
                    if (e1 > e0) g = gather_bits(e0, e1, eb_r, d.in_w);
                    if (!bulk_r) flag_r[i] = 0;
                }
            } else {
                if (live && (bulk_r || flag_r[i])) {
                    const int32_t e0 = d.in_off[i], e1 = d.in_off[i + 1];
                    if (e1 > e0) clear_bits(e0, e1, eb_r);
                    if (!bulk_r) flag_r[i] = 0;
                }
                uint64_t hm = __ballot(live);
                while (hm) {
                    const int l = __ffsll((unsigned long long)hm) - 1;
                    hm &= hm - 1;
                    const float acc = wave_pull_one<float>(wave0 + l, d.in_off, d.in_src, d.in_w, bits_r);
                    if (lane == l) g = acc;
                }
            }
            // RK4 update and spike
            bool spike = false;
            if (live) {
                const float v0 = v;
                rk4(p, v, m, h, ng, s, ie, g);
                d.V[i] = v;
                d.M[i] = m;
                d.H[i] = h;
                d.NG[i] = ng;
                d.S[i] = s;
                spike = (v0 < p.v_thresh) && (v >= p.v_thresh);
            }
            const uint64_t word = __ballot(spike);
            // a wavefront past the tile's end (last chunk shorter than B) has no
            // word of its own: the next word may belong to another tile or lie
            // past the bitmap
            if (lane == 0 && wave0 < tl.y) bits_w[wave0 >> 6] = word;
            if (!word) continue;
            const int k = __popcll(word);
            bool do_push = push != 0;
            if (d.burst >= 0 || d.bulk >= 0) {
                int old = 0;
                if (lane == 0) old = atomicAdd(&d.spk3[cur3], k);
                old = __shfl(old, 0, 64);
                if (d.burst >= 0) do_push = do_push && old + k <= d.burst;
            }
            // push the new spikes into the next step's buckets, one source at
            // a time across the wavefront
            uint64_t sm = do_push ? word : 0ull;
            while (sm) {
                const int l = __ffsll((unsigned long long)sm) - 1;
                sm &= sm - 1;
                const int32_t j = wave0 + l;
                const int32_t e1 = d.out_off[j + 1];
                for (int32_t e = d.out_off[j] + lane; e < e1; e += 64) {
                    const int32_t pos = d.out_to_in[e];
                    atomicOr(&eb_w[pos >> 5], 1u << (pos & 31));
                    if (!bulk_w) flag_w[d.out_tgt[e]] = 1;
                }
            }
            if (spike && d.voxel_of) atomicAdd(&d.voxel_counts[d.voxel_of[i]], 1);
            if (d.net_spikes && lane == 0) atomicAdd(&d.net_spikes[tl.z], k);
            if (d.raster) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(d.raster_len, (unsigned long long)k);
                const uint32_t lo = __shfl((uint32_t)(base & 0xffffffffull), 0, 64);
                const uint32_t hi = __shfl((uint32_t)(base >> 32), 0, 64);
                base = ((unsigned long long)hi << 32) | lo;
                if (spike) {
                    const int64_t pos = (int64_t)base + __popcll(word & ((1ull << lane) - 1ull));
                    if (pos < d.raster_cap) {
                        d.raster[2 * pos] = (int32_t)step;
                        d.raster[2 * pos + 1] = (int32_t)i;
                    }
                }
            }
        }
    }
}

// events of every spike of `list` (all `count` of them), one wavefront per
// source
__global__ __launch_bounds__(256) void k_push_bits(const int32_t* __restrict__ list, const int32_t* __restrict__ count,
                                                   const int32_t* __restrict__ out_off,
                                                   const int32_t* __restrict__ out_tgt,
                                                   const int32_t* __restrict__ out_to_in, uint32_t* __restrict__ eb,
                                                   int32_t* __restrict__ flag) {
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    const int32_t K = *count;
    for (int64_t k = blockIdx.x * 4ll + (threadIdx.x >> 6); k < K; k += nwaves) {
        const int32_t j = list[k];
        const int32_t e1 = out_off[j + 1];
        for (int32_t e = out_off[j] + lane; e < e1; e += 64) {
            const int32_t pos = out_to_in[e];
            atomicOr(&eb[pos >> 5], 1u << (pos & 31));
            flag[out_tgt[e]] = 1;
        }
    }
}

// run set-up: the step counter of the step before the first; fresh = 1
// (buckets empty or about to be filled by a full push of the previous
// step's list) also marks the previous step as no burst
__global__ void k_ens_begin(int64_t* stepctr, int q, int64_t prev, int32_t* spk3, int32_t* evmode, int fresh) {
    stepctr[q] = prev;
    if (fresh) {
        const int p3 = (int)((prev % 3 + 3) % 3);
        spk3[p3] = 0;
        spk3[(p3 + 1) % 3] = 0;
        evmode[q] = 0;  // pending events (if any) carry target flags
    }
}

}  // namespace nnx

using namespace nnx;

namespace {

bool valid_block(int b) { return b == 64 || b == 128 || b == 256 || b == 512; }

// Spikes in a step above which the next step pushes its events without
// target flags, so that every target scans its event bits (`req` > 0: that
// many; < 0: never; 0: the library's rule).  The rule for one network: when
// the events would exceed one per 64 targets, spikes * nnz / n > n / 64
// (C2, 1M neurons and 2.1e8 edges: 81 spikes; 6% off its 10,000-step
// window).  For an ensemble: never - its busy steps are mostly a few
// networks firing while the others are quiet, and scanning every target of
// the launch then costs more than the flags (C4: 77.1 against 76.0 us per
// step).  Either way the sums are the same.
int32_t bulk_threshold(int32_t req, int64_t n, int64_t nnz, bool ensemble) {
    if (req != 0) return req < 0 ? -1 : req;
    if (ensemble || nnz <= 0) return -1;
    const double t = (double)n * (double)n / (64.0 * (double)nnz);
    return (int32_t)std::min(t, 1e9);
}

nnx_status check_ens(nnx_ctx* c, const nnx_ens_plan* p) {
    NNX_CHECK(c, p, "null plan");
    const nnx_hh_state& s = p->state;
    NNX_CHECK(c, s.dtype == NNX_F32, "the ensemble runs in fp32");
    NNX_CHECK(c, s.v && s.m && s.h && s.ng && s.s && s.i_ext, "state needs v, m, h, ng, s, i_ext");
    NNX_CHECK(c, s.n > 0 && s.n % 64 == 0, "state n must be a positive multiple of 64");
    NNX_CHECK(c, p->graph.n == s.n, "graph n != state n");
    NNX_CHECK(c, p->graph.out_offsets && p->graph.out_targets && p->graph.in_offsets && p->graph.in_sources &&
                     p->graph.in_weights && p->graph.out_to_in,
              "graph needs CSR, CSC and out_to_in");
    NNX_CHECK(c, p->bits[0] && p->bits[1], "plan needs bits[2]");
    NNX_CHECK(c, p->nnet > 0 && p->net_begin && p->net_size, "plan needs networks");
    int64_t end = 0;
    for (int i = 0; i < p->nnet; ++i) {
        NNX_CHECK(c, p->net_begin[i] % 64 == 0 && p->net_begin[i] >= end && p->net_size[i] >= 0,
                  "network %d: begin must be a multiple of 64 after the previous network", i);
        end = p->net_begin[i] + p->net_size[i];
    }
    NNX_CHECK(c, end <= s.n, "networks exceed the state");
    if (p->raster) NNX_CHECK(c, p->raster_len && p->raster_cap >= 0, "raster needs raster_len");
    if (p->voxel_of) NNX_CHECK(c, p->voxel_counts, "voxel_of needs voxel_counts");
    NNX_TRY(para_check(c, para_io(p)));
    if (p->params.noise_std != 0.0)
        return fail(c, NNX_ERR_UNSUPPORTED, "white noise (noise_std != 0) is drawn by the leapfrog step only");
    return NNX_OK;
}

nnx_status check_cfg(nnx_ctx* c, const nnx_ens_config* f) {
    NNX_CHECK(c, valid_block(f->block), "block must be 64, 128, 256 or 512");
    NNX_CHECK(c, f->tile >= 64 && f->tile % 64 == 0 && f->tile <= (1 << 24), "tile must be a multiple of 64");
    NNX_CHECK(c, f->grid >= 1 && f->grid <= 65536, "grid must be in [1, 65536]");
    return NNX_OK;
}

// tiles of `tile` neurons, largest first (ties: network order), in a buffer
// that belongs to this tile size and `kind` (captured graphs keep its
// address); the upload is skipped when the buffer already holds this list
nnx_status build_tiles(nnx_ctx* c, const char* kind, int64_t n, int nnet, const int64_t* begin,
                       const int64_t* size, int tile, const int4** out, int* ntiles) {
    std::vector<int4> t;
    for (int i = 0; i < nnet; ++i)
        for (int64_t a = begin[i]; a < begin[i] + size[i]; a += tile)
            t.push_back(make_int4((int)a, (int)std::min<int64_t>(a + tile, begin[i] + size[i]), i, 0));
    std::stable_sort(t.begin(), t.end(), [](const int4& a, const int4& b) { return (a.y - a.x) > (b.y - b.x); });
    char name[48];
    std::snprintf(name, sizeof name, "%s_tiles_%d", kind, tile);
    const size_t cap = (size_t)(n / 64 + nnet + 1) * sizeof(int4);
    void* buf;
    NNX_TRY(scratch(c, name, cap, false, &buf));
    uint64_t h = 0x7469006c65730000ull ^ t.size();
    for (const int4& v : t) h = (h ^ ((uint64_t)(uint32_t)v.x << 32 | (uint32_t)v.y)) * 1099511628211ull;
    h |= 1;  // never 0 (0 marks an empty buffer)
    nnx_ctx::Buf& b = c->bufs[name];
    if (b.tag != h && !t.empty()) {
        NNX_HIP(c, hipMemcpyAsync(buf, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
        NNX_HIP(c, hipStreamSynchronize(c->stream));  // t is a host temporary
    }
    b.tag = h;
    *out = (const int4*)buf;
    *ntiles = (int)t.size();
    return NNX_OK;
}

struct EnsHost {
    EnsK k;
    HHConst<float> hc;
    size_t ebits_bytes;
};

// kernel arguments shared by the ensemble and the fused single-network step;
// `kind` names the bucket, counter and tile buffers of each
nnx_status prepare(nnx_ctx* c, const char* kind, const nnx_hh_state& s, const nnx_hh_params& prm,
                   const nnx_graph& g, double dt, uint64_t* const bits[2], EnsHost* h) {
    const int64_t n = s.n, nnz = g.nnz;
    char nc[48], ns[48], nt[48];
    std::snprintf(nc, sizeof nc, "%s_flag", kind);
    std::snprintf(ns, sizeof ns, "%s_ebits", kind);
    std::snprintf(nt, sizeof nt, "%s_ctr", kind);
    void *flag, *eb, *ctr;
    const int64_t nw = nnz / 32 + 2;  // event-bit words per parity
    NNX_TRY(scratch(c, nc, (size_t)2 * n * 4, true, &flag));
    NNX_TRY(scratch(c, ns, (size_t)2 * nw * 4, true, &eb));
    NNX_TRY(scratch(c, nt, 64, true, &ctr));
    EnsK& k = h->k;
    k = EnsK{};
    k.V = (float*)s.v, k.M = (float*)s.m, k.H = (float*)s.h, k.NG = (float*)s.ng, k.S = (float*)s.s;
    k.iext = (const float*)s.i_ext;
    k.out_off = g.out_offsets, k.out_tgt = g.out_targets, k.out_to_in = g.out_to_in;
    k.in_off = g.in_offsets, k.in_src = g.in_sources, k.in_w = g.in_weights;
    k.flag0 = (int32_t*)flag, k.flag1 = (int32_t*)flag + n;
    k.eb0 = (uint32_t*)eb, k.eb1 = (uint32_t*)eb + nw;
    h->ebits_bytes = (size_t)2 * nw * 4;
    k.bits0 = bits[0], k.bits1 = bits[1];
    k.stepctr = (int64_t*)ctr;                       // bytes 0..15
    k.spk3 = (int32_t*)((char*)ctr + 16);            // bytes 16..27
    k.evmode = (int32_t*)((char*)ctr + 32);          // bytes 32..39
    k.burst = -1;  // no dense fallback unless asked for (event pushes are cheaper up to near-synchronous firing)
    k.bulk = -1;
    h->hc = make_const<float>(&prm, dt);
    return NNX_OK;
}

nnx_status ens_launch(nnx_ctx* c, const EnsK& k, const HHConst<float>& hc, const nnx_ens_config& f, int par,
                      int push) {
    switch (f.block) {
        case 64: k_ens_step<64><<<f.grid, 64, 0, c->stream>>>(k, hc, par, push); break;
        case 128: k_ens_step<128><<<f.grid, 128, 0, c->stream>>>(k, hc, par, push); break;
        case 256: k_ens_step<256><<<f.grid, 256, 0, c->stream>>>(k, hc, par, push); break;
        default: k_ens_step<512><<<f.grid, 512, 0, c->stream>>>(k, hc, par, push); break;
    }
    NNX_LAUNCHED(c);
    return NNX_OK;
}

uint64_t ens_key(const nnx_ens_plan* p, const nnx_ens_config& f) {
    uint64_t h = 0x9e3779b97f4a7c15ull;
    auto mix = [&](const void* b, size_t n) {
        const unsigned char* u = (const unsigned char*)b;
        for (size_t i = 0; i < n; ++i) h = (h ^ u[i]) * 1099511628211ull;
    };
    mix(p, sizeof(*p));
    mix(p->net_begin, sizeof(int64_t) * p->nnet);
    mix(p->net_size, sizeof(int64_t) * p->nnet);
    mix(&f, sizeof f);
    return h ^ 0x454e53454d424c45ull;  // distinct from the single-network step keys
}

// one 256-neuron tile per work item, enough workgroups to keep every CU
// busy through burst steps (measured over 5,000 C4 steps: 82 us/step
// against 116 us for 2048-neuron tiles on 1024 workgroups)
const nnx_ens_config kDefaultCfg = {256, 256, 8192, 0};
const nnx_ens_config kFusedDefault = {256, 256, 4096, 0};

}  // namespace

// The WEIGHTED_FUSED timestep of nnx_run_steps: one network stepped by the
// ensemble kernel.  Each call is self-contained: the previous step's spike
// list is pushed into the buckets first, the last step pushes nothing, and
// the last step's bitmap is compacted into list/count at the end.
nnx_status nnx::fused_run_steps(nnx_ctx* c, const nnx_step_plan* p, uint64_t first_step, int n_steps,
                                int use_graph) {
    const nnx_hh_state& s = p->state;
    NNX_CHECK(c, s.dtype == NNX_F32 && p->integrator == NNX_INT_RK4 && p->spike_rule == NNX_SPIKE_CROSSING,
              "WEIGHTED_FUSED needs an fp32 RK4 plan with crossing spikes");
    NNX_CHECK(c, s.v && s.m && s.h && s.ng && s.s && s.i_ext, "state needs v, m, h, ng, s, i_ext");
    nnx_ens_config f = kFusedDefault;
    if (p->fused_block) f.block = p->fused_block;
    if (p->fused_tile) f.tile = p->fused_tile;
    if (p->fused_grid) f.grid = p->fused_grid;
    NNX_TRY(check_cfg(c, &f));
    EnsHost h;
    uint64_t* const bits[2] = {p->bits[0], p->bits[1]};
    NNX_TRY(prepare(c, "fstep", s, p->params, p->graph, p->dt, bits, &h));
    h.k.burst = p->fused_burst > 0 ? p->fused_burst : -1;
    h.k.bulk = bulk_threshold(p->fused_bulk, s.n, p->graph.nnz, false);
    if (p->voxel_of) h.k.voxel_of = p->voxel_of, h.k.voxel_counts = p->voxel_counts;
    if (p->record) {
        h.k.raster = p->raster;
        h.k.raster_len = (unsigned long long*)p->raster_len;
        h.k.raster_cap = p->raster_cap;
    }
    const int64_t zero = 0;
    NNX_TRY(build_tiles(c, "fstep", s.n, 1, &zero, &s.n, f.tile, &h.k.tiles, &h.k.ntiles));
    const int q0 = (int)((first_step + 1) & 1);
    int32_t* flag_q0 = q0 ? h.k.flag1 : h.k.flag0;
    uint32_t* eb_q0 = q0 ? h.k.eb1 : h.k.eb0;
    // the step number is set outside the captured sequence (a captured
    // kernel argument would replay the first capture's number)
    k_ens_begin<<<1, 1, 0, c->stream>>>(h.k.stepctr, q0, (int64_t)first_step - 1, h.k.spk3, h.k.evmode, 1);
    NNX_LAUNCHED(c);
    NNX_TRY(readout_prepare(c, p));
    NNX_TRY(para_prepare(c, para_io(p)));
    auto body = [&]() -> nnx_status {
        k_push_bits<<<2048, 256, 0, c->stream>>>(p->list, p->count, p->graph.out_offsets, p->graph.out_targets,
                                                 p->graph.out_to_in, eb_q0, flag_q0);
        NNX_LAUNCHED(c);
        for (int i = 0; i < n_steps; ++i) {
            const uint64_t st = first_step + i;
            // print step (as in nnx_run_steps): the step counter slot of the
            // previous step holds st - 1
            if (p->readout_every > 0 && st % (uint64_t)p->readout_every == 0)
                NNX_TRY(readout_launch(c, p, h.k.stepctr + ((st + 1) & 1), 1, p->bits[st & 1]));
            if (p->hebbian.window > 0)
                NNX_TRY(hebbian_launch(c, &p->graph, &p->hebbian, h.k.stepctr + ((st + 1) & 1), 1, 0,
                                       p->bits[(st + 1) & 1]));
            // paracrine solve on its cadence (releases of the steps before st)
            if (p->paracrine.every > 0 && st % (uint64_t)p->paracrine.every == 0) NNX_TRY(para_launch(c, para_io(p)));
            NNX_TRY(ens_launch(c, h.k, h.hc, f, (int)(st & 1), i + 1 < n_steps));
        }
        CompactExtras ex{};
        return compact_launch(c, s.n, p->bits[(first_step + n_steps - 1) & 1], p->list, p->count, ex);
    };
    if (!use_graph) return body();
    uint64_t key = 0xcbf29ce484222325ull;
    const unsigned char* u = (const unsigned char*)p;
    for (size_t i = 0; i < sizeof(*p); ++i) key = (key ^ u[i]) * 1099511628211ull;
    key ^= 0x4655534544535450ull ^ phase_key(p, first_step);
    return capture_or_replay(c, {(const void*)(uintptr_t)key, n_steps * 2 + (int)(first_step & 1)}, body);
}

extern "C" {

nnx_status nnx_ensemble_run(nnx_ctx* c, const nnx_ens_plan* p, const nnx_ens_config* cfg, uint64_t first_step,
                            int n_steps, int use_graph) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(check_ens(c, p));
    const nnx_ens_config f = cfg ? *cfg : kDefaultCfg;
    NNX_TRY(check_cfg(c, &f));
    NNX_CHECK(c, n_steps >= 0, "n_steps < 0");
    if (n_steps == 0) return NNX_OK;
    EnsHost h;
    uint64_t* const bits[2] = {p->bits[0], p->bits[1]};
    NNX_TRY(prepare(c, "ens", p->state, p->params, p->graph, p->dt, bits, &h));
    h.k.voxel_of = p->voxel_of, h.k.voxel_counts = p->voxel_counts, h.k.net_spikes = p->net_spikes;
    h.k.bulk = bulk_threshold(f.bulk, p->state.n, p->graph.nnz, true);
    h.k.raster = p->raster, h.k.raster_len = (unsigned long long*)p->raster_len, h.k.raster_cap = p->raster_cap;
    NNX_TRY(build_tiles(c, "ens", p->state.n, p->nnet, p->net_begin, p->net_size, f.tile, &h.k.tiles,
                        &h.k.ntiles));
    const bool fresh = first_step == 0;
    if (fresh) {
        // a new run: no pending bucket entries, no previous spikes
        NNX_HIP(c, hipMemsetAsync(h.k.flag0, 0, (size_t)2 * p->state.n * 4, c->stream));
        NNX_HIP(c, hipMemsetAsync(h.k.eb0, 0, h.ebits_bytes, c->stream));
        NNX_HIP(c, hipMemsetAsync(p->bits[0], 0, (size_t)p->state.n / 8, c->stream));
        NNX_HIP(c, hipMemsetAsync(p->bits[1], 0, (size_t)p->state.n / 8, c->stream));
    }
    // a continuing run reads the buckets and burst counter the previous
    // call's last step left
    const int q0 = (int)((first_step + 1) & 1);
    k_ens_begin<<<1, 1, 0, c->stream>>>(h.k.stepctr, q0, (int64_t)first_step - 1, h.k.spk3, h.k.evmode,
                                         fresh ? 1 : 0);
    NNX_LAUNCHED(c);
    const int64_t every = p->paracrine.every;
    NNX_TRY(para_prepare(c, para_io(p)));  // buffers exist before any capture
    auto body = [&]() -> nnx_status {
        for (int i = 0; i < n_steps; ++i) {
            // the paracrine solve on its cadence (releases of the steps before this one)
            if (every > 0 && (first_step + i) % (uint64_t)every == 0) NNX_TRY(para_launch(c, para_io(p)));
            NNX_TRY(ens_launch(c, h.k, h.hc, f, (int)((first_step + i) & 1), 1));
        }
        return NNX_OK;
    };
    if (!use_graph) return body();
    uint64_t key = ens_key(p, f);
    if (every > 0) key ^= (first_step % (uint64_t)every + 1) * 0xc2b2ae3d27d4eb4full;
    return capture_or_replay(c, {(const void*)(uintptr_t)key, n_steps * 2 + (int)(first_step & 1)}, body);
}

nnx_status nnx_ensemble_autotune(nnx_ctx* c, const nnx_ens_plan* p, uint64_t first_step, int steps_per_candidate,
                                 nnx_ens_config* best, double* best_us, int* steps_run) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(check_ens(c, p));
    NNX_CHECK(c, best && steps_per_candidate >= 1, "need best and steps_per_candidate >= 1");
    NNX_CHECK(c, !c->capturing, "nnx_ensemble_autotune synchronizes: not under capture");
    // Every candidate runs the same window [first_step, first_step + steps)
    // from the same state: the state the steps change (neuron state, pending
    // events, spike bitmaps, step counters) is saved once and restored before
    // each run and after the search.  The candidates run on a copy of the
    // plan without the accumulating outputs (voxel deposit, spike counts,
    // raster) and without the paracrine solve, which the search therefore
    // leaves untouched.
    const int64_t n = p->state.n;
    EnsHost h;
    uint64_t* const bits[2] = {p->bits[0], p->bits[1]};
    NNX_TRY(prepare(c, "ens", p->state, p->params, p->graph, p->dt, bits, &h));
    struct Piece {
        void* p;
        size_t bytes;
    };
    const Piece pieces[] = {{p->state.v, (size_t)n * 4},   {p->state.m, (size_t)n * 4},
                            {p->state.h, (size_t)n * 4},   {p->state.ng, (size_t)n * 4},
                            {p->state.s, (size_t)n * 4},   {h.k.flag0, (size_t)2 * n * 4},
                            {h.k.eb0, h.ebits_bytes},      {h.k.stepctr, 64},
                            {p->bits[0], (size_t)n / 8},   {p->bits[1], (size_t)n / 8}};
    size_t total = 0;
    for (const Piece& q : pieces) total += (q.bytes + 255) & ~(size_t)255;
    void* snap;
    NNX_TRY(scratch(c, "ens_tune_snap", total, false, &snap));
    auto copy_state = [&](bool save) -> nnx_status {
        char* at = (char*)snap;
        for (const Piece& q : pieces) {
            NNX_HIP(c, hipMemcpyAsync(save ? at : q.p, save ? q.p : at, q.bytes, hipMemcpyDeviceToDevice,
                                      c->stream));
            at += (q.bytes + 255) & ~(size_t)255;
        }
        return NNX_OK;
    };
    NNX_TRY(copy_state(true));
    nnx_ens_plan tp = *p;
    tp.voxel_of = nullptr, tp.voxel_counts = nullptr, tp.net_spikes = nullptr;
    tp.raster = nullptr, tp.raster_len = nullptr, tp.raster_cap = 0;
    tp.paracrine.every = 0;
    hipEvent_t e0, e1;
    NNX_HIP(c, hipEventCreate(&e0));
    NNX_HIP(c, hipEventCreate(&e1));
    // the best of two timed runs, after an untimed one that uploads the
    // candidate's tile list and warms the caches
    auto time_cfg = [&](const nnx_ens_config& f, double* us) -> nnx_status {
        NNX_TRY(nnx_ensemble_run(c, &tp, &f, first_step, steps_per_candidate, 0));
        *us = 1e30;
        for (int rep = 0; rep < 2; ++rep) {
            NNX_TRY(copy_state(false));
            NNX_HIP(c, hipEventRecord(e0, c->stream));
            NNX_TRY(nnx_ensemble_run(c, &tp, &f, first_step, steps_per_candidate, 0));
            NNX_HIP(c, hipEventRecord(e1, c->stream));
            NNX_HIP(c, hipEventSynchronize(e1));
            float ms = 0;
            NNX_HIP(c, hipEventElapsedTime(&ms, e0, e1));
            *us = std::min(*us, 1e3 * ms / steps_per_candidate);
        }
        return copy_state(false);
    };
    nnx_ens_config cur = kDefaultCfg;
    double cur_us;
    nnx_status st = time_cfg(cur, &cur_us);
    const int blocks[] = {64, 128, 256, 512};
    const int tiles[] = {256, 512, 1024, 2048, 4096, 8192};
    const int grids[] = {256, 512, 1024, 2048, 4096, 8192};
    // coordinate search: block, then tile, then grid
    for (int dim = 0; dim < 3 && st == NNX_OK; ++dim) {
        const int* vals = dim == 0 ? blocks : dim == 1 ? tiles : grids;
        const int nv = dim == 0 ? 4 : 6;
        const nnx_ens_config base = cur;
        for (int i = 0; i < nv && st == NNX_OK; ++i) {
            nnx_ens_config f = base;
            (dim == 0 ? f.block : dim == 1 ? f.tile : f.grid) = vals[i];
            if (f.block == base.block && f.tile == base.tile && f.grid == base.grid) continue;
            if (f.tile < f.block) continue;
            double us;
            st = time_cfg(f, &us);
            if (st == NNX_OK && us < cur_us) {
                cur = f;
                cur_us = us;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    NNX_TRY(st);
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    *best = cur;
    if (best_us) *best_us = cur_us;
    if (steps_run) *steps_run = 0;
    return NNX_OK;
}

}  // extern "C"

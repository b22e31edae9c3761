// coupling.hip - spike detection, ascending spike list, deterministic spike
// propagation.
//
// Replaces AtomicCoupling::operator() (src/coupling.cu:111-178) and
// PerformCoupling (src_wholebrain/coupling.cu:43-69).  The reference flags
// firing neurons, compacts them with thrust::copy_if, reads the count back to
// the host and scatters with fp32 atomicAdd (order-nondeterministic).  Here:
//   - spikes are a bitmap (one ballot word per 64 neurons);
//   - the ascending list is built by one workgroup with a block scan, the
//     count stays on the device;
//   - every target's input is summed over its spiking sources in ascending
//     source order, so all modes are deterministic and WEIGHTED_PULL and
//     WEIGHTED_EVENT are bit-identical.
#include "coupling_dev.h"
#include "internal.h"

namespace nnx {


// ------------------------------------------------------------------------
// level rule (CheckFiringOperator, src/coupling.cu:103-109)
// ------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_threshold(int64_t n, const T* __restrict__ out, double thr,
                                                   uint64_t* __restrict__ bits, int64_t nwords) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    const bool f = i < n && (double)out[i] > thr;  // Float > 0.1 compares in double
    const uint64_t w = __ballot(f);
    if ((threadIdx.x & 63) == 0 && (i >> 6) < nwords) bits[i >> 6] = w;
}

// ------------------------------------------------------------------------
// compaction: one workgroup of 1024 threads walks the bitmap in tiles of
// 4096 words; each thread owns 4 consecutive words of a tile, a block scan
// of the per-thread popcounts gives every thread its output offset.
// Optionally appends (step, neuron) raster pairs and counts releases per
// voxel (integer atomics: order-independent).
// ------------------------------------------------------------------------
// Multi-block compaction: block b owns words [1024b, 1024b + 1024), four
// consecutive words per thread (65,536 neurons per block).  Its output offset
// is the popcount of every preceding word, summed by the block itself
// (integers: exact, no inter-block communication; 16-byte loads, four
// independent accumulators); a block scan orders its own words.  The last
// block finalises the count, the raster length and the step counter.  Raster
// length and step live in parity pairs: the launch of parity s reads slot s
// and its last block writes slot s^1, so no block reads a value that another
// block of the same launch updates.
constexpr int kCB = 256, kCW = 4, kCBW = kCB * kCW;

__device__ __forceinline__ int block_scan256(int v, int* total) {
    __shared__ int sh[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    __syncthreads();
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += sh[w];
    *total = sh[0] + sh[1] + sh[2] + sh[3];
    return before + x - v;
}

__global__ __launch_bounds__(kCB) void k_compact(int64_t nwords, const uint64_t* __restrict__ bits,
                                                 int32_t* __restrict__ list,
                                                 int32_t* __restrict__ count, CompactExtras ex) {
    const int64_t b0 = blockIdx.x * (int64_t)kCBW;  // first word of this block (< nwords)
    const bool a16 = ((uintptr_t)bits & 15) == 0;
    int p0 = 0, p1 = 0, p2 = 0, p3 = 0;
    if (a16) {
        // b0 is a multiple of 1024 words: pairs of words are 16-byte aligned
        const ulonglong2* B2 = reinterpret_cast<const ulonglong2*>(bits);
        const int64_t n2 = b0 / 2;
        int64_t q = threadIdx.x;
        for (; q + 3 * kCB < n2; q += 4 * kCB) {
            const ulonglong2 u0 = B2[q], u1 = B2[q + kCB], u2 = B2[q + 2 * kCB], u3 = B2[q + 3 * kCB];
            p0 += __popcll(u0.x) + __popcll(u0.y);
            p1 += __popcll(u1.x) + __popcll(u1.y);
            p2 += __popcll(u2.x) + __popcll(u2.y);
            p3 += __popcll(u3.x) + __popcll(u3.y);
        }
        for (; q < n2; q += kCB) {
            const ulonglong2 u = B2[q];
            p0 += __popcll(u.x) + __popcll(u.y);
        }
    } else {
        for (int64_t q = threadIdx.x; q < b0; q += kCB) p0 += __popcll(bits[q]);
    }
    int pre_tot;
    block_scan256((p0 + p1) + (p2 + p3), &pre_tot);
    const int64_t w0 = b0 + (int64_t)threadIdx.x * kCW;
    uint64_t w[kCW];
#pragma unroll
    for (int q = 0; q < kCW; ++q) w[q] = w0 + q < nwords ? bits[w0 + q] : 0ull;
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < kCW; ++q) cnt += __popcll(w[q]);
    int tot;
    const int off = block_scan256(cnt, &tot);
    const int par = ex.parity;
    const int64_t rbase = ex.raster ? ex.len2[par] : 0;
    const int32_t step = ex.step2 ? (int32_t)ex.step2[par] : 0;
    int64_t o = (int64_t)pre_tot + off;
    if (cnt) {
#pragma unroll
        for (int q = 0; q < kCW; ++q) {
            uint64_t x = w[q];
            while (x) {
                const int b = __ffsll((unsigned long long)x) - 1;
                x &= x - 1;
                const int32_t idx = (int32_t)((w0 + q) * 64 + b);
                list[o] = idx;
                if (ex.raster) {
                    const int64_t r = rbase + o;
                    if (r < ex.raster_cap) {
                        ex.raster[2 * r] = step;
                        ex.raster[2 * r + 1] = idx;
                    }
                }
                if (ex.voxel_of) atomicAdd(&ex.voxel_counts[ex.voxel_of[idx]], 1);
                ++o;
            }
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        const int64_t total = (int64_t)pre_tot + tot;
        *count = (int32_t)total;
        if (ex.raster) ex.len2[par ^ 1] = rbase + total;
        if (ex.step2) ex.step2[par ^ 1] = step + 1;
    }
}

// value of lane l (wave-uniform) of x: v_readlane, a scalar read of one lane
__device__ __forceinline__ int rdl(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ float rdl(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double rdl(double x, int l) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// ------------------------------------------------------------------------
// dense pull over the CSC: one wavefront per target.  Each lane reads four
// consecutive in-edges with one 16-byte load (a wavefront covers 256 edges
// per load, from the target's range rounded down to a multiple of four), tests
// their four spike bits, loads the values of the spiking ones, and the
// contributions are added in ascending edge (= source) order - lanes in
// order, the four edges of a lane in order - by every lane in lockstep.  The
// next target's range is loaded while the current one is summed.
// VEC = 0: in_src not 16-byte aligned, one edge per lane at a time.
// mode: 0 weighted, 2 count, 3 signal
// ------------------------------------------------------------------------
template <class T, bool VEC>
__global__ __launch_bounds__(256) void k_pull(int64_t n, int64_t nnz, int mode, const int32_t* __restrict__ in_off,
                                              const int32_t* __restrict__ in_src,
                                              const float* __restrict__ in_w,
                                              const uint64_t* __restrict__ bits,
                                              const T* __restrict__ outputs, T* __restrict__ g) {
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    int64_t t = blockIdx.x * 4ll + (threadIdx.x >> 6);
    int32_t e0 = t < n ? in_off[t] : 0, e1 = t < n ? in_off[t + 1] : 0;
    for (; t < n; t += nwaves) {
        const int64_t tn = t + nwaves;
        const int32_t ne0 = tn < n ? in_off[tn] : 0, ne1 = tn < n ? in_off[tn + 1] : 0;
        T acc = T(0);
        if (VEC) {
            for (int32_t base = e0 & ~3; base < e1; base += 256) {
                const int32_t q = base + 4 * lane;
                int32_t s4[4] = {-1, -1, -1, -1};
                if (q < e1) {
                    if ((int64_t)q + 3 < nnz) {
                        const int4 v = *reinterpret_cast<const int4*>(in_src + q);
                        s4[0] = v.x, s4[1] = v.y, s4[2] = v.z, s4[3] = v.w;
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if ((int64_t)q + j < nnz) s4[j] = in_src[q + j];
                    }
                }
                int am = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int32_t e = q + j;
                    if (e >= e0 && e < e1 && bit_of(bits, s4[j])) am |= 1 << j;
                }
                T v[4] = {T(0), T(0), T(0), T(0)};
                if (am) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if ((am >> j) & 1)
                            v[j] = mode == 0 ? T(in_w[q + j]) : mode == 2 ? T(1) : outputs[s4[j]];
                }
                uint64_t m = __ballot(am != 0);
                while (m) {
                    const int l = __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    const int ml = rdl(am, l);
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if ((ml >> j) & 1) acc = acc + rdl(v[j], l);
                }
            }
        } else {
            for (int32_t b = e0; b < e1; b += 64) {
                const int32_t e = b + lane;
                bool act = false;
                T val = T(0);
                if (e < e1) {
                    const int32_t src = in_src[e];
                    act = bit_of(bits, src);
                    if (act) val = mode == 0 ? T(in_w[e]) : mode == 2 ? T(1) : outputs[src];
                }
                uint64_t m = __ballot(act);
                while (m) {
                    const int l = __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    acc = acc + rdl(val, l);
                }
            }
        }
        if (lane == 0) g[t] = acc;
        e0 = ne0, e1 = ne1;
    }
}

// Networks of up to kLdsBitsMax neurons (and at least 8 in-edges): the same
// sums, with the spike bitmap copied into LDS first (128 KiB at 1,048,576
// neurons) so that the bit test of every in-edge is an LDS read instead of a
// random L2 transaction.  One workgroup of 1024 threads per CU.  A wavefront
// takes 64 consecutive targets, whose in-edges are one contiguous slab of
// the CSC, and streams the slab in chunks of 256 in-edges (a 16-byte load
// per lane).  Every load of the loop is unconditional - addresses past the
// slab are clamped to its last whole 16-byte group, and a weight (or output)
// load of a lane without a spiking edge reads entry 0 - so that the source
// loads of the next three chunks and the value loads of the last two stay in
// flight while the current chunk is tested and the chunk before last is
// summed.  The sums run over the spiking in-edges in ascending
// position, lanes in order and a lane's four edges in order, so each target
// is summed in ascending source order exactly as in k_pull; target j's sum
// is kept by lane j and the 64 sums are stored together.  Entries of the
// last partial 16-byte group of the CSC (nnz not a multiple of four) are read
// once at the start.
constexpr int64_t kLdsBitsMax = 1 << 20;

template <class T, int MODE>
__global__ __launch_bounds__(1024) void k_pull_lds(int64_t n, int64_t nnz, const int32_t* __restrict__ in_off,
                                                   const int32_t* __restrict__ in_src,
                                                   const float* __restrict__ in_w,
                                                   const uint64_t* __restrict__ bits,
                                                   const T* __restrict__ outputs, T* __restrict__ g) {
    // the value type each chunk loads: weights are float, outputs are T
    using V = typename std::conditional<MODE == 3, T, float>::type;
    __shared__ uint32_t sb[kLdsBitsMax / 32];
    // the bitmap copy: eight 8-byte loads per thread in flight at once
    const int64_t nw64 = (n + 63) >> 6;
#pragma unroll 8
    for (int64_t i = threadIdx.x; i < nw64; i += 1024) {
        const uint64_t w = bits[i];
        sb[2 * i] = (uint32_t)w, sb[2 * i + 1] = (uint32_t)(w >> 32);
    }
    const int32_t tail0 = (int32_t)(nnz & ~3ll), qmax = tail0 - 4;
    const int32_t tv0 = tail0 < nnz ? in_src[tail0] : 0, tv1 = tail0 + 1 < nnz ? in_src[tail0 + 1] : 0,
                  tv2 = tail0 + 2 < nnz ? in_src[tail0 + 2] : 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * 16;
    for (int64_t t0 = (blockIdx.x * 16ll + (threadIdx.x >> 6)) * 64; t0 < n; t0 += nwaves * 64) {
        const int cnt = (int)(n - t0 < 64 ? n - t0 : 64);
        // o1: end of lane j's target range (past the group: never reached)
        const int32_t o1 = lane < cnt ? in_off[t0 + lane + 1] : INT32_MAX;
        const int32_t E0 = in_off[t0], E1 = rdl(o1, cnt - 1), S = E0 & ~3;
        const int32_t nch = E1 > S ? (E1 - S + 255) >> 8 : 0;
        int32_t qlim = (E1 - 1) & ~3;
        qlim = qlim < qmax ? qlim : qmax;
        qlim = qlim > 0 ? qlim : 0;
        const bool tailgrp = E1 > tail0;
        auto load = [&](int32_t c) {
            int32_t q = S + c * 256 + 4 * lane;
            q = q < qlim ? q : qlim;
            q = q > 0 ? q : 0;
            return *reinterpret_cast<const int4*>(in_src + q);
        };
        // the loop starts at chunk -4 (no edges: its tests fail and its value
        // loads read entry 0), so that it enters its steady state - three
        // chunks of sources and two of values in flight - before chunk 0
        int4 ring[4];
        int msk[4];
        V wv[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) ring[u] = make_int4(0, 0, 0, 0), msk[u] = 0;
        T acc = T(0), res = T(0);
        int cur_t = 0;
        int32_t cur_end = INT32_MIN;
        for (int32_t c0 = -4; c0 < nch + 2; c0 += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int32_t c = c0 + u;
                ring[(u + 3) & 3] = load(c + 3);
                // chunk c: spike bits of its in-edges, then their values
                const int32_t q = S + c * 256 + 4 * lane;
                int32_t src[4] = {ring[u].x, ring[u].y, ring[u].z, ring[u].w};
                if (tailgrp) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int32_t e = q + k;
                        if (e >= tail0) src[k] = e == tail0 ? tv0 : e == tail0 + 1 ? tv1 : tv2;
                    }
                }
                // every source read here is an entry of the CSC, so the LDS
                // read is in range for the edges outside the slab too
                unsigned am = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int32_t e = q + k;
                    const unsigned in = (unsigned)(e >= E0) & (unsigned)(e < E1);
                    am |= (in & (sb[src[k] >> 5] >> (src[k] & 31))) << k;
                }
                msk[u] = (int)am;
                if (MODE != 2) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const bool f = (am >> k) & 1;
                        if (MODE == 0) wv[u][k] = in_w[f ? q + k : 0];
                        else wv[u][k] = outputs[f ? src[k] : 0];
                    }
                }
                // chunk c - 2: ordered sums
                const int ps = (u + 2) & 3;
                const int32_t pq = S + (c - 2) * 256;
                uint64_t m = __ballot(msk[ps] != 0);
                while (m) {
                    const int l = __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    const int ml = rdl(msk[ps], l);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        if ((ml >> k) & 1) {
                            const int32_t e = pq + 4 * l + k;
                            if (e >= cur_end) {  // the first spiking in-edge of another target
                                if (lane == cur_t) res = acc;
                                acc = T(0);
                                cur_t = __popcll(__ballot(o1 <= e));
                                cur_end = rdl(o1, cur_t);
                            }
                            acc = acc + (MODE == 2 ? T(1) : T(rdl(wv[ps][k], l)));
                        }
                    }
                }
            }
        }
        if (lane == cur_t) res = acc;
        if (lane < cnt) g[t0 + lane] = res;
    }
}

// ------------------------------------------------------------------------
// event-driven weighted coupling.
// push: for each spiking source (wavefront per source, lanes over its
// out-edges) reserve a slot in the target's bucket and store the CSC
// position of the edge; buckets live at the target's CSC range, so they can
// never overflow.  gather: each target sums the weights of its bucket in
// ascending CSC position, i.e. ascending source order - the same sequence of
// additions as the dense pull - and resets its count.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_push_event(const int32_t* __restrict__ list,
                                                    const int32_t* __restrict__ count,
                                                    const int32_t* __restrict__ out_off,
                                                    const int32_t* __restrict__ out_tgt,
                                                    const int32_t* __restrict__ out_to_in,
                                                    const int32_t* __restrict__ in_off,
                                                    int32_t* __restrict__ cnt,
                                                    int32_t* __restrict__ slot,
                                                    int32_t burst, int32_t* __restrict__ dense) {
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    const int32_t K = *count;
    // a burst step skips the push: the ephys kernel pulls every target
    const bool is_burst = burst >= 0 && K > burst;
    if (blockIdx.x == 0 && threadIdx.x == 0) *dense = is_burst;
    if (is_burst) return;
    for (int64_t k = blockIdx.x * 4ll + (threadIdx.x >> 6); k < K; k += nwaves) {
        const int32_t j = list[k];
        const int32_t e1 = out_off[j + 1];
        for (int32_t e = out_off[j] + lane; e < e1; e += 64) {
            const int32_t t = out_tgt[e];
            const int32_t pos = out_to_in[e];
            const int32_t q = atomicAdd(&cnt[t], 1);
            const int32_t b = in_off[t];
            if (q < in_off[t + 1] - b) slot[b + q] = pos;  // always true for a consistent CSR/CSC
        }
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_gather_event(int64_t n, int32_t* __restrict__ cnt,
                                                      const int32_t* __restrict__ slot,
                                                      const int32_t* __restrict__ in_off,
                                                      const int32_t* __restrict__ in_src,
                                                      const float* __restrict__ in_w,
                                                      const uint64_t* __restrict__ bits,
                                                      T* __restrict__ g) {
    const int64_t t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= n) return;
    const int32_t c = cnt[t];
    g[t] = gather_one<T>((int32_t)t, c, in_off, in_src, in_w, bits, slot);
    if (c) cnt[t] = 0;
}

}  // namespace nnx

using namespace nnx;

namespace nnx {

nnx_status compact_launch(nnx_ctx* c, int64_t n, const uint64_t* bits, int32_t* list,
                          int32_t* count, const CompactExtras& ex) {
    const int64_t nwords = ceil_div(n, 64);
    k_compact<<<(unsigned)std::max<int64_t>(1, ceil_div(nwords, kCBW)), kCB, 0, c->stream>>>(
        nwords, bits, list, count, ex);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

nnx_status event_scratch(nnx_ctx* c, int64_t n, int64_t nnz, int32_t** cnt, int32_t** slot) {
    void *a, *b;
    NNX_TRY(scratch(c, "event_cnt", (size_t)(n + 1) * 4, true, &a));
    NNX_TRY(scratch(c, "event_slot", (size_t)(nnz + 1) * 4, false, &b));
    void* dn;
    NNX_TRY(scratch(c, "event_dense", 256, true, &dn));  // burst flag
    *cnt = (int32_t*)a;
    *slot = (int32_t*)b;
    return NNX_OK;
}

nnx_status push_event_launch(nnx_ctx* c, const nnx_graph* g, const int32_t* list,
                             const int32_t* count, EventIn* ev, const uint64_t* prev_bits) {
    int32_t *cnt, *slot;
    NNX_TRY(event_scratch(c, g->n, g->nnz, &cnt, &slot));
    void* dn;
    NNX_TRY(scratch(c, "event_dense", 256, true, &dn));
    // above n/16 spikes a dense pull (~8 B per edge) beats pushing each
    // active edge through an atomic slot reservation
    const int32_t burst = (int32_t)std::max<int64_t>(1024, g->n / 16);
    k_push_event<<<2048, 256, 0, c->stream>>>(list, count, g->out_offsets, g->out_targets,
                                              g->out_to_in, g->in_offsets, cnt, slot, burst,
                                              (int32_t*)dn);
    NNX_LAUNCHED(c);
    ev->dense = (const int32_t*)dn;
    ev->cnt = cnt;
    ev->slot = slot;
    ev->in_off = g->in_offsets;
    ev->in_src = g->in_sources;
    ev->in_w = g->in_weights;
    ev->prev_bits = prev_bits;
    return NNX_OK;
}

nnx_status couple_launch(nnx_ctx* c, const nnx_graph* g, int mode, int dtype, const uint64_t* bits,
                         const int32_t* list, const int32_t* count, const void* outputs,
                         void* g_in) {
    const int64_t n = g->n;
    if (n == 0) return NNX_OK;
    if (mode == NNX_COUPLE_WEIGHTED_EVENT) {
        int32_t *cnt, *slot;
        NNX_TRY(event_scratch(c, n, g->nnz, &cnt, &slot));
        void* dn;
        NNX_TRY(scratch(c, "event_dense", 256, true, &dn));
        k_push_event<<<2048, 256, 0, c->stream>>>(list, count, g->out_offsets, g->out_targets,
                                                  g->out_to_in, g->in_offsets, cnt, slot, -1,
                                                  (int32_t*)dn);
        NNX_LAUNCHED(c);
        const unsigned blocks = (unsigned)ceil_div(n, 256);
        if (dtype == NNX_F32)
            k_gather_event<float><<<blocks, 256, 0, c->stream>>>(
                n, cnt, slot, g->in_offsets, g->in_sources, g->in_weights, bits, (float*)g_in);
        else
            k_gather_event<double><<<blocks, 256, 0, c->stream>>>(
                n, cnt, slot, g->in_offsets, g->in_sources, g->in_weights, bits, (double*)g_in);
        NNX_LAUNCHED(c);
        return NNX_OK;
    }
    const int m = mode == NNX_COUPLE_WEIGHTED_PULL ? 0 : mode == NNX_COUPLE_COUNT ? 2 : 3;
    const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 4), 65536);
    const bool vec = ((uintptr_t)g->in_sources & 15) == 0;
    if (vec && n <= kLdsBitsMax && g->nnz >= 8 && g->nnz < (1ll << 31) - 4096) {
        // one 1024-thread workgroup per CU (the bitmap copy takes 128 KiB of LDS)
        const unsigned nb = (unsigned)std::min<int64_t>(ceil_div(n, 1024), 256);
#define NNX_PULL_LDS(T, M)                                                                                  \
    k_pull_lds<T, M><<<nb, 1024, 0, c->stream>>>(n, g->nnz, g->in_offsets, g->in_sources, g->in_weights, bits, \
                                                 (const T*)outputs, (T*)g_in)
        if (dtype == NNX_F32) {
            if (m == 0) NNX_PULL_LDS(float, 0);
            else if (m == 2) NNX_PULL_LDS(float, 2);
            else NNX_PULL_LDS(float, 3);
        } else {
            if (m == 0) NNX_PULL_LDS(double, 0);
            else if (m == 2) NNX_PULL_LDS(double, 2);
            else NNX_PULL_LDS(double, 3);
        }
#undef NNX_PULL_LDS
        NNX_LAUNCHED(c);
        return NNX_OK;
    }
#define NNX_PULL(T, V)                                                                                    \
    k_pull<T, V><<<blocks, 256, 0, c->stream>>>(n, g->nnz, m, g->in_offsets, g->in_sources, g->in_weights, \
                                                bits, (const T*)outputs, (T*)g_in)
    if (dtype == NNX_F32) {
        if (vec) NNX_PULL(float, true);
        else NNX_PULL(float, false);
    } else {
        if (vec) NNX_PULL(double, true);
        else NNX_PULL(double, false);
    }
#undef NNX_PULL
    NNX_LAUNCHED(c);
    return NNX_OK;
}

nnx_status check_graph(nnx_ctx* c, const nnx_graph* g, int mode) {
    NNX_CHECK(c, g, "null graph");
    NNX_CHECK(c, g->n >= 0 && g->n < (int64_t)1 << 31, "graph n out of range");
    NNX_CHECK(c, g->nnz >= 0 && g->nnz < (int64_t)1 << 31, "graph nnz out of int32 range");
    NNX_CHECK(c, g->in_offsets && (g->in_sources || g->nnz == 0), "graph needs the CSC (in_offsets, in_sources)");
    const bool event = mode == NNX_COUPLE_WEIGHTED_EVENT || mode == NNX_COUPLE_WEIGHTED_FUSED;
    if (mode == NNX_COUPLE_WEIGHTED_PULL || event)
        NNX_CHECK(c, g->in_weights || g->nnz == 0, "weighted coupling needs in_weights");
    if (event)
        NNX_CHECK(c, g->out_offsets && (g->nnz == 0 || (g->out_targets && g->out_to_in)),
                  "event coupling needs out_offsets, out_targets and out_to_in");
    return NNX_OK;
}

}  // namespace nnx

extern "C" {

nnx_status nnx_spike_threshold(nnx_ctx* c, int dtype, int64_t n, const void* out, double thr,
                               uint64_t* bits) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, n >= 0 && (n == 0 || (out && bits)), "bad arguments");
    if (n == 0) return NNX_OK;
    const int64_t nwords = ceil_div(n, 64);
    const unsigned blocks = (unsigned)ceil_div(n, 256);
    if (dtype == NNX_F32)
        k_threshold<float><<<blocks, 256, 0, c->stream>>>(n, (const float*)out, thr, bits, nwords);
    else if (dtype == NNX_F64)
        k_threshold<double><<<blocks, 256, 0, c->stream>>>(n, (const double*)out, thr, bits, nwords);
    else
        return fail(c, NNX_ERR_ARG, "bad dtype");
    NNX_LAUNCHED(c);
    return NNX_OK;
}

nnx_status nnx_spike_compact(nnx_ctx* c, int64_t n, const uint64_t* bits, int32_t* list,
                             int32_t* count) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, n >= 0 && count && (n == 0 || (bits && list)), "bad arguments");
    CompactExtras ex{};
    return compact_launch(c, n, bits, list, count, ex);
}

nnx_status nnx_couple(nnx_ctx* c, const nnx_graph* g, int mode, int dtype, const uint64_t* bits,
                      const int32_t* list, const int32_t* count, const void* outputs, void* g_in) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(check_graph(c, g, mode));
    NNX_CHECK(c, mode >= NNX_COUPLE_WEIGHTED_PULL && mode <= NNX_COUPLE_SIGNAL, "bad mode %d", mode);
    NNX_CHECK(c, dtype == NNX_F32 || dtype == NNX_F64, "bad dtype");
    NNX_CHECK(c, g->n == 0 || (bits && g_in), "null bits or g_in");
    if (mode == NNX_COUPLE_WEIGHTED_EVENT) NNX_CHECK(c, list && count, "event coupling needs the spike list and count");
    if (mode == NNX_COUPLE_SIGNAL) NNX_CHECK(c, outputs, "signal coupling needs the outputs");
    return couple_launch(c, g, mode, dtype, bits, list, count, outputs, g_in);
}

}  // extern "C"

// ------------------------------------------------------------------------
// CSR -> CSC on the device (nnx_graph_transpose).
// 1. in-degree by integer atomics, 2. exclusive scan into in_offsets,
// 3. atomic fill (arbitrary order inside a target), 4. per-target rank sort
// by (source, CSR position) - keys are unique, so the result is the same
// deterministic layout the host builder produces.
// ------------------------------------------------------------------------
namespace nnx {

// flag[0] |= 1 if offsets are not non-decreasing from 0 to nnz, 2 if a
// target is outside [0, n)
__global__ void k_validate_csr(int64_t n, int64_t nnz, const int32_t* __restrict__ off,
                               const int32_t* __restrict__ tgt, int32_t* flag) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += stride)
        if (off[i + 1] < off[i]) atomicOr(flag, 1);
    if (blockIdx.x == 0 && threadIdx.x == 0 && (off[0] != 0 || off[n] != nnz)) atomicOr(flag, 1);
    for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < nnz; e += stride) {
        const int32_t t = tgt[e];
        if (t < 0 || t >= n) atomicOr(flag, 2);
    }
}

__global__ void k_indeg(int64_t nnz, const int32_t* __restrict__ tgt, int32_t* __restrict__ deg) {
    for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * 256)
        atomicAdd(&deg[tgt[e]], 1);
}

constexpr int kScanThreads = 1024, kScanItems = 4, kScanTile = kScanThreads * kScanItems;

// exclusive scan over a 1024-thread block (16 waves), total in *total
__device__ __forceinline__ int block_exclusive_scan(int v, int* sh_wave, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh_wave[wave] = x;
    __syncthreads();
    if (wave == 0) {
        int s = lane < kScanThreads / 64 ? sh_wave[lane] : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(s, o, 64);
            if (lane >= o) s += y;
        }
        if (lane < kScanThreads / 64) sh_wave[lane] = s;  // inclusive per wave
    }
    __syncthreads();
    const int before = wave ? sh_wave[wave - 1] : 0;
    *total = sh_wave[kScanThreads / 64 - 1];
    __syncthreads();
    return before + x - v;
}

// per-tile sums
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(int64_t n, const int32_t* __restrict__ in,
                                                             int32_t* __restrict__ tile_sum) {
    __shared__ int sh_wave[kScanThreads / 64];
    const int64_t base = blockIdx.x * (int64_t)kScanTile + threadIdx.x * kScanItems;
    int v = 0;
    for (int q = 0; q < kScanItems; ++q) v += base + q < n ? in[base + q] : 0;
    int tot;
    block_exclusive_scan(v, sh_wave, &tot);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}
// exclusive scan of the tile sums in one block (up to kScanTile tiles)
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(int64_t ntiles, int32_t* __restrict__ s) {
    __shared__ int sh_wave[kScanThreads / 64];
    int loc[kScanItems];
    int v = 0;
    for (int q = 0; q < kScanItems; ++q) {
        const int64_t i = threadIdx.x * (int64_t)kScanItems + q;
        loc[q] = i < ntiles ? s[i] : 0;
        v += loc[q];
    }
    int tot;
    int off = block_exclusive_scan(v, sh_wave, &tot);
    for (int q = 0; q < kScanItems; ++q) {
        const int64_t i = threadIdx.x * (int64_t)kScanItems + q;
        if (i < ntiles) s[i] = off;
        off += loc[q];
    }
}
// out[i] = exclusive prefix; out[n] = total
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(int64_t n, const int32_t* __restrict__ in,
                                                             const int32_t* __restrict__ tile_off,
                                                             int32_t* __restrict__ out) {
    __shared__ int sh_wave[kScanThreads / 64];
    const int64_t base = blockIdx.x * (int64_t)kScanTile + threadIdx.x * kScanItems;
    int loc[kScanItems];
    int v = 0;
    for (int q = 0; q < kScanItems; ++q) {
        loc[q] = base + q < n ? in[base + q] : 0;
        v += loc[q];
    }
    int tot;
    int off = block_exclusive_scan(v, sh_wave, &tot) + tile_off[blockIdx.x];
    for (int q = 0; q < kScanItems; ++q) {
        if (base + q < n) out[base + q] = off;
        off += loc[q];
        if (base + q == n - 1) out[n] = off;
    }
}

__global__ void k_fill_csc(int64_t n, const int32_t* __restrict__ off, const int32_t* __restrict__ tgt,
                           int32_t* __restrict__ cursor, int32_t* __restrict__ in_src,
                           int32_t* __restrict__ tmp_e) {
    for (int64_t s = blockIdx.x * 256ll + threadIdx.x; s < n; s += (int64_t)gridDim.x * 256)
        for (int32_t e = off[s]; e < off[s + 1]; ++e) {
            const int32_t p = atomicAdd(&cursor[tgt[e]], 1);
            in_src[p] = (int32_t)s;
            tmp_e[p] = e;
        }
}

// one wavefront per target: rank of each (src, e) key within the segment
__global__ __launch_bounds__(256) void k_sort_csc(int64_t n, const int32_t* __restrict__ in_off,
                                                  const int32_t* __restrict__ src_unsorted,
                                                  const int32_t* __restrict__ e_unsorted,
                                                  const float* __restrict__ w,
                                                  int32_t* __restrict__ in_src, float* __restrict__ in_w,
                                                  int32_t* __restrict__ out_to_in) {
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    for (int64_t t = blockIdx.x * 4ll + (threadIdx.x >> 6); t < n; t += nwaves) {
        const int32_t b = in_off[t], L = in_off[t + 1] - b;
        for (int32_t i = lane; i < L; i += 64) {
            const uint64_t ki = ((uint64_t)(uint32_t)src_unsorted[b + i] << 32) | (uint32_t)e_unsorted[b + i];
            int32_t rank = 0;
            for (int32_t j = 0; j < L; ++j) {
                const uint64_t kj = ((uint64_t)(uint32_t)src_unsorted[b + j] << 32) | (uint32_t)e_unsorted[b + j];
                rank += kj < ki;
            }
            const int32_t p = b + rank, e = (int32_t)(ki & 0xffffffffu);
            in_src[p] = (int32_t)(ki >> 32);
            if (in_w) in_w[p] = w ? w[e] : 0.0f;
            if (out_to_in) out_to_in[e] = p;
        }
    }
}

template <class T> __global__ void k_deg_out(int64_t n, const int32_t* __restrict__ in_off, T* __restrict__ d) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        d[i] = T(in_off[i + 1] - in_off[i]);
}

}  // namespace nnx

extern "C" nnx_status nnx_graph_transpose(nnx_ctx* c, int64_t n, int64_t nnz, const int32_t* off,
                                          const int32_t* tgt, const float* w, int32_t* in_off,
                                          int32_t* in_src, float* in_w, int32_t* out_to_in,
                                          int deg_dtype, void* in_degree) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, n >= 0 && n < ((int64_t)1 << 31) && nnz >= 0 && nnz < ((int64_t)1 << 31), "size out of range");
    NNX_CHECK(c, off && in_off && (nnz == 0 || (tgt && in_src)), "null argument");
    const int64_t ntiles = ceil_div(n, kScanTile);
    NNX_CHECK(c, ntiles <= kScanTile, "n too large for the single-level tile scan");
    void *deg, *tsum, *cur, *su, *eu;
    NNX_TRY(scratch(c, "tr_deg", (n + 1) * 4, false, &deg));
    NNX_TRY(scratch(c, "tr_tsum", (ntiles + 1) * 4, false, &tsum));
    NNX_TRY(scratch(c, "tr_cur", (n + 1) * 4, false, &cur));
    NNX_TRY(scratch(c, "tr_src", (nnz + 1) * 4, false, &su));
    NNX_TRY(scratch(c, "tr_e", (nnz + 1) * 4, false, &eu));
    {
        void* flag;
        NNX_TRY(scratch(c, "tr_flag", 256, true, &flag));
        NNX_TRY(fill_i32(c, (int32_t*)flag, 1, 0));
        const unsigned vb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(std::max(n, nnz), 256), 4096));
        k_validate_csr<<<vb, 256, 0, c->stream>>>(n, nnz, off, tgt, (int32_t*)flag);
        NNX_LAUNCHED(c);
        NNX_HIP(c, hipMemcpyAsync(c->pinned, flag, 4, hipMemcpyDeviceToHost, c->stream));
        NNX_HIP(c, hipStreamSynchronize(c->stream));
        const int32_t f = *(int32_t*)c->pinned;
        if (f & 1) return fail(c, NNX_ERR_ARG, "CSR offsets are not non-decreasing from 0 to nnz");
        if (f & 2) return fail(c, NNX_ERR_ARG, "CSR target index outside [0, n)");
    }
    NNX_TRY(fill_i32(c, (int32_t*)deg, n + 1, 0));
    const unsigned fb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nnz, 256), 8192));
    if (nnz) {
        k_indeg<<<fb, 256, 0, c->stream>>>(nnz, tgt, (int32_t*)deg);
        NNX_LAUNCHED(c);
    }
    if (n) {
        k_scan_tiles<<<(unsigned)ntiles, kScanThreads, 0, c->stream>>>(n, (int32_t*)deg, (int32_t*)tsum);
        k_scan_sums<<<1, kScanThreads, 0, c->stream>>>(ntiles, (int32_t*)tsum);
        k_scan_apply<<<(unsigned)ntiles, kScanThreads, 0, c->stream>>>(n, (int32_t*)deg, (int32_t*)tsum, in_off);
        NNX_LAUNCHED(c);
    } else {
        NNX_HIP(c, hipMemsetAsync(in_off, 0, 4, c->stream));
    }
    if (nnz) {
        NNX_HIP(c, hipMemcpyAsync(cur, in_off, n * 4, hipMemcpyDeviceToDevice, c->stream));
        const unsigned sb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 8192));
        k_fill_csc<<<sb, 256, 0, c->stream>>>(n, off, tgt, (int32_t*)cur, (int32_t*)su, (int32_t*)eu);
        NNX_LAUNCHED(c);
        const unsigned wb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 4), 65536));
        k_sort_csc<<<wb, 256, 0, c->stream>>>(n, in_off, (int32_t*)su, (int32_t*)eu, w, in_src, in_w, out_to_in);
        NNX_LAUNCHED(c);
    }
    if (in_degree && n) {
        const unsigned db = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 4096));
        if (deg_dtype == NNX_F64)
            k_deg_out<double><<<db, 256, 0, c->stream>>>(n, in_off, (double*)in_degree);
        else
            k_deg_out<float><<<db, 256, 0, c->stream>>>(n, in_off, (float*)in_degree);
        NNX_LAUNCHED(c);
    }
    return NNX_OK;
}

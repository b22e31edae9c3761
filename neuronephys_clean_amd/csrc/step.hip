// step.hip - the composite timestep and its hipGraph.
//
// One step in the reference's order (src/main.cu:166-181,
// src_wholebrain/mouse_cortex.cu:104-120): coupling on the previous step's
// spikes, release accumulation for the paracrine grid, the ephys update,
// then the ascending spike list (with raster and release counts folded into
// the same workgroup).  The reference synchronises the device after every
// launch and reads the firing count back each step; here a whole block of
// steps is enqueued (or replayed from a captured graph) without a host sync.
#include <cstring>

#include "internal.h"

namespace nnx {

// run bookkeeping, outside the captured steps: the parity pairs of the
// raster length and step counter are seeded from the caller's raster length
// and the first step, and the caller's raster length is updated at the end
__global__ void k_run_begin(int64_t* len2, int64_t* step2, const int64_t* raster_len, int par,
                            int64_t first) {
    len2[par] = raster_len ? *raster_len : 0;
    step2[par] = first;
}
__global__ void k_run_end(const int64_t* len2, int64_t* raster_len, int par) { *raster_len = len2[par]; }

static uint64_t plan_hash(const nnx_step_plan* p) {
    const unsigned char* b = (const unsigned char*)p;
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < sizeof(*p); ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

// The captured sequence fixes which steps print and which solve the
// paracrine field: graphs are keyed on the phase of both cadences.
uint64_t phase_key(const nnx_step_plan* p, uint64_t first_step) {
    uint64_t k = 0;
    if (p->readout_every > 0) k ^= (first_step % (uint64_t)p->readout_every) * 0x9e3779b97f4a7c15ull;
    if (p->paracrine.every > 0) k ^= (first_step % (uint64_t)p->paracrine.every + 1) * 0xc2b2ae3d27d4eb4full;
    return k;
}

static nnx_status check_plan(nnx_ctx* c, const nnx_step_plan* p) {
    NNX_CHECK(c, p, "null plan");
    const nnx_hh_state& s = p->state;
    NNX_CHECK(c, s.n == p->graph.n, "state n (%lld) != graph n (%lld)", (long long)s.n,
              (long long)p->graph.n);
    NNX_TRY(check_graph(c, &p->graph, p->couple_mode));
    NNX_CHECK(c, p->bits[0] && p->bits[1] && p->list && p->count, "plan needs bits[2], list, count");
    if (p->integrator == NNX_INT_RK4) {
        NNX_CHECK(c, p->couple_mode == NNX_COUPLE_WEIGHTED_PULL || p->couple_mode == NNX_COUPLE_WEIGHTED_EVENT ||
                         p->couple_mode == NNX_COUPLE_WEIGHTED_FUSED,
                  "RK4 plan needs weighted coupling");
        NNX_CHECK(c, p->g_in || p->couple_mode == NNX_COUPLE_WEIGHTED_FUSED, "RK4 plan needs g_in");
    } else {
        NNX_CHECK(c, s.in_local, "leapfrog plan writes the coupling into state.in_local");
        NNX_CHECK(c, p->couple_mode == NNX_COUPLE_COUNT || p->couple_mode == NNX_COUPLE_SIGNAL,
                  "leapfrog plan needs COUNT or SIGNAL coupling");
        if (p->couple_mode == NNX_COUPLE_SIGNAL) NNX_CHECK(c, s.out_local, "SIGNAL needs out_local");
    }
    if (p->record) NNX_CHECK(c, p->raster && p->raster_len && p->raster_cap >= 0, "record needs raster buffers");
    if (p->voxel_of) NNX_CHECK(c, p->voxel_counts, "voxel_of needs voxel_counts");
    if (p->params.noise_std != 0.0 && p->integrator != NNX_INT_LEAPFROG)
        return fail(c, NNX_ERR_UNSUPPORTED, "white noise (noise_std != 0) is drawn by the leapfrog step only");
    NNX_TRY(check_hebbian_plan(c, p));
    NNX_TRY(para_check(c, para_io(p)));
    return check_readout(c, p);
}

static nnx_status enqueue_step(nnx_ctx* c, const nnx_step_plan* p, uint64_t s, int64_t* ctr) {
    const uint64_t* prev = p->bits[(s + 1) & 1];
    uint64_t* cur = p->bits[s & 1];
    const bool rk4 = p->integrator == NNX_INT_RK4;
    // print step: the state entering step s; cur still holds the spikes of
    // step s-2, which the coupling of step s-1 read (the reference's
    // is_firing at its measurement); the row comes from the device step
    if (p->readout_every > 0 && s % (uint64_t)p->readout_every == 0)
        NNX_TRY(readout_launch(c, p, ctr + 2 + (s & 1), 0, cur));
    // Hebbian learning on the spikes this step's coupling reads
    if (p->hebbian.window > 0) NNX_TRY(hebbian_launch(c, &p->graph, &p->hebbian, ctr + 2 + (s & 1), 0, 0, prev));
    // paracrine solve on its cadence, from the releases of the steps before s
    if (p->paracrine.every > 0 && s % (uint64_t)p->paracrine.every == 0) NNX_TRY(para_launch(c, para_io(p)));
    if (rk4 && p->couple_mode == NNX_COUPLE_WEIGHTED_EVENT) {
        // push the previous step's spikes into the targets' buckets; the
        // ephys kernel sums each bucket in source order before its update
        EventIn ev;
        NNX_TRY(push_event_launch(c, &p->graph, p->list, p->count, &ev, prev));
        NNX_TRY(hh_rk4_event_launch(c, &p->state, &p->params, p->dt, ev, cur));
    } else {
        void* target = rk4 ? p->g_in : (void*)p->state.in_local;
        NNX_TRY(couple_launch(c, &p->graph, p->couple_mode, p->state.dtype, prev, p->list,
                              p->count, p->state.out_local, target));
        // the step number of the noise from the device counter (a captured
        // sequence is replayed at other steps)
        NNX_TRY(hh_step_launch(c, &p->state, &p->params, p->integrator, p->spike_rule, p->dt, s, ctr + 2 + (s & 1),
                               rk4 ? p->g_in : nullptr, cur));
    }
    CompactExtras ex{};
    if (p->record) {
        ex.raster = p->raster;
        ex.raster_cap = p->raster_cap;
    }
    ex.len2 = ctr;
    ex.step2 = ctr + 2;
    ex.parity = (int)(s & 1);
    ex.voxel_of = p->voxel_of;
    ex.voxel_counts = p->voxel_counts;
    return compact_launch(c, p->state.n, cur, p->list, p->count, ex);
}

}  // namespace nnx

using namespace nnx;

static nnx_status replay(nnx_ctx* c, nnx_step_plan* p, uint64_t first_step, int n_steps, int64_t* ctr);

extern "C" nnx_status nnx_run_steps(nnx_ctx* c, nnx_step_plan* p, uint64_t first_step, int n_steps,
                                    int use_graph) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(check_plan(c, p));
    NNX_CHECK(c, n_steps >= 0, "n_steps < 0");
    if (n_steps == 0 || p->state.n == 0) return NNX_OK;
    if (p->couple_mode == NNX_COUPLE_WEIGHTED_FUSED) return fused_run_steps(c, p, first_step, n_steps, use_graph);
    // everything a step touches is allocated before any capture
    if (p->couple_mode == NNX_COUPLE_WEIGHTED_EVENT) {
        int32_t *a, *b;
        NNX_TRY(event_scratch(c, p->graph.n, p->graph.nnz, &a, &b));
    }
    NNX_TRY(readout_prepare(c, p));
    NNX_TRY(para_prepare(c, para_io(p)));
    void* ctrp;
    NNX_TRY(scratch(c, "step_ctr", 64, true, &ctrp));
    int64_t* ctr = (int64_t*)ctrp;  // [0..1] raster length, [2..3] step, by parity
    const int par0 = (int)(first_step & 1), par1 = (int)((first_step + n_steps) & 1);
    k_run_begin<<<1, 1, 0, c->stream>>>(ctr, ctr + 2, p->record ? p->raster_len : nullptr, par0,
                                        (int64_t)first_step);
    NNX_LAUNCHED(c);
    if (!use_graph) {
        for (int i = 0; i < n_steps; ++i) NNX_TRY(enqueue_step(c, p, first_step + i, ctr));
    } else {
        NNX_TRY(replay(c, p, first_step, n_steps, ctr));
    }
    if (p->record) {
        k_run_end<<<1, 1, 0, c->stream>>>(ctr, p->raster_len, par1);
        NNX_LAUNCHED(c);
    }
    return NNX_OK;
}

static nnx_status replay(nnx_ctx* c, nnx_step_plan* p, uint64_t first_step, int n_steps, int64_t* ctr) {
    const GraphKey key((const void*)(uintptr_t)(plan_hash(p) ^ phase_key(p, first_step)),
                       n_steps * 2 + (int)(first_step & 1));
    return capture_or_replay(c, key, [&]() -> nnx_status {
        for (int i = 0; i < n_steps; ++i) NNX_TRY(enqueue_step(c, p, first_step + i, ctr));
        return NNX_OK;
    });
}

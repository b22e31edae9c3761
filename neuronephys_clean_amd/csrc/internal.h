// internal.h - context and shared helpers of libnnx.so (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>

#include "../../include/nnx.h"
#include "hh_math.h"

struct nnx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int owns_stream = 0;
    std::string err;
    // pinned host words for count / flag read-backs
    int64_t* pinned = nullptr;
    // named device scratch buffers, grown on demand, owned by the context
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
        uint64_t tag = 0;  // hash of host data last uploaded into it (0: none)
    };
    std::map<std::string, Buf> bufs;
    // captured step graphs keyed by (plan hash, n_steps and parity); they
    // embed scratch addresses, so growing any scratch buffer drops them all;
    // at most NNX_GRAPH_CACHE are kept (least recently launched evicted)
    struct Graph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        uint64_t last_use = 0;
    };
    std::map<std::pair<const void*, int>, Graph> graphs;
    uint64_t graph_clock = 0;
    int64_t graphs_instantiated = 0;
    int capturing = 0;
};

namespace nnx {

inline nnx_status fail(nnx_ctx* c, nnx_status s, const char* fmt, ...) {
    char msg[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof msg, fmt, ap);
    va_end(ap);
    if (c) c->err = msg;
    return s;
}

#define NNX_HIP(ctx, call)                                                                      \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return nnx::fail((ctx), NNX_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                               \
    } while (0)

#define NNX_LAUNCHED(ctx) NNX_HIP(ctx, hipGetLastError())

#define NNX_CHECK(ctx, cond, ...)                                  \
    do {                                                           \
        if (!(cond)) return nnx::fail((ctx), NNX_ERR_ARG, __VA_ARGS__); \
    } while (0)

#define NNX_TRY(expr)                      \
    do {                                   \
        nnx_status s_ = (expr);            \
        if (s_ != NNX_OK) return s_;       \
    } while (0)

// Device scratch buffer `name` of at least `bytes`; grows (never shrinks).
// New memory is zero-filled when `zero` is set.  Not callable while a graph
// is being captured if the buffer would have to grow.
nnx_status scratch(nnx_ctx* c, const char* name, size_t bytes, bool zero, void** out);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Destroys every cached graph (waits for the stream first).
nnx_status graph_cache_clear(nnx_ctx* c);

// Launches the graph cached under `key`, capturing `body` (which enqueues on
// c->stream) into a new one first if there is none; evicts the least
// recently launched graph beyond NNX_GRAPH_CACHE.
typedef std::pair<const void*, int> GraphKey;
template <class F> nnx_status capture_or_replay(nnx_ctx* c, GraphKey key, F body);
nnx_status graph_insert_launch(nnx_ctx* c, GraphKey key, hipGraph_t graph);
nnx_status graph_launch_cached(nnx_ctx* c, GraphKey key, bool* found);

template <class F> nnx_status capture_or_replay(nnx_ctx* c, GraphKey key, F body) {
    NNX_CHECK(c, c->stream != nullptr, "graph replay needs a non-null stream");
    bool found = false;
    NNX_TRY(graph_launch_cached(c, key, &found));
    if (found) return NNX_OK;
    hipGraph_t graph = nullptr;
    NNX_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = 1;
    const nnx_status st = body();
    c->capturing = 0;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (st != NNX_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return fail(c, NNX_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    return graph_insert_launch(c, key, graph);
}

template <class T> HHConst<T> make_const(const nnx_hh_params* q, double dt);

// fill kernel (graph-safe replacement for hipMemsetAsync)
nnx_status fill_i32(nnx_ctx* c, int32_t* p, int64_t n, int32_t v);

// optional work folded into the spike compaction
struct CompactExtras {
    int32_t* raster;      // (step, neuron) pairs
    int64_t* len2;        // raster length, parity pair (read [parity], write [parity^1])
    int64_t raster_cap;
    int64_t* step2;       // step number, parity pair (raster step numbers)
    int32_t parity;
    const int32_t* voxel_of;
    int32_t* voxel_counts;  // +1 per spike at voxel_of[neuron]
};

nnx_status compact_launch(nnx_ctx* c, int64_t n, const uint64_t* bits, int32_t* list,
                          int32_t* count, const CompactExtras& ex);
nnx_status couple_launch(nnx_ctx* c, const nnx_graph* g, int mode, int dtype, const uint64_t* bits,
                         const int32_t* list, const int32_t* count, const void* outputs,
                         void* g_in);
nnx_status check_graph(nnx_ctx* c, const nnx_graph* g, int mode);

// event-driven coupling folded into the ephys step
struct EventIn {
    int32_t* cnt;               // pending contributions per target (cleared by the step)
    const int32_t* slot;        // bucket entries: CSC positions
    const int32_t* in_off;
    const int32_t* in_src;
    const float* in_w;
    const uint64_t* prev_bits;  // spikes of the previous step (dense fallback)
    const int32_t* dense;       // set by the push when the step is a burst: pull every target
};
nnx_status push_event_launch(nnx_ctx* c, const nnx_graph* g, const int32_t* list,
                             const int32_t* count, EventIn* ev, const uint64_t* prev_bits);
nnx_status hh_rk4_event_launch(nnx_ctx* c, const nnx_hh_state* st, const nnx_hh_params* q,
                               double dt, const EventIn& ev, uint64_t* bits);
nnx_status event_scratch(nnx_ctx* c, int64_t n, int64_t nnz, int32_t** cnt, int32_t** slot);
// nnx_run_steps for WEIGHTED_FUSED plans (ensemble.hip); plan already checked
nnx_status fused_run_steps(nnx_ctx* c, const nnx_step_plan* p, uint64_t first_step, int n_steps,
                           int use_graph);

// electrode readout inside nnx_run_steps (readout.hip): plan fields checked,
// scratch allocated before a capture, and the launches of one print step
// (row = (*stepp + add) / readout_every, from the device step counter)
nnx_status check_readout(nnx_ctx* c, const nnx_step_plan* p);
nnx_status readout_prepare(nnx_ctx* c, const nnx_step_plan* p);
nnx_status readout_launch(nnx_ctx* c, const nnx_step_plan* p, const int64_t* stepp, int64_t add,
                          const uint64_t* bits);

// Hebbian learning inside nnx_run_steps (hebbian.hip): step = *stepp + add
// (stepp NULL: step)
nnx_status hebbian_launch(nnx_ctx* c, const nnx_graph* g, const nnx_hebbian* h, const int64_t* stepp, int64_t add,
                          int64_t step, const uint64_t* fire);
nnx_status check_hebbian_plan(nnx_ctx* c, const nnx_step_plan* p);

// Device memory of one CG working set (paracrine.hip): four n-vectors (r,
// p ping-pong, q), the Jacobi inverse diagonal (n floats, variable
// coefficients with Jacobi only), 64 scalars, 64 flags and the reduction
// partials (kCgPartDoubles).
struct CgMem {
    float *r, *pa, *pb, *q, *dinv;
    double* sc;
    int32_t* fl;
    double* part;
};
constexpr size_t kCgPartDoubles = (size_t)1024 * 16;
// bytes of a working set for grid g (256-byte aligned pieces) and its
// layout inside `base` (256-byte aligned, cg_mem_bytes long)
size_t cg_mem_bytes(const nnx_grid* g, int precond);
CgMem cg_mem_carve(void* base, const nnx_grid* g, int precond);
// argument checks and the launches of nnx_cg_solve_async on a working set
nnx_status cg_async_check(nnx_ctx* c, const nnx_grid* g, const float* b, const float* x, const nnx_cg_opts* o,
                          const double* d_result);
nnx_status cg_async_launch(nnx_ctx* c, const nnx_grid* g, const float* b, float* x, const nnx_cg_opts* o,
                           double* d_result, const CgMem& m);

// the working memory of nnx_cg_solve_async_noalloc (noalloc.hip): looks up
// the device address of its array on `device` (the current device) once
nnx_status noalloc_init(nnx_ctx* c, int device);

// nnx_hh_step with the step number read from the device (stepp, when not
// NULL) - the white noise of a captured step sequence (ephys.hip)
nnx_status hh_step_launch(nnx_ctx* c, const nnx_hh_state* st, const nnx_hh_params* q, int integrator, int rule,
                          double dt, uint64_t step, const int64_t* stepp, const void* g_in, uint64_t* bits);

// the paracrine solve inside nnx_run_steps (paracrine.hip): plan check,
// scratch set-up before a capture, and the launches of one solve
// the paracrine solve's inputs: the block, the release deposit, the neurons
struct ParaIo {
    const nnx_paracrine* q;
    const int32_t* voxel_of;
    int32_t* voxel_counts;
    int64_t n;
};
inline ParaIo para_io(const nnx_step_plan* p) { return {&p->paracrine, p->voxel_of, p->voxel_counts, p->state.n}; }
inline ParaIo para_io(const nnx_ens_plan* p) { return {&p->paracrine, p->voxel_of, p->voxel_counts, p->state.n}; }
nnx_status para_check(nnx_ctx* c, const ParaIo& io);
nnx_status para_prepare(nnx_ctx* c, const ParaIo& io);
nnx_status para_launch(nnx_ctx* c, const ParaIo& io);
// graph key component for the print and paracrine phases of first_step
uint64_t phase_key(const nnx_step_plan* p, uint64_t first_step);

}  // namespace nnx

// z = M(r) of the U-Net preconditioner (unet.hip); dims already checked
nnx_status nnx_unet_forward_internal(nnx_ctx* c, nnx_unet* u, int nx, int ny, int nz, int batch,
                                     const float* r, const float* feat, float* z);

// ctx.hip - context lifetime, errors, memory helpers, parameter presets.
#include <cmath>

#include "internal.h"

namespace nnx {

nnx_status scratch(nnx_ctx* c, const char* name, size_t bytes, bool zero, void** out) {
    nnx_ctx::Buf& b = c->bufs[name];
    if (b.bytes < bytes) {
        if (c->capturing)
            return fail(c, NNX_ERR_STATE, "scratch '%s' must grow to %zu bytes during graph capture",
                        name, bytes);
        if (b.p) {
            // cached graphs hold the old address: drop them (re-captured on use)
            NNX_TRY(graph_cache_clear(c));
            NNX_HIP(c, hipFree(b.p));
            b.p = nullptr;
            b.bytes = 0;
        }
        size_t alloc = bytes < 256 ? 256 : bytes;
        NNX_HIP(c, hipMalloc(&b.p, alloc));
        b.bytes = alloc;
        b.tag = 0;
        if (zero) NNX_HIP(c, hipMemsetAsync(b.p, 0, alloc, c->stream));
    }
    *out = b.p;
    return NNX_OK;
}

nnx_status graph_cache_clear(nnx_ctx* c) {
    if (c->graphs.empty()) return NNX_OK;
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& g : c->graphs) {
        if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
        if (g.second.graph) (void)hipGraphDestroy(g.second.graph);
    }
    c->graphs.clear();
    return NNX_OK;
}

nnx_status graph_launch_cached(nnx_ctx* c, GraphKey key, bool* found) {
    auto it = c->graphs.find(key);
    *found = it != c->graphs.end();
    if (!*found) return NNX_OK;
    it->second.last_use = ++c->graph_clock;
    NNX_HIP(c, hipGraphLaunch(it->second.exec, c->stream));
    return NNX_OK;
}

nnx_status graph_insert_launch(nnx_ctx* c, GraphKey key, hipGraph_t graph) {
    // evict the least recently launched graphs first (after the stream has
    // drained: an evicted graph may still be running)
    if ((int)c->graphs.size() >= NNX_GRAPH_CACHE) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return fail(c, NNX_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
        }
        while ((int)c->graphs.size() >= NNX_GRAPH_CACHE) {
            auto lru = c->graphs.begin();
            for (auto it = c->graphs.begin(); it != c->graphs.end(); ++it)
                if (it->second.last_use < lru->second.last_use) lru = it;
            if (lru->second.exec) (void)hipGraphExecDestroy(lru->second.exec);
            if (lru->second.graph) (void)hipGraphDestroy(lru->second.graph);
            c->graphs.erase(lru);
        }
    }
    nnx_ctx::Graph g;
    g.graph = graph;
    hipError_t e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(c, NNX_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    ++c->graphs_instantiated;
    g.last_use = ++c->graph_clock;
    c->graphs[key] = g;
    NNX_HIP(c, hipGraphLaunch(g.exec, c->stream));
    return NNX_OK;
}

template <class T> HHConst<T> make_const(const nnx_hh_params* q, double dt) {
    HHConst<T> p;
    p.C = T(q->C);
    p.invC = T(1.0 / q->C);
    p.gna = T(q->gna);
    p.gk = T(q->gk);
    p.gl = T(q->gl);
    p.ena = T(q->ena);
    p.ek = T(q->ek);
    p.el = T(q->el);
    p.vshift = T(q->v_shift);
    p.esyn = T(q->e_syn);
    p.tmax = T(q->tmax);
    p.vt = T(q->vt);
    p.kp = T(q->kp);
    p.ar = T(q->ar);
    p.ad = T(q->ad);
    p.g_local = T(q->g_local);
    p.e_local = T(q->e_local);
    p.g_i = T(q->g_i);
    p.e_i = T(q->e_i);
    p.g_e = T(q->g_e);
    p.e_e = T(q->e_e);
    p.noise_mean = T(q->noise_mean);
    p.noise_std = T(q->noise_std);
    p.v_thresh = T(q->v_thresh);
    p.out_thresh = q->out_thresh;
    p.dt = T(dt);
    p.hdt = T(dt / 2.0);
    p.dt6 = T(dt / 6.0);
    const double tau = q->tau_syn > 0 ? q->tau_syn : 1.0;
    p.eh = T(std::exp(-dt / (2.0 * tau)));
    p.ef = T(std::exp(-dt / tau));
    return p;
}
template HHConst<float> make_const<float>(const nnx_hh_params*, double);
template HHConst<double> make_const<double>(const nnx_hh_params*, double);

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

nnx_status fill_i32(nnx_ctx* c, int32_t* p, int64_t n, int32_t v) {
    if (n <= 0) return NNX_OK;
    int64_t blocks = ceil_div(n, 256);
    if (blocks > 4096) blocks = 4096;
    k_fill_i32<<<(unsigned)blocks, 256, 0, c->stream>>>(p, n, v);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

}  // namespace nnx

using namespace nnx;

extern "C" {

const char* nnx_version(void) { return "nnx 0.1 gfx950"; }

nnx_status nnx_create(nnx_ctx** out, int device, void* hip_stream) {
    if (!out) return NNX_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return NNX_ERR_HIP;
    if (device < 0 || device >= count) return NNX_ERR_ARG;
    nnx_ctx* c = new nnx_ctx();
    c->device = device;
    c->stream = (hipStream_t)hip_stream;
    if (hipSetDevice(device) != hipSuccess ||
        hipHostMalloc((void**)&c->pinned, 64 * sizeof(int64_t), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return NNX_ERR_HIP;
    }
    // the fixed working memory of the allocation-free entry point on this
    // device (a failed lookup is reported by that entry point)
    (void)noalloc_init(c, device);
    c->err.clear();
    // without a caller stream the context owns a blocking stream (ordered
    // with the legacy null stream, so work other libraries put there stays
    // ordered with ours)
    if (!c->stream) {
        if (hipStreamCreate(&c->stream) != hipSuccess) {
            (void)hipHostFree(c->pinned);
            delete c;
            return NNX_ERR_HIP;
        }
        c->owns_stream = 1;
    }
    *out = c;
    return NNX_OK;
}

nnx_status nnx_destroy(nnx_ctx* c) {
    if (!c) return NNX_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)graph_cache_clear(c);
    for (auto& b : c->bufs)
        if (b.second.p) (void)hipFree(b.second.p);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->owns_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return NNX_OK;
}

nnx_status nnx_graph_cache_clear(nnx_ctx* c) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, !c->capturing, "graph cache cleared during a capture");
    return graph_cache_clear(c);
}

nnx_status nnx_graph_cache_stats(const nnx_ctx* c, int* cached, int64_t* instantiated) {
    if (!c) return NNX_ERR_ARG;
    if (cached) *cached = (int)c->graphs.size();
    if (instantiated) *instantiated = c->graphs_instantiated;
    return NNX_OK;
}

const char* nnx_last_error(const nnx_ctx* c) { return c ? c->err.c_str() : "null context"; }

nnx_status nnx_set_stream(nnx_ctx* c, void* s) {
    if (!c || !s) return NNX_ERR_ARG;
    if (c->owns_stream) {
        NNX_HIP(c, hipStreamSynchronize(c->stream));
        (void)hipStreamDestroy(c->stream);
        c->owns_stream = 0;
    }
    c->stream = (hipStream_t)s;
    return NNX_OK;
}

nnx_status nnx_synchronize(nnx_ctx* c) {
    if (!c) return NNX_ERR_ARG;
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    return NNX_OK;
}

nnx_status nnx_malloc(nnx_ctx* c, void** p, size_t bytes) {
    if (!c || !p) return NNX_ERR_ARG;
    NNX_HIP(c, hipMalloc(p, bytes ? bytes : 1));
    return NNX_OK;
}
nnx_status nnx_free(nnx_ctx* c, void* p) {
    if (!c) return NNX_ERR_ARG;
    if (p) NNX_HIP(c, hipFree(p));
    return NNX_OK;
}
nnx_status nnx_memcpy_h2d(nnx_ctx* c, void* d, const void* h, size_t bytes) {
    if (!c) return NNX_ERR_ARG;
    NNX_HIP(c, hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    return NNX_OK;
}
nnx_status nnx_memcpy_d2h(nnx_ctx* c, void* h, const void* d, size_t bytes) {
    if (!c) return NNX_ERR_ARG;
    NNX_HIP(c, hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream));
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    return NNX_OK;
}
nnx_status nnx_memcpy_d2d(nnx_ctx* c, void* d, const void* s, size_t bytes) {
    if (!c) return NNX_ERR_ARG;
    NNX_HIP(c, hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, c->stream));
    return NNX_OK;
}
nnx_status nnx_memset(nnx_ctx* c, void* d, int v, size_t bytes) {
    if (!c) return NNX_ERR_ARG;
    NNX_HIP(c, hipMemsetAsync(d, v, bytes, c->stream));
    return NNX_OK;
}

nnx_status nnx_event_create(nnx_ctx* c, void** ev) {
    if (!c || !ev) return NNX_ERR_ARG;
    NNX_HIP(c, hipEventCreate((hipEvent_t*)ev));
    return NNX_OK;
}
nnx_status nnx_event_destroy(nnx_ctx* c, void* ev) {
    if (!c) return NNX_ERR_ARG;
    if (ev) NNX_HIP(c, hipEventDestroy((hipEvent_t)ev));
    return NNX_OK;
}
nnx_status nnx_event_record(nnx_ctx* c, void* ev) {
    if (!c || !ev) return NNX_ERR_ARG;
    NNX_HIP(c, hipEventRecord((hipEvent_t)ev, c->stream));
    return NNX_OK;
}
nnx_status nnx_event_elapsed(nnx_ctx* c, void* a, void* b, float* ms) {
    if (!c || !a || !b || !ms) return NNX_ERR_ARG;
    NNX_HIP(c, hipEventSynchronize((hipEvent_t)b));
    NNX_HIP(c, hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return NNX_OK;
}
void* nnx_get_stream(const nnx_ctx* c) { return c ? (void*)c->stream : nullptr; }

nnx_status nnx_read_count(nnx_ctx* c, const int32_t* d_count, int32_t* h_count) {
    if (!c || !d_count || !h_count) return NNX_ERR_ARG;
    NNX_HIP(c, hipMemcpyAsync(c->pinned, d_count, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    *h_count = *(int32_t*)c->pinned;
    return NNX_OK;
}

void nnx_hh_params_reference(nnx_hh_params* p) {
    // src/defines.cuh:87-102
    *p = nnx_hh_params{};
    p->C = 1;
    p->gna = 120;
    p->ena = 55;
    p->gk = 36;
    p->ek = -72;
    p->gl = 0.3;
    p->el = -50;
    p->v_shift = 0;
    p->tau_syn = 2;
    p->e_syn = 0;
    p->tmax = 1.0;
    p->vt = -20.0;
    p->kp = 3.0;
    p->ar = 5.0;
    p->ad = 0.18;
    p->g_local = 0.1;  // g_ampa
    p->e_local = 0.0;  // e_ampa
    p->g_i = 0.1;      // g_gaba
    p->e_i = -70.0;    // e_gaba
    p->g_e = 0.1;      // g_ampa
    p->e_e = 0.0;
    p->noise_mean = 3.0;
    p->noise_std = 1.5;
    p->noise_seed = 0;
    p->v_thresh = 0.0;
    p->out_thresh = 0.1;  // CheckFiringOperator, src/coupling.cu:107
}

void nnx_hh_params_squid(nnx_hh_params* p) {
    nnx_hh_params_reference(p);
    p->ena = 50;
    p->ek = -77;
    p->el = -54.4;
    p->v_shift = 5;
    p->tau_syn = 2;
    p->e_syn = 0;
    p->noise_mean = 0;
    p->noise_std = 0;
    p->v_thresh = 0;
}

}  // extern "C"

// noalloc.hip - nnx_cg_solve_async without any allocation, for callers
// that cannot allocate device memory (code under hipGraph capture, tight
// serving loops, hosts that forbid hipMalloc).
//
// The CG working set lives in `scratch`, a fixed array of this library's
// code object: every GPU the library runs on holds one copy of it until the
// process ends.  Its device address on each GPU is looked up once (when the
// first context of that GPU is created) and kept per device; host code only
// ever passes that address, never the array's name (which is a host-side
// stand-in).  Every call on one GPU uses the same array: it keeps the last
// call's contents, and calls that overlap on one GPU (two streams or two
// threads) race.  Nothing in it has to start at zero: the CG init writes
// every scalar, flag and the initial direction before they are read.
#include <mutex>

#include "internal.h"

// 48 MiB: a 128^3 grid with variable coefficients and Jacobi needs 40 MiB
// plus 132 KiB (five n-vectors and the reduction partials)
constexpr size_t kNoallocBytes = (size_t)48 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];         // device address of `scratch` on each GPU (nullptr: not looked up)
hipError_t g_err[kMaxDevices];    // the lookup's error, if it failed
}  // namespace

// Called by nnx_create with `device` current.  A failed lookup does not fail
// the context (no other entry point needs the array); it is reported by
// nnx_cg_solve_async_noalloc.
nnx_status nnx::noalloc_init(nnx_ctx* c, int device) {
    if (device < 0 || device >= kMaxDevices) return fail(c, NNX_ERR_ARG, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return NNX_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || !ptr) {
        g_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        return fail(c, NNX_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                    hipGetErrorString(g_err[device]));
    }
    g_ptr[device] = ptr;
    return NNX_OK;
}

extern "C" nnx_status nnx_cg_solve_async_noalloc(nnx_ctx* c, const nnx_grid* g, const float* b, float* x,
                                                 const nnx_cg_opts* o, double* d_result) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(nnx::cg_async_check(c, g, b, x, o, d_result));
    const size_t need = nnx::cg_mem_bytes(g, o->precond);
    if (need > kNoallocBytes)
        return nnx::fail(c, NNX_ERR_ARG,
                         "nnx_cg_solve_async_noalloc: a %dx%dx%d grid needs %zu bytes of working memory, "
                         "the fixed array holds %zu",
                         g->nx, g->ny, g->nz, need, kNoallocBytes);
    void* base = nullptr;
    hipError_t err = hipSuccess;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (c->device >= 0 && c->device < kMaxDevices) base = g_ptr[c->device], err = g_err[c->device];
    }
    if (!base)
        return nnx::fail(c, NNX_ERR_STATE, "no working memory on device %d: the address lookup %s", c->device,
                         err != hipSuccess ? hipGetErrorString(err) : "did not run");
    return nnx::cg_async_launch(c, g, b, x, o, d_result, nnx::cg_mem_carve(base, g, o->precond));
}

// paracrine.hip - grid operators, preconditioned CG with FP64 reductions,
// FP64 iterative refinement, the reference Crank-Nicolson update, release
// deposit and CSR SpMV.
//
// Replaces Density_on_grid_3d::update_density (src/paracrine.cu:63-134: FFT
// convolutions through ArrayFire, every scalar read back to the host each
// iteration) and the cuSPARSE SpMV of Scatter_grid_converter_3d
// (src/paracrine.cu:276-337).  Here the stencil is applied directly, every
// CG scalar (alpha, beta, norms, the stop flag) is produced on the device by
// the last workgroup of the reducing kernel (fixed summation order, so runs
// are reproducible), and the host only reads a flag every check_every
// iterations.
#include <chrono>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace nnx {

struct GridK {
    int32_t nx, ny, nz, kind;  // kind: 0 const 7-pt, 1 variable 7-pt, 2 27-pt
    float c, diag, a0, w;      // fp32 constants (cast once on the host)
    double cd, diagd, a0d, wd; // fp64 constants
    const float *kx, *ky, *kz;
};

static GridK make_grid(const nnx_grid* g) {
    GridK k;
    k.nx = g->nx;
    k.ny = g->ny;
    k.nz = g->nz;
    k.kind = g->stencil == NNX_STENCIL_27 ? 2 : (g->kx ? 1 : 0);
    k.cd = g->c;
    k.diagd = 1.0 + 6.0 * g->c;
    k.a0d = 1.0 + 0.5 * g->decay_dt;
    k.wd = g->c / 60.0;
    k.c = (float)k.cd;
    k.diag = (float)k.diagd;
    k.a0 = (float)k.a0d;
    k.w = (float)k.wd;
    k.kx = g->kx;
    k.ky = g->ky;
    k.kz = g->kz;
    return k;
}

template <class T> __device__ __forceinline__ T cst(const GridK& g, float f, double d);
template <> __device__ __forceinline__ float cst<float>(const GridK&, float f, double) { return f; }
template <> __device__ __forceinline__ double cst<double>(const GridK&, float, double d) { return d; }

template <class T>
__device__ __forceinline__ T at(const T* __restrict__ x, const GridK& g, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return T(0);
    return x[i + (int64_t)g.nx * (j + (int64_t)g.ny * k)];
}

// (A x)_ijk in T with the operation order of the CPU restatement.
template <class T, class U>
__device__ __forceinline__ T apply_at(const U* __restrict__ x, const GridK& g, int i, int j, int k) {
    const int64_t id = i + (int64_t)g.nx * (j + (int64_t)g.ny * k);
    const T xc = T(x[id]);
    if (g.kind == 0) {
        T s = T(at(x, g, i - 1, j, k)) + T(at(x, g, i + 1, j, k));
        s = s + T(at(x, g, i, j - 1, k));
        s = s + T(at(x, g, i, j + 1, k));
        s = s + T(at(x, g, i, j, k - 1));
        s = s + T(at(x, g, i, j, k + 1));
        return cst<T>(g, g.diag, g.diagd) * xc - cst<T>(g, g.c, g.cd) * s;
    }
    if (g.kind == 1) {
        const int64_t nx = g.nx, ny = g.ny;
        const int64_t fx = i + (nx + 1) * (j + ny * k), fy = i + nx * (j + (ny + 1) * k);
        const T kxm = T(g.kx[fx]), kxp = T(g.kx[fx + 1]);
        const T kym = T(g.ky[fy]), kyp = T(g.ky[fy + nx]);
        const T kzm = T(g.kz[id]), kzp = T(g.kz[id + nx * ny]);
        const T d = T(1) + kxm + kxp + kym + kyp + kzm + kzp;
        T s = kxm * T(at(x, g, i - 1, j, k)) + kxp * T(at(x, g, i + 1, j, k));
        s = s + kym * T(at(x, g, i, j - 1, k));
        s = s + kyp * T(at(x, g, i, j + 1, k));
        s = s + kzm * T(at(x, g, i, j, k - 1));
        s = s + kzp * T(at(x, g, i, j, k + 1));
        return d * xc - s;
    }
    T corners = T(0), edges = T(0), faces = T(0);
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                const int nzc = (di != 0) + (dj != 0) + (dk != 0);
                if (nzc == 0) continue;
                const T v = T(at(x, g, i + di, j + dj, k + dk));
                if (nzc == 3) corners = corners + v;
                else if (nzc == 2) edges = edges + v;
                else faces = faces + v;
            }
    const T s27 = ((corners + T(3) * edges) + T(14) * faces) - T(128) * xc;
    return cst<T>(g, g.a0, g.a0d) * xc - cst<T>(g, g.w, g.wd) * s27;
}

// 1/diag for Jacobi
__device__ __forceinline__ float dinv_at(const GridK& g, int64_t id) {
    if (g.kind == 0) return (float)(1.0 / g.diagd);
    if (g.kind == 2) return (float)(1.0 / (g.a0d + 128.0 * g.wd));
    const int64_t nx = g.nx, ny = g.ny;
    const int64_t i = id % nx, j = (id / nx) % ny, k = id / (nx * ny);
    const int64_t fx = i + (nx + 1) * (j + ny * k), fy = i + nx * (j + (ny + 1) * k);
    const double d = 1.0 + (double)g.kx[fx] + g.kx[fx + 1] + g.ky[fy] + g.ky[fy + nx] + g.kz[id] +
                     g.kz[id + nx * ny];
    return (float)(1.0 / d);
}

template <class T, class U>
__global__ __launch_bounds__(256) void k_apply(GridK g, const U* __restrict__ x, T* __restrict__ y) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    y[i + (int64_t)g.nx * (j + (int64_t)g.ny * k)] = apply_at<T, U>(x, g, i, j, k);
}

// ---------------------------------------------------------------------------
// deterministic block reductions and the last-block pattern
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sums NV doubles over a 256-thread block; result valid in thread 0
template <int NV> __device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[4][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) sh[wave][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = ((sh[0][q] + sh[1][q]) + sh[2][q]) + sh[3][q];
    __syncthreads();
}

__device__ __forceinline__ double load_coherent(const double* p) {
    unsigned long long b = __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(double, b);
}

// Every block deposits its NV partial sums; the last block to arrive sums all
// partials in block order and returns true with the totals in `tot` (all
// threads of that block).  The counter is reset for the next launch.
template <int NV>
__device__ __forceinline__ bool last_block(double (&v)[NV], double* partials, unsigned* counter,
                                           double (&tot)[NV]) {
    __shared__ int am_last;
    block_sum<NV>(v);
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partials[(size_t)b * NV + q] = v[q];
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        am_last = prev == nb - 1;
    }
    __syncthreads();
    if (!am_last) return false;
    __threadfence();
    double s[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] = 0;
    for (unsigned t = threadIdx.x; t < nb; t += 256)
#pragma unroll
        for (int q = 0; q < NV; ++q) s[q] += load_coherent(partials + (size_t)t * NV + q);
    block_sum<NV>(s);
    __shared__ double bc[NV];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) bc[q] = s[q];
        *counter = 0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) tot[q] = bc[q];
    return true;
}

// ---------------------------------------------------------------------------
// CG state on the device
// ---------------------------------------------------------------------------
enum { SC_RZ = 0, SC_RR, SC_PQ, SC_ALPHA, SC_BETA, SC_BB, SC_THR, SC_TRUE, SC_RR0, SC_N };
enum { FL_DONE = 0, FL_ITERS, FL_MAXIT, FL_ABS, FL_INITDONE, FL_N };

struct CgDev {
    float *r, *p, *q, *dinv;  // dinv NULL: constant (scalar dinv_c)
    float dinv_c;
    double* sc;
    int32_t* fl;
    double* partials;
    unsigned* counter;  // 4 slots
};

__device__ __forceinline__ bool cg_done(const CgDev& d) {
    return __hip_atomic_load(&d.fl[FL_DONE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}
__device__ __forceinline__ float dv(const CgDev& d, int64_t i) { return d.dinv ? d.dinv[i] : d.dinv_c; }

// init: r = b - A x (x0) or r = b with x = 0; p = 0; beta = 0; norms.
__global__ __launch_bounds__(256) void k_cg_init(GridK g, CgDev d, const float* __restrict__ b,
                                                 float* __restrict__ x, int use_x0, double tol,
                                                 int maxit, int abs_mode) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    double v[3] = {0, 0, 0};  // bb, rr, rz
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float bi = b[i];
        float ri = bi;
        if (use_x0) {
            const int ii = (int)(i % g.nx), jj = (int)((i / g.nx) % g.ny), kk = (int)(i / ((int64_t)g.nx * g.ny));
            ri = bi - apply_at<float, float>(x, g, ii, jj, kk);
        } else {
            x[i] = 0.0f;
        }
        d.r[i] = ri;
        d.p[i] = 0.0f;
        const float zi = ri * dv(d, i);
        v[0] += (double)bi * bi;
        v[1] += (double)ri * ri;
        v[2] += (double)ri * zi;
    }
    double t[3];
    if (last_block<3>(v, d.partials, d.counter + 0, t) && threadIdx.x == 0) {
        d.sc[SC_BB] = t[0];
        d.sc[SC_RR] = t[1];
        d.sc[SC_RR0] = t[1];
        d.sc[SC_RZ] = t[2];
        d.sc[SC_BETA] = 0.0;
        const double thr = abs_mode ? tol : tol * tol * t[0];
        d.sc[SC_THR] = thr;
        const int done = abs_mode ? (t[1] < thr) : (t[1] <= thr);
        d.fl[FL_DONE] = done || maxit <= 0;
        d.fl[FL_INITDONE] = done;
        d.fl[FL_ITERS] = 0;
        d.fl[FL_MAXIT] = maxit;
        d.fl[FL_ABS] = abs_mode;
    }
}

// p = z + beta p, z = dinv r
__global__ __launch_bounds__(256) void k_cg_dir(int64_t n, CgDev d) {
    if (cg_done(d)) return;
    const float beta = (float)d.sc[SC_BETA];
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        d.p[i] = d.r[i] * dv(d, i) + beta * d.p[i];
}

// q = A p (z-marching, 8 planes per thread for the constant 7-point stencil),
// pq partials, alpha = rz / pq by the last block.
constexpr int kZc = 8;
__global__ __launch_bounds__(256) void k_cg_spmv7(GridK g, CgDev d) {
    if (cg_done(d)) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k0 = blockIdx.z * kZc;
    const bool in = i < g.nx && j < g.ny;
    const int64_t sxy = (int64_t)g.nx * g.ny;
    const float* __restrict__ p = d.p;
    double v[1] = {0};
    int64_t id = i + (int64_t)g.nx * j + sxy * k0;
    float pm = (in && k0 > 0) ? p[id - sxy] : 0.0f;
    float pc = in ? p[id] : 0.0f;
    for (int kk = 0; kk < kZc; ++kk, id += sxy) {
        const int k = k0 + kk;
        if (k >= g.nz) break;
        const float pp = (in && k + 1 < g.nz) ? p[id + sxy] : 0.0f;
        if (in) {
            float s = (i > 0 ? p[id - 1] : 0.0f) + (i + 1 < g.nx ? p[id + 1] : 0.0f);
            s = s + (j > 0 ? p[id - g.nx] : 0.0f);
            s = s + (j + 1 < g.ny ? p[id + g.nx] : 0.0f);
            s = s + pm;
            s = s + pp;
            const float qv = g.diag * pc - g.c * s;
            d.q[id] = qv;
            v[0] += (double)pc * qv;
        }
        pm = pc;
        pc = pp;
    }
    double t[1];
    if (last_block<1>(v, d.partials, d.counter + 1, t) && threadIdx.x == 0) {
        d.sc[SC_PQ] = t[0];
        d.sc[SC_ALPHA] = d.sc[SC_RZ] / t[0];
    }
}

// generic operator (variable 7-point, 27-point), one voxel per thread
__global__ __launch_bounds__(256) void k_cg_spmv_gen(GridK g, CgDev d) {
    if (cg_done(d)) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    double v[1] = {0};
    if (i < g.nx && j < g.ny) {
        const int64_t id = i + (int64_t)g.nx * (j + (int64_t)g.ny * k);
        const float qv = apply_at<float, float>(d.p, g, i, j, k);
        d.q[id] = qv;
        v[0] = (double)d.p[id] * qv;
    }
    double t[1];
    if (last_block<1>(v, d.partials, d.counter + 1, t) && threadIdx.x == 0) {
        d.sc[SC_PQ] = t[0];
        d.sc[SC_ALPHA] = d.sc[SC_RZ] / t[0];
    }
}

// x += alpha p; r -= alpha q; rz, rr; beta; stop test (last block)
__global__ __launch_bounds__(256) void k_cg_update(int64_t n, CgDev d, float* __restrict__ x) {
    if (cg_done(d)) return;
    const float alpha = (float)d.sc[SC_ALPHA];
    double v[2] = {0, 0};
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        x[i] = x[i] + alpha * d.p[i];
        const float ri = d.r[i] - alpha * d.q[i];
        d.r[i] = ri;
        v[0] += (double)ri * (ri * dv(d, i));
        v[1] += (double)ri * ri;
    }
    double t[2];
    if (last_block<2>(v, d.partials, d.counter + 2, t) && threadIdx.x == 0) {
        d.sc[SC_BETA] = t[0] / d.sc[SC_RZ];
        d.sc[SC_RZ] = t[0];
        d.sc[SC_RR] = t[1];
        const int it = d.fl[FL_ITERS] + 1;
        d.fl[FL_ITERS] = it;
        const double thr = d.sc[SC_THR];
        const int conv = d.fl[FL_ABS] ? (t[1] < thr) : (t[1] <= thr);
        if (conv || it >= d.fl[FL_MAXIT]) d.fl[FL_DONE] = 1;
    }
}

__global__ void k_dinv(GridK g, float* dinv) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        dinv[i] = dinv_at(g, i);
}

// ||b - A x||^2 and ||b||^2 in FP64 (x float or double), into sc[SC_TRUE], sc[SC_BB]
template <class U, class B>
__global__ __launch_bounds__(256) void k_true_res(GridK g, const U* __restrict__ x,
                                                  const B* __restrict__ b, double* partials,
                                                  unsigned* counter, double* out2,
                                                  double* __restrict__ r_out) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    double v[2] = {0, 0};
    if (i < g.nx && j < g.ny) {
        const int64_t id = i + (int64_t)g.nx * (j + (int64_t)g.ny * k);
        const double ax = apply_at<double, U>(x, g, i, j, k);
        const double r = (double)b[id] - ax;
        if (r_out) r_out[id] = r;
        v[0] = r * r;
        v[1] = (double)b[id] * (double)b[id];
    }
    double t[2];
    if (last_block<2>(v, partials, counter, t) && threadIdx.x == 0) {
        out2[0] = t[0];
        out2[1] = t[1];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static dim3 grid3(const GridK& g) {
    return dim3((unsigned)ceil_div(g.nx, 64), (unsigned)ceil_div(g.ny, 4), (unsigned)g.nz);
}
static unsigned flat_blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 2048)); }

static nnx_status check_grid(nnx_ctx* c, const nnx_grid* g) {
    NNX_CHECK(c, g, "null grid");
    NNX_CHECK(c, g->nx > 0 && g->ny > 0 && g->nz > 0, "grid dims must be positive");
    NNX_CHECK(c, (int64_t)g->nx * g->ny * g->nz < ((int64_t)1 << 31), "grid too large");
    NNX_CHECK(c, g->stencil == NNX_STENCIL_7 || g->stencil == NNX_STENCIL_27, "bad stencil %d", g->stencil);
    NNX_CHECK(c, g->stencil == NNX_STENCIL_7 || !g->kx, "variable coefficients need STENCIL_7");
    NNX_CHECK(c, (!g->kx && !g->ky && !g->kz) || (g->kx && g->ky && g->kz), "kx, ky, kz all or none");
    return NNX_OK;
}

struct CgHost {
    CgDev d;
    double* out2;
    unsigned* cnt_true;
};

static nnx_status cg_buffers(nnx_ctx* c, const GridK& gk, int precond, CgHost* h) {
    const int64_t n = (int64_t)gk.nx * gk.ny * gk.nz;
    void *r, *p, *q, *sc, *fl, *part, *cnt;
    NNX_TRY(scratch(c, "cg_r", n * 4, false, &r));
    NNX_TRY(scratch(c, "cg_p", n * 4, false, &p));
    NNX_TRY(scratch(c, "cg_q", n * 4, false, &q));
    NNX_TRY(scratch(c, "cg_sc", 64 * 8, true, &sc));
    NNX_TRY(scratch(c, "cg_fl", 64 * 4, true, &fl));
    const int64_t nb = std::max<int64_t>((int64_t)ceil_div(gk.nx, 64) * ceil_div(gk.ny, 4) * gk.nz, 4096);
    NNX_TRY(scratch(c, "cg_part", (size_t)nb * 3 * 8, false, &part));
    NNX_TRY(scratch(c, "cg_cnt", 64, true, &cnt));
    h->d.r = (float*)r;
    h->d.p = (float*)p;
    h->d.q = (float*)q;
    h->d.sc = (double*)sc;
    h->d.fl = (int32_t*)fl;
    h->d.partials = (double*)part;
    h->d.counter = (unsigned*)cnt;
    h->d.dinv = nullptr;
    h->d.dinv_c = 1.0f;
    h->out2 = (double*)sc + 32;
    h->cnt_true = (unsigned*)cnt + 8;
    if (precond == NNX_PRECOND_JACOBI) {
        if (gk.kind == 1) {
            void* dv;
            NNX_TRY(scratch(c, "cg_dinv", n * 4, false, &dv));
            h->d.dinv = (float*)dv;
            k_dinv<<<flat_blocks(n), 256, 0, c->stream>>>(gk, h->d.dinv);
            NNX_LAUNCHED(c);
        } else {
            h->d.dinv_c = gk.kind == 0 ? (float)(1.0 / gk.diagd) : (float)(1.0 / (gk.a0d + 128.0 * gk.wd));
        }
    }
    return NNX_OK;
}

// Runs CG on the prepared buffers; x holds the guess when use_x0.
static nnx_status cg_run(nnx_ctx* c, const GridK& gk, CgHost& h, const float* b, float* x,
                         int use_x0, double tol, int maxit, int abs_mode, int check_every,
                         nnx_cg_result* res) {
    const int64_t n = (int64_t)gk.nx * gk.ny * gk.nz;
    const unsigned fb = flat_blocks(n);
    k_cg_init<<<fb, 256, 0, c->stream>>>(gk, h.d, b, x, use_x0, tol, maxit, abs_mode);
    NNX_LAUNCHED(c);
    const dim3 g7((unsigned)ceil_div(gk.nx, 64), (unsigned)ceil_div(gk.ny, 4), (unsigned)ceil_div(gk.nz, kZc));
    if (check_every < 1) check_every = 1;
    int iters = 0, done = 0;
    while (!done) {
        for (int q = 0; q < check_every; ++q) {
            k_cg_dir<<<fb, 256, 0, c->stream>>>(n, h.d);
            if (gk.kind == 0)
                k_cg_spmv7<<<g7, 256, 0, c->stream>>>(gk, h.d);
            else
                k_cg_spmv_gen<<<grid3(gk), 256, 0, c->stream>>>(gk, h.d);
            k_cg_update<<<fb, 256, 0, c->stream>>>(n, h.d, x);
        }
        NNX_LAUNCHED(c);
        NNX_HIP(c, hipMemcpyAsync(c->pinned, h.d.fl, 4 * FL_N, hipMemcpyDeviceToHost, c->stream));
        NNX_HIP(c, hipStreamSynchronize(c->stream));
        const int32_t* f = (const int32_t*)c->pinned;
        done = f[FL_DONE];
        iters = f[FL_ITERS];
        if (iters >= maxit) done = 1;
    }
    double sc[SC_N];
    NNX_HIP(c, hipMemcpyAsync(c->pinned, h.d.sc, 8 * SC_N, hipMemcpyDeviceToHost, c->stream));
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    std::memcpy(sc, c->pinned, 8 * SC_N);
    const int32_t initdone = 0;
    (void)initdone;
    if (res) {
        res->iters = iters;
        const double bb = sc[SC_BB] > 0 ? sc[SC_BB] : 1.0;
        res->relres = abs_mode ? sc[SC_RR] : std::sqrt(sc[SC_RR] / bb);
        res->converged = abs_mode ? (sc[SC_RR] < sc[SC_THR]) : (sc[SC_RR] <= sc[SC_THR]);
    }
    return NNX_OK;
}

static nnx_status true_residual(nnx_ctx* c, const GridK& gk, CgHost& h, const float* x,
                                const float* b, double* rel) {
    k_true_res<float, float><<<grid3(gk), 256, 0, c->stream>>>(gk, x, b, h.d.partials, h.cnt_true,
                                                              h.out2, nullptr);
    NNX_LAUNCHED(c);
    NNX_HIP(c, hipMemcpyAsync(c->pinned, h.out2, 16, hipMemcpyDeviceToHost, c->stream));
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    const double* o = (const double*)c->pinned;
    *rel = std::sqrt(o[0] / (o[1] > 0 ? o[1] : 1.0));
    return NNX_OK;
}

// ---- refinement kernels ----
__global__ void k_f64_to_f32_scaled(int64_t n, const double* __restrict__ r, double s,
                                    float* __restrict__ out) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        out[i] = (float)(r[i] * s);
}
__global__ void k_axpy_f64(int64_t n, double s, const float* __restrict__ d, double* __restrict__ x) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        x[i] = x[i] + s * (double)d[i];
}

// ---- reference CN update kernels (src/paracrine.cu:70-105) ----
__global__ __launch_bounds__(256) void k_ref_init(GridK g, const float* __restrict__ data,
                                                  float inv30h2, float c_xg, float d_xg, float c_b,
                                                  float d_b, float* __restrict__ xg,
                                                  float* __restrict__ bb) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int64_t id = i + (int64_t)g.nx * (j + (int64_t)g.ny * k);
    float corners = 0.f, edges = 0.f, faces = 0.f;
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                const int nzc = (di != 0) + (dj != 0) + (dk != 0);
                if (nzc == 0) continue;
                const float v = at(data, g, i + di, j + dj, k + dk);
                if (nzc == 3) corners = corners + v;
                else if (nzc == 2) edges = edges + v;
                else faces = faces + v;
            }
    const float u = data[id];
    const float lap = (((corners + 3.0f * edges) + 14.0f * faces) - 128.0f * u) * inv30h2;
    xg[id] = c_xg * u + d_xg * lap;
    bb[id] = c_b * u + d_b * lap;
}

// ---- deposit / rhs / spmv ----
__global__ void k_deposit(const int32_t* __restrict__ voxel_of, const int32_t* __restrict__ list,
                          const int32_t* __restrict__ count, int32_t* __restrict__ vc) {
    const int32_t K = *count;
    for (int64_t q = blockIdx.x * 256ll + threadIdx.x; q < K; q += (int64_t)gridDim.x * 256)
        atomicAdd(&vc[voxel_of[list[q]]], 1);
}
__global__ void k_rhs_counts(int64_t n, const float* __restrict__ cprev, int32_t* __restrict__ vc,
                             float amount, float* __restrict__ b) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int32_t k = vc[i];
        b[i] = (cprev ? cprev[i] : 0.0f) + amount * (float)k;
        if (k) vc[i] = 0;
    }
}
__global__ void k_spmv_csr(int64_t rows, const int32_t* __restrict__ off,
                           const int32_t* __restrict__ cols, const float* __restrict__ vals,
                           const float* __restrict__ x, float* __restrict__ y, int acc) {
    for (int64_t r = blockIdx.x * 256ll + threadIdx.x; r < rows; r += (int64_t)gridDim.x * 256) {
        float s = 0.0f;
        for (int32_t e = off[r]; e < off[r + 1]; ++e) s = s + vals[e] * x[cols[e]];
        y[r] = acc ? y[r] + s : s;
    }
}

}  // namespace nnx

using namespace nnx;

extern "C" {

nnx_status nnx_stencil_apply(nnx_ctx* c, const nnx_grid* g, const float* x, float* y) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(check_grid(c, g));
    NNX_CHECK(c, x && y && x != y, "x and y must be distinct device arrays");
    const GridK gk = make_grid(g);
    k_apply<float, float><<<grid3(gk), 256, 0, c->stream>>>(gk, x, y);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

nnx_status nnx_cg_solve(nnx_ctx* c, const nnx_grid* g, const float* b, float* x,
                        const nnx_cg_opts* o, nnx_cg_result* res) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(check_grid(c, g));
    NNX_CHECK(c, b && x && o, "null b, x or opts");
    NNX_CHECK(c, o->rtol >= 0 && o->maxit >= 0, "bad rtol/maxit");
    if (o->precond == NNX_PRECOND_LEARNED)
        return fail(c, NNX_ERR_UNSUPPORTED, "learned preconditioner: use nnx_fcg_solve");
    NNX_CHECK(c, o->precond == NNX_PRECOND_NONE || o->precond == NNX_PRECOND_JACOBI, "bad precond");
    const GridK gk = make_grid(g);
    auto t0 = std::chrono::steady_clock::now();
    CgHost h;
    NNX_TRY(cg_buffers(c, gk, o->precond, &h));
    nnx_cg_result r{};
    NNX_TRY(cg_run(c, gk, h, b, x, o->use_x0, o->rtol, o->maxit, 0, o->check_every, &r));
    auto t1 = std::chrono::steady_clock::now();
    NNX_TRY(true_residual(c, gk, h, x, b, &r.true_relres));
    r.ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    r.precond_used = o->precond;
    if (res) *res = r;
    return r.converged ? NNX_OK : fail(c, NNX_ERR_NOCONV, "CG: %d iterations, relres %.3e", r.iters, r.relres);
}

nnx_status nnx_cg_solve_refine(nnx_ctx* c, const nnx_grid* g, const double* b, double* x,
                               const nnx_cg_opts* o, double inner_rtol, int max_outer,
                               nnx_cg_result* res) {
    if (!c) return NNX_ERR_ARG;
    NNX_TRY(check_grid(c, g));
    NNX_CHECK(c, b && x && o, "null b, x or opts");
    const GridK gk = make_grid(g);
    const int64_t n = (int64_t)gk.nx * gk.ny * gk.nz;
    auto t0 = std::chrono::steady_clock::now();
    CgHost h;
    NNX_TRY(cg_buffers(c, gk, o->precond, &h));
    void *r64, *rhs, *d;
    NNX_TRY(scratch(c, "ir_r64", n * 8, false, &r64));
    NNX_TRY(scratch(c, "ir_rhs", n * 4, false, &rhs));
    NNX_TRY(scratch(c, "ir_d", n * 4, false, &d));
    if (!o->use_x0) NNX_HIP(c, hipMemsetAsync(x, 0, n * 8, c->stream));
    nnx_cg_result out{};
    int outer = 0, inner_total = 0;
    double rel = 0;
    for (;;) {
        k_true_res<double, double><<<grid3(gk), 256, 0, c->stream>>>(gk, x, b, h.d.partials, h.cnt_true,
                                                                    h.out2, (double*)r64);
        NNX_LAUNCHED(c);
        NNX_HIP(c, hipMemcpyAsync(c->pinned, h.out2, 16, hipMemcpyDeviceToHost, c->stream));
        NNX_HIP(c, hipStreamSynchronize(c->stream));
        const double rr = ((const double*)c->pinned)[0], bb = ((const double*)c->pinned)[1];
        rel = std::sqrt(rr / (bb > 0 ? bb : 1.0));
        if (rel <= o->rtol || outer >= max_outer || rr == 0) break;
        // inner solve on the residual scaled to unit norm
        const double s = 1.0 / std::sqrt(rr);
        k_f64_to_f32_scaled<<<flat_blocks(n), 256, 0, c->stream>>>(n, (const double*)r64, s, (float*)rhs);
        NNX_LAUNCHED(c);
        nnx_cg_result ir{};
        NNX_TRY(cg_run(c, gk, h, (const float*)rhs, (float*)d, 0, inner_rtol, o->maxit, 0,
                       o->check_every, &ir));
        inner_total += ir.iters;
        k_axpy_f64<<<flat_blocks(n), 256, 0, c->stream>>>(n, 1.0 / s, (const float*)d, x);
        NNX_LAUNCHED(c);
        ++outer;
    }
    auto t1 = std::chrono::steady_clock::now();
    out.iters = inner_total;
    out.outer = outer;
    out.relres = rel;
    out.true_relres = rel;
    out.converged = rel <= o->rtol;
    out.precond_used = o->precond;
    out.ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (res) *res = out;
    return out.converged ? NNX_OK : fail(c, NNX_ERR_NOCONV, "refinement: %d outer steps, relres %.3e", outer, rel);
}

nnx_status nnx_update_density_reference(nnx_ctx* c, int nx, int ny, int nz, double h, double D,
                                        double k, double dt, double eb, int max_iter, float* data,
                                        nnx_cg_result* res) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, data && h > 0, "bad arguments");
    nnx_grid g{nx, ny, nz, NNX_STENCIL_27, dt * D / (h * h), dt * k, nullptr, nullptr, nullptr};
    NNX_TRY(check_grid(c, &g));
    const GridK gk = make_grid(&g);
    const int64_t n = (int64_t)nx * ny * nz;
    auto t0 = std::chrono::steady_clock::now();
    CgHost hh;
    NNX_TRY(cg_buffers(c, gk, NNX_PRECOND_NONE, &hh));
    void *xg, *bb;
    NNX_TRY(scratch(c, "ref_xg", n * 4, false, &xg));
    NNX_TRY(scratch(c, "ref_b", n * 4, false, &bb));
    k_ref_init<<<grid3(gk), 256, 0, c->stream>>>(gk, data, (float)(1.0 / (h * h * 30.0)),
                                                 (float)(1.0 - dt * k), (float)(dt * D),
                                                 (float)(1.0 - h * k / 2.0), (float)(h * D / 2.0),
                                                 (float*)xg, (float*)bb);
    NNX_LAUNCHED(c);
    nnx_cg_result r{};
    NNX_TRY(cg_run(c, gk, hh, (const float*)bb, (float*)xg, 1, eb, max_iter, 1, 4, &r));
    NNX_HIP(c, hipMemcpyAsync(c->pinned, hh.d.fl, 4 * FL_N, hipMemcpyDeviceToHost, c->stream));
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    const int initdone = ((const int32_t*)c->pinned)[FL_INITDONE];
    // src/paracrine.cu:114-117: an initial residual below the bound returns
    // without writing the explicit guess; otherwise data = the iterate.
    if (!initdone) NNX_HIP(c, hipMemcpyAsync(data, xg, n * 4, hipMemcpyDeviceToDevice, c->stream));
    auto t1 = std::chrono::steady_clock::now();
    r.ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (res) *res = r;
    return NNX_OK;
}

nnx_status nnx_deposit_count(nnx_ctx* c, const int32_t* voxel_of, const int32_t* list,
                             const int32_t* count, int32_t* vc) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, voxel_of && list && count && vc, "null argument");
    k_deposit<<<256, 256, 0, c->stream>>>(voxel_of, list, count, vc);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

nnx_status nnx_rhs_from_counts(nnx_ctx* c, int64_t nvox, const float* cprev, int32_t* vc,
                               double amount, float* b) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, nvox >= 0 && vc && b, "null argument");
    if (!nvox) return NNX_OK;
    k_rhs_counts<<<flat_blocks(nvox), 256, 0, c->stream>>>(nvox, cprev, vc, (float)amount, b);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

nnx_status nnx_csr_spmv(nnx_ctx* c, int64_t rows, const int32_t* off, const int32_t* cols,
                        const float* vals, const float* x, float* y, int acc) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, rows >= 0 && off && cols && vals && x && y, "null argument");
    if (!rows) return NNX_OK;
    k_spmv_csr<<<flat_blocks(rows), 256, 0, c->stream>>>(rows, off, cols, vals, x, y, acc);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

}  // extern "C"

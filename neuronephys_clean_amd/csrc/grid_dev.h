// grid_dev.h - the paracrine grid operators shared by the solvers:
// constant 7-point, variable-coefficient 7-point and the reference's 27-point
// stencil, evaluated per voxel in a fixed operation order that the CPU
// restatement repeats.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nnx.h"

namespace nnx {

struct GridK {
    int32_t nx, ny, nz, kind;  // kind: 0 const 7-pt, 1 variable 7-pt, 2 27-pt
    int32_t block;             // kind 0: Dirichlet faces every `block` voxels (0: none)
    float c, diag, a0, w;      // fp32 constants (cast once on the host)
    double cd, diagd, a0d, wd; // fp64 constants
    const float *kx, *ky, *kz;
};

inline GridK make_grid(const nnx_grid* g) {
    GridK k;
    k.nx = g->nx;
    k.ny = g->ny;
    k.nz = g->nz;
    k.kind = g->stencil == NNX_STENCIL_27 ? 2 : (g->kx ? 1 : 0);
    k.block = g->block;
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
        if (g.block > 0) {
            // a neighbour across a block face is outside that block's grid
            const int B = g.block;
            const bool xm = i % B == 0, xp = (i + 1) % B == 0, ym = j % B == 0, yp = (j + 1) % B == 0,
                       zm = k % B == 0, zp = (k + 1) % B == 0;
            T s = (xm ? T(0) : T(at(x, g, i - 1, j, k))) + (xp ? T(0) : T(at(x, g, i + 1, j, k)));
            s = s + (ym ? T(0) : T(at(x, g, i, j - 1, k)));
            s = s + (yp ? T(0) : T(at(x, g, i, j + 1, k)));
            s = s + (zm ? T(0) : T(at(x, g, i, j, k - 1)));
            s = s + (zp ? T(0) : T(at(x, g, i, j, k + 1)));
            return cst<T>(g, g.diag, g.diagd) * xc - cst<T>(g, g.c, g.cd) * s;
        }
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

// Visits every voxel of an nx x ny x nz grid once: wavefronts take whole
// x-rows (row = j + ny k, dealt round-robin over the launch), lanes the
// voxels of a row, so (i, j, k) come without a division per voxel.
// f(i, j, k, id) with id = i + nx (j + ny k).
template <class F> __device__ __forceinline__ void for_rows(int nx, int ny, int nz, F&& f) {
    const int lane = threadIdx.x & 63;
    const int wpb = (int)(blockDim.x >> 6);
    const int rows = ny * nz;
    for (int row = (int)blockIdx.x * wpb + (int)(threadIdx.x >> 6); row < rows; row += (int)gridDim.x * wpb) {
        const int j = row % ny, k = row / ny;
        const int64_t base = (int64_t)row * nx;
        for (int i = lane; i < nx; i += 64) f(i, j, k, base + i);
    }
}

}  // namespace nnx

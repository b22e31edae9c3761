// mg.hip - geometric multigrid V-cycle for the variable-coefficient 7-point
// backward-Euler operator, the preconditioner of flexible CG for the
// variable-diffusivity paracrine solve and the base the learned
// preconditioner corrects.
//
// Cell-centred coarsening by 2 x 2 x 2 blocks while every dimension is even
// and at least 8.  A coarse face carries 1/16 of the sum of the 4 fine faces
// it covers: the operator rediscretised on the coarse grid (a face twice as
// wide and a distance twice as long: (4 h^2 / (2h)) / h^3 per unit volume
// of the 8 times larger cell), with the 4 parallel fine conductances added.
// Restriction is the block mean of the residual, prolongation trilinear
// between cell centres with zero beyond the boundary (the fine operator's
// own boundary treatment), smoothing nu sweeps before and after (3 by
// default) of a Chebyshev polynomial in D^-1 A on [1/8, 2] (2 bounds the
// spectrum of D^-1 A for these diagonally dominant operators) or of damped
// Jacobi, the coarsest level 16 damped-Jacobi sweeps (one workgroup when
// the level is small).  Block-mean restriction with trilinear prolongation
// is not a symmetric pair, so the cycle is used inside flexible CG.
//
// No counterpart in the reference, whose CG is unpreconditioned
// (src/paracrine.cu:106-129).
#include <algorithm>
#include <cstdio>

#include "mg.h"

namespace nnx {

namespace {

unsigned grid_blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }
// row kernels: enough workgroups for one wavefront per x-row on small levels
unsigned row_blocks(int ny, int nz) { return (unsigned)std::max(1, std::min((ny * nz + 3) / 4, 4096)); }
constexpr int kCoarseOneBlock = 4096;  // coarsest levels up to this size are solved by one workgroup

__global__ __launch_bounds__(256) void k_mg_dinv(GridK g, float* __restrict__ dinv) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dinv[i] = dinv_at(g, i);
}

// coarse face coefficients from the fine ones: x faces, then y, then z
__global__ __launch_bounds__(256) void k_mg_coarsen(GridK f, float* __restrict__ kx, float* __restrict__ ky,
                                                    float* __restrict__ kz, int NX, int NY, int NZ) {
    const int64_t nfx = (int64_t)(NX + 1) * NY * NZ, nfy = (int64_t)NX * (NY + 1) * NZ,
                  nfz = (int64_t)NX * NY * (NZ + 1);
    const int64_t nx = f.nx, ny = f.ny;
    for (int64_t q = blockIdx.x * 256ll + threadIdx.x; q < nfx + nfy + nfz; q += (int64_t)gridDim.x * 256) {
        float s;
        if (q < nfx) {
            const int64_t I = q % (NX + 1), J = (q / (NX + 1)) % NY, K = q / ((int64_t)(NX + 1) * NY);
            const float* a = f.kx + 2 * I + (nx + 1) * (2 * J + ny * 2 * K);
            s = (a[0] + a[nx + 1]) + (a[(nx + 1) * ny] + a[(nx + 1) * (ny + 1)]);
            kx[q] = s * (1.0f / 16.0f);
        } else if (q < nfx + nfy) {
            const int64_t r = q - nfx;
            const int64_t I = r % NX, J = (r / NX) % (NY + 1), K = r / ((int64_t)NX * (NY + 1));
            const float* a = f.ky + 2 * I + nx * (2 * J + (ny + 1) * 2 * K);
            s = (a[0] + a[1]) + (a[nx * (ny + 1)] + a[nx * (ny + 1) + 1]);
            ky[r] = s * (1.0f / 16.0f);
        } else {
            const int64_t r = q - nfx - nfy;
            const int64_t I = r % NX, J = (r / NX) % NY, K = r / ((int64_t)NX * NY);
            const float* a = f.kz + 2 * I + nx * (2 * J + ny * 2 * K);
            s = (a[0] + a[1]) + (a[nx] + a[nx + 1]);
            kz[r] = s * (1.0f / 16.0f);
        }
    }
}

// xout = zero ? w D^-1 b : xin + w D^-1 (b - A xin)
__global__ __launch_bounds__(256) void k_mg_smooth(GridK g, const float* __restrict__ dinv,
                                                   const float* __restrict__ b, const float* __restrict__ xin,
                                                   float* __restrict__ xout, float w, int zero) {
    if (zero) {
        const int64_t n = (int64_t)g.nx * g.ny * g.nz;
        for (int64_t id = blockIdx.x * 256ll + threadIdx.x; id < n; id += (int64_t)gridDim.x * 256)
            xout[id] = w * dinv[id] * b[id];
        return;
    }
    for_rows(g.nx, g.ny, g.nz, [&](int i, int j, int k, int64_t id) {
        xout[id] = xin[id] + w * dinv[id] * (b[id] - apply_at<float, float>(xin, g, i, j, k));
    });
}

// Chebyshev sweep: d = c1 d + c2 D^-1 (b - A xin), xout = xin + d; xin NULL
// means x = 0 (and c1 = 0 means d is not read)
__global__ __launch_bounds__(256) void k_mg_cheb(GridK g, const float* __restrict__ dinv, const float* __restrict__ b,
                                                 const float* __restrict__ xin, float* __restrict__ xout,
                                                 float* __restrict__ d, float c1, float c2) {
    if (!xin) {
        const int64_t n = (int64_t)g.nx * g.ny * g.nz;
        for (int64_t id = blockIdx.x * 256ll + threadIdx.x; id < n; id += (int64_t)gridDim.x * 256) {
            const float dn = c2 * dinv[id] * b[id];
            d[id] = dn;
            xout[id] = dn;
        }
        return;
    }
    for_rows(g.nx, g.ny, g.nz, [&](int i, int j, int k, int64_t id) {
        const float res = b[id] - apply_at<float, float>(xin, g, i, j, k);
        const float t = c2 * dinv[id] * res;
        const float dn = c1 != 0.0f ? c1 * d[id] + t : t;
        d[id] = dn;
        xout[id] = xin[id] + dn;
    });
}

// bc = block mean of (b - A x) over each coarse cell's 8 children
__global__ __launch_bounds__(256) void k_mg_restrict(GridK g, const float* __restrict__ b,
                                                     const float* __restrict__ x, float* __restrict__ bc, int NX,
                                                     int NY, int NZ) {
    const int64_t nx = g.nx, ny = g.ny;
    for_rows(NX, NY, NZ, [&](int I, int J, int K, int64_t q) {
        float s = 0.f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int i = 2 * I + (c & 1), j = 2 * J + ((c >> 1) & 1), k = 2 * K + (c >> 2);
            const int64_t id = i + nx * (j + ny * k);
            s = s + (b[id] - apply_at<float, float>(x, g, i, j, k));
        }
        bc[q] = 0.125f * s;
    });
}

// all `sweeps` damped-Jacobi sweeps of a small level from zero in one
// workgroup (x and t ping-pong, the result ends in x or t as returned by
// the host's parity bookkeeping: sweep s writes x for even s, t for odd s)
__global__ __launch_bounds__(1024) void k_mg_coarse(GridK g, const float* __restrict__ dinv,
                                                    const float* __restrict__ b, float* __restrict__ x,
                                                    float* __restrict__ t, float w, int sweeps) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int n = nx * ny * nz;
    for (int s = 0; s < sweeps; ++s) {
        const float* src = (s & 1) ? x : t;
        float* dst = (s & 1) ? t : x;
        for (int id = threadIdx.x; id < n; id += blockDim.x) {
            if (s == 0) {
                dst[id] = w * dinv[id] * b[id];
            } else {
                const int i = id % nx, j = (id / nx) % ny, k = id / (nx * ny);
                dst[id] = src[id] + w * dinv[id] * (b[id] - apply_at<float, float>(src, g, i, j, k));
            }
        }
        __syncthreads();
    }
}

// cell-centred linear weights along one axis: fine cell i sits at coarse
// coordinate i/2 - 1/4: 3/4 on its parent, 1/4 on the parent's neighbour
// toward it (none beyond the boundary)
__device__ __forceinline__ void lin_w(int i, int N, int& a, int& b, float& wb) {
    a = i >> 1;
    b = (i & 1) ? a + 1 : a - 1;
    wb = (b >= 0 && b < N) ? 0.25f : 0.0f;
    if (b < 0 || b >= N) b = a;
}

// x += P e (trilinear, zero outside)
__global__ __launch_bounds__(256) void k_mg_prolong_add(int nx, int ny, int nz, const float* __restrict__ e, int NX,
                                                        int NY, int NZ, float* __restrict__ x) {
    for_rows(nx, ny, nz, [&](int i, int j, int k, int64_t id) {
        int ia, ib, ja, jb, ka, kb;
        float wi, wj, wk;
        lin_w(i, NX, ia, ib, wi);
        lin_w(j, NY, ja, jb, wj);
        lin_w(k, NZ, ka, kb, wk);
        auto E = [&](int I, int J, int K) { return e[I + (int64_t)NX * (J + (int64_t)NY * K)]; };
        const float ai = 0.75f, aj = 0.75f, ak = 0.75f;
        const float v0 = ai * E(ia, ja, ka) + wi * E(ib, ja, ka);
        const float v1 = ai * E(ia, jb, ka) + wi * E(ib, jb, ka);
        const float v2 = ai * E(ia, ja, kb) + wi * E(ib, ja, kb);
        const float v3 = ai * E(ia, jb, kb) + wi * E(ib, jb, kb);
        const float u0 = aj * v0 + wj * v1;
        const float u1 = aj * v2 + wj * v3;
        x[id] = x[id] + (ak * u0 + wk * u1);
    });
}

}  // namespace

nnx_status mg_setup(nnx_ctx* c, const GridK& g0, int nu, int smoother, Mg* m) {
    NNX_CHECK(c, g0.kind == 1, "multigrid needs the variable-coefficient 7-point operator");
    NNX_CHECK(c, nu >= 0 && nu <= 16, "nu must be in [0, 16]");
    NNX_CHECK(c, smoother == NNX_MG_CHEBYSHEV || smoother == NNX_MG_JACOBI, "unknown smoother");
    m->nu = nu > 0 ? nu : 3;
    m->smoother = smoother;
    // Chebyshev polynomial on [lmax/16, lmax] with lmax = 2, the bound of
    // D^-1 A for these diagonally dominant operators (Gershgorin)
    {
        const double lmax = 2.0, lmin = lmax / 16.0;
        const double th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin), sig = th / de;
        double rho = 1.0 / sig;
        m->c1[0] = 0.0f;
        m->c2[0] = (float)(1.0 / th);
        for (int k = 1; k < 16; ++k) {
            const double r1 = 1.0 / (2.0 * sig - rho);
            m->c1[k] = (float)(r1 * rho);
            m->c2[k] = (float)(2.0 * r1 / de);
            rho = r1;
        }
    }
    m->g[0] = g0;
    int l = 0;
    while (l + 1 < kMgMaxLevels) {
        const GridK& f = m->g[l];
        if ((f.nx & 1) || (f.ny & 1) || (f.nz & 1) || std::min(f.nx, std::min(f.ny, f.nz)) < 8) break;
        GridK& g = m->g[l + 1];
        g = f;
        g.nx = f.nx / 2, g.ny = f.ny / 2, g.nz = f.nz / 2;
        const int64_t nfx = (int64_t)(g.nx + 1) * g.ny * g.nz, nfy = (int64_t)g.nx * (g.ny + 1) * g.nz,
                      nfz = (int64_t)g.nx * g.ny * (g.nz + 1);
        char name[32];
        void* p;
        std::snprintf(name, sizeof name, "mg_k%d", l + 1);
        NNX_TRY(scratch(c, name, (size_t)(nfx + nfy + nfz) * 4, false, &p));
        g.kx = (float*)p, g.ky = (float*)p + nfx, g.kz = (float*)p + nfx + nfy;
        ++l;
    }
    m->nlev = l + 1;
    for (int q = 0; q < m->nlev; ++q) {
        const int64_t n = (int64_t)m->g[q].nx * m->g[q].ny * m->g[q].nz;
        char name[32];
        void* p;
        std::snprintf(name, sizeof name, "mg_v%d", q);
        NNX_TRY(scratch(c, name, (size_t)(q == 0 ? 3 : 5) * n * 4, false, &p));
        float* f = (float*)p;
        m->dinv[q] = f;
        m->t[q] = f + n;
        m->d[q] = f + 2 * n;
        m->x[q] = q ? f + 3 * n : nullptr;
        m->b[q] = q ? f + 4 * n : nullptr;
    }
    for (int q = 1; q < m->nlev; ++q) {
        const GridK& f = m->g[q - 1];
        GridK& g = m->g[q];
        const int64_t nf = (int64_t)(g.nx + 1) * g.ny * g.nz + (int64_t)g.nx * (g.ny + 1) * g.nz +
                           (int64_t)g.nx * g.ny * (g.nz + 1);
        k_mg_coarsen<<<grid_blocks(nf), 256, 0, c->stream>>>(f, (float*)g.kx, (float*)g.ky, (float*)g.kz, g.nx,
                                                             g.ny, g.nz);
        NNX_LAUNCHED(c);
    }
    for (int q = 0; q < m->nlev; ++q) {
        const int64_t n = (int64_t)m->g[q].nx * m->g[q].ny * m->g[q].nz;
        k_mg_dinv<<<grid_blocks(n), 256, 0, c->stream>>>(m->g[q], m->dinv[q]);
        NNX_LAUNCHED(c);
    }
    return NNX_OK;
}

nnx_status mg_vcycle(nnx_ctx* c, Mg& m, const float* r, float* z) {
    const float* B[kMgMaxLevels];
    float* cur[kMgMaxLevels];
    auto n_of = [&](int l) { return (int64_t)m.g[l].nx * m.g[l].ny * m.g[l].nz; };
    // one smoothing sweep number s of a phase: xin NULL means x = 0
    auto sweep = [&](int l, int s, const float* xin, float* xout, bool jacobi) -> nnx_status {
        const unsigned blocks = xin ? row_blocks(m.g[l].ny, m.g[l].nz) : grid_blocks(n_of(l));
        if (jacobi)
            k_mg_smooth<<<blocks, 256, 0, c->stream>>>(m.g[l], m.dinv[l], B[l], xin, xout, m.w, xin == nullptr);
        else
            k_mg_cheb<<<blocks, 256, 0, c->stream>>>(m.g[l], m.dinv[l], B[l], xin, xout, m.d[l], m.c1[s], m.c2[s]);
        NNX_LAUNCHED(c);
        return NNX_OK;
    };
    // sweeps from zero: the first writes `first`, later ones ping-pong
    auto sweeps = [&](int l, int count, float* first, float* other, bool jacobi) -> nnx_status {
        float *src = nullptr, *dst = first;
        for (int s = 0; s < count; ++s) {
            NNX_TRY(sweep(l, s, src, dst, jacobi));
            src = dst;
            dst = (dst == first) ? other : first;
        }
        cur[l] = src;
        return NNX_OK;
    };
    // sweeps from the current iterate
    auto more = [&](int l, int count) -> nnx_status {
        float* other = (cur[l] == m.t[l]) ? (l ? m.x[l] : z) : m.t[l];
        for (int s = 0; s < count; ++s) {
            NNX_TRY(sweep(l, s, cur[l], other, m.smoother == NNX_MG_JACOBI));
            std::swap(cur[l], other);
        }
        return NNX_OK;
    };
    B[0] = r;
    for (int l = 1; l < m.nlev; ++l) B[l] = m.b[l];
    const int L = m.nlev - 1;
    // down: at level 0 start in t so that 2 nu - 1 swaps end in z
    for (int l = 0; l < L; ++l) {
        NNX_TRY(sweeps(l, m.nu, m.t[l], l ? m.x[l] : z, m.smoother == NNX_MG_JACOBI));
        const GridK& g = m.g[l];
        k_mg_restrict<<<row_blocks(m.g[l + 1].ny, m.g[l + 1].nz), 256, 0, c->stream>>>(
            g, B[l], cur[l], m.b[l + 1], m.g[l + 1].nx, m.g[l + 1].ny, m.g[l + 1].nz);
        NNX_LAUNCHED(c);
    }
    if (L > 0 && n_of(L) <= kCoarseOneBlock) {
        k_mg_coarse<<<1, 1024, 0, c->stream>>>(m.g[L], m.dinv[L], B[L], m.x[L], m.t[L], m.w, m.coarse_sweeps);
        NNX_LAUNCHED(c);
        cur[L] = (m.coarse_sweeps & 1) ? m.x[L] : m.t[L];
    } else {
        NNX_TRY(sweeps(L, m.coarse_sweeps, m.x[L] ? m.x[L] : z, m.t[L], true));
    }
    // up
    for (int l = L - 1; l >= 0; --l) {
        const GridK& g = m.g[l];
        k_mg_prolong_add<<<row_blocks(g.ny, g.nz), 256, 0, c->stream>>>(g.nx, g.ny, g.nz, cur[l + 1], m.g[l + 1].nx,
                                                                      m.g[l + 1].ny, m.g[l + 1].nz, cur[l]);
        NNX_LAUNCHED(c);
        NNX_TRY(more(l, m.nu));
    }
    if (cur[0] != z) NNX_HIP(c, hipMemcpyAsync(z, cur[0], n_of(0) * 4, hipMemcpyDeviceToDevice, c->stream));
    return NNX_OK;
}

}  // namespace nnx

using namespace nnx;

extern "C" nnx_status nnx_mg_apply(nnx_ctx* c, const nnx_grid* g, int nu, int smoother, const float* d_r,
                                   float* d_z) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, g && d_r && d_z, "null argument");
    NNX_CHECK(c, g->stencil == NNX_STENCIL_7 && g->kx && g->ky && g->kz, "multigrid needs a 7-point grid with kx, ky, kz");
    NNX_CHECK(c, g->nx > 0 && g->ny > 0 && g->nz > 0, "bad dims");
    Mg m;
    NNX_TRY(mg_setup(c, make_grid(g), nu, smoother, &m));
    return mg_vcycle(c, m, d_r, d_z);
}

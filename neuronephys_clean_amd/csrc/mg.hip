// mg.hip - geometric multigrid for the variable-coefficient 7-point
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
// Jacobi, the coarsest level 16 damped-Jacobi sweeps.  A W-cycle (each
// coarser level visited twice, the second time on the residual the first
// visit leaves) on the levels of more than 4096 cells, a V-cycle below.
// Block-mean restriction with trilinear prolongation is not a symmetric
// pair, so the cycle is used inside flexible CG.
//
// One generator produces the cycle as a sequence of steps (sweep, restrict,
// residual, prolong-add, add).  Steps on large levels are kernel launches;
// the steps of a visit of the first level of <= 4096 cells and below are
// built once per hierarchy and run by a single workgroup in one launch
// (tens of steps of a few microseconds' latency each would otherwise be
// tens of launches).
//
// No counterpart in the reference, whose CG is unpreconditioned
// (src/paracrine.cu:106-129).
#include <algorithm>
#include <cstdio>
#include <vector>

#include "mg.h"

namespace nnx {

namespace {

unsigned grid_blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }
// row kernels: enough workgroups for one wavefront per x-row on small levels
unsigned row_blocks(int ny, int nz) { return (unsigned)std::max(1, std::min((ny * nz + 3) / 4, 4096)); }

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

// ---- the steps of a cycle, per cell ----
// (every load of a step precedes its stores: the buffers of one step never
// alias, but the compiler cannot know that)
__device__ __forceinline__ void op_cheb(const GridK& g, const MgOp& o, int i, int j, int k, int64_t id) {
    const float* __restrict__ xin = o.xin;
    const float xc = xin ? xin[id] : 0.0f;
    const float res = xin ? o.rhs[id] - apply_at<float, float>(xin, g, i, j, k) : o.rhs[id];
    const float dold = o.c1 != 0.0f ? o.dvec[id] : 0.0f;
    const float t = o.c2 * o.dinv[id] * res;
    const float dn = o.c1 != 0.0f ? o.c1 * dold + t : t;
    o.dvec[id] = dn;
    o.xout[id] = xin ? xc + dn : dn;
}
__device__ __forceinline__ void op_jac(const GridK& g, const MgOp& o, int i, int j, int k, int64_t id) {
    const float* __restrict__ xin = o.xin;
    o.xout[id] = xin ? xin[id] + o.c1 * o.dinv[id] * (o.rhs[id] - apply_at<float, float>(xin, g, i, j, k))
                     : o.c1 * o.dinv[id] * o.rhs[id];
}
// coarse cell (I, J, K): block mean of the fine residual over its 8 children
__device__ __forceinline__ void op_restrict(const GridK& g, const MgOp& o, int I, int J, int K, int64_t q) {
    const int64_t nx = g.nx, ny = g.ny;
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int i = 2 * I + (c & 1), j = 2 * J + ((c >> 1) & 1), k = 2 * K + (c >> 2);
        const int64_t id = i + nx * (j + ny * k);
        s = s + (o.rhs[id] - apply_at<float, float>(o.xin, g, i, j, k));
    }
    o.xout[q] = 0.125f * s;
}
__device__ __forceinline__ void op_resid(const GridK& g, const MgOp& o, int i, int j, int k, int64_t id) {
    o.xout[id] = o.xout[id] - apply_at<float, float>(o.xin, g, i, j, k);
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
// x += P e (trilinear, zero outside); g fine, (NX, NY, NZ) coarse dims
__device__ __forceinline__ void op_prolong(const MgOp& o, int NX, int NY, int NZ, int i, int j, int k, int64_t id) {
    int ia, ib, ja, jb, ka, kb;
    float wi, wj, wk;
    lin_w(i, NX, ia, ib, wi);
    lin_w(j, NY, ja, jb, wj);
    lin_w(k, NZ, ka, kb, wk);
    const float* e = o.xin;
    auto E = [&](int I, int J, int K) { return e[I + (int64_t)NX * (J + (int64_t)NY * K)]; };
    const float ai = 0.75f, aj = 0.75f, ak = 0.75f;
    const float v0 = ai * E(ia, ja, ka) + wi * E(ib, ja, ka);
    const float v1 = ai * E(ia, jb, ka) + wi * E(ib, jb, ka);
    const float v2 = ai * E(ia, ja, kb) + wi * E(ib, ja, kb);
    const float v3 = ai * E(ia, jb, kb) + wi * E(ib, jb, kb);
    const float u0 = aj * v0 + wj * v1;
    const float u1 = aj * v2 + wj * v3;
    o.xout[id] = o.xout[id] + (ak * u0 + wk * u1);
}

template <int T> __device__ __forceinline__ void op_at(const GridK& g, const GridK& gc, const MgOp& o, int i, int j,
                                                       int k, int64_t id) {
    if (T == MG_CHEB) op_cheb(g, o, i, j, k, id);
    if (T == MG_JAC) op_jac(g, o, i, j, k, id);
    if (T == MG_RESTRICT) op_restrict(g, o, i, j, k, id);
    if (T == MG_RESID) op_resid(g, o, i, j, k, id);
    if (T == MG_PROLONG) op_prolong(o, gc.nx, gc.ny, gc.nz, i, j, k, id);
    if (T == MG_ADD) o.xout[id] = o.xout[id] + o.xin[id];
}

// one step over a large level: a wavefront per x-row (RESTRICT: per coarse row)
template <int T> __global__ __launch_bounds__(256) void k_mg_op(MgOp o, GridK g, GridK gc) {
    const GridK& d = T == MG_RESTRICT ? gc : g;
    for_rows(d.nx, d.ny, d.nz, [&](int i, int j, int k, int64_t id) { op_at<T>(g, gc, o, i, j, k, id); });
}

// a whole sequence of steps over small levels by one workgroup
__global__ __launch_bounds__(1024) void k_mg_sub(const MgOp* __restrict__ ops, int nops, const GridK* __restrict__ gk) {
    for (int q = 0; q < nops; ++q) {
        const MgOp o = ops[q];
        const GridK g = gk[o.level];
        const GridK gc = gk[o.level + 1 < kMgMaxLevels ? o.level + 1 : o.level];
        const GridK& d = o.type == MG_RESTRICT ? gc : g;
        const int nx = d.nx, ny = d.ny, n = d.nx * d.ny * d.nz;
        for (int id = threadIdx.x; id < n; id += blockDim.x) {
            const int i = id % nx, j = (id / nx) % ny, k = id / (nx * ny);
            switch (o.type) {
                case MG_CHEB: op_at<MG_CHEB>(g, gc, o, i, j, k, id); break;
                case MG_JAC: op_at<MG_JAC>(g, gc, o, i, j, k, id); break;
                case MG_RESTRICT: op_at<MG_RESTRICT>(g, gc, o, i, j, k, id); break;
                case MG_RESID: op_at<MG_RESID>(g, gc, o, i, j, k, id); break;
                case MG_PROLONG: op_at<MG_PROLONG>(g, gc, o, i, j, k, id); break;
                default: op_at<MG_ADD>(g, gc, o, i, j, k, id); break;
            }
        }
        __syncthreads();
    }
}

// The steps of one visit of level l (right-hand side rhs, buffer set `set`),
// appended to ops; returns the buffer that holds the visit's result.  With
// `host`, a visit of the level cycled in one workgroup is one MG_SUB marker
// (type -1, c1 = set).  Level 0 starts its sweeps in t so that an even
// number of sweeps ends in `z0`.
float* gen_visit(const Mg& m, int l, const float* rhs, int set, std::vector<MgOp>& ops, bool host, float* z0) {
    if (host && l == m.sub) {
        MgOp o{};
        o.type = -1, o.level = l, o.c1 = (float)set;
        ops.push_back(o);
        return m.sub_res[set];
    }
    float* X = l == 0 ? z0 : m.x[set][l];
    float* T = m.t[set][l];
    float *cur = nullptr, *oth = nullptr;
    auto sweep = [&](int s, bool jacobi) {
        MgOp o{};
        o.level = l, o.rhs = rhs, o.dinv = m.dinv[l];
        if (!cur) {
            o.xin = nullptr;
            o.xout = l == 0 ? T : X;
            oth = l == 0 ? X : T;
        } else {
            o.xin = cur;
            o.xout = oth;
            oth = cur;
        }
        if (jacobi) {
            o.type = MG_JAC;
            o.c1 = m.w;
        } else {
            o.type = MG_CHEB;
            o.dvec = m.d[l];
            o.c1 = m.c1[s];
            o.c2 = m.c2[s];
        }
        ops.push_back(o);
        cur = o.xout;
    };
    if (l == m.nlev - 1) {
        for (int s = 0; s < m.coarse_sweeps; ++s) sweep(s, true);
        return cur;
    }
    const bool jac = m.smoother == NNX_MG_JACOBI;
    for (int s = 0; s < m.nu; ++s) sweep(s, jac);
    {
        MgOp o{};
        o.type = MG_RESTRICT, o.level = l, o.rhs = rhs, o.xin = cur, o.xout = m.b[l + 1];
        ops.push_back(o);
    }
    float* e = gen_visit(m, l + 1, m.b[l + 1], 0, ops, host, z0);
    // the W-cycle's second visit only for levels above the one-workgroup ones
    const int gam = (m.sub < 0 || l + 1 < m.sub) ? m.gamma : 1;
    for (int v = 1; v < gam; ++v) {
        MgOp r{};
        r.type = MG_RESID, r.level = l + 1, r.xin = e, r.xout = m.b[l + 1];
        ops.push_back(r);
        float* e2 = gen_visit(m, l + 1, m.b[l + 1], 1, ops, host, z0);
        MgOp a{};
        a.type = MG_ADD, a.level = l + 1, a.xin = e2, a.xout = e;
        ops.push_back(a);
    }
    {
        MgOp o{};
        o.type = MG_PROLONG, o.level = l, o.xin = e, o.xout = cur;
        ops.push_back(o);
    }
    for (int s = 0; s < m.nu; ++s) sweep(s, jac);
    return cur;
}

}  // namespace

nnx_status mg_setup(nnx_ctx* c, const GridK& g0, int nu, int smoother, int gamma, Mg* m) {
    NNX_CHECK(c, g0.kind == 1, "multigrid needs the variable-coefficient 7-point operator");
    NNX_CHECK(c, nu >= 0 && nu <= 16, "nu must be in [0, 16]");
    NNX_CHECK(c, smoother == NNX_MG_CHEBYSHEV || smoother == NNX_MG_JACOBI, "unknown smoother");
    NNX_CHECK(c, gamma >= 0 && gamma <= 2, "cycle must be 0 (W), 1 (V) or 2 (W)");
    m->nu = nu > 0 ? nu : 3;
    m->smoother = smoother;
    m->gamma = gamma > 0 ? gamma : 2;
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
    auto n_of = [&](int q) { return (int64_t)m->g[q].nx * m->g[q].ny * m->g[q].nz; };
    for (int q = 0; q < m->nlev; ++q) {
        const int64_t n = n_of(q);
        char name[32];
        void* p;
        std::snprintf(name, sizeof name, "mg_v%d", q);
        NNX_TRY(scratch(c, name, (size_t)(q == 0 ? 3 : 7) * n * 4, false, &p));
        float* f = (float*)p;
        m->dinv[q] = f;
        m->d[q] = f + n;
        m->t[0][q] = f + 2 * n;
        m->x[0][q] = q ? f + 3 * n : nullptr;
        m->b[q] = q ? f + 4 * n : nullptr;
        m->x[1][q] = q ? f + 5 * n : nullptr;
        m->t[1][q] = q ? f + 6 * n : nullptr;
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
        k_mg_dinv<<<grid_blocks(n_of(q)), 256, 0, c->stream>>>(m->g[q], m->dinv[q]);
        NNX_LAUNCHED(c);
    }
    // the small levels: the first level >= 1 with at most kMgSubCells cells
    m->sub = -1;
    for (int q = 1; q < m->nlev; ++q)
        if (n_of(q) <= kMgSubCells) {
            m->sub = q;
            break;
        }
    if (m->sub > 0) {
        std::vector<MgOp> lists[2];
        for (int set = 0; set < 2; ++set)
            m->sub_res[set] = gen_visit(*m, m->sub, m->b[m->sub], set, lists[set], false, nullptr);
        const size_t n0 = lists[0].size(), n1 = lists[1].size();
        void *ops, *gk;
        NNX_TRY(scratch(c, "mg_ops", (n0 + n1) * sizeof(MgOp), false, &ops));
        NNX_TRY(scratch(c, "mg_gk", sizeof(m->g), false, &gk));
        std::vector<MgOp> all(lists[0]);
        all.insert(all.end(), lists[1].begin(), lists[1].end());
        NNX_HIP(c, hipMemcpyAsync(ops, all.data(), all.size() * sizeof(MgOp), hipMemcpyHostToDevice, c->stream));
        NNX_HIP(c, hipMemcpyAsync(gk, m->g, sizeof(m->g), hipMemcpyHostToDevice, c->stream));
        NNX_HIP(c, hipStreamSynchronize(c->stream));  // host temporaries
        m->sub_ops[0] = (const MgOp*)ops;
        m->sub_ops[1] = (const MgOp*)ops + n0;
        m->sub_nops[0] = (int)n0;
        m->sub_nops[1] = (int)n1;
        m->d_gk = (const GridK*)gk;
    }
    return NNX_OK;
}

nnx_status mg_vcycle(nnx_ctx* c, Mg& m, const float* r, float* z) {
    std::vector<MgOp> ops;
    ops.reserve(256);
    float* res = gen_visit(m, 0, r, 0, ops, true, z);
    for (const MgOp& o : ops) {
        if (o.type < 0) {
            const int set = (int)o.c1;
            k_mg_sub<<<1, 1024, 0, c->stream>>>(m.sub_ops[set], m.sub_nops[set], m.d_gk);
            NNX_LAUNCHED(c);
            continue;
        }
        const GridK& g = m.g[o.level];
        const GridK& gc = m.g[o.level + 1 < m.nlev ? o.level + 1 : o.level];
        const GridK& d = o.type == MG_RESTRICT ? gc : g;
        const unsigned blocks = row_blocks(d.ny, d.nz);
        switch (o.type) {
            case MG_CHEB: k_mg_op<MG_CHEB><<<blocks, 256, 0, c->stream>>>(o, g, gc); break;
            case MG_JAC: k_mg_op<MG_JAC><<<blocks, 256, 0, c->stream>>>(o, g, gc); break;
            case MG_RESTRICT: k_mg_op<MG_RESTRICT><<<blocks, 256, 0, c->stream>>>(o, g, gc); break;
            case MG_RESID: k_mg_op<MG_RESID><<<blocks, 256, 0, c->stream>>>(o, g, gc); break;
            case MG_PROLONG: k_mg_op<MG_PROLONG><<<blocks, 256, 0, c->stream>>>(o, g, gc); break;
            default: k_mg_op<MG_ADD><<<blocks, 256, 0, c->stream>>>(o, g, gc); break;
        }
        NNX_LAUNCHED(c);
    }
    if (res != z) NNX_HIP(c, hipMemcpyAsync(z, res, m.g[0].nx * (size_t)m.g[0].ny * m.g[0].nz * 4,
                                            hipMemcpyDeviceToDevice, c->stream));
    return NNX_OK;
}

}  // namespace nnx

using namespace nnx;

extern "C" nnx_status nnx_mg_apply(nnx_ctx* c, const nnx_grid* g, int nu, int smoother, int cycle,
                                   const float* d_r, float* d_z) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, g && d_r && d_z, "null argument");
    NNX_CHECK(c, g->stencil == NNX_STENCIL_7 && g->kx && g->ky && g->kz, "multigrid needs a 7-point grid with kx, ky, kz");
    NNX_CHECK(c, g->nx > 0 && g->ny > 0 && g->nz > 0, "bad dims");
    Mg m;
    NNX_TRY(mg_setup(c, make_grid(g), nu, smoother, cycle, &m));
    return mg_vcycle(c, m, d_r, d_z);
}

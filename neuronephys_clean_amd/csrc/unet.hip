// unet.hip - the learned preconditioner: a 3-level 3-D convolutional U-Net
// evaluated and trained with BF16 MFMA implicit-GEMM kernels (FP32
// accumulation), used as the variable preconditioner of flexible CG.
// No counterpart in the reference (its paracrine solve is plain CG,
// src/paracrine.cu:106-129).
//
// Network (x = [q, q * feature], q = r / rms(r), channels-last BF16
// activations; no biases: they are held at zero so that N(0) = 0 and
// N(a r) = a N(r) for a > 0, as a preconditioner needs):
//   e0 = act(conv3(x; 2->16))            level 0    p0 = avgpool2(e0)
//   e1 = act(conv3(p0; 16->32))          level 1    p1 = avgpool2(e1)
//   e2 = act(conv3(p1; 32->32))          level 2
//   d1 = act(conv3(e1 + up(e2); 32->16)) level 1    (up: nearest, x2)
//   d0 = act(conv3(e0 + up(d1); 16->16)) level 0
//   z  = rms(r) * (w_h . d0 + b_h)        1x1 head, fp32
// conv3 is a 3x3x3 convolution with zero padding, act the leaky ReLU
// (slope 0.1 below zero).  63,201 parameters.
//
// Convolution as implicit GEMM on v_mfma_f32_16x16x32_bf16: D[co][voxel] +=
// W[co][k] * X[k][voxel], k = tap*Cin + ci (taps padded to a multiple of 32).
// A workgroup owns a 16 x TY x TZ output tile; the input tile with its
// one-voxel halo and the whole weight matrix sit in LDS; each wave keeps
// the accumulators of its output rows in registers and reuses every weight
// fragment across them.  Pooling, the skip addition with upsampling, ReLU,
// bias and the 1x1 head are fused into the staging and epilogue.
//
// Backward: data gradients are the same kernel with the taps flipped and
// the channel roles exchanged; weight gradients are a second MFMA GEMM
// (D[co][k] += G[co][voxel] * X[voxel][k], K running over voxels) with
// per-workgroup partials reduced in a fixed order.
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace nnx {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bf2f(uint32_t b16) { return __uint_as_float(b16 << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ uint32_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ f32x4_t mfma16(u32x4_t a, u32x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                   __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
}
constexpr int kpad(int cin) { return (27 * cin + 31) / 32 * 32; }
constexpr float kLeak = 0.1f;  // slope of the activation below zero (leaky ReLU)
__device__ __forceinline__ int tap_dz(int t) { return t / 9; }
__device__ __forceinline__ int tap_dy(int t) { return (t / 3) % 3; }
__device__ __forceinline__ int tap_dx(int t) { return t % 3; }

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
// two floats to bf16 (round to nearest even) in one instruction
__device__ __forceinline__ uint32_t pack2bf(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{lo, hi}, bf16x2_t));
}
// 8 bf16 of a + s * c, each one fma rounded once to bf16
__device__ __forceinline__ u32x4_t fma_scaled8(u32x4_t a, u32x4_t c, float s) {
    u32x4_t o;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float lo = __builtin_fmaf(s, __uint_as_float(c[q] << 16), __uint_as_float(a[q] << 16));
        const float hi = __builtin_fmaf(s, __uint_as_float(c[q] & 0xffff0000u), __uint_as_float(a[q] & 0xffff0000u));
        o[q] = pack2bf(lo, hi);
    }
    return o;
}

// 8 bf16 of a + s * c, rounded once
__device__ __forceinline__ u32x4_t add_scaled8(u32x4_t a, u32x4_t c, float s) {
    u32x4_t o;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float lo = bf2f(a[q] & 0xffffu) + s * bf2f(c[q] & 0xffffu);
        const float hi = bf2f(a[q] >> 16) + s * bf2f(c[q] >> 16);
        o[q] = f2bf(lo) | (f2bf(hi) << 16);
    }
    return o;
}

// ---------------------------------------------------------------------------
// forward convolution
// ---------------------------------------------------------------------------
struct ConvArgs {
    int nx, ny, nz, batch;
    const uint16_t* in;      // [b][z][y][x][CIN]
    const uint16_t* coarse;  // optional: in + cscale * coarse[parent voxel] (half resolution)
    float cscale;
    const uint16_t* w;       // [COUT][kpad(CIN)], zero padded
    const float* bias;       // [COUT] or null
    int relu;
    uint16_t* out;           // [b][z][y][x][COUT] or null
    uint16_t* pool;          // 2x2x2 average of the activated output, or null
    const float* head;       // [COUT] weights + bias of the 1x1 head, or null
    const float* hscale;     // [batch] output scale of the head, or null (1)
    float* hout;             // [b][z][y][x] fp32 head output
};

// output tile 16 x TY x TZ.  TZ = 4 keeps the 16-channel layers at three
// workgroups per CU (input halo 34.5 KiB + weights 14.3 KiB of LDS) and the
// 2-channel layer at half the registers of TZ = 8; the kernel is bound by
// each workgroup's staging -> MFMA -> epilogue latency, so occupancy pays.
template <int CIN> struct ConvTile {
    static constexpr int TY = CIN == 32 ? 4 : 8, TZ = 4;
};

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3(ConvArgs a) {
    constexpr int TY = ConvTile<CIN>::TY, TZ = ConvTile<CIN>::TZ;
    constexpr int KP = kpad(CIN), NKB = KP / 32, KS = KP + 8;
    constexpr int HX = 18, HY = TY + 2, HZ = TZ + 2, HV = HX * HY * HZ;
    constexpr int MT = COUT / 16;
    constexpr int QPW = (TY / 2) * (TZ / 2) / 4;  // 2x2 (y,z) row quads per wave
    constexpr int RPW = 4 * QPW;
    __shared__ __attribute__((aligned(16))) uint16_t s_in[HV * CIN];
    __shared__ __attribute__((aligned(16))) uint16_t s_w[COUT * KS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, xl = lane & 15;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int nzt = nz / TZ;
    const int b = blockIdx.z / nzt, z0 = (blockIdx.z % nzt) * TZ, y0 = blockIdx.y * TY,
              x0 = blockIdx.x * 16;
    // Staging: every load of a thread (weights, input halo, coarse input) is
    // issued before the first one is used, so a workgroup pays one memory
    // latency, not one per chunk.
    constexpr int NW = COUT * (KP / 8), WI = (NW + 255) / 256;
    u32x4_t wv[WI];
#pragma unroll
    for (int u = 0; u < WI; ++u) {
        const int i = min(tid + 256 * u, NW - 1);
        const int co = i / (KP / 8), k8 = i - co * (KP / 8);
        wv[u] = *(const u32x4_t*)&a.w[(size_t)co * KP + 8 * k8];
    }
    // The input halo.  Every load is unconditional (a voxel outside the grid
    // reads the nearest inside one and is zeroed when stored): a load under a
    // condition made the compiler wait for it at the end of the condition,
    // one memory latency per chunk.
    {
        constexpr int CH = CIN == 2 ? 1 : CIN / 8, NCH = HV * CH, IT = (NCH + 255) / 256;
        using SV = typename std::conditional<CIN == 2, uint32_t, u32x4_t>::type;
        SV v[IT], cv[IT];
        uint32_t inside = 0;
        size_t fo[IT];
        int coff[IT];
#pragma unroll
        for (int u = 0; u < IT; ++u) {
            int i = tid + 256 * u;
            i = i < NCH ? i : NCH - 1;
            const int hv = i / CH, c8 = i - hv * CH;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
            if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) inside |= 1u << u;
            const int cx = min(max(gx, 0), nx - 1), cy = min(max(gy, 0), ny - 1), cz = min(max(gz, 0), nz - 1);
            fo[u] = ((((size_t)b * nz + cz) * ny + cy) * nx + cx) * CIN + 8 * c8;
            coff[u] = (((cz >> 1) * (ny >> 1) + (cy >> 1)) * (nx >> 1) + (cx >> 1)) * CIN + 8 * c8;
            if constexpr (CIN == 2)
                v[u] = *(const uint32_t*)&a.in[fo[u]];
            else
                v[u] = *(const u32x4_t*)&a.in[fo[u]];
        }
        if constexpr (CIN != 2) {
            if (a.coarse) {
                const size_t cb = (size_t)b * (nz >> 1) * (ny >> 1) * (nx >> 1) * CIN;
#pragma unroll
                for (int u = 0; u < IT; ++u) cv[u] = *(const u32x4_t*)&a.coarse[cb + coff[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < WI; ++u) {
            const int i = tid + 256 * u;
            if (i < NW) {
                const int co = i / (KP / 8), k8 = i - co * (KP / 8);
                *(u32x4_t*)&s_w[co * KS + 8 * k8] = wv[u];
            }
        }
#pragma unroll
        for (int u = 0; u < IT; ++u) {
            const int i = tid + 256 * u;
            if (i < NCH) {
                const bool in = (inside >> u) & 1u;
                if constexpr (CIN == 2) {
                    *(uint32_t*)&s_in[i * 2] = in ? v[u] : 0u;
                } else {
                    const u32x4_t z = u32x4_t{0u, 0u, 0u, 0u};
                    *(u32x4_t*)&s_in[i * 8] = !in ? z : a.coarse ? add_scaled8(v[u], cv[u], a.cscale) : v[u];
                }
            }
        }
    }
    __syncthreads();
    // Rows of this wave: quads of 2x2 (y,z) rows.  Row r sits at
    // (wy + ryc(r), wz + rzc(r)) in the tile, with ryc / rzc compile-time
    // constants, so its LDS address is a per-lane base plus an immediate.
    // a wave owns RY x 2 (y, z) rows; the four waves tile TY x TZ
    constexpr int RY = RPW / 2, NYB = TY / RY;
    static_assert(NYB * (TZ / 2) == 4, "four waves per tile");
    const int wy = RY * (wave % NYB), wz = 2 * (wave / NYB);
    const int vbase = ((wz * HY + wy) * HX + xl) * CIN;
    auto ryc = [](int r) { return 2 * (r >> 2) + (r & 1); };
    auto rzc = [](int r) { return (r >> 1) & 1; };
    auto toff = [&](int tap, int coff) {
        tap = tap > 26 ? 26 : tap;  // padded k: zero weights
        return ((tap_dz(tap) * HY + tap_dy(tap)) * HX + tap_dx(tap)) * CIN + coff;
    };
    f32x4_t acc[MT][RPW];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < RPW; ++r) acc[mt][r] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int kb = 0; kb < NKB; ++kb) {
        u32x4_t af[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) af[mt] = *(const u32x4_t*)&s_w[(16 * mt + xl) * KS + 32 * kb + 8 * g];
        u32x4_t bv[RPW];
        if (CIN == 2) {
            int o[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) o[u] = vbase + toff(16 * kb + 4 * g + u, 0);
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int ro = (rzc(r) * HY + ryc(r)) * HX * CIN;
#pragma unroll
                for (int u = 0; u < 4; ++u) bv[r][u] = *(const uint32_t*)&s_in[o[u] + ro];
            }
        } else {
            const int o = vbase + (CIN == 16 ? toff(2 * kb + (g >> 1), 8 * (g & 1)) : toff(kb, 8 * g));
#pragma unroll
            for (int r = 0; r < RPW; ++r)
                bv[r] = *(const u32x4_t*)&s_in[o + (rzc(r) * HY + ryc(r)) * HX * CIN];
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[mt][r] = mfma16(af[mt], bv[r], acc[mt][r]);
    }
    // epilogue: lane holds voxel x0 + xl, channels 16 mt + 4 g + q
    float bias_r[MT][4], head_r[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int co = 16 * mt + 4 * g + q;
            bias_r[mt][q] = a.bias ? a.bias[co] : 0.f;
            head_r[mt][q] = a.head ? a.head[co] : 0.f;
        }
    const float hb = a.head ? a.head[COUT] : 0.f, hs = a.hscale ? a.hscale[b] : 1.f;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int y = y0 + wy + ryc(r), z = z0 + wz + rzc(r);
        const size_t vox = (((size_t)b * nz + z) * ny + y) * nx + x0 + xl;
        float hsum = 0.f;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            f32x4_t v = acc[mt][r];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float s = v[q] + bias_r[mt][q];
                if (a.relu) s = s > 0.f ? s : kLeak * s;
                v[q] = s;
                hsum += head_r[mt][q] * s;
            }
            acc[mt][r] = v;
            if (a.out) {
                u32x2_t o;
                o[0] = f2bf(v[0]) | (f2bf(v[1]) << 16);
                o[1] = f2bf(v[2]) | (f2bf(v[3]) << 16);
                *(u32x2_t*)&a.out[vox * COUT + 16 * mt + 4 * g] = o;
            }
        }
        if (a.head) {
            hsum += __shfl_xor(hsum, 16, 64);
            hsum += __shfl_xor(hsum, 32, 64);
            if (g == 0) a.hout[vox] = hs * (hsum + hb);
        }
    }
    if (a.pool) {
        const int px = nx >> 1, py = ny >> 1, pz = nz >> 1;
#pragma unroll
        for (int j = 0; j < QPW; ++j) {
            const int y = (y0 + wy + ryc(4 * j)) >> 1, z = (z0 + wz + rzc(4 * j)) >> 1;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                f32x4_t s = ((acc[mt][4 * j] + acc[mt][4 * j + 1]) + acc[mt][4 * j + 2]) + acc[mt][4 * j + 3];
#pragma unroll
                for (int q = 0; q < 4; ++q) s[q] = (s[q] + __shfl_xor(s[q], 1, 64)) * 0.125f;
                if ((xl & 1) == 0) {
                    u32x2_t o;
                    o[0] = f2bf(s[0]) | (f2bf(s[1]) << 16);
                    o[1] = f2bf(s[2]) | (f2bf(s[3]) << 16);
                    const size_t pv = (((size_t)b * pz + z) * py + y) * px + ((x0 + xl) >> 1);
                    *(u32x2_t*)&a.pool[pv * COUT + 16 * mt + 4 * g] = o;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// forward convolution, z-marching form
// ---------------------------------------------------------------------------
// A workgroup owns a (16 WX) x (RY 4 / WX) column of output rows over a
// chunk of zc planes and walks up z one input plane at a time.  The plane,
// with its one-voxel x/y halo and the upsampled coarse term added, is staged
// in LDS in two buffers, its loads issued two planes ahead into one of two
// register sets.  K is ordered plane by plane (a k-block holds taps (dy, dx) of one
// input plane x channels), so the B fragment read for an output row serves
// the three output planes that read this input plane (dz = 0, 1, 2), and
// three accumulator sets rotate as output planes complete.  The weights
// stay in registers as A fragments, loaded once per workgroup.  An input
// plane is read from memory once per chunk (zc + 2 planes for zc outputs);
// a 16-channel plane's 9 taps take 5 k-blocks (one half-block of zero
// weights), a 32-channel plane's 9, the 2-channel input's 1 (7 taps of
// zero weights).
template <int CIN> struct ZK {
    static constexpr int TPK = 32 / CIN;             // taps per k-block
    static constexpr int KBP = (9 + TPK - 1) / TPK;  // k-blocks per input plane
    static constexpr bool WIDE = CIN >= 8;           // staged in 16-byte chunks of 8 channels
    static constexpr int NC8 = WIDE ? CIN / 8 : 1;
};

template <int CIN, int COUT, int RY, int MT, bool CO, int WX = 1>
__global__ __launch_bounds__(256, 2) void k_conv3z(ConvArgs a, int zc) {
    static_assert(!CO || ZK<CIN>::WIDE, "the 2-channel input takes no upsampled term");
    static_assert(RY % 2 == 0, "row pairs for the pooled output");
    // MT m-tiles (16 output channels each) per workgroup: channel group
    // blockIdx.y % NG of NG.  WX of the four waves side by side in x: a column is 16 WX wide and
    // RY (4 / WX) rows high (the 2-channel input takes WX = 4, so that its
    // staged rows of 66 voxels x 4 bytes fill whole cache lines)
    static_assert(WX == 1 || WX == 2 || WX == 4, "waves along x");
    constexpr int TX = 16 * WX, HX = TX + 2;
    constexpr int KBP = ZK<CIN>::KBP, NC8 = ZK<CIN>::NC8, NG = COUT / (16 * MT), TY = (4 / WX) * RY, HR = TY + 2;
    constexpr int KP = kpad(CIN);
    constexpr bool WIDE = ZK<CIN>::WIDE;
    constexpr int RS = WIDE ? NC8 * HX * 8 : HX * 2;     // bf16 elements per staged row
    constexpr int PE = HR * RS;                          // bf16 elements per staged plane
    constexpr int NCH = WIDE ? HR * HX * NC8 : HR * HX;  // load chunks (16 B or 4 B) per plane
    constexpr int IT = (NCH + 255) / 256;
    using SV = typename std::conditional<WIDE, u32x4_t, uint32_t>::type;
    __shared__ __attribute__((aligned(16))) uint16_t s_pl[2][PE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, xl = lane & 15;
    const int nx = a.nx, ny = a.ny, nz = a.nz, ntx = nx / TX;
    const int x0 = (blockIdx.x % ntx) * TX, y0 = (blockIdx.x / ntx) * TY;
    const int xw = 16 * (wave % WX), wy = RY * (wave / WX);  // this wave's 16 x RY rows in the column
    const int b = blockIdx.z, z0 = (blockIdx.y / NG) * zc, cb = 16 * MT * (blockIdx.y % NG);
    const int zn = min(zc, nz - z0), nsteps = zn + 2;
    const u32x4_t zero4 = u32x4_t{0u, 0u, 0u, 0u};
    // A fragments: lane (g, xl) holds W[co = 16 mt + xl][k = 8g .. 8g + 7] of
    // k-block j of input plane dz; taps past the ninth have zero weights
    u32x4_t wa[3][KBP][MT];
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
        for (int j = 0; j < KBP; ++j)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const uint16_t* wr = a.w + (size_t)(cb + 16 * mt + xl) * KP + dz * 9 * CIN;
                if constexpr (CIN == 2) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int s = 4 * g + u;
                        wa[dz][j][mt][u] = s < 9 ? *(const uint32_t*)&wr[2 * s] : 0u;
                    }
                } else if constexpr (CIN == 16) {
                    const int s = 2 * j + (g >> 1);
                    wa[dz][j][mt] = s < 9 ? *(const u32x4_t*)&wr[16 * s + 8 * (g & 1)] : zero4;
                } else {
                    wa[dz][j][mt] = *(const u32x4_t*)&wr[32 * j + 8 * g];
                }
            }
    // B fragment offsets in a staged plane (row 0 = y0 - 1, column 0 = x0 - 1):
    // lane (g, xl) reads tap (dy, dx) of its k-slice at voxel x0 + xw + xl
    constexpr int NF = WIDE ? KBP : 4;
    int foff[NF];
#pragma unroll
    for (int j = 0; j < NF; ++j) {
        if constexpr (WIDE) {
            const int s = CIN == 32 ? j : min(2 * j + (g >> 1), 8), c8 = CIN == 32 ? g : (g & 1);
            foff[j] = (s / 3) * RS + c8 * HX * 8 + (s % 3 + xw + xl) * 8;
        } else {
            const int s = min(4 * g + j, 8);
            foff[j] = (s / 3) * RS + (s % 3 + xw + xl) * 2;
        }
    }
    // staging: chunk i of a plane is (row, x, 8-channel group) in memory
    // order; a voxel outside the grid reads the nearest inside one and is
    // stored as zero (unconditional loads)
    const int64_t pstride = (int64_t)ny * nx * CIN, cpstride = (int64_t)(ny >> 1) * (nx >> 1) * CIN;
    const uint16_t* inb = a.in + (size_t)b * nz * pstride;
    const uint16_t* cinb = CO ? a.coarse + (size_t)b * (nz >> 1) * cpstride : nullptr;
    int gofs[IT], cofs[IT], lofs[IT];
    uint32_t inside = 0;
#pragma unroll
    for (int u = 0; u < IT; ++u) {
        int i = tid + 256 * u;
        i = i < NCH ? i : NCH - 1;
        const int c8 = i % NC8, vx = (i / NC8) % HX, row = i / (NC8 * HX);
        const int gx = x0 - 1 + vx, gy = y0 - 1 + row;
        if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) inside |= 1u << u;
        const int cx = min(max(gx, 0), nx - 1), cy = min(max(gy, 0), ny - 1);
        gofs[u] = (cy * nx + cx) * CIN + 8 * c8;
        cofs[u] = ((cy >> 1) * (nx >> 1) + (cx >> 1)) * CIN + 8 * c8;
        lofs[u] = WIDE ? ((row * NC8 + c8) * HX + vx) * 8 : (row * HX + vx) * 2;
    }
    // two register sets: plane s + 2 is loaded during step s while plane
    // s + 1, loaded during step s - 1, is stored to LDS at its end
    SV v[2][IT];
    u32x4_t cv[2][CO ? IT : 1];
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int u = 0; u < IT; ++u) v[e][u] = SV{};
    auto load = [&](auto ec, int iz) __attribute__((always_inline)) {
        constexpr int E = decltype(ec)::value;
        const uint16_t* p = inb + (size_t)iz * pstride;
#pragma unroll
        for (int u = 0; u < IT; ++u) v[E][u] = *(const SV*)&p[gofs[u]];
        if constexpr (CO) {
            const uint16_t* q = cinb + (size_t)(iz >> 1) * cpstride;
#pragma unroll
            for (int u = 0; u < IT; ++u) cv[E][u] = *(const u32x4_t*)&q[cofs[u]];
        }
    };
    auto store = [&](auto ec, uint16_t* pl, bool zero) __attribute__((always_inline)) {
        constexpr int E = decltype(ec)::value;
#pragma unroll
        for (int u = 0; u < IT; ++u) {
            if (tid + 256 * u < NCH) {
                const bool in = !zero && ((inside >> u) & 1u);
                if constexpr (WIDE) {
                    u32x4_t o = v[E][u];
                    if constexpr (CO) o = fma_scaled8(v[E][u], cv[E][u], a.cscale);
                    *(u32x4_t*)&pl[lofs[u]] = in ? o : zero4;
                } else {
                    *(uint32_t*)&pl[lofs[u]] = in ? v[E][u] : 0u;
                }
            }
        }
    };
    // epilogue constants: lane holds voxel x0 + xw + xl, channels cb + 16 mt + 4 g + q
    // (the 1x1 head needs every channel in one workgroup: NG = 1)
    float bias_r[MT][4], head_r[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int co = cb + 16 * mt + 4 * g + q;
            bias_r[mt][q] = a.bias ? a.bias[co] : 0.f;
            head_r[mt][q] = a.head ? a.head[co] : 0.f;
        }
    const float hb = a.head ? a.head[COUT] : 0.f, hs = a.hscale ? a.hscale[b] : 1.f;
    const int yw = y0 + wy;
    f32x4_t pacc[RY / 2][MT];
    f32x4_t acc[3][RY][MT];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < RY; ++r)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[p][r][mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    // (lambdas are forced inline: an accumulator set passed out of line
    // would live in scratch memory)
    // per-lane offsets within an output plane (voxels); the activation's
    // slope below zero (1: identity); lane-exchange addresses
    const int lvox = yw * nx + x0 + xw + xl;
    const float leak = a.relu ? kLeak : 1.f;
    const int sw16 = (lane ^ 16) << 2, sw32 = (lane ^ 32) << 2;
    auto xch = [](int addr, float v) __attribute__((always_inline)) {
        return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
    };
    auto epilogue = [&](auto sc, int z) __attribute__((always_inline)) {
        auto& A = acc[decltype(sc)::value];
        const int64_t zb = ((int64_t)b * nz + z) * ny * nx;  // first voxel of output plane z
        uint16_t* outz = a.out ? a.out + zb * COUT : nullptr;
        float* houtz = a.head ? a.hout + zb : nullptr;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int vo = lvox + r * nx;
            float hsum = 0.f;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                f32x4_t v4 = A[r][mt];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float s = v4[q];  // the bias is in the accumulator's start value
                    s = fmaxf(s, leak * s);  // leaky ReLU (slope leak <= 1)
                    v4[q] = s;
                    hsum = __builtin_fmaf(head_r[mt][q], s, hsum);
                }
                A[r][mt] = v4;
                if (outz) {
                    u32x2_t o;
                    o[0] = pack2bf(v4[0], v4[1]);
                    o[1] = pack2bf(v4[2], v4[3]);
                    *(u32x2_t*)&outz[vo * COUT + cb + 16 * mt + 4 * g] = o;
                }
            }
            if (houtz) {
                hsum += xch(sw16, hsum);
                hsum += xch(sw32, hsum);
                if (g == 0) houtz[vo] = hs * (hsum + hb);
            }
        }
        if (a.pool) {
            // 2x2x2 average: (y, y+1) pairs here, (z, z+1) across two
            // planes, (x, x+1) across lanes
            const int px = nx >> 1, py = ny >> 1, pz = nz >> 1;
#pragma unroll
            for (int j = 0; j < RY / 2; ++j)
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    const f32x4_t sm = A[2 * j][mt] + A[2 * j + 1][mt];
                    if ((z & 1) == 0) {
                        pacc[j][mt] = sm;
                    } else {
                        f32x4_t s = pacc[j][mt] + sm;
                        // x pairs: lanes (2i, 2i+1) swapped by a quad permutation [1,0,3,2]
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            s[q] = (s[q] + __int_as_float(__builtin_amdgcn_update_dpp(
                                               0, __float_as_int(s[q]), 0xB1, 0xF, 0xF, false))) *
                                   0.125f;
                        if ((xl & 1) == 0) {
                            u32x2_t o;
                            o[0] = pack2bf(s[0], s[1]);
                            o[1] = pack2bf(s[2], s[3]);
                            const int64_t pv =
                                (((int64_t)b * pz + (z >> 1)) * py + ((yw >> 1) + j)) * px + ((x0 + xw + xl) >> 1);
                            *(u32x2_t*)&a.pool[pv * COUT + cb + 16 * mt + 4 * g] = o;
                        }
                    }
                }
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    // input plane z0 - 1 (zero below the grid) to LDS; plane z0 in flight
    if (z0 > 0) load(I0{}, z0 - 1);
    store(I0{}, s_pl[0], z0 == 0);
    load(I1{}, z0);
    bool zpl[2] = {false, false};  // register set holds a plane outside the grid
    // every load so far (weights, epilogue constants) complete before the
    // loop: otherwise the first use of a weight inside the loop carries a
    // memory-count wait that, on every later step, also waits for the
    // plane just prefetched
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // Step s reads input plane z0 - 1 + s from buffer s & 1: output plane
    // z0 + s - dz takes tap plane dz, so accumulator set (s - dz) mod 3 is
    // updated (set s mod 3 starts afresh) and output plane z0 + s - 2 is
    // complete at the end of the step.  Updates of planes outside the chunk
    // land in sets that are restarted before use or never stored.
    auto step = [&](auto pc, auto ec, int s) __attribute__((always_inline)) {
        constexpr int P = decltype(pc)::value, E = decltype(ec)::value;
        const uint16_t* pl = s_pl[E] + wy * RS;
        // input plane of step s + 2, loaded unconditionally (a plane past the
        // grid reads the last one and is stored as zeros; one past the chunk
        // is never stored): a load under a condition makes the compiler's
        // memory-count wait at the end of the step cover the loads just
        // issued, not only the set being stored
        const int iz = z0 + s + 1;
        zpl[E] = iz >= nz;
        load(ec, min(iz, nz - 1));
#pragma unroll
        for (int j = 0; j < KBP; ++j)
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                u32x4_t bf;
                if constexpr (WIDE) {
                    bf = *(const u32x4_t*)&pl[r * RS + foff[j]];
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) bf[u] = *(const uint32_t*)&pl[r * RS + foff[u]];
                }
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    // a plane's first contribution starts from the bias
                    const f32x4_t c0 = j == 0 ? f32x4_t{bias_r[mt][0], bias_r[mt][1], bias_r[mt][2], bias_r[mt][3]}
                                              : acc[P][r][mt];
                    acc[P][r][mt] = mfma16(wa[0][j][mt], bf, c0);
                    acc[(P + 2) % 3][r][mt] = mfma16(wa[1][j][mt], bf, acc[(P + 2) % 3][r][mt]);
                    acc[(P + 1) % 3][r][mt] = mfma16(wa[2][j][mt], bf, acc[(P + 1) % 3][r][mt]);
                }
            }
        // the staging store before the epilogue: its memory-count wait then
        // follows this step's loads and the previous step's output stores,
        // issued a whole MFMA phase earlier, not the stores just issued
        store(std::integral_constant<int, E ^ 1>{}, s_pl[E ^ 1], zpl[E ^ 1]);
        if (s >= 2 && s - 2 < zn) epilogue(std::integral_constant<int, (P + 1) % 3>{}, z0 + s - 2);
        __syncthreads();
    };
    // whole groups of six steps (the rotation periods 3 and 2): steps past
    // the chunk's last only load and stage clamped planes and store nothing,
    // so no step sits under a condition (a skipped step would merge register
    // and memory-count states at the loop head)
    for (int s0 = 0; s0 < nsteps; s0 += 6) {
        step(I0{}, I0{}, s0);
        step(I1{}, I1{}, s0 + 1);
        step(I2{}, I0{}, s0 + 2);
        step(I0{}, I1{}, s0 + 3);
        step(I1{}, I0{}, s0 + 4);
        step(I2{}, I1{}, s0 + 5);
    }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[co][tap*CIN + ci] = sum_v G[v][co] X[v + off(tap)][ci]
// ---------------------------------------------------------------------------
struct WgArgs {
    int nx, ny, nz, batch, ntiles;
    const uint16_t* in;
    const uint16_t* coarse;
    float cscale;
    const uint16_t* g;  // [b][z][y][x][COUT]
    float* part;        // [gridDim.x][COUT][kpad(CIN)]
    float* bpart;       // [gridDim.x][COUT]
};

constexpr int kWgRows = 264;  // LDS row stride of the gradient tile (256 voxels + pad)

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad(WgArgs a) {
    constexpr int KP = kpad(CIN), NT = KP / 16, MT = COUT / 16, NTW = (NT + 3) / 4;
    constexpr int PLANE = 6 * 6 * 16 + 8;  // one (dx, ci) copy of the 18x6x6 tile, x shifted by dx
    __shared__ __attribute__((aligned(16))) uint16_t s_in[3 * CIN * PLANE];
    __shared__ __attribute__((aligned(16))) uint16_t s_g[COUT * kWgRows];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, xl = lane & 15;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int tx = nx / 16, ty = ny / 4, tz = nz / 4;
    f32x4_t acc[NTW][MT];
#pragma unroll
    for (int j = 0; j < NTW; ++j)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[j][mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const int bx = tile % tx, by = (tile / tx) % ty, rest = tile / (tx * ty);
        const int bz = rest % tz, b = rest / tz;
        const int x0 = 16 * bx, y0 = 4 * by, z0 = 4 * bz;
        __syncthreads();
        constexpr int CH = CIN == 2 ? 1 : CIN / 8;
        for (int i = tid; i < 18 * 36 * CH; i += 256) {
            const int hv = i / CH, c8 = i - hv * CH;
            const int hx = hv % 18, hy = (hv / 18) % 6, hz = hv / 108;
            const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
            const bool in = gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz;
            const size_t gv = (((size_t)b * nz + gz) * ny + gy) * nx + gx;
            uint16_t val[8];
            if (CIN == 2) {
                const uint32_t v = in ? *(const uint32_t*)&a.in[gv * 2] : 0u;
                val[0] = (uint16_t)(v & 0xffffu);
                val[1] = (uint16_t)(v >> 16);
            } else {
                u32x4_t v = {0u, 0u, 0u, 0u};
                if (in) {
                    v = *(const u32x4_t*)&a.in[gv * CIN + 8 * c8];
                    if (a.coarse) {
                        const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
                        const u32x4_t cv = *(const u32x4_t*)&a.coarse
                            [((((size_t)b * cz + (gz >> 1)) * cy + (gy >> 1)) * cx + (gx >> 1)) * CIN + 8 * c8];
                        v = add_scaled8(v, cv, a.cscale);
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    val[2 * q] = (uint16_t)(v[q] & 0xffffu);
                    val[2 * q + 1] = (uint16_t)(v[q] >> 16);
                }
            }
            constexpr int NC = CIN == 2 ? 2 : 8;
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int xs = hx - dx;
                if (xs < 0 || xs >= 16) continue;
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    s_in[(dx * CIN + (CIN == 2 ? c : 8 * c8 + c)) * PLANE + (hz * 6 + hy) * 16 + xs] = val[c];
            }
        }
        for (int i = tid; i < 256 * (COUT / 8); i += 256) {
            const int v = i / (COUT / 8), c8 = i - v * (COUT / 8);
            const int xs = v % 16, rr = v / 16;
            const size_t gv = (((size_t)b * nz + z0 + (rr >> 2)) * ny + y0 + (rr & 3)) * nx + x0 + xs;
            const u32x4_t gg = *(const u32x4_t*)&a.g[gv * COUT + 8 * c8];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                s_g[(8 * c8 + 2 * q) * kWgRows + v] = (uint16_t)(gg[q] & 0xffffu);
                s_g[(8 * c8 + 2 * q + 1) * kWgRows + v] = (uint16_t)(gg[q] >> 16);
            }
        }
        __syncthreads();
        if (tid < COUT)
            for (int v = 0; v < 256; ++v) bsum += bf2f(s_g[tid * kWgRows + v]);
        for (int ch = 0; ch < 8; ++ch) {  // 32 voxels: rows 2ch, 2ch+1 of the 16 (y,z) rows
            const int rr = 2 * ch + (g >> 1), xs0 = 8 * (g & 1);
            const int ry = rr & 3, rz = rr >> 2;
            u32x4_t af[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) af[mt] = *(const u32x4_t*)&s_g[(16 * mt + xl) * kWgRows + rr * 16 + xs0];
#pragma unroll
            for (int j = 0; j < NTW; ++j) {
                const int nt = wave + 4 * j;
                if (nt >= NT) break;
                const int n = 16 * nt + xl;
                int tap = n / CIN;
                const int ci = n - tap * CIN;
                tap = tap > 26 ? 26 : tap;  // padded columns are discarded
                const u32x4_t bv = *(const u32x4_t*)&s_in[(tap_dx(tap) * CIN + ci) * PLANE +
                                                          ((rz + tap_dz(tap)) * 6 + ry + tap_dy(tap)) * 16 + xs0];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) acc[j][mt] = mfma16(af[mt], bv, acc[j][mt]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NTW; ++j) {
        const int nt = wave + 4 * j;
        if (nt >= NT) break;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                a.part[((size_t)blockIdx.x * COUT + 16 * mt + 4 * g + q) * KP + 16 * nt + xl] = acc[j][mt][q];
    }
    if (tid < COUT) a.bpart[blockIdx.x * COUT + tid] = bsum;
}

// dW[co][k] (k < 27 CIN) and db[co]: partials added in workgroup order
__global__ void k_wgrad_reduce(const float* __restrict__ part, const float* __restrict__ bpart, int nwg,
                               int cout, int kp, int kreal, float* __restrict__ dw, float* __restrict__ db) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < cout * kreal) {
        const int co = i / kreal, k = i - co * kreal;
        float s = 0.f;
        for (int w = 0; w < nwg; ++w) s += part[((size_t)w * cout + co) * kp + k];
        dw[i] = s;
    } else if (i < cout * kreal + cout) {
        const int co = i - cout * kreal;
        float s = 0.f;
        for (int w = 0; w < nwg; ++w) s += bpart[w * cout + co];
        db[co] = s;
    }
}

// ---------------------------------------------------------------------------
// weights, elementwise and reduction kernels
// ---------------------------------------------------------------------------
// fp32 [cout][27 cin] -> bf16 [cout][kpad(cin)] (flip = 0), or the flipped
// transpose for the data gradient: bf16 [cin][kpad(cout)] with
// out[ci][(26 - tap) cout + co] = w[co][tap cin + ci] (flip = 1)
__global__ void k_prep_w(const float* __restrict__ w, int cin, int cout, int flip, uint16_t* __restrict__ out) {
    const int ci_out = flip ? cout : cin, rows = flip ? cin : cout;
    const int kp = kpad(ci_out);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= rows * kp) return;
    const int row = i / kp, k = i - row * kp;
    float v = 0.f;
    if (k < 27 * ci_out) {
        const int tap = k / ci_out, c = k - tap * ci_out;
        v = flip ? w[(size_t)c * 27 * cin + (26 - tap) * cin + row] : w[(size_t)row * 27 * cin + k];
    }
    out[i] = (uint16_t)f2bf(v);
}

// per-batch sum of squares of a fp32 field: partials [batch][256]
// per-block FP64 partial sums of x^2 over batch member blockIdx.y (nper a
// multiple of 4: the network's grids are)
__global__ __launch_bounds__(256) void k_sumsq(const float* __restrict__ x, int64_t nper, double* __restrict__ part) {
    __shared__ double sh[4];
    const int b = blockIdx.y;
    const float4* xb = (const float4*)(x + (size_t)b * nper);
    double s = 0;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nper / 4; i += (int64_t)gridDim.x * 256) {
        const float4 v = xb[i];
        s += ((double)v.x * v.x + (double)v.y * v.y) + ((double)v.z * v.z + (double)v.w * v.w);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}
// s[b] = sqrt(sum / nper); inv[b] = 1/s (0 for a zero field); sum[b] kept.
// One 256-thread block per batch member, fixed summation order.
__global__ __launch_bounds__(256) void k_rms_fin(const double* __restrict__ part, int nblk, int64_t nper,
                                                 int batch, float* __restrict__ s, float* __restrict__ inv,
                                                 double* __restrict__ sum) {
    __shared__ double sh[4];
    const int b = blockIdx.x;
    double t = 0;
    for (int i = threadIdx.x; i < nblk; i += 256) t += part[b * nblk + i];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        t = ((sh[0] + sh[1]) + sh[2]) + sh[3];
        const double r = std::sqrt(t / (double)nper);
        s[b] = (float)r;
        inv[b] = r > 0 ? (float)(1.0 / r) : 0.f;
        if (sum) sum[b] = t;
    }
}

// x[v] = (bf16(q), bf16(q feat[v])) with q = r[v] inv[b]: both channels are
// linear in r, so the bias-free network satisfies N(0) = 0
// x = [q, q * feat] in bf16, q = r / rms(r), batch member blockIdx.y (nper a
// multiple of 4)
__global__ void k_pack(const float* __restrict__ r, const float* __restrict__ feat, const float* __restrict__ inv,
                       int64_t nper, uint32_t* __restrict__ x) {
    const int b = blockIdx.y;
    const float ib = inv[b];
    const size_t o = (size_t)b * nper;
    const float4* r4 = (const float4*)(r + o);
    const float4* f4 = feat ? (const float4*)(feat + o) : nullptr;
    u32x4_t* x4 = (u32x4_t*)(x + o);
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nper / 4; i += (int64_t)gridDim.x * 256) {
        const float4 rv = r4[i];
        const float4 fv = f4 ? f4[i] : float4{0.f, 0.f, 0.f, 0.f};
        const float q0 = rv.x * ib, q1 = rv.y * ib, q2 = rv.z * ib, q3 = rv.w * ib;
        x4[i] = u32x4_t{pack2bf(q0, q0 * fv.x), pack2bf(q1, q1 * fv.y), pack2bf(q2, q2 * fv.z),
                        pack2bf(q3, q3 * fv.w)};
    }
}

// G = (skip + cscale * coarse[parent]) * (act > 0 ? 1 : kLeak), channels C (multiple of 8)
__global__ void k_grad_combine(int nx, int ny, int nz, int batch, int C, const uint16_t* __restrict__ act,
                               const uint16_t* __restrict__ skip, const uint16_t* __restrict__ coarse, float cscale,
                               uint16_t* __restrict__ out) {
    const int64_t nv = (int64_t)nx * ny * nz * batch, nch = nv * (C / 8);
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nch; i += (int64_t)gridDim.x * 256) {
        const int64_t v = i / (C / 8);
        const int c8 = (int)(i - v * (C / 8));
        u32x4_t s = skip ? *(const u32x4_t*)&skip[v * C + 8 * c8] : u32x4_t{0u, 0u, 0u, 0u};
        if (coarse) {
            const int x = (int)(v % nx), y = (int)((v / nx) % ny);
            const int64_t zb = v / ((int64_t)nx * ny);
            const int z = (int)(zb % nz), b = (int)(zb / nz);
            const size_t pv = (((size_t)b * (nz >> 1) + (z >> 1)) * (ny >> 1) + (y >> 1)) * (nx >> 1) + (x >> 1);
            s = add_scaled8(s, *(const u32x4_t*)&coarse[pv * C + 8 * c8], cscale);
        }
        const u32x4_t m = *(const u32x4_t*)&act[v * C + 8 * c8];
        u32x4_t o;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float lo = bf2f(s[q] & 0xffffu), hi = bf2f(s[q] >> 16);
            o[q] = f2bf(bf2f(m[q] & 0xffffu) > 0.f ? lo : kLeak * lo) |
                   (f2bf(bf2f(m[q] >> 16) > 0.f ? hi : kLeak * hi) << 16);
        }
        *(u32x4_t*)&out[v * C + 8 * c8] = o;
    }
}

// out[u] = sum of the 8 children of coarse voxel u (adjoint of nearest upsampling)
__global__ void k_sumpool8(int cx, int cy, int cz, int batch, int C, const uint16_t* __restrict__ in,
                           uint16_t* __restrict__ out) {
    const int64_t nv = (int64_t)cx * cy * cz * batch;
    const int fx = 2 * cx, fy = 2 * cy, fz = 2 * cz;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nv * C; i += (int64_t)gridDim.x * 256) {
        const int64_t u = i / C;
        const int c = (int)(i - u * C);
        const int x = (int)(u % cx), y = (int)((u / cx) % cy);
        const int64_t zb = u / ((int64_t)cx * cy);
        const int z = (int)(zb % cz), b = (int)(zb / cz);
        float s = 0.f;
        for (int d = 0; d < 8; ++d) {
            const size_t fv = (((size_t)b * fz + 2 * z + (d >> 2)) * fy + 2 * y + ((d >> 1) & 1)) * fx + 2 * x + (d & 1);
            s += bf2f(in[fv * C + c]);
        }
        out[i] = (uint16_t)f2bf(s);
    }
}

// target t = e / s[b]; per batch T2[b] = sum t^2 (partials [batch][256])
__global__ __launch_bounds__(256) void k_t2(const float* __restrict__ e, const float* __restrict__ inv, int64_t nper,
                                            double* __restrict__ part) {
    __shared__ double sh[4];
    const int b = blockIdx.y;
    double s = 0;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nper; i += (int64_t)gridDim.x * 256) {
        const double t = (double)(e[(size_t)b * nper + i] * inv[b]);
        s += t * t;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// Two-grid mode: the network runs on the 2x coarser grid.  Input: the mean
// over each 2x2x2 block of q = r inv[b] and of q feat; output: cell-centred
// trilinear prolongation (weights 3/4, 1/4 per axis, zero outside the box).
__global__ void k_restrict_pack(const float* __restrict__ r, const float* __restrict__ feat,
                                const float* __restrict__ inv, int nx, int ny, int nz, int batch,
                                uint32_t* __restrict__ x) {
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
    const int64_t nc = (int64_t)cx * cy * cz, nf = (int64_t)nx * ny * nz;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nc * batch; i += (int64_t)gridDim.x * 256) {
        const int b = (int)(i / nc);
        const int64_t v = i - b * nc;
        const int X = (int)(v % cx), Y = (int)((v / cx) % cy), Z = (int)(v / ((int64_t)cx * cy));
        float q = 0.f, qf = 0.f;
        for (int d = 0; d < 8; ++d) {
            const int64_t f = (int64_t)b * nf + ((int64_t)(2 * Z + (d >> 2)) * ny + 2 * Y + ((d >> 1) & 1)) * nx +
                              2 * X + (d & 1);
            const float qq = r[f] * inv[b];
            q += qq;
            qf += feat ? qq * feat[f] : 0.f;
        }
        x[i] = f2bf(0.125f * q) | (f2bf(0.125f * qf) << 16);
    }
}

// 1-D prolongation stencil of fine index i: coarse indices and weights
__device__ __forceinline__ void prol1(int i, int nc, int& i0, float& w0, int& i1, float& w1) {
    const int I = i >> 1;
    if (i & 1) { i0 = I; w0 = 0.75f; i1 = I + 1; w1 = 0.25f; }
    else { i0 = I - 1; w0 = 0.25f; i1 = I; w1 = 0.75f; }
    if (i0 < 0 || i0 >= nc) { i0 = 0; w0 = 0.f; }
    if (i1 < 0 || i1 >= nc) { i1 = 0; w1 = 0.f; }
}

__global__ void k_prolong(const float* __restrict__ zc, int nx, int ny, int nz, int batch,
                          const float* __restrict__ scale, float* __restrict__ z) {
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
    const int64_t nc = (int64_t)cx * cy * cz, nf = (int64_t)nx * ny * nz;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nf * batch; i += (int64_t)gridDim.x * 256) {
        const int b = (int)(i / nf);
        const int64_t v = i - b * nf;
        const int x = (int)(v % nx), y = (int)((v / nx) % ny), zz = (int)(v / ((int64_t)nx * ny));
        int xi[2], yi[2], zi[2];
        float xw[2], yw[2], zw[2];
        prol1(x, cx, xi[0], xw[0], xi[1], xw[1]);
        prol1(y, cy, yi[0], yw[0], yi[1], yw[1]);
        prol1(zz, cz, zi[0], zw[0], zi[1], zw[1]);
        const float* zb = zc + b * nc;
        float acc = 0.f;
        for (int c = 0; c < 2; ++c)
            for (int bb = 0; bb < 2; ++bb)
                for (int a = 0; a < 2; ++a)
                    acc += (zw[c] * yw[bb] * xw[a]) * zb[((int64_t)zi[c] * cy + yi[bb]) * cx + xi[a]];
        z[i] = (scale ? scale[b] : 1.f) * acc;
    }
}

// adjoint of k_prolong (without the scale): dzc[V] = sum_v w(v, V) dz[v]
__global__ void k_prolong_adj(const float* __restrict__ dz, int nx, int ny, int nz, int batch,
                              float* __restrict__ dzc) {
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
    const int64_t nc = (int64_t)cx * cy * cz, nf = (int64_t)nx * ny * nz;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nc * batch; i += (int64_t)gridDim.x * 256) {
        const int b = (int)(i / nc);
        const int64_t v = i - b * nc;
        const int X = (int)(v % cx), Y = (int)((v / cx) % cy), Z = (int)(v / ((int64_t)cx * cy));
        // fine indices 2I-1 .. 2I+2 with weights 1/4, 3/4, 3/4, 1/4
        const float w4[4] = {0.25f, 0.75f, 0.75f, 0.25f};
        const float* db = dz + b * nf;
        float acc = 0.f;
        for (int c = 0; c < 4; ++c) {
            const int fz = 2 * Z - 1 + c;
            if (fz < 0 || fz >= nz) continue;
            for (int bb = 0; bb < 4; ++bb) {
                const int fy = 2 * Y - 1 + bb;
                if (fy < 0 || fy >= ny) continue;
                for (int a = 0; a < 4; ++a) {
                    const int fx = 2 * X - 1 + a;
                    if (fx < 0 || fx >= nx) continue;
                    acc += (w4[c] * w4[bb] * w4[a]) * db[((int64_t)fz * ny + fy) * nx + fx];
                }
            }
        }
        dzc[i] = acc;
    }
}

// A-norm loss: Ad = A_b (z - t) for every voxel of sample b (variable
// 7-point operator with the sample's own faces, nnx_grid layout, Dirichlet
// 0), and partials of t . A_b t per sample ([batch][gridDim.x]).
__global__ __launch_bounds__(256) void k_anorm(int nx, int ny, int nz, const float* __restrict__ z,
                                               const float* __restrict__ e, const float* __restrict__ inv,
                                               const float* __restrict__ kx, const float* __restrict__ ky,
                                               const float* __restrict__ kz, float* __restrict__ Ad,
                                               double* __restrict__ part) {
    __shared__ double sh[4];
    const int b = blockIdx.y;
    const int64_t nper = (int64_t)nx * ny * nz;
    const float sb = inv[b];
    const float* zb = z + b * nper;
    const float* eb = e + b * nper;
    const float* fx = kx + (int64_t)b * (nx + 1) * ny * nz;
    const float* fy = ky + (int64_t)b * nx * (ny + 1) * nz;
    const float* fz = kz + (int64_t)b * nx * ny * (nz + 1);
    double acc = 0;
    for (int64_t id = blockIdx.x * 256ll + threadIdx.x; id < nper; id += (int64_t)gridDim.x * 256) {
        const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((int64_t)nx * ny));
        const int64_t ix = i + (int64_t)(nx + 1) * (j + (int64_t)ny * k), iy = i + (int64_t)nx * (j + (int64_t)(ny + 1) * k);
        const float kxm = fx[ix], kxp = fx[ix + 1], kym = fy[iy], kyp = fy[iy + nx], kzm = fz[id],
                    kzp = fz[id + (int64_t)nx * ny];
        const float dg = 1.f + kxm + kxp + kym + kyp + kzm + kzp;
        auto tv = [&](int64_t q) { return eb[q] * sb; };
        auto dv = [&](int64_t q) { return zb[q] - eb[q] * sb; };
        float sd = 0.f, st = 0.f;
        if (i > 0) { sd += kxm * dv(id - 1); st += kxm * tv(id - 1); }
        if (i < nx - 1) { sd += kxp * dv(id + 1); st += kxp * tv(id + 1); }
        if (j > 0) { sd += kym * dv(id - nx); st += kym * tv(id - nx); }
        if (j < ny - 1) { sd += kyp * dv(id + nx); st += kyp * tv(id + nx); }
        if (k > 0) { sd += kzm * dv(id - (int64_t)nx * ny); st += kzm * tv(id - (int64_t)nx * ny); }
        if (k < nz - 1) { sd += kzp * dv(id + (int64_t)nx * ny); st += kzp * tv(id + (int64_t)nx * ny); }
        Ad[b * nper + id] = dg * dv(id) - sd;
        const float t = tv(id);
        acc += (double)t * (dg * t - st);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// Per-sample operator pieces for the smoothed A-norm loss (faces of sample
// b in nnx_grid layout, Dirichlet 0, D = 1 + the six faces):
//   mode 0: out = A x;  mode 1: out = x - w D^-1 A x  (S x);
//   mode 2: out = D^-1 x;  mode 3: out = x2 - w A x  (with x = D^-1 x2: S^T x2).
constexpr float kSmoothWTrain = 2.0f / 3.0f;
__global__ __launch_bounds__(256) void k_sample_op(int nx, int ny, int nz, int batch, int mode,
                                                   const float* __restrict__ x, const float* __restrict__ x2,
                                                   const float* __restrict__ kx, const float* __restrict__ ky,
                                                   const float* __restrict__ kz, float* __restrict__ out) {
    const int64_t nper = (int64_t)nx * ny * nz;
    for (int64_t g = blockIdx.x * 256ll + threadIdx.x; g < nper * batch; g += (int64_t)gridDim.x * 256) {
        const int b = (int)(g / nper);
        const int64_t id = g - b * nper;
        const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / ((int64_t)nx * ny));
        const float* fx = kx + (int64_t)b * (nx + 1) * ny * nz;
        const float* fy = ky + (int64_t)b * nx * (ny + 1) * nz;
        const float* fz = kz + (int64_t)b * nx * ny * (nz + 1);
        const int64_t ix = i + (int64_t)(nx + 1) * (j + (int64_t)ny * k), iy = i + (int64_t)nx * (j + (int64_t)(ny + 1) * k);
        const float kxm = fx[ix], kxp = fx[ix + 1], kym = fy[iy], kyp = fy[iy + nx], kzm = fz[id],
                    kzp = fz[id + (int64_t)nx * ny];
        const float dg = 1.f + kxm + kxp + kym + kyp + kzm + kzp;
        const float* xb = x + b * nper;
        if (mode == 2) {
            out[g] = xb[id] / dg;
            continue;
        }
        float sn = 0.f;
        if (i > 0) sn += kxm * xb[id - 1];
        if (i < nx - 1) sn += kxp * xb[id + 1];
        if (j > 0) sn += kym * xb[id - nx];
        if (j < ny - 1) sn += kyp * xb[id + nx];
        if (k > 0) sn += kzm * xb[id - (int64_t)nx * ny];
        if (k < nz - 1) sn += kzp * xb[id + (int64_t)nx * ny];
        const float ax = dg * xb[id] - sn;
        if (mode == 0) out[g] = ax;
        else if (mode == 1) out[g] = xb[id] - kSmoothWTrain * ax / dg;
        else out[g] = x2[g] - kSmoothWTrain * ax;
    }
}
// d = z - e inv[b]
__global__ void k_diff(const float* __restrict__ z, const float* __restrict__ e, const float* __restrict__ inv,
                       int64_t nper, int batch, float* __restrict__ d) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nper * batch; i += (int64_t)gridDim.x * 256)
        d[i] = z[i] - e[i] * inv[i / nper];
}

// Loss (1/B) sum_b |d_b|^2 / T2_b (or, with Ad, d_b.A_b d_b / t_b.A_b t_b)
// with d = z - e inv[b], and dZ = 2 d / (B T2_b) (or 2 A_b d_b / (B t_b.A_b t_b));
// one loss partial per block.
constexpr int kHeadBlocks = 512;
constexpr int kSumBlocks = 1024;  // blocks per batch member of the sum-of-squares reductions
__global__ __launch_bounds__(256) void k_loss_grad(const float* __restrict__ z, const float* __restrict__ e,
                                                   const float* __restrict__ inv, const double* __restrict__ t2,
                                                   const float* __restrict__ Ad, int64_t nper, int batch,
                                                   float* __restrict__ dz, double* __restrict__ lpart) {
    __shared__ double sh[4];
    double acc = 0;
    const int64_t n = nper * batch;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int b = (int)(i / nper);
        const float diff = z[i] - e[i] * inv[b];
        const float gd = Ad ? Ad[i] : diff;
        dz[i] = (float)(2.0 * gd / ((double)batch * t2[b]));
        acc += (double)diff * gd / t2[b];
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) lpart[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// Through the head: G0 = hw dZ act'(d0); partials of dhw (16) and dhb.
__global__ __launch_bounds__(256) void k_head_bwd(const float* __restrict__ dzv, int64_t n,
                                                  const uint16_t* __restrict__ d0, const float* __restrict__ head,
                                                  uint16_t* __restrict__ g0, double* __restrict__ part) {
    __shared__ double sh[4][17];
    double acc[17];
    for (int q = 0; q < 17; ++q) acc[q] = 0;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float dz = dzv[i];
        acc[16] += dz;
        uint32_t o[8];
        for (int c8 = 0; c8 < 2; ++c8) {
            const u32x4_t a4 = *(const u32x4_t*)&d0[i * 16 + 8 * c8];
            for (int q = 0; q < 4; ++q) {
                const float lo = bf2f(a4[q] & 0xffffu), hi = bf2f(a4[q] >> 16);
                const int c = 8 * c8 + 2 * q;
                acc[c] += (double)lo * dz;
                acc[c + 1] += (double)hi * dz;
                const float gl = head[c] * dz, gh = head[c + 1] * dz;
                const uint32_t glo = f2bf(lo > 0.f ? gl : kLeak * gl);
                const uint32_t ghi = f2bf(hi > 0.f ? gh : kLeak * gh);
                o[4 * c8 + q] = glo | (ghi << 16);
            }
        }
        *(u32x4_t*)&g0[i * 16] = u32x4_t{o[0], o[1], o[2], o[3]};
        *(u32x4_t*)&g0[i * 16 + 8] = u32x4_t{o[4], o[5], o[6], o[7]};
    }
    for (int q = 0; q < 17; ++q) {
        double s = acc[q];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        acc[q] = s;
    }
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 17; ++q) sh[threadIdx.x >> 6][q] = acc[q];
    __syncthreads();
    if (threadIdx.x < 17) {
        const int q = threadIdx.x;
        part[blockIdx.x * 17 + q] = ((sh[0][q] + sh[1][q]) + sh[2][q]) + sh[3][q];
    }
}
// head gradients (dhw[16], dhb) into grad[], the loss into loss[0]
__global__ void k_head_fin(const double* __restrict__ part, const double* __restrict__ lpart, int nblk,
                           float* __restrict__ ghead, double* loss, int batch) {
    const int q = threadIdx.x;
    if (q < 17) {
        double s = 0;
        for (int i = 0; i < nblk; ++i) s += part[i * 17 + q];
        ghead[q] = (float)s;
    } else if (q == 17) {
        double s = 0;
        for (int i = 0; i < nblk; ++i) s += lpart[i];
        *loss = s / batch;
    }
}

__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

}  // namespace nnx

using namespace nnx;

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

struct Layer {
    int cin, cout;
    int64_t w_off, b_off;  // offsets in the parameter vector
};
// e0, e1, e2, d1, d0, then the head (16 weights + bias)
const Layer kLayers[5] = {{2, 16, 0, 0}, {16, 32, 0, 0}, {32, 32, 0, 0}, {32, 16, 0, 0}, {16, 16, 0, 0}};

struct Plan {
    Layer L[5];
    int64_t head_off, total;
    Plan() {
        int64_t o = 0;
        for (int i = 0; i < 5; ++i) {
            L[i] = kLayers[i];
            L[i].w_off = o;
            o += (int64_t)L[i].cout * 27 * L[i].cin;
            L[i].b_off = o;
            o += L[i].cout;
        }
        head_off = o;
        total = o + 17;
    }
};
const Plan& plan() {
    static Plan p;
    return p;
}

}  // namespace

struct nnx_unet {
    float *p, *grad, *m, *v;     // fp32 parameters, gradient, Adam moments
    uint16_t* wf[5];             // bf16 forward weights [cout][kpad(cin)]
    uint16_t* wb[5];             // bf16 flipped weights [cin][kpad(cout)]
    int64_t step;
    bool dirty;                  // bf16 copies out of date
    int coarse;                  // 1: two-grid mode (network on the 2x coarser grid)
};

namespace {

// workgroups of 256 threads of kernel k resident at once on the device
template <class K> int64_t resident_wgs(nnx_ctx* c, K k) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1)
        cus = 1;
    return (int64_t)per_cu * cus;
}

// Chunk length zc (even: chunks start on even planes, as the pooled pairs
// need) minimising rounds x steps: a launch of W workgroups per chunk row
// runs ceil(W nch / resident) rounds of workgroups that each walk zc + 2
// input planes in whole groups of six steps.
inline int conv_chunk(int nz, int64_t per_chunk, int64_t resident) {
    int best = nz + (nz & 1);
    double best_t = 1e300;
    for (int n = 1; n <= std::max(1, nz / 2); ++n) {
        int zc = (int)ceil_div(nz, n);
        zc += zc & 1;
        if (ceil_div(nz, zc) != n) continue;
        const double t = (double)ceil_div(per_chunk * n, resident) * (6 * ceil_div(zc + 2, 6));
        if (t < best_t) best_t = t, best = zc;
    }
    return best;
}

template <int CIN, int COUT, int RY, int MT, int WX>
nnx_status launch_conv_z(nnx_ctx* c, const ConvArgs& a) {
    constexpr int TX = 16 * WX, TY = (4 / WX) * RY, NG = COUT / (16 * MT);
    NNX_CHECK(c, NG == 1 || !a.head, "the 1x1 head needs every channel in one workgroup");
    const bool co = ZK<CIN>::WIDE && a.coarse;
    // residency of each form, measured once
    static const int64_t res0 = resident_wgs(c, k_conv3z<CIN, COUT, RY, MT, false, WX>);
    static const int64_t res1 =
        ZK<CIN>::WIDE ? resident_wgs(c, k_conv3z<CIN, COUT, RY, MT, ZK<CIN>::WIDE, WX>) : res0;
    const int64_t cols = (int64_t)(a.nx / TX) * (a.ny / TY);
    const int zc = conv_chunk(a.nz, cols * a.batch * NG, co ? res1 : res0);
    const int nch = (int)ceil_div(a.nz, zc);
    const dim3 grid((unsigned)cols, (unsigned)(nch * NG), (unsigned)a.batch);
    if (co)
        k_conv3z<CIN, COUT, RY, MT, ZK<CIN>::WIDE, WX><<<grid, 256, 0, c->stream>>>(a, zc);
    else
        k_conv3z<CIN, COUT, RY, MT, false, WX><<<grid, 256, 0, c->stream>>>(a, zc);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

template <int CIN, int COUT>
nnx_status launch_conv(nnx_ctx* c, const ConvArgs& a) {
    constexpr int TY = ConvTile<CIN>::TY, TZ = ConvTile<CIN>::TZ;
    // the z-marching form where its 16 x 4RY columns tile the grid (RY and
    // the m-tiles per workgroup keep weights, accumulators and the staged
    // plane within 256 registers), else one 16 x TY x TZ tile per workgroup
    constexpr int RY = CIN == 32 ? 2 : 4, MTZ = 1;
    // the 2-channel input in 64-voxel-wide, 4-row columns where they tile
    // the grid (its rows are 4 bytes a voxel: 16-voxel columns would read
    // each input cache line three to four times)
    constexpr int WXW = CIN == 2 ? 4 : 1;
    const bool wide = WXW > 1 && a.nx % (16 * WXW) == 0 && a.ny % ((4 / WXW) * RY) == 0;
    const bool zform = wide || a.ny % (4 * RY) == 0;
    NNX_CHECK(c, a.nx > 0 && a.ny > 0 && a.nz > 0 && a.nx % 16 == 0 && a.batch >= 1 &&
                     (zform || (a.ny % TY == 0 && a.nz % TZ == 0)),
              "conv %d->%d: dims (%d,%d,%d) must be multiples of (16,%d,1) or (16,%d,%d)", CIN, COUT, a.nx, a.ny,
              a.nz, 4 * RY, TY, TZ);
    if (a.pool || a.coarse)
        NNX_CHECK(c, a.nx % 32 == 0 && a.ny % 2 == 0 && a.nz % 2 == 0, "pooled / upsampled conv needs even dims");
    if (wide) return launch_conv_z<CIN, COUT, RY, MTZ, WXW>(c, a);
    if (zform) return launch_conv_z<CIN, COUT, RY, MTZ, 1>(c, a);
    dim3 grid(a.nx / 16, a.ny / TY, (a.nz / TZ) * a.batch);
    k_conv3<CIN, COUT><<<grid, 256, 0, c->stream>>>(a);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

nnx_status conv_dispatch(nnx_ctx* c, int cin, int cout, const ConvArgs& a) {
    if (cin == 2 && cout == 16) return launch_conv<2, 16>(c, a);
    if (cin == 16 && cout == 16) return launch_conv<16, 16>(c, a);
    if (cin == 16 && cout == 32) return launch_conv<16, 32>(c, a);
    if (cin == 32 && cout == 16) return launch_conv<32, 16>(c, a);
    if (cin == 32 && cout == 32) return launch_conv<32, 32>(c, a);
    return fail(c, NNX_ERR_UNSUPPORTED, "conv %d->%d not instantiated", cin, cout);
}

constexpr int kWgMax = 512;

template <int CIN, int COUT>
nnx_status launch_wgrad(nnx_ctx* c, WgArgs a, float* dw, float* db) {
    NNX_CHECK(c, a.nx % 16 == 0 && a.ny % 4 == 0 && a.nz % 4 == 0, "wgrad dims must be multiples of (16,4,4)");
    if (a.coarse) NNX_CHECK(c, a.nx % 32 == 0 && a.ny % 8 == 0 && a.nz % 8 == 0, "upsampled wgrad needs even dims");
    a.ntiles = (a.nx / 16) * (a.ny / 4) * (a.nz / 4) * a.batch;
    const int nwg = std::min(a.ntiles, kWgMax);
    void *part, *bpart;
    NNX_TRY(scratch(c, "un_wg_part", (size_t)nwg * COUT * kpad(CIN) * 4, false, &part));
    NNX_TRY(scratch(c, "un_wg_bpart", (size_t)nwg * COUT * 4, false, &bpart));
    a.part = (float*)part;
    a.bpart = (float*)bpart;
    k_wgrad<CIN, COUT><<<nwg, 256, 0, c->stream>>>(a);
    const int tot = COUT * 27 * CIN + COUT;
    k_wgrad_reduce<<<ceil_div(tot, 256), 256, 0, c->stream>>>(a.part, a.bpart, nwg, COUT, kpad(CIN), 27 * CIN,
                                                                 dw, db);
    NNX_LAUNCHED(c);
    return NNX_OK;
}

nnx_status wgrad_dispatch(nnx_ctx* c, int cin, int cout, const WgArgs& a, float* dw, float* db) {
    if (cin == 2 && cout == 16) return launch_wgrad<2, 16>(c, a, dw, db);
    if (cin == 16 && cout == 16) return launch_wgrad<16, 16>(c, a, dw, db);
    if (cin == 16 && cout == 32) return launch_wgrad<16, 32>(c, a, dw, db);
    if (cin == 32 && cout == 16) return launch_wgrad<32, 16>(c, a, dw, db);
    if (cin == 32 && cout == 32) return launch_wgrad<32, 32>(c, a, dw, db);
    return fail(c, NNX_ERR_UNSUPPORTED, "wgrad %d->%d not instantiated", cin, cout);
}

unsigned eblocks(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 4096)); }

nnx_status prep_weights(nnx_ctx* c, nnx_unet* u) {
    if (!u->dirty) return NNX_OK;
    const Plan& P = plan();
    for (int i = 0; i < 5; ++i) {
        const Layer& L = P.L[i];
        const int nf = L.cout * kpad(L.cin), nb = L.cin * kpad(L.cout);
        k_prep_w<<<ceil_div(nf, 256), 256, 0, c->stream>>>(u->p + L.w_off, L.cin, L.cout, 0, u->wf[i]);
        k_prep_w<<<ceil_div(nb, 256), 256, 0, c->stream>>>(u->p + L.w_off, L.cin, L.cout, 1, u->wb[i]);
    }
    NNX_LAUNCHED(c);
    u->dirty = false;
    return NNX_OK;
}

struct Acts {
    uint16_t *x, *e0, *p0, *e1, *p1, *e2, *d1, *d0;
    float *s, *inv, *z, *zc;
    double* part;
};

// n0: voxels at the network's finest level; nf: voxels of the caller's grid
nnx_status acts(nnx_ctx* c, int64_t n0, int64_t nf, int batch, bool train, bool coarse, Acts* A) {
    const int64_t n1 = n0 / 8, n2 = n0 / 64;
    void* p;
    NNX_TRY(scratch(c, "un_x", n0 * 4, false, &p));
    A->x = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_e0", n0 * 32, false, &p));
    A->e0 = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_p0", n1 * 32, false, &p));
    A->p0 = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_e1", n1 * 64, false, &p));
    A->e1 = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_p1", n2 * 64, false, &p));
    A->p1 = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_e2", n2 * 64, false, &p));
    A->e2 = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_d1", n1 * 32, false, &p));
    A->d1 = (uint16_t*)p;
    A->d0 = nullptr;
    A->z = nullptr;
    A->zc = nullptr;
    if (train) {
        NNX_TRY(scratch(c, "un_d0", n0 * 32, false, &p));
        A->d0 = (uint16_t*)p;
        NNX_TRY(scratch(c, "un_z", nf * 4, false, &p));
        A->z = (float*)p;
    }
    if (coarse) {
        NNX_TRY(scratch(c, "un_zc", n0 * 4, false, &p));
        A->zc = (float*)p;
    }
    NNX_TRY(scratch(c, "un_s", 4096, false, &p));
    A->s = (float*)p;
    A->inv = (float*)p + 512;
    NNX_TRY(scratch(c, "un_part", (size_t)std::max(kSumBlocks * batch, kHeadBlocks * 18) * 8 + 4096, false, &p));
    A->part = (double*)p;
    return NNX_OK;
}

nnx_status check_dims(nnx_ctx* c, const nnx_unet* u, int nx, int ny, int nz, int batch) {
    NNX_CHECK(c, nx > 0 && ny > 0 && nz > 0 && batch >= 1 && batch <= 256, "bad dims or batch");
    const int f = u->coarse ? 2 : 1;
    NNX_CHECK(c, nx % (64 * f) == 0 && ny % (32 * f) == 0 && nz % (32 * f) == 0,
              "U-Net grids need nx %% %d == 0 and ny, nz %% %d == 0 (got %d, %d, %d)", 64 * f, 32 * f, nx, ny,
              nz);
    NNX_CHECK(c, (int64_t)nx * ny * nz * batch < ((int64_t)1 << 31), "grid too large");
    return NNX_OK;
}

// Forward pass on the caller's grid (nx, ny, nz); z = rms(r) * N(r / rms(r))
// (train: z unscaled).  In two-grid mode the network runs on the 2x coarser
// grid between restriction and prolongation.
nnx_status forward(nnx_ctx* c, nnx_unet* u, int nx, int ny, int nz, int batch, const float* r, const float* feat,
                   float* z, bool train, Acts& A) {
    const Plan& P = plan();
    const int64_t nper = (int64_t)nx * ny * nz;
    const int f = u->coarse ? 2 : 1;
    const int mx = nx / f, my = ny / f, mz = nz / f;  // the network's finest level
    const int64_t n0 = (int64_t)mx * my * mz * batch;
    NNX_TRY(prep_weights(c, u));
    k_sumsq<<<dim3(kSumBlocks, batch), 256, 0, c->stream>>>(r, nper, A.part);
    k_rms_fin<<<batch, 256, 0, c->stream>>>(A.part, kSumBlocks, nper, batch, A.s, A.inv, nullptr);
    if (u->coarse)
        k_restrict_pack<<<eblocks(n0), 256, 0, c->stream>>>(r, feat, A.inv, nx, ny, nz, batch, (uint32_t*)A.x);
    else
        k_pack<<<dim3((unsigned)std::min<int64_t>(ceil_div(nper / 4, 256), 2048), batch), 256, 0, c->stream>>>(
            r, feat, A.inv, nper, (uint32_t*)A.x);
    NNX_LAUNCHED(c);
    ConvArgs a{};
    a.batch = batch;
    a.relu = 1;
    // e0 (+ p0)
    a.nx = mx, a.ny = my, a.nz = mz;
    a.in = A.x, a.w = u->wf[0], a.bias = u->p + P.L[0].b_off, a.out = A.e0, a.pool = A.p0;
    NNX_TRY(conv_dispatch(c, 2, 16, a));
    // e1 (+ p1)
    a.nx = mx / 2, a.ny = my / 2, a.nz = mz / 2;
    a.in = A.p0, a.w = u->wf[1], a.bias = u->p + P.L[1].b_off, a.out = A.e1, a.pool = A.p1;
    NNX_TRY(conv_dispatch(c, 16, 32, a));
    // e2
    a.nx = mx / 4, a.ny = my / 4, a.nz = mz / 4;
    a.in = A.p1, a.w = u->wf[2], a.bias = u->p + P.L[2].b_off, a.out = A.e2, a.pool = nullptr;
    NNX_TRY(conv_dispatch(c, 32, 32, a));
    // d1 = act(conv(e1 + up(e2)))
    a.nx = mx / 2, a.ny = my / 2, a.nz = mz / 2;
    a.in = A.e1, a.coarse = A.e2, a.cscale = 1.f, a.w = u->wf[3], a.bias = u->p + P.L[3].b_off, a.out = A.d1;
    NNX_TRY(conv_dispatch(c, 32, 16, a));
    // d0 = act(conv(e0 + up(d1))) -> head
    a.nx = mx, a.ny = my, a.nz = mz;
    a.in = A.e0, a.coarse = A.d1, a.w = u->wf[4], a.bias = u->p + P.L[4].b_off, a.out = train ? A.d0 : nullptr;
    a.head = u->p + P.head_off;
    a.hscale = (train || u->coarse) ? nullptr : A.s;
    a.hout = u->coarse ? A.zc : z;
    NNX_TRY(conv_dispatch(c, 16, 16, a));
    if (u->coarse) {
        k_prolong<<<eblocks(nper * batch), 256, 0, c->stream>>>(A.zc, nx, ny, nz, batch, train ? nullptr : A.s, z);
        NNX_LAUNCHED(c);
    }
    return NNX_OK;
}

}  // namespace

nnx_status nnx_unet_forward_internal(nnx_ctx* c, nnx_unet* u, int nx, int ny, int nz, int batch, const float* r,
                                     const float* feat, float* z) {
    NNX_TRY(check_dims(c, u, nx, ny, nz, batch));
    const int f = u->coarse ? 2 : 1;
    const int64_t nf = (int64_t)nx * ny * nz * batch;
    Acts A;
    NNX_TRY(acts(c, nf / (f * f * f), nf, batch, false, u->coarse != 0, &A));
    return forward(c, u, nx, ny, nz, batch, r, feat, z, false, A);
}

extern "C" {

int64_t nnx_unet_param_count(void) { return plan().total; }

nnx_status nnx_unet_create(nnx_ctx* c, const float* h_params, uint64_t seed, nnx_unet** out) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, out, "null out");
    NNX_HIP(c, hipSetDevice(c->device));
    const Plan& P = plan();
    std::vector<float> h(P.total, 0.f);
    if (h_params) {
        std::memcpy(h.data(), h_params, P.total * 4);
    } else {
        // He-uniform weights from a counter-based hash of (seed, index); zero biases
        auto unif = [&](uint64_t i) {
            uint64_t x = (seed + 0x9e3779b97f4a7c15ull) ^ (i * 0xbf58476d1ce4e5b9ull);
            x ^= x >> 31;
            x *= 0x94d049bb133111ebull;
            x ^= x >> 29;
            x *= 0xbf58476d1ce4e5b9ull;
            x ^= x >> 32;
            return (double)(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
        };
        for (int l = 0; l < 5; ++l) {
            const Layer& L = P.L[l];
            const double bound = std::sqrt(6.0 / (27.0 * L.cin));
            for (int64_t i = 0; i < (int64_t)L.cout * 27 * L.cin; ++i) h[L.w_off + i] = (float)(bound * unif(L.w_off + i));
        }
        // a small head: the untrained network's output starts well below the target
        const double bound = 0.1 * std::sqrt(3.0 / 16.0);
        for (int i = 0; i < 16; ++i) h[P.head_off + i] = (float)(bound * unif(P.head_off + i));
    }
    nnx_unet* u = new nnx_unet();
    size_t bytes = (size_t)P.total * 4;
    hipError_t e = hipSuccess;
    e = e ? e : hipMalloc(&u->p, bytes);
    e = e ? e : hipMalloc(&u->grad, bytes);
    e = e ? e : hipMalloc(&u->m, bytes);
    e = e ? e : hipMalloc(&u->v, bytes);
    for (int l = 0; l < 5 && !e; ++l) {
        const Layer& L = P.L[l];
        e = hipMalloc(&u->wf[l], (size_t)L.cout * kpad(L.cin) * 2);
        if (!e) e = hipMalloc(&u->wb[l], (size_t)L.cin * kpad(L.cout) * 2);
    }
    if (!e) e = hipMemcpyAsync(u->p, h.data(), bytes, hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipMemsetAsync(u->m, 0, bytes, c->stream);
    if (!e) e = hipMemsetAsync(u->v, 0, bytes, c->stream);
    if (!e) e = hipMemsetAsync(u->grad, 0, bytes, c->stream);
    if (!e) e = hipStreamSynchronize(c->stream);
    if (e) {
        delete u;
        return fail(c, NNX_ERR_HIP, "nnx_unet_create: %s", hipGetErrorString(e));
    }
    u->step = 0;
    u->dirty = true;
    u->coarse = 0;
    *out = u;
    return NNX_OK;
}

nnx_status nnx_unet_destroy(nnx_ctx* c, nnx_unet* u) {
    if (!u) return NNX_OK;
    if (c) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(u->p);
    (void)hipFree(u->grad);
    (void)hipFree(u->m);
    (void)hipFree(u->v);
    for (int l = 0; l < 5; ++l) {
        (void)hipFree(u->wf[l]);
        (void)hipFree(u->wb[l]);
    }
    delete u;
    return NNX_OK;
}

nnx_status nnx_unet_get_params(nnx_ctx* c, const nnx_unet* u, float* h_params) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, u && h_params, "null argument");
    NNX_HIP(c, hipMemcpyAsync(h_params, u->p, plan().total * 4, hipMemcpyDeviceToHost, c->stream));
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    return NNX_OK;
}

nnx_status nnx_unet_apply(nnx_ctx* c, nnx_unet* u, int nx, int ny, int nz, int batch, const float* d_r,
                          const float* d_feat, float* d_z) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, u && d_r && d_z, "null argument");
    return nnx_unet_forward_internal(c, u, nx, ny, nz, batch, d_r, d_feat, d_z);
}

nnx_status nnx_unet_set_coarse(nnx_ctx* c, nnx_unet* u, int coarse) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, u, "null network");
    u->coarse = coarse ? 1 : 0;
    return NNX_OK;
}

nnx_status nnx_unet_train_step(nnx_ctx* c, nnx_unet* u, int nx, int ny, int nz, int batch, const float* d_r,
                               const float* d_feat, const float* d_e, const float* d_kx, const float* d_ky,
                               const float* d_kz, int smooth, double lr, double* h_loss) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, u && d_r && d_e, "null argument");
    NNX_TRY(check_dims(c, u, nx, ny, nz, batch));
    const Plan& P = plan();
    const int64_t nper = (int64_t)nx * ny * nz, nf = nper * batch;
    // the network's levels (the caller's grid, or the 2x coarser one)
    const int f = u->coarse ? 2 : 1;
    const int mx = nx / f, my = ny / f, mz = nz / f;
    const int64_t n0 = nf / (f * f * f), n1 = n0 / 8, n2 = n0 / 64;
    Acts A;
    NNX_TRY(acts(c, n0, nf, batch, true, u->coarse != 0, &A));
    NNX_TRY(forward(c, u, nx, ny, nz, batch, d_r, d_feat, A.z, true, A));
    void *p, *lossp;
    NNX_TRY(scratch(c, "un_g0", n0 * 32, false, &p));
    uint16_t* G0 = (uint16_t*)p;  // later: gradient of e0
    NNX_TRY(scratch(c, "un_u0", n0 * 32, false, &p));
    uint16_t* U0 = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_l1a", n1 * 64, false, &p));
    uint16_t* L1a = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_l1b", n1 * 64, false, &p));
    uint16_t* L1b = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_l1c", n1 * 64, false, &p));
    uint16_t* L1c = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_l2a", n2 * 64, false, &p));
    uint16_t* L2a = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_l2b", n2 * 64, false, &p));
    uint16_t* L2b = (uint16_t*)p;
    NNX_TRY(scratch(c, "un_loss", 64, false, &lossp));
    NNX_TRY(scratch(c, "un_t2", 256 * 8, false, &p));
    double* t2 = (double*)p;
    // loss and head
    const bool anorm = d_kx != nullptr;
    NNX_CHECK(c, !anorm || (d_ky && d_kz), "A-norm loss needs kx, ky and kz");
    float* Ad = nullptr;
    if (anorm) {
        NNX_TRY(scratch(c, "un_ad", nf * 4, false, &p));
        Ad = (float*)p;
        k_anorm<<<dim3(kSumBlocks, batch), 256, 0, c->stream>>>(nx, ny, nz, A.z, d_e, A.inv, d_kx, d_ky, d_kz, Ad,
                                                                A.part);
        if (smooth > 0) {
            // smoothed A-norm: |S^m d|_A^2 with S = I - w D^-1 A; gradient
            // 2 (S^T)^m A S^m d (replaces Ad; the loss kernel's d . Ad is
            // then |S^m d|_A^2)
            NNX_CHECK(c, smooth <= 4, "at most 4 smoothing sweeps in the loss");
            void *q1, *q2;
            NNX_TRY(scratch(c, "un_sm1", nf * 4, false, &q1));
            NNX_TRY(scratch(c, "un_sm2", nf * 4, false, &q2));
            float *U = (float*)q1, *V = (float*)q2;
            const unsigned eb = eblocks(nf);
            k_diff<<<eb, 256, 0, c->stream>>>(A.z, d_e, A.inv, nper, batch, Ad);
            float *src = Ad, *dst = U;
            for (int m = 0; m < smooth; ++m) {  // src <- S src
                k_sample_op<<<eb, 256, 0, c->stream>>>(nx, ny, nz, batch, 1, src, nullptr, d_kx, d_ky, d_kz, dst);
                std::swap(src, dst);
            }
            k_sample_op<<<eb, 256, 0, c->stream>>>(nx, ny, nz, batch, 0, src, nullptr, d_kx, d_ky, d_kz, dst);
            std::swap(src, dst);  // src = A S^m d
            for (int m = 0; m < smooth; ++m) {  // src <- S^T src
                k_sample_op<<<eb, 256, 0, c->stream>>>(nx, ny, nz, batch, 2, src, nullptr, d_kx, d_ky, d_kz, dst);
                k_sample_op<<<eb, 256, 0, c->stream>>>(nx, ny, nz, batch, 3, dst, src, d_kx, d_ky, d_kz, V);
                std::swap(src, V);
                if (V == Ad) std::swap(V, dst);
            }
            if (src != Ad) NNX_HIP(c, hipMemcpyAsync(Ad, src, nf * 4, hipMemcpyDeviceToDevice, c->stream));
            NNX_LAUNCHED(c);
        }
    } else {
        k_t2<<<dim3(kSumBlocks, batch), 256, 0, c->stream>>>(d_e, A.inv, nper, A.part);
    }
    k_rms_fin<<<batch, 256, 0, c->stream>>>(A.part, kSumBlocks, 1, batch, A.s + 256, A.inv + 256, t2);
    // dZ on the caller's grid, then (two-grid) through the prolongation
    void *dzp, *dzcp, *lpart;
    NNX_TRY(scratch(c, "un_dz", nf * 4, false, &dzp));
    NNX_TRY(scratch(c, "un_lpart", kHeadBlocks * 8, false, &lpart));
    k_loss_grad<<<kHeadBlocks, 256, 0, c->stream>>>(A.z, d_e, A.inv, t2, Ad, nper, batch, (float*)dzp,
                                                    (double*)lpart);
    float* dzn = (float*)dzp;
    if (u->coarse) {
        NNX_TRY(scratch(c, "un_dzc", n0 * 4, false, &dzcp));
        k_prolong_adj<<<eblocks(n0), 256, 0, c->stream>>>((const float*)dzp, nx, ny, nz, batch, (float*)dzcp);
        dzn = (float*)dzcp;
    }
    k_head_bwd<<<kHeadBlocks, 256, 0, c->stream>>>(dzn, n0, A.d0, u->p + P.head_off, G0, A.part);
    k_head_fin<<<1, 64, 0, c->stream>>>(A.part, (const double*)lpart, kHeadBlocks, u->grad + P.head_off,
                                        (double*)lossp, batch);
    NNX_LAUNCHED(c);
    const int nx1 = mx / 2, ny1 = my / 2, nz1 = mz / 2, nx2 = mx / 4, ny2 = my / 4, nz2 = mz / 4;
    auto WG = [&](int nxl, int nyl, int nzl, const uint16_t* in, const uint16_t* coarse, float cs, const uint16_t* g) {
        WgArgs w{};
        w.nx = nxl, w.ny = nyl, w.nz = nzl, w.batch = batch, w.in = in, w.coarse = coarse, w.cscale = cs, w.g = g;
        return w;
    };
    auto BD = [&](int nxl, int nyl, int nzl, const uint16_t* g, const uint16_t* w, uint16_t* out) {
        ConvArgs a{};
        a.nx = nxl, a.ny = nyl, a.nz = nzl, a.batch = batch, a.in = g, a.w = w, a.out = out;
        return a;
    };
    // d0: input e0 + up(d1), gradient G0
    NNX_TRY(wgrad_dispatch(c, 16, 16, WG(mx, my, mz, A.e0, A.d1, 1.f, G0), u->grad + P.L[4].w_off,
                           u->grad + P.L[4].b_off));
    NNX_TRY(conv_dispatch(c, 16, 16, BD(mx, my, mz, G0, u->wb[4], U0)));  // dU0 (skip part of e0's gradient)
    k_sumpool8<<<eblocks(n1 * 16), 256, 0, c->stream>>>(nx1, ny1, nz1, batch, 16, U0, L1a);  // d(d1)
    k_grad_combine<<<eblocks(n1 * 2), 256, 0, c->stream>>>(nx1, ny1, nz1, batch, 16, A.d1, L1a, nullptr, 0.f, L1b);
    NNX_LAUNCHED(c);
    // d1: input e1 + up(e2), gradient L1b (16 ch)
    NNX_TRY(wgrad_dispatch(c, 32, 16, WG(nx1, ny1, nz1, A.e1, A.e2, 1.f, L1b), u->grad + P.L[3].w_off,
                           u->grad + P.L[3].b_off));
    NNX_TRY(conv_dispatch(c, 16, 32, BD(nx1, ny1, nz1, L1b, u->wb[3], L1c)));  // dU1 (32 ch)
    k_sumpool8<<<eblocks(n2 * 32), 256, 0, c->stream>>>(nx2, ny2, nz2, batch, 32, L1c, L2a);  // d(e2)
    k_grad_combine<<<eblocks(n2 * 4), 256, 0, c->stream>>>(nx2, ny2, nz2, batch, 32, A.e2, L2a, nullptr, 0.f, L2b);
    NNX_LAUNCHED(c);
    // e2: input p1, gradient L2b
    NNX_TRY(wgrad_dispatch(c, 32, 32, WG(nx2, ny2, nz2, A.p1, nullptr, 0.f, L2b), u->grad + P.L[2].w_off,
                           u->grad + P.L[2].b_off));
    NNX_TRY(conv_dispatch(c, 32, 32, BD(nx2, ny2, nz2, L2b, u->wb[2], L2a)));  // d(p1)
    // e1 gradient: skip part dU1 + pooled part d(p1)/8
    k_grad_combine<<<eblocks(n1 * 4), 256, 0, c->stream>>>(nx1, ny1, nz1, batch, 32, A.e1, L1c, L2a, 0.125f, L1a);
    NNX_LAUNCHED(c);
    NNX_TRY(wgrad_dispatch(c, 16, 32, WG(nx1, ny1, nz1, A.p0, nullptr, 0.f, L1a), u->grad + P.L[1].w_off,
                           u->grad + P.L[1].b_off));
    NNX_TRY(conv_dispatch(c, 32, 16, BD(nx1, ny1, nz1, L1a, u->wb[1], L1b)));  // d(p0) (16 ch)
    // e0 gradient: dU0 + d(p0)/8
    k_grad_combine<<<eblocks(n0 * 2), 256, 0, c->stream>>>(mx, my, mz, batch, 16, A.e0, U0, L1b, 0.125f, G0);
    NNX_LAUNCHED(c);
    NNX_TRY(wgrad_dispatch(c, 2, 16, WG(mx, my, mz, A.x, nullptr, 0.f, G0), u->grad + P.L[0].w_off,
                           u->grad + P.L[0].b_off));
    // the biases are held at zero (the network must map r = 0 to 0)
    for (int l = 0; l < 5; ++l)
        NNX_HIP(c, hipMemsetAsync(u->grad + P.L[l].b_off, 0, P.L[l].cout * 4, c->stream));
    NNX_HIP(c, hipMemsetAsync(u->grad + P.head_off + 16, 0, 4, c->stream));
    // Adam
    u->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 - std::pow((double)b1, (double)u->step));
    const float c2 = (float)(1.0 - std::pow((double)b2, (double)u->step));
    if (lr > 0) {
        k_adam<<<eblocks(P.total), 256, 0, c->stream>>>(P.total, u->p, u->grad, u->m, u->v, (float)lr, b1, b2, 1e-8f,
                                                        c1, c2);
        NNX_LAUNCHED(c);
        u->dirty = true;
    }
    if (h_loss) {
        NNX_HIP(c, hipMemcpyAsync(c->pinned, lossp, 8, hipMemcpyDeviceToHost, c->stream));
        NNX_HIP(c, hipStreamSynchronize(c->stream));
        std::memcpy(h_loss, c->pinned, 8);
    }
    return NNX_OK;
}

nnx_status nnx_unet_get_grad(nnx_ctx* c, const nnx_unet* u, float* h_grad) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, u && h_grad, "null argument");
    NNX_HIP(c, hipMemcpyAsync(h_grad, u->grad, plan().total * 4, hipMemcpyDeviceToHost, c->stream));
    NNX_HIP(c, hipStreamSynchronize(c->stream));
    return NNX_OK;
}

// single-layer primitives (weights given in fp32 [cout][27 cin])
nnx_status nnx_conv3_bf16(nnx_ctx* c, int cin, int cout, int nx, int ny, int nz, int batch, const uint16_t* d_in,
                          const uint16_t* d_coarse, double coarse_scale, const float* d_w, const float* d_bias,
                          int relu, uint16_t* d_out, uint16_t* d_pool) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, d_in && d_w && (d_out || d_pool), "null argument");
    NNX_CHECK(c, cin != 2 || !d_coarse, "the 2-channel input layer takes no upsampled term");
    void* wp;
    NNX_TRY(scratch(c, "un_prim_w", (size_t)cout * kpad(cin) * 2, false, &wp));
    k_prep_w<<<ceil_div(cout * kpad(cin), 256), 256, 0, c->stream>>>(d_w, cin, cout, 0, (uint16_t*)wp);
    NNX_LAUNCHED(c);
    ConvArgs a{};
    a.nx = nx, a.ny = ny, a.nz = nz, a.batch = batch, a.in = d_in, a.coarse = d_coarse, a.cscale = (float)coarse_scale;
    a.w = (const uint16_t*)wp, a.bias = d_bias, a.relu = relu, a.out = d_out, a.pool = d_pool;
    return conv_dispatch(c, cin, cout, a);
}

nnx_status nnx_conv3_wgrad_bf16(nnx_ctx* c, int cin, int cout, int nx, int ny, int nz, int batch,
                                const uint16_t* d_in, const uint16_t* d_coarse, double coarse_scale,
                                const uint16_t* d_g, float* d_dw, float* d_db) {
    if (!c) return NNX_ERR_ARG;
    NNX_CHECK(c, d_in && d_g && d_dw && d_db, "null argument");
    WgArgs w{};
    w.nx = nx, w.ny = ny, w.nz = nz, w.batch = batch, w.in = d_in, w.coarse = d_coarse,
    w.cscale = (float)coarse_scale, w.g = d_g;
    return wgrad_dispatch(c, cin, cout, w, d_dw, d_db);
}

}  // extern "C"

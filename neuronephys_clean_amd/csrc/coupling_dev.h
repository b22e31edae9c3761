// coupling_dev.h - device helpers of the deterministic coupling, shared by
// the coupling kernels and the ephys kernel that folds the gather in.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nnx {

__device__ __forceinline__ bool bit_of(const uint64_t* bits, int32_t s) {
    return (bits[s >> 6] >> (s & 63)) & 1ull;
}

// Sum of target t's pending event-driven contributions.  Its c bucket
// entries (CSC positions of the edges pushed this step) sit at
// slot[in_off[t] ..]; they are added in ascending CSC position, i.e. ascending
// source order - the order of the dense pull.  With more than kSortMax
// entries (or an inconsistent count) the dense pull over t's in-edges with
// the previous step's spike bitmap gives the same sum.
constexpr int kSortMax = 8;

template <class T>
__device__ __forceinline__ T gather_one(int32_t t, int32_t c, const int32_t* __restrict__ in_off,
                                        const int32_t* __restrict__ in_src,
                                        const float* __restrict__ in_w,
                                        const uint64_t* __restrict__ bits,
                                        const int32_t* __restrict__ slot) {
    T acc = T(0);
    if (c == 0) return acc;
    const int32_t b = in_off[t];
    if (c <= kSortMax && c <= in_off[t + 1] - b) {
        // ascending order by repeated minimum (positions are distinct)
        int32_t last = -1;
        for (int r = 0; r < c; ++r) {
            int32_t mn = 0x7fffffff;
            for (int q = 0; q < c; ++q) {
                const int32_t p = slot[b + q];
                mn = (p > last && p < mn) ? p : mn;
            }
            acc = acc + T(in_w[mn]);
            last = mn;
        }
    } else {
        const int32_t e1 = in_off[t + 1];
        for (int32_t e = b; e < e1; ++e)
            if (bit_of(bits, in_src[e])) acc = acc + T(in_w[e]);
    }
    return acc;
}


// Wavefront-cooperative dense pull of target t (same value on every lane):
// lanes read 64 in-edges at a time and the spiking lanes' weights are added
// in ascending lane (= source) order.  All 64 lanes must call it together.
template <class T>
__device__ __forceinline__ T wave_pull_one(int64_t t, const int32_t* __restrict__ in_off,
                                           const int32_t* __restrict__ in_src,
                                           const float* __restrict__ in_w,
                                           const uint64_t* __restrict__ bits) {
    const int lane = threadIdx.x & 63;
    const int32_t e0 = in_off[t], e1 = in_off[t + 1];
    T acc = T(0);
    for (int32_t b = e0; b < e1; b += 64) {
        const int32_t e = b + lane;
        bool act = false;
        T val = T(0);
        if (e < e1) {
            act = bit_of(bits, in_src[e]);
            if (act) val = T(in_w[e]);
        }
        uint64_t m = __ballot(act);
        while (m) {
            const int l = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            acc = acc + __shfl(val, l, 64);
        }
    }
    return acc;
}

}  // namespace nnx

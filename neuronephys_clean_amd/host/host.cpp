// host.cpp - host-side builders (libnnx_host.so): Philox streams, synthetic
// networks, CSR/CSC layouts, trilinear transfer matrices.  Plain C++/OpenMP.
//
// Network construction stays on the host as in the reference
// (src/network_factory.cu, src/paracrine.cu:136-274); these are new builders
// for the benchmark configurations.
#include "../../include/nnx_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr uint64_t kM0 = 0xD2E7470EE14C6C93ULL, kM1 = 0xCA5A826395121157ULL;
constexpr uint64_t kW0 = 0x9E3779B97F4A7C15ULL, kW1 = 0xBB67AE8584CAA73BULL;

inline void mulhilo(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
    __uint128_t p = (__uint128_t)a * b;
    lo = (uint64_t)p;
    hi = (uint64_t)(p >> 64);
}

// Philox4x64 with 10 rounds (Salmon, Moraes, Dror, Shaw, SC'11): each round
// multiplies words 0 and 2 by the round constants and mixes hi/lo halves with
// words 1, 3 and the key; the key is bumped by the Weyl constants between
// rounds.
inline void philox(const uint64_t c[4], const uint64_t k[2], uint64_t o[4]) {
    uint64_t x0 = c[0], x1 = c[1], x2 = c[2], x3 = c[3], k0 = k[0], k1 = k[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += kW0;
            k1 += kW1;
        }
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo(kM0, x0, hi0, lo0);
        mulhilo(kM1, x2, hi1, lo1);
        uint64_t y0 = hi1 ^ x1 ^ k0, y2 = hi0 ^ x3 ^ k1;
        x0 = y0;
        x1 = lo1;
        x2 = y2;
        x3 = lo0;
    }
    o[0] = x0;
    o[1] = x1;
    o[2] = x2;
    o[3] = x3;
}

// Counter stream for one (seed, purpose, item).
struct Stream {
    uint64_t key[2], ctr[4], buf[4];
    int pos;
    Stream(uint64_t seed, uint64_t purpose, uint64_t item) {
        key[0] = seed;
        key[1] = 0;
        ctr[0] = 0;
        ctr[1] = purpose;
        ctr[2] = item;
        ctr[3] = 0;
        pos = 4;
    }
    uint64_t next() {
        if (pos == 4) {
            philox(ctr, key, buf);
            ++ctr[0];
            pos = 0;
        }
        return buf[pos++];
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double uniform_open0() { return (double)((next() >> 11) + 1) * 0x1.0p-53; }  // (0, 1]
    double normal() {
        double u1 = uniform_open0(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// value i of the flat stream (seed, purpose): word i%4 of block i/4, item 0
inline uint64_t flat_word(uint64_t seed, uint64_t purpose, uint64_t i) {
    uint64_t c[4] = {i / 4, purpose, 0, 0}, k[2] = {seed, 0}, o[4];
    philox(c, k, o);
    return o[i % 4];
}

constexpr uint64_t kPurposeEr = 0x4552, kPurposeDeg = 0x444547, kPurposeSpatial = 0x535041,
                   kPurposePois = 0x504f4953;

inline double rate_am(double u) {
    double a = 3.5 + 0.1 * u;
    return std::fabs(a) < 1e-9 ? 1.0 : a / (1.0 - std::exp(-a));
}
inline double rate_an(double u) {
    double b = 5.0 + 0.1 * u;
    return std::fabs(b) < 1e-9 ? 0.1 : 0.1 * b / (1.0 - std::exp(-b));
}

// open-addressing set of int32 for duplicate checks
struct SmallSet {
    std::vector<int32_t> t;
    uint32_t mask = 0;
    void reset(int64_t k) {
        uint32_t cap = 16;
        while (cap < 4 * (uint64_t)k) cap <<= 1;
        t.assign(cap, -1);
        mask = cap - 1;
    }
    bool insert(int32_t v) {  // false if present
        uint32_t h = (uint32_t)v * 2654435761u & mask;
        while (t[h] != -1) {
            if (t[h] == v) return false;
            h = (h + 1) & mask;
        }
        t[h] = v;
        return true;
    }
};

}  // namespace

extern "C" {

void nnx_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
    philox(ctr, key, out);
}

void nnx_gen_uniform(uint64_t seed, uint64_t purpose, int64_t n, double lo, double hi, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        out[i] = lo + (hi - lo) * ((double)(flat_word(seed, purpose, i) >> 11) * 0x1.0p-53);
}

void nnx_gen_normal(uint64_t seed, uint64_t purpose, int64_t n, double mu, double sigma, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double u1 = (double)((flat_word(seed, purpose, 2 * i) >> 11) + 1) * 0x1.0p-53;
        double u2 = (double)(flat_word(seed, purpose, 2 * i + 1) >> 11) * 0x1.0p-53;
        out[i] = mu + sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
}

void nnx_gen_lognormal(uint64_t seed, uint64_t purpose, int64_t n, double mu, double sigma,
                       double* out) {
    nnx_gen_normal(seed, purpose, n, mu, sigma, out);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = std::exp(out[i]);
}

int nnx_gen_er_fixed(uint64_t seed, int64_t n, int32_t k, int allow_self, int32_t* offsets,
                     int32_t* targets) {
    int64_t cand = allow_self ? n : n - 1;
    if (k < 0 || k > cand) return -1;
    for (int64_t i = 0; i <= n; ++i) offsets[i] = (int32_t)(i * k);
#pragma omp parallel
    {
        SmallSet set;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            Stream rng(seed, kPurposeEr, (uint64_t)i);
            set.reset(k);
            int32_t* out = targets + i * k;
            int32_t got = 0;
            while (got < k) {
                int64_t j = (int64_t)(rng.uniform() * (double)cand);
                if (j >= cand) j = cand - 1;
                if (!allow_self && j >= i) ++j;
                if (set.insert((int32_t)j)) out[got++] = (int32_t)j;
            }
            std::sort(out, out + k);
        }
    }
    return 0;
}

int nnx_gen_uniform_targets(uint64_t seed, int64_t n, const int32_t* offsets, int allow_self,
                             int32_t* targets) {
    const int64_t cand = allow_self ? n : n - 1;
    int bad = 0;
#pragma omp parallel reduction(| : bad)
    {
        SmallSet set;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            const int32_t k = offsets[i + 1] - offsets[i];
            if (k > cand) {
                bad = 1;
                continue;
            }
            Stream rng(seed, kPurposeEr, (uint64_t)i);
            set.reset(k);
            int32_t* out = targets + offsets[i];
            int32_t got = 0;
            while (got < k) {
                int64_t j = (int64_t)(rng.uniform() * (double)cand);
                if (j >= cand) j = cand - 1;
                if (!allow_self && j >= i) ++j;
                if (set.insert((int32_t)j)) out[got++] = (int32_t)j;
            }
            std::sort(out, out + k);
        }
    }
    return bad ? -1 : 0;
}

int64_t nnx_gen_outdeg_lognormal(uint64_t seed, int64_t n, double mu, double sigma, int32_t cap,
                                 int32_t* outdeg) {
    std::vector<double> z(n);
    nnx_gen_normal(seed, kPurposeDeg, n, mu, sigma, z.data());
    int64_t tot = 0;
    for (int64_t i = 0; i < n; ++i) {
        double v = std::nearbyint(std::exp(z[i]));
        int64_t k = v < 1 ? 1 : (int64_t)v;
        if (k > cap) k = cap;
        outdeg[i] = (int32_t)k;
        tot += k;
    }
    return tot;
}

int64_t nnx_gen_outdeg_poisson(uint64_t seed, int64_t n, double lambda, int32_t* outdeg) {
    int64_t tot = 0;
    const double p0 = std::exp(-lambda);
#pragma omp parallel for schedule(static) reduction(+ : tot)
    for (int64_t i = 0; i < n; ++i) {
        // inversion of the Poisson CDF
        double u = (double)(flat_word(seed, kPurposePois, i) >> 11) * 0x1.0p-53;
        double p = p0, c = p0;
        int32_t k = 0;
        while (u > c && k < 100000) {
            ++k;
            p *= lambda / k;
            c += p;
        }
        outdeg[i] = k;
        tot += k;
    }
    return tot;
}

int64_t nnx_gen_spatial_targets(uint64_t seed, int64_t n, const float* pos, double box,
                                double lambda, double cutoff, int32_t* offsets, int32_t* targets) {
    // cells of edge ~4 (about 2 neurons per cell at 0.5 neurons / unit^3)
    int nc = std::max(1, (int)std::floor(box / 4.0));
    const double cs = box / nc;
    auto cell_of = [&](double v) {
        int c = (int)std::floor(v / cs);
        return c < 0 ? 0 : c >= nc ? nc - 1 : c;
    };
    const int64_t ncell = (int64_t)nc * nc * nc;
    std::vector<int64_t> cstart(ncell + 1, 0);
    std::vector<int32_t> cid(n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t c = cell_of(pos[3 * i]) + nc * (cell_of(pos[3 * i + 1]) + (int64_t)nc * cell_of(pos[3 * i + 2]));
        cid[i] = (int32_t)c;
        cstart[c + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    std::vector<int32_t> items(n);
    {
        std::vector<int64_t> cur(cstart.begin(), cstart.end() - 1);
        for (int64_t i = 0; i < n; ++i) items[cur[cid[i]]++] = (int32_t)i;  // ascending per cell
    }
    // cell offsets within the cutoff with the upper bound exp(-dmin/lambda)
    struct Off {
        int dx, dy, dz;
        double u;
    };
    std::vector<Off> offs;
    const int R = (int)std::ceil(cutoff / cs) + 1;
    for (int dz = -R; dz <= R; ++dz)
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) {
                double gx = std::max(std::abs(dx) - 1, 0) * cs, gy = std::max(std::abs(dy) - 1, 0) * cs,
                       gz = std::max(std::abs(dz) - 1, 0) * cs;
                double dmin = std::sqrt(gx * gx + gy * gy + gz * gz);
                if (dmin <= cutoff) offs.push_back({dx, dy, dz, std::exp(-dmin / lambda)});
            }
    int64_t failed = 0;
#pragma omp parallel reduction(+ : failed)
    {
        std::vector<double> cum;
        std::vector<int64_t> ccell;
        std::vector<double> cu;
        SmallSet set;
#pragma omp for schedule(dynamic, 16)
        for (int64_t c = 0; c < ncell; ++c) {
            if (cstart[c] == cstart[c + 1]) continue;
            int cx = (int)(c % nc), cy = (int)((c / nc) % nc), cz = (int)(c / ((int64_t)nc * nc));
            cum.clear();
            ccell.clear();
            cu.clear();
            double acc = 0;
            for (const Off& o : offs) {
                int x = cx + o.dx, y = cy + o.dy, z = cz + o.dz;
                if (x < 0 || y < 0 || z < 0 || x >= nc || y >= nc || z >= nc) continue;
                int64_t t = x + nc * (y + (int64_t)nc * z);
                int64_t cnt = cstart[t + 1] - cstart[t];
                if (!cnt) continue;
                acc += (double)cnt * o.u;
                cum.push_back(acc);
                ccell.push_back(t);
                cu.push_back(o.u);
            }
            for (int64_t q = cstart[c]; q < cstart[c + 1]; ++q) {
                const int32_t i = items[q];
                const int32_t k = offsets[i + 1] - offsets[i];
                int32_t* out = targets + offsets[i];
                Stream rng(seed, kPurposeSpatial, (uint64_t)i);
                set.reset(k);
                int32_t got = 0;
                int64_t attempts = 0, limit = 2000LL * k + 100000;
                const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
                while (got < k && attempts < limit) {
                    ++attempts;
                    double r = rng.uniform() * acc;
                    size_t b = std::upper_bound(cum.begin(), cum.end(), r) - cum.begin();
                    if (b >= cum.size()) b = cum.size() - 1;
                    int64_t t = ccell[b];
                    int64_t cnt = cstart[t + 1] - cstart[t];
                    int64_t pick = (int64_t)(rng.uniform() * (double)cnt);
                    if (pick >= cnt) pick = cnt - 1;
                    int32_t j = items[cstart[t] + pick];
                    double acc_u = rng.uniform();
                    if (j == i) continue;
                    double dx = pos[3 * j] - px, dy = pos[3 * j + 1] - py, dz = pos[3 * j + 2] - pz;
                    double d = std::sqrt(dx * dx + dy * dy + dz * dz);
                    if (d > cutoff) continue;
                    if (acc_u * cu[b] >= std::exp(-d / lambda)) continue;
                    if (set.insert(j)) out[got++] = j;
                }
                if (got < k) {
                    ++failed;
                    for (int32_t g = got; g < k; ++g) out[g] = -1;
                }
                std::sort(out, out + got);
            }
        }
    }
    return failed ? -failed : 0;
}

void nnx_csr_to_csc(int64_t n, int64_t nnz, const int32_t* off, const int32_t* tgt, const float* w,
                    int32_t* in_off, int32_t* in_src, float* in_w, int32_t* out_to_in,
                    float* in_degree) {
    std::vector<int64_t> cnt(n + 1, 0);
    for (int64_t e = 0; e < nnz; ++e) cnt[tgt[e] + 1]++;
    for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    for (int64_t i = 0; i <= n; ++i) in_off[i] = (int32_t)cnt[i];
    if (in_degree)
        for (int64_t i = 0; i < n; ++i) in_degree[i] = (float)(cnt[i + 1] - cnt[i]);
    std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
    for (int64_t s = 0; s < n; ++s)
        for (int64_t e = off[s]; e < off[s + 1]; ++e) {
            int64_t p = cur[tgt[e]]++;
            in_src[p] = (int32_t)s;
            if (in_w) in_w[p] = w ? w[e] : 0.0f;
            if (out_to_in) out_to_in[e] = (int32_t)p;
        }
    (void)nnz;
}

void nnx_hh_steady_state(int64_t n, const double* V, double v_shift, double* m, double* h,
                         double* ng) {
    for (int64_t i = 0; i < n; ++i) {
        double u = V[i] + v_shift;
        double am = rate_am(u), bm = 4.0 * std::exp(-(u + 60.0) / 18.0);
        double ah = 0.07 * std::exp(-u / 20.0 - 3.0), bh = 1.0 / (std::exp(-3.0 - 0.1 * u) + 1.0);
        double an = rate_an(u), bn = 0.125 * std::exp(-(u + 60.0) / 80.0);
        m[i] = am / (am + bm);
        h[i] = ah / (ah + bh);
        ng[i] = an / (an + bn);
    }
}

void nnx_voxel_of(int64_t n, const float* pos, double h, int32_t nx, int32_t ny, int32_t nz,
                  int32_t* voxel) {
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3];
        const int32_t dims[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a) {
            int64_t v = (int64_t)std::floor(pos[3 * i + a] / h);
            c[a] = v < 0 ? 0 : v >= dims[a] ? dims[a] - 1 : v;
        }
        voxel[i] = (int32_t)(c[0] + (int64_t)nx * (c[1] + (int64_t)ny * c[2]));
    }
}

int64_t nnx_trilinear_matrix(int64_t n, const float* x, const float* y, const float* z, int32_t nx,
                             int32_t ny, int32_t nz, double h, int32_t* offsets, int32_t* cols,
                             float* vals) {
    // float arithmetic as in the reference (Float = float)
    const float hs = (float)h;
    const float sx = (float)(nx - 1) * hs, sy = (float)(ny - 1) * hs, sz = (float)(nz - 1) * hs;
    int64_t outside = 0;
    for (int64_t i = 0; i < n; ++i) {
        offsets[i] = (int32_t)(8 * i);
        float rx = x[i] / sx + 0.5f, ry = y[i] / sy + 0.5f, rz = z[i] / sz + 0.5f;
        int m = (int)((float)(nx - 1) * rx), nn = (int)((float)(ny - 1) * ry), p = (int)((float)(nz - 1) * rz);
        float fx = (float)(nx - 1) * rx - (float)m, fy = (float)(ny - 1) * ry - (float)nn,
              fz = (float)(nz - 1) * rz - (float)p;
        float wx0 = 1.0f - fx, wx1 = fx, wy0 = 1.0f - fy, wy1 = fy, wz0 = 1.0f - fz, wz1 = fz;
        // corner order of the reference: +0, +1, +nx, +nx+1, then the same +nx*ny
        const int di[8] = {0, 1, 0, 1, 0, 1, 0, 1}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                  dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
        float wxy[4] = {wx0 * wy0, wx1 * wy0, wx0 * wy1, wx1 * wy1};
        bool out = false;
        for (int c = 0; c < 8; ++c) {
            int ii = m + di[c], jj = nn + dj[c], kk = p + dk[c];
            float wv = wxy[c & 3] * (c < 4 ? wz0 : wz1);
            if (ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz) {
                out = true;
                ii = std::min(std::max(ii, 0), nx - 1);
                jj = std::min(std::max(jj, 0), ny - 1);
                kk = std::min(std::max(kk, 0), nz - 1);
                wv = 0.0f;
            }
            cols[8 * i + c] = ii + nx * (jj + ny * kk);
            vals[8 * i + c] = wv;
        }
        outside += out;
    }
    offsets[n] = (int32_t)(8 * n);
    return outside;
}

void nnx_csr_transpose(int64_t rows, int64_t cols, const int32_t* off, const int32_t* idx,
                       const float* val, int32_t* t_off, int32_t* t_idx, float* t_val) {
    std::vector<int64_t> cnt(cols + 1, 0);
    for (int64_t e = off[0]; e < off[rows]; ++e) cnt[idx[e] + 1]++;
    for (int64_t c = 0; c < cols; ++c) cnt[c + 1] += cnt[c];
    for (int64_t c = 0; c <= cols; ++c) t_off[c] = (int32_t)cnt[c];
    std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t e = off[r]; e < off[r + 1]; ++e) {
            int64_t p = cur[idx[e]]++;
            t_idx[p] = (int32_t)r;
            t_val[p] = val[e];
        }
}

// Gaussian random field on an nx*ny*nz grid (x fastest): standard normal
// white noise (Philox stream `purpose`) on a grid padded by R = ceil(3 s) on
// every side, filtered by the separable kernel exp(-d^2 / (2 s^2)) with
// s = corr / sqrt(2), and scaled by the kernel's L2 norm, so that the field
// has unit variance and covariance exp(-r^2 / (2 corr^2)).
void nnx_gen_grf(uint64_t seed, uint64_t purpose, int32_t nx, int32_t ny, int32_t nz, double corr,
                 float* out) {
    const double s = corr / std::sqrt(2.0);
    const int R = (int)std::ceil(3.0 * s);
    std::vector<double> k(2 * R + 1);
    double k2 = 0;
    for (int d = -R; d <= R; ++d) {
        k[d + R] = std::exp(-(double)d * d / (2.0 * s * s));
        k2 += k[d + R] * k[d + R];
    }
    const double norm = 1.0 / std::pow(k2, 1.5);
    const int64_t px = nx + 2 * R, py = ny + 2 * R, pz = nz + 2 * R;
    std::vector<double> w((size_t)(px * py * pz));
    nnx_gen_normal(seed, purpose, (int64_t)w.size(), 0.0, 1.0, w.data());
    // x pass: (px, py, pz) -> (nx, py, pz)
    std::vector<double> a((size_t)(nx * py * pz));
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < py * pz; ++row)
        for (int i = 0; i < nx; ++i) {
            double t = 0;
            for (int d = 0; d <= 2 * R; ++d) t += k[d] * w[(size_t)(row * px + i + d)];
            a[(size_t)(row * nx + i)] = t;
        }
    std::vector<double>().swap(w);
    // y pass: (nx, py, pz) -> (nx, ny, pz)
    std::vector<double> b((size_t)(nx * ny * pz));
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < pz; ++z)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double t = 0;
                for (int d = 0; d <= 2 * R; ++d) t += k[d] * a[(size_t)((z * py + j + d) * nx + i)];
                b[(size_t)((z * ny + j) * nx + i)] = t;
            }
    std::vector<double>().swap(a);
    // z pass: (nx, ny, pz) -> (nx, ny, nz)
    const int64_t plane = (int64_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < nz; ++z)
        for (int64_t e = 0; e < plane; ++e) {
            double t = 0;
            for (int d = 0; d <= 2 * R; ++d) t += k[d] * b[(size_t)((z + d) * plane + e)];
            out[z * plane + e] = (float)(t * std::sqrt(norm));
        }
}

// Face coefficients of the variable 7-point operator from per-voxel
// coefficients c (= dt D / h^2): interior faces take the harmonic mean of
// their two voxels, boundary faces the voxel's own value.  Layout as
// nnx_grid's kx / ky / kz.
void nnx_face_coefficients(int32_t nx, int32_t ny, int32_t nz, const float* c, float* kx, float* ky,
                           float* kz) {
    auto hm = [](float a, float b) { return (float)(2.0 * (double)a * b / ((double)a + b)); };
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j) {
            const float* row = c + (k * ny + j) * nx;
            float* fx = kx + (k * ny + j) * (nx + 1);
            fx[0] = row[0];
            for (int i = 1; i < nx; ++i) fx[i] = hm(row[i - 1], row[i]);
            fx[nx] = row[nx - 1];
            for (int i = 0; i < nx; ++i) {
                const int64_t id = (k * ny + j) * nx + i;
                const float cy = j == 0 ? c[id] : hm(c[id - nx], c[id]);
                ky[i + nx * (j + (int64_t)(ny + 1) * k)] = cy;
                if (j == ny - 1) ky[i + nx * (ny + (int64_t)(ny + 1) * k)] = c[id];
                const float cz = k == 0 ? c[id] : hm(c[id - (int64_t)nx * ny], c[id]);
                kz[id] = cz;
                if (k == nz - 1) kz[id + (int64_t)nx * ny] = c[id];
            }
        }
}

// k nearest neurons of each query point by squared Euclidean distance
// (FP64 from the float coordinates), ties by lower index, nearest first.
// A uniform cell grid over the bounding box (about 4 neurons per cell) is
// searched in Chebyshev shells around the query's cell until the k-th
// distance is below the distance to the searched block's boundary.
int nnx_spherical_regions(int64_t n, const float* pos, int64_t n_regions, const int32_t* center_idx,
                          const float* center_pos, int32_t k, int32_t* offsets, int32_t* idx) {
    if (k < 1 || k > n || n_regions < 0 || (!center_idx && !center_pos && n_regions > 0)) return -1;
    if (center_idx)
        for (int64_t r = 0; r < n_regions; ++r)
            if (center_idx[r] < 0 || center_idx[r] >= n) return -1;
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = pos[d];
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], (double)pos[3 * i + d]);
            hi[d] = std::max(hi[d], (double)pos[3 * i + d]);
        }
    double ext = 0;
    for (int d = 0; d < 3; ++d) ext = std::max(ext, hi[d] - lo[d]);
    ext = ext > 0 ? ext : 1.0;
    // cell edge from the box volume (degenerate axes count as one cell)
    double vol = 1;
    for (int d = 0; d < 3; ++d) vol *= std::max(hi[d] - lo[d], ext * 1e-6);
    const double cs = std::max(std::cbrt(vol * 4.0 / (double)n), ext * 1e-6);
    int64_t nc[3];
    for (int d = 0; d < 3; ++d) nc[d] = std::min<int64_t>((int64_t)((hi[d] - lo[d]) / cs) + 1, 1024);
    auto cell_of = [&](double x, int d) {
        int64_t c = (int64_t)std::floor((x - lo[d]) / cs);
        return std::min<int64_t>(std::max<int64_t>(c, 0), nc[d] - 1);
    };
    const int64_t ncell = nc[0] * nc[1] * nc[2];
    std::vector<int64_t> start(ncell + 1, 0);
    std::vector<int32_t> member(n);
    std::vector<int64_t> cid(n);
    for (int64_t i = 0; i < n; ++i) {
        cid[i] = cell_of(pos[3 * i], 0) + nc[0] * (cell_of(pos[3 * i + 1], 1) + nc[1] * cell_of(pos[3 * i + 2], 2));
        ++start[cid[i] + 1];
    }
    for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    {
        std::vector<int64_t> fillp(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < n; ++i) member[fillp[cid[i]]++] = (int32_t)i;  // ascending within a cell
    }
    for (int64_t r = 0; r <= n_regions; ++r) offsets[r] = (int32_t)(r * k);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < n_regions; ++r) {
        double q[3];
        for (int d = 0; d < 3; ++d)
            q[d] = center_idx ? (double)pos[3 * (int64_t)center_idx[r] + d] : (double)center_pos[3 * r + d];
        int64_t qc[3];
        for (int d = 0; d < 3; ++d) qc[d] = cell_of(q[d], d);
        // max-heap of the best k as (distance, index), lexicographic
        std::vector<std::pair<double, int32_t>> heap;
        heap.reserve(k + 1);
        for (int64_t ring = 0;; ++ring) {
            bool any = false;
            for (int64_t cz = qc[2] - ring; cz <= qc[2] + ring; ++cz) {
                if (cz < 0 || cz >= nc[2]) continue;
                for (int64_t cy = qc[1] - ring; cy <= qc[1] + ring; ++cy) {
                    if (cy < 0 || cy >= nc[1]) continue;
                    for (int64_t cx = qc[0] - ring; cx <= qc[0] + ring; ++cx) {
                        if (cx < 0 || cx >= nc[0]) continue;
                        const int64_t cheb = std::max({std::llabs(cx - qc[0]), std::llabs(cy - qc[1]),
                                                       std::llabs(cz - qc[2])});
                        if (cheb != ring) continue;
                        any = true;
                        const int64_t c = cx + nc[0] * (cy + nc[1] * cz);
                        for (int64_t t = start[c]; t < start[c + 1]; ++t) {
                            const int32_t i = member[t];
                            double d2 = 0;
                            for (int d = 0; d < 3; ++d) {
                                const double dd = (double)pos[3 * (int64_t)i + d] - q[d];
                                d2 += dd * dd;
                            }
                            const std::pair<double, int32_t> e(d2, i);
                            if ((int32_t)heap.size() < k) {
                                heap.push_back(e);
                                std::push_heap(heap.begin(), heap.end());
                            } else if (e < heap.front()) {
                                std::pop_heap(heap.begin(), heap.end());
                                heap.back() = e;
                                std::push_heap(heap.begin(), heap.end());
                            }
                        }
                    }
                }
            }
            // every point outside the searched block is farther than this
            double bound = INFINITY;
            bool whole = true;
            for (int d = 0; d < 3; ++d) {
                if (qc[d] - ring > 0) bound = std::min(bound, q[d] - (lo[d] + (qc[d] - ring) * cs)), whole = false;
                if (qc[d] + ring < nc[d] - 1)
                    bound = std::min(bound, lo[d] + (qc[d] + ring + 1) * cs - q[d]), whole = false;
            }
            if (whole || (!any && ring > 2 * 1024)) break;
            if ((int32_t)heap.size() == k && bound > 0 && heap.front().first < bound * bound) break;
        }
        std::sort_heap(heap.begin(), heap.end());
        for (int32_t t = 0; t < k; ++t) idx[r * k + t] = heap[t].second;
    }
    return 0;
}

}  // extern "C"

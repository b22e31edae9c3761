// capi.cpp - C entry points of the CPU oracle (liboracle.so), for ctypes.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and bench.py's
// CPU-baseline leg, never by the product.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <vector>

#include "oracle.hpp"

using namespace oracle;

extern "C" {

// Same field order as the product's parameter block, kept separately so the
// oracle depends on nothing in the product.
struct orc_params {
    double C, gna, gk, gl, ena, ek, el, v_shift, tau_syn, e_syn;
    double tmax, vt, kp, ar, ad, g_local, e_local, g_i, e_i, g_e, e_e;
    double noise_mean, noise_std;
    uint64_t noise_seed;
    double v_thresh, out_thresh;
};

struct orc_grid {
    int32_t nx, ny, nz, stencil;
    double c, decay_dt;
    const float *kx, *ky, *kz;
};
}

template <class T> static P<T> cast(const orc_params* q) {
    P<T> p;
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
    p.v_thresh = T(q->v_thresh);
    p.out_thresh = q->out_thresh;
    return p;
}

static Grid grid_of(const orc_grid* g) {
    return Grid{g->nx, g->ny, g->nz, g->stencil, g->c, g->decay_dt, g->kx, g->ky, g->kz};
}

struct RkConst {
    double dt, hdt, dt6, eh, ef;
};
static RkConst rk_const(const orc_params* q, double dt) {
    return RkConst{dt, dt / 2.0, dt / 6.0, std::exp(-dt / (2.0 * q->tau_syn)),
                   std::exp(-dt / q->tau_syn)};
}

template <class T>
static void rk4_arrays(int64_t n, const orc_params* q, double dt, T* V, T* m, T* h, T* ng, T* s,
                       const T* iext, const T* gin, uint8_t* spiked) {
    P<T> p = cast<T>(q);
    RkConst k = rk_const(q, dt);
    const T tdt = T(k.dt), thdt = T(k.hdt), tdt6 = T(k.dt6), teh = T(k.eh), tef = T(k.ef);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        T v0 = V[i];
        rk4_step(p, tdt, thdt, tdt6, teh, tef, V[i], m[i], h[i], ng[i], s[i], iext[i],
                 gin ? gin[i] : T(0));
        if (spiked) spiked[i] = (v0 < p.v_thresh && V[i] >= p.v_thresh) ? 1 : 0;
    }
}

// white noise of neuron i at `step`: noise_std z + noise_mean in T (the
// mean alone when noise_std = 0), z the standard normal of (seed, step, i)
template <class T> static inline T white_noise(const orc_params* q, const P<T>& p, int64_t step, int64_t i) {
    if (q->noise_std == 0.0) return p.noise_mean;
    return T(q->noise_std) * T(noise::normal(q->noise_seed, (uint64_t)step, (uint64_t)i)) + p.noise_mean;
}

template <class T>
static void leapfrog_arrays(int64_t n, const orc_params* q, double dt, T* const* sv,
                            uint8_t* spiked, int64_t step) {
    // sv: the reference's 17 SoA arrays in SIM_PARAMS order (src/defines.cuh:57-78)
    P<T> p = cast<T>(q);
    const T tdt = T(dt), thdt = T(dt) / T(2);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        LeapfrogIO<T> s;
        s.V = sv[0][i];
        s.M = sv[1][i];
        s.N = sv[2][i];
        s.H = sv[3][i];
        s.Yl = sv[4][i];
        s.Yi = sv[5][i];
        s.Ye = sv[6][i];
        s.in_l = sv[8][i];
        s.cnt_l = sv[9][i];
        s.in_e = sv[11][i];
        s.cnt_e = sv[12][i];
        s.in_i = sv[14][i];
        s.cnt_i = sv[15][i];
        s.appcur = sv[16][i];
        leapfrog_step(p, tdt, thdt, white_noise(q, p, step, i), s);
        sv[0][i] = s.V;
        sv[1][i] = s.M;
        sv[2][i] = s.N;
        sv[3][i] = s.H;
        sv[4][i] = s.Yl;
        sv[5][i] = s.Yi;
        sv[6][i] = s.Ye;
        sv[7][i] = s.out;
        sv[10][i] = s.out;
        sv[13][i] = s.out;
        if (spiked) spiked[i] = (double)s.out > p.out_thresh ? 1 : 0;
    }
}

extern "C" {

void orc_exp_f32(const float* x, float* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = oexp(x[i]);
}
void orc_exp_f64(const double* x, double* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = oexp(x[i]);
}
void orc_rates_f64(const double* u, double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        Rates<double> R = rates(u[i]);
        double* o = out + 6 * i;
        o[0] = R.am; o[1] = R.bm; o[2] = R.ah; o[3] = R.bh; o[4] = R.an; o[5] = R.bn;
    }
}
void orc_rates_f32(const float* u, float* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        Rates<float> R = rates(u[i]);
        float* o = out + 6 * i;
        o[0] = R.am; o[1] = R.bm; o[2] = R.ah; o[3] = R.bh; o[4] = R.an; o[5] = R.bn;
    }
}

// One RK4 step over arrays (state updated in place).
void orc_hh_rk4_f32(int64_t n, const orc_params* q, double dt, float* V, float* m, float* h,
                    float* ng, float* s, const float* iext, const float* gin, uint8_t* spiked) {
    rk4_arrays<float>(n, q, dt, V, m, h, ng, s, iext, gin, spiked);
}
void orc_hh_rk4_f64(int64_t n, const orc_params* q, double dt, double* V, double* m, double* h,
                    double* ng, double* s, const double* iext, const double* gin,
                    uint8_t* spiked) {
    rk4_arrays<double>(n, q, dt, V, m, h, ng, s, iext, gin, spiked);
}

// One reference leapfrog step over the 17 SoA arrays (in place).
void orc_hh_leapfrog_f32(int64_t n, const orc_params* q, double dt, float* const* sv,
                         uint8_t* spiked, int64_t step) {
    leapfrog_arrays<float>(n, q, dt, sv, spiked, step);
}
void orc_hh_leapfrog_f64(int64_t n, const orc_params* q, double dt, double* const* sv,
                         uint8_t* spiked, int64_t step) {
    leapfrog_arrays<double>(n, q, dt, sv, spiked, step);
}

// the white-noise generator's parts (for its pins)
void orc_philox4x64_10(const uint64_t* c, const uint64_t* k, uint64_t* out) {
    noise::philox(c[0], c[1], c[2], c[3], k[0], k[1], out);
}
void orc_noise_log(const double* u, double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = noise::log_u(u[i]);
}
void orc_noise_cos2pi(const double* u, double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = noise::cos2pi(u[i]);
}
// standard normals of neurons first .. first+n-1 at `step`
void orc_noise_normal(uint64_t seed, uint64_t step, int64_t first, int64_t n, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = noise::normal(seed, step, (uint64_t)(first + i));
}

// Coupling from a 0/1 firing array; g overwritten (see couple_push).
void orc_couple_f32(int64_t n, const int32_t* off, const int32_t* tgt, const float* w,
                    const uint8_t* fire, int mode, const float* out, float* g) {
    std::vector<int32_t> list;
    compact(n, fire, list);
    couple_push<float>(n, off, tgt, w, list, mode, out, g);
}
void orc_couple_f64(int64_t n, const int32_t* off, const int32_t* tgt, const float* w,
                    const uint8_t* fire, int mode, const double* out, double* g) {
    std::vector<int32_t> list;
    compact(n, fire, list);
    couple_push<double>(n, off, tgt, w, list, mode, out, g);
}

void orc_apply_f32(const orc_grid* g, const float* x, float* y) { apply(grid_of(g), x, y); }
void orc_apply_f64(const orc_grid* g, const double* x, double* y) { apply(grid_of(g), x, y); }

int orc_cg_f32(const orc_grid* g, const float* b, float* x, double rtol, int maxit, int jacobi,
               int use_x0, double* relres) {
    return cg<float>(grid_of(g), b, x, rtol, maxit, jacobi != 0, use_x0 != 0, relres);
}
int orc_cg_f64(const orc_grid* g, const double* b, double* x, double rtol, int maxit, int jacobi,
               int use_x0, double* relres) {
    return cg<double>(grid_of(g), b, x, rtol, maxit, jacobi != 0, use_x0 != 0, relres);
}
int orc_refine(const orc_grid* g, const double* b, double* x, double rtol, double inner_rtol,
               int maxit, int max_outer, double* relres, int* inner_total) {
    return refine(grid_of(g), b, x, rtol, inner_rtol, maxit, max_outer, relres, inner_total);
}
int orc_update_density_reference(int nx, int ny, int nz, double h, double D, double k, double dt,
                                 double eb, int max_iter, float* data, double* rr) {
    return update_density_reference(nx, ny, nz, h, D, k, dt, eb, max_iter, data, rr);
}

}  // extern "C"

// Full RK4 run in the reference's step order (src/main.cu:166-181):
// coupling on the previous step's spikes -> release deposit -> ephys.
// Raster pairs (step, neuron) are appended to `raster` (capacity cap pairs);
// returns the number of pairs (may exceed cap: only cap are written).
// voxel_of/voxel_counts: optional release counting (+1 per spike).
template <class T>
static int64_t run_rk4(int64_t n, const orc_params* q, double dt, int64_t steps, const int32_t* off,
                       const int32_t* tgt, const float* w, T* V, T* m, T* h, T* ng, T* s,
                       const T* iext, int32_t* raster, int64_t cap, const int32_t* voxel_of,
                       int32_t* voxel_counts, double* seconds, uint8_t* fire_state = nullptr) {
    std::vector<T> g(n, T(0));
    // fire_state (optional, n bytes): the spikes of the step before the first
    // (their coupling is the first step's input), and on return those of the
    // last step, so that a run can continue in the next call
    std::vector<uint8_t> fire(n, 0);
    if (fire_state) std::copy(fire_state, fire_state + n, fire.begin());
    std::vector<int32_t> list;
    int64_t len = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int64_t k = 0; k < steps; ++k) {
        compact(n, fire.data(), list);
        couple_push<T>(n, off, tgt, w, list, 0, nullptr, g.data());
        rk4_arrays<T>(n, q, dt, V, m, h, ng, s, iext, g.data(), fire.data());
        for (int64_t i = 0; i < n; ++i)
            if (fire[i]) {
                if (len < cap) {
                    raster[2 * len] = (int32_t)k;
                    raster[2 * len + 1] = (int32_t)i;
                }
                ++len;
                if (voxel_of) voxel_counts[voxel_of[i]] += 1;
            }
    }
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (fire_state) std::copy(fire.begin(), fire.end(), fire_state);
    return len;
}

extern "C" {
int64_t orc_run_rk4_cont_f32(int64_t n, const orc_params* q, double dt, int64_t steps,
                             const int32_t* off, const int32_t* tgt, const float* w, float* V, float* m,
                             float* h, float* ng, float* s, const float* iext, int32_t* raster,
                             int64_t cap, const int32_t* voxel_of, int32_t* voxel_counts,
                             double* seconds, uint8_t* fire_state) {
    return run_rk4<float>(n, q, dt, steps, off, tgt, w, V, m, h, ng, s, iext, raster, cap,
                          voxel_of, voxel_counts, seconds, fire_state);
}
int64_t orc_run_rk4_cont_f64(int64_t n, const orc_params* q, double dt, int64_t steps,
                             const int32_t* off, const int32_t* tgt, const float* w, double* V,
                             double* m, double* h, double* ng, double* s, const double* iext,
                             int32_t* raster, int64_t cap, const int32_t* voxel_of,
                             int32_t* voxel_counts, double* seconds, uint8_t* fire_state) {
    return run_rk4<double>(n, q, dt, steps, off, tgt, w, V, m, h, ng, s, iext, raster, cap,
                           voxel_of, voxel_counts, seconds, fire_state);
}
int64_t orc_run_rk4_f32(int64_t n, const orc_params* q, double dt, int64_t steps,
                        const int32_t* off, const int32_t* tgt, const float* w, float* V, float* m,
                        float* h, float* ng, float* s, const float* iext, int32_t* raster,
                        int64_t cap, const int32_t* voxel_of, int32_t* voxel_counts,
                        double* seconds) {
    return run_rk4<float>(n, q, dt, steps, off, tgt, w, V, m, h, ng, s, iext, raster, cap,
                          voxel_of, voxel_counts, seconds);
}
int64_t orc_run_rk4_f64(int64_t n, const orc_params* q, double dt, int64_t steps,
                        const int32_t* off, const int32_t* tgt, const float* w, double* V,
                        double* m, double* h, double* ng, double* s, const double* iext,
                        int32_t* raster, int64_t cap, const int32_t* voxel_of,
                        int32_t* voxel_counts, double* seconds) {
    return run_rk4<double>(n, q, dt, steps, off, tgt, w, V, m, h, ng, s, iext, raster, cap,
                           voxel_of, voxel_counts, seconds);
}

}  // extern "C"

// Reference-mode run: one coupling structure (LOCAL_IN / LOCAL_COUNT) with
// COUNT (+1, src/coupling.cu:46) or SIGNAL (+output, src_wholebrain/coupling.cu:31)
// coupling on the neurons whose LOCAL_OUTPUT exceeded the level (computed
// from the outputs the previous step produced, the initial outputs for the
// first step), then the leapfrog step.  Raster pair (k, i): step k's update
// left neuron i above the level.
template <class T>
static int64_t run_leapfrog(int64_t n, const orc_params* q, double dt, int64_t steps,
                            const int32_t* off, const int32_t* tgt, int mode, T* const* sv,
                            int32_t* raster, int64_t cap, int64_t first_step) {
    std::vector<uint8_t> fire(n, 0);
    std::vector<int32_t> list;
    int64_t len = 0;
    // the first coupling reads the firing of the initial outputs
    for (int64_t i = 0; i < n; ++i) fire[i] = (double)sv[7][i] > q->out_thresh ? 1 : 0;
    compact(n, fire.data(), list);
    for (int64_t k = 0; k < steps; ++k) {
        couple_push<T>(n, off, tgt, nullptr, list, mode, sv[7], sv[8]);
        leapfrog_arrays<T>(n, q, dt, sv, fire.data(), first_step + k);
        compact(n, fire.data(), list);
        for (int32_t i : list) {
            if (len < cap) {
                raster[2 * len] = (int32_t)k;
                raster[2 * len + 1] = i;
            }
            ++len;
        }
    }
    return len;
}
extern "C" {
int64_t orc_run_leapfrog_f32(int64_t n, const orc_params* q, double dt, int64_t steps,
                             const int32_t* off, const int32_t* tgt, int mode, float* const* sv,
                             int32_t* raster, int64_t cap, int64_t first_step) {
    return run_leapfrog<float>(n, q, dt, steps, off, tgt, mode, sv, raster, cap, first_step);
}
int64_t orc_run_leapfrog_f64(int64_t n, const orc_params* q, double dt, int64_t steps,
                             const int32_t* off, const int32_t* tgt, int mode, double* const* sv,
                             int32_t* raster, int64_t cap, int64_t first_step) {
    return run_leapfrog<double>(n, q, dt, steps, off, tgt, mode, sv, raster, cap, first_step);
}

// Electrode readout, network mean and synchrony (see oracle.hpp).
void orc_electrode_measure_f32(int64_t nreg, int64_t nnz, const int32_t* off, const int32_t* idx, int64_t n,
                               const float* V, const int32_t* fire, float* avg, int32_t* cnt) {
    electrode_measure<float>(nreg, nnz, off, idx, n, V, fire, avg, cnt);
}
void orc_electrode_measure_f64(int64_t nreg, int64_t nnz, const int32_t* off, const int32_t* idx, int64_t n,
                               const double* V, const int32_t* fire, double* avg, int32_t* cnt) {
    electrode_measure<double>(nreg, nnz, off, idx, n, V, fire, avg, cnt);
}
float orc_network_mean_f32(int64_t n, const float* V) { return network_mean<float>(n, V); }
double orc_network_mean_f64(int64_t n, const double* V) { return network_mean<double>(n, V); }
void orc_sync_accumulate_f32(int64_t n, const float* V, int64_t n_idx, const int32_t* idx, int n_sub,
                             const int32_t* counts, double* vi, double* vi2, double* va, double* va2) {
    synchrony_accumulate<float>(n, V, n_idx, idx, n_sub, counts, vi, vi2, va, va2);
}
void orc_sync_accumulate_f64(int64_t n, const double* V, int64_t n_idx, const int32_t* idx, int n_sub,
                             const int32_t* counts, double* vi, double* vi2, double* va, double* va2) {
    synchrony_accumulate<double>(n, V, n_idx, idx, n_sub, counts, vi, vi2, va, va2);
}
void orc_sync_measures(int64_t n_idx, int n_sub, const int32_t* counts, const double* vi, const double* vi2,
                       const double* va, const double* va2, int64_t steps, double* chi) {
    synchrony_measures(n_idx, n_sub, counts, vi, vi2, va, va2, steps, chi);
}

// One Hebbian::UpdateWeights call (see oracle.hpp).
void orc_hebbian_update(int64_t n, const int32_t* off, const int32_t* tgt, const uint8_t* fire, int32_t window,
                        float heb_inc, int32_t* timers, float* w) {
    hebbian_update(n, off, tgt, fire, window, heb_inc, timers, w);
}

}  // extern "C"

// Diekman step over arrays: x = 12 state arrays (DIEKMAN_VAR order V..Cac),
// in = 5 input arrays (Tin, num_in, egaba, gkca, gkleak), out = IOUTPUT;
// c = 27 constants in nnx_diekman_params order; spiked (optional):
// IOUTPUT > thr with the output widened to double.
template <class T>
static void diekman_arrays(int64_t n, const double* c, double dt, double thr, T* const* x, const T* const* in, T* out,
                           uint8_t* spiked) {
    DkC<T> k;
    T* f = &k.gna;
    for (int i = 0; i < 27; ++i) f[i] = (T)c[i];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        T s[12], v[5];
        for (int q = 0; q < 12; ++q) s[q] = x[q][i];
        for (int q = 0; q < 5; ++q) v[q] = in[q][i];
        out[i] = diekman_step<T>(k, (T)dt, s, v);
        for (int q = 0; q < 12; ++q) x[q][i] = s[q];
        if (spiked) spiked[i] = (double)out[i] > thr ? 1 : 0;
    }
}
extern "C" {
void orc_diekman_f32(int64_t n, const double* c, double dt, double thr, float* const* x, const float* const* in,
                     float* out, uint8_t* spiked) {
    diekman_arrays<float>(n, c, dt, thr, x, in, out, spiked);
}
void orc_diekman_f64(int64_t n, const double* c, double dt, double thr, double* const* x, const double* const* in,
                     double* out, uint8_t* spiked) {
    diekman_arrays<double>(n, c, dt, thr, x, in, out, spiked);
}

}  // extern "C"

// SAN step over arrays: x = 5 state arrays (V, nK, Ca, Ye, Yi), in = 4
// input arrays (Ein, Ecount, Iin, Icount), out = the output; c = 24
// constants in nnx_san_params order; spiked (optional): out > thr widened.
template <class T>
static void san_arrays(int64_t n, const double* c, double dt, double thr, T* const* x, const T* const* in, T* out,
                       uint8_t* spiked) {
    T k[24];
    for (int i = 0; i < 24; ++i) k[i] = (T)c[i];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        T s[5], v[4];
        for (int q = 0; q < 5; ++q) s[q] = x[q][i];
        for (int q = 0; q < 4; ++q) v[q] = in[q][i];
        out[i] = san_step<T>(k, (T)dt, s, v);
        for (int q = 0; q < 5; ++q) x[q][i] = s[q];
        if (spiked) spiked[i] = (double)out[i] > thr ? 1 : 0;
    }
}

extern "C" {
void orc_san_f32(int64_t n, const double* c, double dt, double thr, float* const* x, const float* const* in, float* out,
                 uint8_t* spiked) {
    san_arrays<float>(n, c, dt, thr, x, in, out, spiked);
}
void orc_san_f64(int64_t n, const double* c, double dt, double thr, double* const* x, const double* const* in,
                 double* out, uint8_t* spiked) {
    san_arrays<double>(n, c, dt, thr, x, in, out, spiked);
}

}  // extern "C"

extern "C" {
// OpenMP threads of the oracle's parallel loops (n <= 0: the default)
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : omp_get_num_procs()); }
int orc_threads(void) { return omp_get_max_threads(); }
}  // extern "C"

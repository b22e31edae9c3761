// nnx_compat.hpp - the reference's C++ entry points of the per-timestep hot
// path (ephys, coupling, paracrine), with their names and signatures, over
// the C ABI of include/nnx.h.  Header-only, plain C++17: a caller compiles it
// with any host compiler and links libnnx.so (+ libnnx_host.so for the
// trilinear transfer matrices); no HIP toolchain is needed.
//
// Replaced interfaces (neuronephys_clean sources):
//   Ephys / HHEphys::SimulateEphys(Float**)     src/ephys.cuh:6-31, src/ephys.cu:98-103
//   DiekmanEphys::SimulateEphys(Float**)        src/ephys.cu:250-353
//   SANEphys::SimulateEphys(Float**)            src/ephys.cu:203-248
//   AtomicCoupling(int), operator(), GetFiringNeurons()   src/coupling.cuh:17-38
//   AtomicCoupling(int, AdjacencyList&, Float*, Float*), PerformCoupling()
//                                               src_wholebrain/coupling.cuh:26-43
//   AdjacencyList                               src/coupling.cuh:4-8, src_wholebrain/coupling.cuh:11-16
//   ChunkedAdjacencyList, chunked operator()    src/coupling.cuh:10-14,37, src/coupling.cu:143-178
//   Network::ToChunkedAdjacencyList(Float*&)    src/network_factory.cuh:38, src/network_factory.cu:29-85
//   Density_on_grid_3d                          src/paracrine.cuh:121-173
//   Scatter_grid_converter_3d                   src/paracrine.cuh:184-209
//   ParacrineModule                             src/paracrine.cuh:212-267
//   ElectrodeNet, SynchronyNet                  src/electrode_net.cuh:6-38,
//                                               src_wholebrain/electrode_net.cuh:6-92
//   Morphology::GenSphericalRegions             src/morphology.cpp:48-97
//   Hebbian                                     src_wholebrain/hebbian.cuh:3-27
//   Network, FileNetwork                        src/network_factory.cuh:5-40,
//                                               src/network_factory.cu:5-27,155-212
//   FileMorphology                              src_wholebrain/morphology.cpp:255-370
//
// Deliberate differences (also listed in the design notes):
//   - thrust::device_vector members become nnx_compat::DeviceVector (same
//     data()/size() use); ParacrineModule takes coordinate arrays as device
//     pointers; density() returns the device pointer of the field.
//   - errors throw nnx_compat::Error (the reference prints and exits).
//   - the white noise of Leapfrog_HH (cuRAND XORWOW, src/ephys.cu:82) is
//     drawn by the library's counter-based generator (Philox4x64-10 keyed by
//     seed, call number and neuron; Box-Muller): same distribution, other
//     numbers; noise_std = 0 gives the deterministic mode.
//   - the chunked coupling is computed without chunks (same counts; see
//     AtomicCoupling), so it has no N_PER_CHUNK shared-memory limit of its
//     own, but ToChunkedAdjacencyList keeps the reference's layout and cap.
//   - update_on_fine_timescale adds the release to the firing neurons (the
//     reference's transform_if uses the compacted index list as its stencil,
//     src/paracrine.cuh:250-252, which selects list positions, not neurons).
//   - AtomicCoupling::is_firing is the spike bitmap (one bit per neuron), not
//     an int per neuron; ElectrodeNet's wholebrain overload takes it as is.
//   - host_vector arguments are std::vector; electrode positions are
//     std::array<float, 3> (no glm); Morphology holds positions only.
//   - FileNetwork / FileMorphology open the path as given (the reference
//     prepends "../"); FileNetwork keeps exactly N rows and only the kept
//     edges (the reference keeps numpnts+1 offsets and zero-filled index
//     slots, which ToAdjacencyList then counts as in-edges of neuron 0).
#pragma once

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nnx.h"
#include "../../include/nnx_host.h"

namespace nnx_compat {

using Float = float;  // USE_DOUBLE_PRECISION 0 (src/defines.cuh:50-55)

// SoA indices of the HH state (src/defines.cuh:57-78)
enum SimParams {
    SIM_V = 0, SIM_M = 1, SIM_N = 2, SIM_H = 3, SIM_LOCAL_Y = 4, SIM_IY = 5, SIM_EY = 6,
    SIM_LOCAL_OUTPUT = 7, SIM_LOCAL_IN = 8, SIM_LOCAL_COUNT = 9, SIM_EOUTPUT = 10, SIM_EIN = 11,
    SIM_ECOUNT = 12, SIM_IOUTPUT = 13, SIM_IIN = 14, SIM_ICOUNT = 15, SIM_APPCUR = 16,
    NUM_SIM = 17
};

struct Error : std::runtime_error {
    nnx_status status;
    Error(nnx_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// One context per process (the reference is single-GPU, single-thread).
inline nnx_ctx* ctx() {
    static nnx_ctx* c = [] {
        nnx_ctx* p = nullptr;
        const nnx_status s = nnx_create(&p, 0, nullptr);
        if (s != NNX_OK) throw Error(s, "nnx_create failed (no GPU?)");
        return p;
    }();
    return c;
}

inline void check(nnx_status s) {
    if (s != NNX_OK) throw Error(s, nnx_last_error(ctx()));
}

// Device array owning its memory (the role thrust::device_vector plays in
// the reference).
template <class T> class DeviceVector {
    T* p_ = nullptr;
    size_t n_ = 0;

   public:
    DeviceVector() = default;
    explicit DeviceVector(size_t n, bool zero = true) : n_(n) {
        void* v = nullptr;
        check(nnx_malloc(ctx(), &v, n * sizeof(T) + 16));
        p_ = static_cast<T*>(v);
        if (zero) check(nnx_memset(ctx(), p_, 0, n * sizeof(T) + 16));
    }
    explicit DeviceVector(const std::vector<T>& h) : DeviceVector(h.size(), false) { upload(h); }
    DeviceVector(const DeviceVector&) = delete;
    DeviceVector& operator=(const DeviceVector&) = delete;
    DeviceVector(DeviceVector&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DeviceVector& operator=(DeviceVector&& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~DeviceVector() {
        if (p_) nnx_free(ctx(), p_);
    }
    T* data() { return p_; }
    const T* data() const { return p_; }
    size_t size() const { return n_; }
    void upload(const std::vector<T>& h) {
        if (!h.empty()) check(nnx_memcpy_h2d(ctx(), p_, h.data(), h.size() * sizeof(T)));
    }
    std::vector<T> host() const {
        std::vector<T> h(n_);
        if (n_) check(nnx_memcpy_d2h(ctx(), h.data(), p_, n_ * sizeof(T)));
        return h;
    }
};

// CSR of out-edges (indices / offset_in_indices) as the reference's
// AdjacencyList; the transposed layout the deterministic coupling reads is
// built on the device on first use.
struct AdjacencyList {
    DeviceVector<int> indices;
    DeviceVector<int> offset_in_indices;
    DeviceVector<Float> weights;  // optional (empty: unweighted)

    struct Transposed {
        DeviceVector<int32_t> in_off, in_src, out_to_in;
        DeviceVector<float> in_w, in_deg;
        bool built = false;
    };
    mutable std::shared_ptr<Transposed> t = std::make_shared<Transposed>();

    AdjacencyList() = default;
    AdjacencyList(const std::vector<int>& offsets, const std::vector<int>& targets)
        : indices(targets), offset_in_indices(offsets) {}

    int64_t n() const { return offset_in_indices.size() ? (int64_t)offset_in_indices.size() - 1 : 0; }
    int64_t nnz() const { return (int64_t)indices.size(); }

    nnx_graph graph() const {
        Transposed& T = *t;
        if (!T.built) {
            const int64_t n = this->n(), e = nnz();
            T.in_off = DeviceVector<int32_t>(n + 1);
            T.in_src = DeviceVector<int32_t>(e ? e : 1);
            T.out_to_in = DeviceVector<int32_t>(e ? e : 1);
            T.in_w = DeviceVector<float>(e ? e : 1);
            T.in_deg = DeviceVector<float>(n ? n : 1);
            check(nnx_graph_transpose(ctx(), n, e, offset_in_indices.data(), indices.data(),
                                      weights.size() ? weights.data() : nullptr, T.in_off.data(),
                                      T.in_src.data(), T.in_w.data(), T.out_to_in.data(), NNX_F32,
                                      T.in_deg.data()));
            T.built = true;
        }
        nnx_graph g{};
        g.n = n();
        g.nnz = nnz();
        g.out_offsets = offset_in_indices.data();
        g.out_targets = indices.data();
        g.out_weights = weights.size() ? weights.data() : nullptr;
        g.in_offsets = T.in_off.data();
        g.in_sources = T.in_src.data();
        g.in_weights = T.in_w.data();
        g.out_to_in = T.out_to_in.data();
        return g;
    }
    // in-degree, as Network::ToAdjacencyList writes it into COUNT
    // (src/network_factory.cu:20-23)
    const float* in_degree() const {
        graph();
        return t->in_deg.data();
    }
};

// Launch and chunking constants of the reference (src/defines.cuh:81-84);
// the chunked list's layout caps N at NCHUNKS * N_PER_CHUNK as there.
constexpr int NBLOCKS = 32, NTHREADS = 64, NCHUNKS = 2, N_PER_CHUNK = 7319;

// The reference's chunked out-edge lists (src/coupling.cuh:10-14): for each
// target chunk c (target / N_PER_CHUNK), a CSR over all sources whose
// entries are the targets of that chunk relative to its first neuron.
// indices_chunks / offset_in_indices_chunks are device tables of device
// pointers, as there.  Network::ToChunkedAdjacencyList builds it and also
// keeps the unchunked list (`list`), which the library's deterministic
// coupling reads.
struct ChunkedAdjacencyList {
    DeviceVector<int*> indices_chunks;
    DeviceVector<int*> offset_in_indices_chunks;
    std::vector<DeviceVector<int>> chunk_indices, chunk_offsets;  // the arrays the tables point to
    AdjacencyList list;
};

// --------------------------------------------------------------- ephys ----
class Ephys {
   protected:
    int const N;

   public:
    explicit Ephys(int const N_) : N{N_} { ctx(); }
    virtual void SimulateEphys(Float** sim_vars) = 0;
    virtual ~Ephys() = default;
};

// Leapfrog_HH (src/ephys.cu:28-96) at hh_params::leapfrog_dt = 0.4 ms.
// sim_vars is a device array of NUM_SIM device pointers, as the reference's
// thrust::raw_pointer_cast(sim_vars.data()).
// The white noise white_noise_intensity * N(0,1) + white_noise_mean
// (src/ephys.cu:82, src/defines.cuh:99-101) is drawn per (seed, call,
// neuron) by the library's counter-based generator; params().noise_std = 0
// gives the deterministic mode.
class HHEphys : public Ephys {
    nnx_hh_params p_{};
    uint64_t step_ = 0;

   public:
    static constexpr Float leapfrog_dt = 0.4f;  // src/defines.cuh:89
    explicit HHEphys(int const N_) : Ephys(N_) { nnx_hh_params_reference(&p_); }
    nnx_hh_params& params() { return p_; }
    uint64_t steps_done() const { return step_; }
    void SimulateEphys(Float** sim_vars) override {
        // the pointer table is read on every call, as the reference's kernel
        // reads it on every launch
        std::vector<Float*> a(NUM_SIM);
        check(nnx_memcpy_d2h(ctx(), a.data(), sim_vars, NUM_SIM * sizeof(Float*)));
        nnx_hh_state st{};
        st.dtype = NNX_F32;
        st.n = N;
        st.v = a[SIM_V];
        st.m = a[SIM_M];
        st.h = a[SIM_H];
        st.ng = a[SIM_N];
        st.y_local = a[SIM_LOCAL_Y];
        st.y_i = a[SIM_IY];
        st.y_e = a[SIM_EY];
        st.in_local = a[SIM_LOCAL_IN];
        st.count_local = a[SIM_LOCAL_COUNT];
        st.in_i = a[SIM_IIN];
        st.count_i = a[SIM_ICOUNT];
        st.in_e = a[SIM_EIN];
        st.count_e = a[SIM_ECOUNT];
        st.appcur = a[SIM_APPCUR];
        st.out_local = a[SIM_LOCAL_OUTPUT];
        st.out_i = a[SIM_IOUTPUT];
        st.out_e = a[SIM_EOUTPUT];
        check(nnx_hh_step(ctx(), &st, &p_, NNX_INT_LEAPFROG, NNX_SPIKE_LEVEL, leapfrog_dt, step_++, nullptr,
                          nullptr));
        check(nnx_synchronize(ctx()));  // the reference synchronises after the launch
    }
};

// SoA indices of the Diekman model (src/defines.cuh:153-173)
enum DiekmanVars {
    DIEKMAN_VAR_V = 0, DIEKMAN_VAR_M, DIEKMAN_VAR_H, DIEKMAN_VAR_N, DIEKMAN_VAR_RL, DIEKMAN_VAR_RNL, DIEKMAN_VAR_FNL,
    DIEKMAN_VAR_S, DIEKMAN_VAR_P, DIEKMAN_VAR_Y, DIEKMAN_VAR_CAS, DIEKMAN_VAR_CAC, DIEKMAN_VAR_G_BIOCHEM,
    DIEKMAN_VAR_IIN, DIEKMAN_VAR_ICOUNT, DIEKMAN_VAR_IOUTPUT, DIEKMAN_VAR_EGABA, DIEKMAN_VAR_PARACRINE,
    DIEKMAN_VAR_GKCA, DIEKMAN_VAR_GKLEAK, NUM_DIEKMAN_VAR
};

// Leapfrog_diekman at hh_params::leapfrog_dt (0.4 ms), as
// DiekmanEphys::SimulateEphys launches it; sim_vars is a device array of
// NUM_DIEKMAN_VAR device pointers.
class DiekmanEphys : public Ephys {
    nnx_diekman_params p_{};

   public:
    explicit DiekmanEphys(int const N_) : Ephys(N_) { nnx_diekman_params_reference(&p_); }
    nnx_diekman_params& params() { return p_; }
    void SimulateEphys(Float** sim_vars) override {
        // the pointer table is read on every call (the reference's kernel
        // reads it on every launch)
        std::vector<Float*> a(NUM_DIEKMAN_VAR);
        check(nnx_memcpy_d2h(ctx(), a.data(), sim_vars, NUM_DIEKMAN_VAR * sizeof(Float*)));
        nnx_diekman_state st{};
        st.dtype = NNX_F32;
        st.n = N;
        st.v = a[DIEKMAN_VAR_V], st.m = a[DIEKMAN_VAR_M], st.h = a[DIEKMAN_VAR_H], st.ng = a[DIEKMAN_VAR_N];
        st.rl = a[DIEKMAN_VAR_RL], st.rnl = a[DIEKMAN_VAR_RNL], st.fnl = a[DIEKMAN_VAR_FNL], st.s = a[DIEKMAN_VAR_S];
        st.p = a[DIEKMAN_VAR_P], st.y = a[DIEKMAN_VAR_Y], st.cas = a[DIEKMAN_VAR_CAS], st.cac = a[DIEKMAN_VAR_CAC];
        st.iin = a[DIEKMAN_VAR_IIN], st.icount = a[DIEKMAN_VAR_ICOUNT], st.ioutput = a[DIEKMAN_VAR_IOUTPUT];
        st.egaba = a[DIEKMAN_VAR_EGABA], st.gkca = a[DIEKMAN_VAR_GKCA], st.gkleak = a[DIEKMAN_VAR_GKLEAK];
        check(nnx_diekman_step(ctx(), &st, &p_, HHEphys::leapfrog_dt, nullptr));
        check(nnx_synchronize(ctx()));
    }
};

// SoA indices of the SAN model (src/defines.cuh:107-117)
enum AnIeVars {
    AN_IE_VAR_V = 0, AN_IE_VAR_NK, AN_IE_VAR_CA, AN_IE_VAR_YE, AN_IE_VAR_YI, AN_IE_VAR_EIN, AN_IE_VAR_ECOUNT,
    AN_IE_VAR_EOUTPUT, AN_IE_VAR_IIN, AN_IE_VAR_ICOUNT, AN_IE_VAR_IOUTPUT, NUM_AN_IE_VAR
};

// Euler_an_ie_model at hh_params::leapfrog_dt (0.4 ms), as
// SANEphys::SimulateEphys launches it (white noise replaced by its mean).
class SANEphys : public Ephys {
    nnx_san_params p_{};

   public:
    explicit SANEphys(int const N_) : Ephys(N_) { nnx_san_params_reference(&p_); }
    nnx_san_params& params() { return p_; }
    void SimulateEphys(Float** sim_vars) override {
        // the pointer table is read on every call (the reference's kernel
        // reads it on every launch)
        std::vector<Float*> a(NUM_AN_IE_VAR);
        check(nnx_memcpy_d2h(ctx(), a.data(), sim_vars, NUM_AN_IE_VAR * sizeof(Float*)));
        nnx_san_state st{};
        st.dtype = NNX_F32;
        st.n = N;
        st.v = a[AN_IE_VAR_V], st.nk = a[AN_IE_VAR_NK], st.ca = a[AN_IE_VAR_CA], st.ye = a[AN_IE_VAR_YE];
        st.yi = a[AN_IE_VAR_YI], st.ein = a[AN_IE_VAR_EIN], st.ecount = a[AN_IE_VAR_ECOUNT], st.iin = a[AN_IE_VAR_IIN];
        st.icount = a[AN_IE_VAR_ICOUNT], st.eoutput = a[AN_IE_VAR_EOUTPUT], st.ioutput = a[AN_IE_VAR_IOUTPUT];
        check(nnx_san_step(ctx(), &st, &p_, HHEphys::leapfrog_dt, nullptr));
        check(nnx_synchronize(ctx()));
    }
};

// ------------------------------------------------------------ coupling ----
// Threshold (output > 0.1) -> ascending firing list -> deterministic
// propagation.  The src/ form adds 1 per firing source; the
// src_wholebrain/ form (constructed with the list and the two arrays) adds
// the source's output.
class AtomicCoupling {
    int N;
    DeviceVector<int32_t> count_;
    Float* downstream_inputs = nullptr;
    Float* neuron_outputs = nullptr;

    int couple(const AdjacencyList& l, Float* spikes, const Float* outputs, int mode) {
        check(nnx_spike_threshold(ctx(), NNX_F32, N, outputs, 0.1, is_firing.data()));
        check(nnx_spike_compact(ctx(), N, is_firing.data(), firing_neurons.data(), count_.data()));
        const nnx_graph g = l.graph();
        check(nnx_couple(ctx(), &g, mode, NNX_F32, is_firing.data(), firing_neurons.data(),
                         count_.data(), outputs, spikes));
        int32_t k = 0;
        check(nnx_read_count(ctx(), count_.data(), &k));  // the reference returns the count
        return k;
    }

   public:
    AdjacencyList list;                    // src_wholebrain form
    DeviceVector<uint64_t> is_firing;      // bit i: neuron i's output exceeded the level
    DeviceVector<int32_t> firing_neurons;  // ascending; first count entries valid

    explicit AtomicCoupling(int const N_)
        : N(N_), count_(1), is_firing((N_ + 63) / 64 + 1), firing_neurons(N_ + 1) {}
    AtomicCoupling(int const N_, AdjacencyList& list_, Float* downstream_inputs_, Float* neuron_outputs_)
        : AtomicCoupling(N_) {
        std::swap(list, list_);  // the reference moves the list in (src_wholebrain/coupling.cu:6-9)
        downstream_inputs = downstream_inputs_;
        neuron_outputs = neuron_outputs_;
    }
    int operator()(const AdjacencyList& connectivity, Float* spikes, const Float* outputs) {
        return couple(connectivity, spikes, outputs, NNX_COUPLE_COUNT);
    }
    // The chunked form (src/coupling.cu:143-178) merges per-block shared-
    // memory histograms of the firing sources' targets into `spikes`: the
    // number of firing sources of each target, the counts the unchunked
    // form produces (integer sums do not depend on the order).  Here they
    // come from the deterministic COUNT coupling over the unchunked list,
    // without the reference's shared-memory race (src/coupling.cu:63-83).
    // `spikes` is overwritten on every call; the reference leaves it
    // untouched until the first call that finds firing neurons and adds that
    // call's counts to it (the same when it starts at zero).
    int operator()(const ChunkedAdjacencyList& connectivity, Float* spikes, const Float* outputs) {
        return couple(connectivity.list, spikes, outputs, NNX_COUPLE_COUNT);
    }
    int PerformCoupling() { return couple(list, downstream_inputs, neuron_outputs, NNX_COUPLE_SIGNAL); }
    const DeviceVector<int32_t>& GetFiringNeurons() const { return firing_neurons; }
    const int32_t* device_count() const { return count_.data(); }
};

// ----------------------------------------------------------- paracrine ----
// The grid (x fastest) with the reference's Crank-Nicolson 27-point update.
struct Density_on_grid_3d {
    Float spatial_gridsize;
    int x_count, y_count, z_count;
    Float diffusion_coefficient, decay_coefficient;
    DeviceVector<Float> data;
    Float* raw_data_d_ptr;
    nnx_cg_result last{};

    Density_on_grid_3d(int x_count_, int y_count_, int z_count_, Float h, Float D, Float decay)
        : spatial_gridsize(h), x_count(x_count_), y_count(y_count_), z_count(z_count_),
          diffusion_coefficient(D), decay_coefficient(decay),
          data((size_t)x_count_ * y_count_ * z_count_), raw_data_d_ptr(data.data()) {}
    void update_density(const Float timestep, const Float error_bound = 1e-6f) {
        check(nnx_update_density_reference(ctx(), x_count, y_count, z_count, spatial_gridsize,
                                           diffusion_coefficient, decay_coefficient, timestep,
                                           error_bound, 20, data.data(), &last));
    }
    void set_delta() {
        std::vector<Float> one(1, 1.0f);
        const size_t i = x_count / 2 + (size_t)x_count * (y_count / 2 + (size_t)y_count * (z_count / 2));
        check(nnx_memcpy_h2d(ctx(), data.data() + i, one.data(), sizeof(Float)));
    }
    const Float* density() const { return data.data(); }
};

// Trilinear neuron <-> grid transfer (src/paracrine.cu:136-296): interpolation
// CSR (N x G, 8 per row) and its transpose for the scatter.
class Scatter_grid_converter_3d {
    DeviceVector<int32_t> w_off_, w_col_, s_off_, s_col_;
    DeviceVector<float> w_val_, s_val_;
    int64_t G_;

   public:
    const int neuron_count;
    int64_t outside = 0;

    Scatter_grid_converter_3d(const Density_on_grid_3d& g, int N, const Float* d_x, const Float* d_y,
                              const Float* d_z)
        : G_((int64_t)g.x_count * g.y_count * g.z_count), neuron_count(N) {
        std::vector<float> x(N), y(N), z(N);
        check(nnx_memcpy_d2h(ctx(), x.data(), d_x, N * sizeof(float)));
        check(nnx_memcpy_d2h(ctx(), y.data(), d_y, N * sizeof(float)));
        check(nnx_memcpy_d2h(ctx(), z.data(), d_z, N * sizeof(float)));
        std::vector<int32_t> off(N + 1), col(8 * (size_t)N), toff(G_ + 1), tcol(8 * (size_t)N);
        std::vector<float> val(8 * (size_t)N), tval(8 * (size_t)N);
        outside = nnx_trilinear_matrix(N, x.data(), y.data(), z.data(), g.x_count, g.y_count,
                                       g.z_count, g.spatial_gridsize, off.data(), col.data(),
                                       val.data());
        nnx_csr_transpose(N, G_, off.data(), col.data(), val.data(), toff.data(), tcol.data(),
                          tval.data());
        w_off_ = DeviceVector<int32_t>(off);
        w_col_ = DeviceVector<int32_t>(col);
        w_val_ = DeviceVector<float>(val);
        s_off_ = DeviceVector<int32_t>(toff);
        s_col_ = DeviceVector<int32_t>(tcol);
        s_val_ = DeviceVector<float>(tval);
    }
    // grid += W^T scatter (rows of the transpose list neurons in ascending order)
    void scatter_to_grid(const Float* scatter, Density_on_grid_3d& grid) {
        check(nnx_csr_spmv(ctx(), G_, s_off_.data(), s_col_.data(), s_val_.data(), scatter,
                           grid.data.data(), 1));
    }
    // out = W grid
    void interpolate_at_location(Float* out, const Density_on_grid_3d& grid) {
        check(nnx_csr_spmv(ctx(), neuron_count, w_off_.data(), w_col_.data(), w_val_.data(),
                           grid.density(), out, 0));
    }
};

class ParacrineModule {
    Density_on_grid_3d grid;
    Scatter_grid_converter_3d grid_converter;

   public:
    int x_gridcount, y_gridcount, z_gridcount, timescale;

   private:
    Float* d_g_paracrine;
    Float par_dt, coup_strength;
    DeviceVector<Float> density_at_neurons, density_generated_at_neurons;

   public:
    ParacrineModule(Float*& d_g_paracrine_, int const neuron_count, int const nx, int const ny,
                    int const nz, Float const par_dt_, int const timescale_, Float const coup_strength_,
                    Float const spatial_gridsize, Float const diffusion_coef, Float const decay_coef,
                    const Float* d_x, const Float* d_y, const Float* d_z)
        : grid(nx, ny, nz, spatial_gridsize, diffusion_coef, decay_coef),
          grid_converter(grid, neuron_count, d_x, d_y, d_z),
          x_gridcount(nx), y_gridcount(ny), z_gridcount(nz), timescale(timescale_),
          d_g_paracrine(d_g_paracrine_), par_dt(par_dt_), coup_strength(coup_strength_),
          density_at_neurons(neuron_count), density_generated_at_neurons(neuron_count) {
        grid_converter.interpolate_at_location(density_at_neurons.data(), grid);
        check(nnx_memcpy_d2d(ctx(), d_g_paracrine, density_at_neurons.data(),
                             neuron_count * sizeof(Float)));
    }
    // release of this step: +coup_strength at every firing neuron
    void update_on_fine_timescale(const AtomicCoupling& coupling) {
        check(nnx_release_add(ctx(), coupling.GetFiringNeurons().data(), coupling.device_count(),
                              coup_strength, density_generated_at_neurons.data()));
    }
    // scatter the release, one Crank-Nicolson update, interpolate back
    // (src/paracrine.cuh:255-265)
    void update_on_own_timescale() {
        grid_converter.scatter_to_grid(density_generated_at_neurons.data(), grid);
        grid.update_density(par_dt);
        grid_converter.interpolate_at_location(density_at_neurons.data(), grid);
        check(nnx_memset(ctx(), density_generated_at_neurons.data(), 0,
                         density_generated_at_neurons.size() * sizeof(Float)));
        check(nnx_memcpy_d2d(ctx(), d_g_paracrine, density_at_neurons.data(),
                             density_at_neurons.size() * sizeof(Float)));
    }
    const Density_on_grid_3d& get_grid() const { return grid; }
};

// ----------------------------------------------------------- readout ----
// Largest index + 1 of a region list: the neuron count the device bounds
// its reads by (the reference's classes are not told N).
inline int64_t index_bound(const std::vector<int>& idx) {
    int64_t m = 0;
    for (int i : idx) m = std::max<int64_t>(m, (int64_t)i + 1);
    return m;
}

// Electrodes measuring the mean voltage (and, wholebrain form, the number
// of firing neurons) of regions stored as an adjacency list.
class ElectrodeNet {
    DeviceVector<int32_t> inds_, offs_;
    int N_regions;
    int64_t nnz_, n_;
    DeviceVector<Float> d_avg_voltages;
    DeviceVector<int32_t> d_firing_per_region;
    std::vector<Float> h_avg_voltages;
    std::vector<int> h_firing_per_region;

    nnx_regions regions() const {
        nnx_regions r{};
        r.n_regions = N_regions;
        r.nnz = nnz_;
        r.offsets = offs_.data();
        r.idx = inds_.data();
        return r;
    }

   public:
    ElectrodeNet(const std::vector<int>& regions_inds, const std::vector<int>& regions_offsets)
        : inds_(regions_inds.empty() ? std::vector<int>(1, 0) : regions_inds), offs_(regions_offsets),
          N_regions((int)regions_offsets.size() - 1), nnz_((int64_t)regions_inds.size()),
          n_(index_bound(regions_inds)), d_avg_voltages(std::max(N_regions, 1)),
          d_firing_per_region(std::max(N_regions, 1)), h_avg_voltages(std::max(N_regions, 0), 0),
          h_firing_per_region(std::max(N_regions, 0), 0) {}

    // mean V per region, copied to the host (src/electrode_net.cu:23-36)
    void DoElectrodeMeasurements(const Float* V_ptr) {
        const nnx_regions r = regions();
        check(nnx_electrode_measure(ctx(), &r, NNX_F32, n_, V_ptr, nullptr, d_avg_voltages.data(), nullptr));
        if (N_regions > 0)
            check(nnx_memcpy_d2h(ctx(), h_avg_voltages.data(), d_avg_voltages.data(), N_regions * sizeof(Float)));
    }
    // and the firing count per region (src_wholebrain/electrode_net.cu:32-49)
    void DoElectrodeMeasurements(const Float* V_ptr, const DeviceVector<uint64_t>& is_firing) {
        const nnx_regions r = regions();
        check(nnx_electrode_measure(ctx(), &r, NNX_F32, n_, V_ptr, is_firing.data(), d_avg_voltages.data(),
                                    d_firing_per_region.data()));
        if (N_regions > 0) {
            check(nnx_memcpy_d2h(ctx(), h_avg_voltages.data(), d_avg_voltages.data(), N_regions * sizeof(Float)));
            check(nnx_memcpy_d2h(ctx(), h_firing_per_region.data(), d_firing_per_region.data(),
                                 N_regions * sizeof(int)));
        }
    }
    static Float MeasureAverageEntireNetwork(const Float* V_ptr, int const N) {
        static DeviceVector<Float> out(1);
        check(nnx_network_mean(ctx(), NNX_F32, N, V_ptr, out.data()));
        Float m = 0;
        check(nnx_memcpy_d2h(ctx(), &m, out.data(), sizeof(Float)));
        return m;
    }
    const Float* ElectrodeOutputHandle() const { return h_avg_voltages.data(); }
    const int* FiringPerRegionHandle() const { return h_firing_per_region.data(); }
};

// chi^2 synchrony of sub-populations (the first sub_sample_counts[s]
// neurons of the region) over the timesteps measured so far.
class SynchronyNet {
    std::vector<int> h_region_inds;
    DeviceVector<int32_t> d_region_inds;
    std::vector<int> sub_sample_counts;
    DeviceVector<int32_t> d_counts;
    DeviceVector<double> d_Vi_sum, d_Vi2_sum, d_Vavg_sum, d_Vavg2_sum, d_chi;
    int64_t n_;
    int N_temporal_steps = 0;

   public:
    SynchronyNet(std::vector<int>&& h_region_inds_, std::vector<int> sub_sample_rates_)
        : h_region_inds(std::move(h_region_inds_)),
          d_region_inds(h_region_inds.empty() ? std::vector<int>(1, 0) : h_region_inds),
          sub_sample_counts(std::move(sub_sample_rates_)),
          d_counts(sub_sample_counts.empty() ? std::vector<int>(1, 0) : sub_sample_counts),
          d_Vi_sum(std::max<size_t>(h_region_inds.size(), 1)), d_Vi2_sum(std::max<size_t>(h_region_inds.size(), 1)),
          d_Vavg_sum(std::max<size_t>(sub_sample_counts.size(), 1)),
          d_Vavg2_sum(std::max<size_t>(sub_sample_counts.size(), 1)),
          d_chi(std::max<size_t>(sub_sample_counts.size(), 1)), n_(index_bound(h_region_inds)) {}

    void MeasureVariances(const Float* V_ptr) {
        check(nnx_synchrony_accumulate(ctx(), NNX_F32, n_, V_ptr, (int64_t)h_region_inds.size(),
                                       d_region_inds.data(), (int32_t)sub_sample_counts.size(), d_counts.data(),
                                       d_Vi_sum.data(), d_Vi2_sum.data(), d_Vavg_sum.data(), d_Vavg2_sum.data()));
        N_temporal_steps += 1;
    }
    std::vector<double> GetSynchronyMeasures() {
        const int32_t ns = (int32_t)sub_sample_counts.size();
        check(nnx_synchrony_measures(ctx(), (int64_t)h_region_inds.size(), ns, d_counts.data(), d_Vi_sum.data(),
                                     d_Vi2_sum.data(), d_Vavg_sum.data(), d_Vavg2_sum.data(), N_temporal_steps,
                                     d_chi.data()));
        std::vector<double> chi(ns);
        if (ns) check(nnx_memcpy_d2h(ctx(), chi.data(), d_chi.data(), ns * sizeof(double)));
        return chi;
    }
};

// ------------------------------------------------------------ hebbian ----
// Hebbian learning of connectivity.weights (CSR order), one call per
// timestep.  The firing set is read from is_firing (the coupling's bitmap);
// firing_inds names the same neurons and is not needed.
class Hebbian {
    int Nconns;
    int64_t calls_ = 0;
    const void* bound_ = nullptr;  // the list the state was built for
    DeviceVector<uint8_t> armed_;
    DeviceVector<int32_t> last_, in_to_out_;
    DeviceVector<uint64_t> ring_;

   public:
    int const window_duration;
    Float const heb_inc;

    Hebbian(int window_duration_, Float heb_inc_, int Nconns_)
        : Nconns(Nconns_), window_duration(window_duration_), heb_inc(heb_inc_) {}

    void UpdateWeights(AdjacencyList& connectivity, const DeviceVector<uint64_t>& is_firing,
                       const DeviceVector<int32_t>& firing_inds) {
        (void)firing_inds;
        const nnx_graph g = connectivity.graph();
        if ((int64_t)connectivity.weights.size() != g.nnz || g.nnz != Nconns)
            throw Error(NNX_ERR_ARG, "Hebbian: the list needs one weight per synapse (Nconns)");
        nnx_hebbian h{};
        h.window = window_duration;
        h.heb_inc = heb_inc;
        h.weights = connectivity.weights.data();
        if (bound_ != &connectivity) {
            armed_ = DeviceVector<uint8_t>(std::max<int64_t>(g.nnz, 1));
            in_to_out_ = DeviceVector<int32_t>(std::max<int64_t>(g.nnz, 1));
            last_ = DeviceVector<int32_t>(std::max<int64_t>(g.n, 1));
            ring_ = DeviceVector<uint64_t>(std::max<int64_t>((int64_t)window_duration * ((g.n + 63) / 64), 1));
        }
        h.armed = armed_.data();
        h.last_fire = last_.data();
        h.ring = ring_.data();
        h.in_to_out = in_to_out_.data();
        if (bound_ != &connectivity) {
            check(nnx_hebbian_init(ctx(), &g, &h));
            bound_ = &connectivity;
            calls_ = 0;
        }
        check(nnx_hebbian_update(ctx(), &g, &h, calls_++, is_firing.data()));
    }
};

// ------------------------------------------------------------- files ----
// Host CSR of out-edges, as the reference's Network base class.
class Network {
   protected:
    std::vector<int> indices, offset_in_indices;

   public:
    Network() = default;
    virtual ~Network() = default;
    virtual bool Generate(int const N) = 0;
    // the list on the device, and the in-degree of each neuron written into
    // the caller's device array Tcount (N floats), as src/network_factory.cu:5-27
    AdjacencyList ToAdjacencyList(Float*& Tcount) const {
        AdjacencyList list(offset_in_indices, indices);
        const int64_t N = (int64_t)offset_in_indices.size() - 1;
        if (N > 0) check(nnx_memcpy_d2d(ctx(), Tcount, list.in_degree(), N * sizeof(Float)));
        return list;
    }
    // The chunked lists of src/network_factory.cu:29-85 (targets split by
    // target / N_PER_CHUNK, each chunk's entries relative to its first
    // neuron, one offset array of N+1 per chunk), the in-degree written into
    // Tcount as ToAdjacencyList does.
    ChunkedAdjacencyList ToChunkedAdjacencyList(Float*& Tcount) const {
        ChunkedAdjacencyList cl;
        const int64_t N = (int64_t)offset_in_indices.size() - 1;
        if (N < 0) {
            std::fprintf(stderr, "Connectivity is empty. Returning blank list.\n");
            return cl;
        }
        std::vector<std::vector<int>> ind(NCHUNKS), off(NCHUNKS, std::vector<int>(N + 1, 0));
        for (int64_t i = 0; i < N; ++i) {
            for (auto& o : off) o[i + 1] = o[i];
            for (int j = offset_in_indices[i]; j < offset_in_indices[i + 1]; ++j) {
                const int t = indices[j], c = t / N_PER_CHUNK;
                if (t < 0 || c >= NCHUNKS)
                    throw Error(NNX_ERR_ARG, "target " + std::to_string(t) + " lies beyond the " +
                                                 std::to_string(NCHUNKS) + " chunks of " + std::to_string(N_PER_CHUNK));
                ind[c].push_back(t % N_PER_CHUNK);
                off[c][i + 1] += 1;
            }
        }
        std::vector<int*> pi(NCHUNKS), po(NCHUNKS);
        for (int c = 0; c < NCHUNKS; ++c) {
            cl.chunk_indices.emplace_back(ind[c].empty() ? std::vector<int>(1, 0) : ind[c]);
            cl.chunk_offsets.emplace_back(off[c]);
            pi[c] = cl.chunk_indices.back().data();
            po[c] = cl.chunk_offsets.back().data();
        }
        cl.indices_chunks = DeviceVector<int*>(pi);
        cl.offset_in_indices_chunks = DeviceVector<int*>(po);
        cl.list = ToAdjacencyList(Tcount);
        return cl;
    }
    const std::vector<int>& GetCpuIndices() const { return indices; }
    const std::vector<int>& GetCpuOffsetsIntoIndices() const { return offset_in_indices; }
};

// Connectivity from an HDF5 file holding /row_ptr and /col_idx, sampled to
// N neurons as FileNetwork::GenerateFromHDF does.
class FileNetwork : public Network {
    std::string fl_name, format;

   public:
    int64_t dropped = 0;  // sampled targets >= N (numpnts not a multiple of N)
    explicit FileNetwork(const std::string& fl_name_)
        : fl_name(fl_name_), format(fl_name_.substr(fl_name_.find_last_of('.') + 1)) {}
    bool Generate(int const N) override {
        if (format != "h5" && format != "H5") {
            std::fprintf(stderr, "Unsupported format %s\n", format.c_str());
            return false;
        }
        char err[512];
        const int64_t nr = nnx_h5_count(fl_name.c_str(), "/row_ptr", err, sizeof err);
        const int64_t nc = nnx_h5_count(fl_name.c_str(), "/col_idx", err, sizeof err);
        if (nr < 2 || nc < 0) {
            std::fprintf(stderr, "ERROR : Failed to read in coupling (%s)\n", err);
            return false;
        }
        std::vector<int32_t> rp(nr), ci(std::max<int64_t>(nc, 1));
        if (nnx_h5_read(fl_name.c_str(), "/row_ptr", NNX_H5_I32, rp.data(), nr, err, sizeof err) < 0 ||
            nnx_h5_read(fl_name.c_str(), "/col_idx", NNX_H5_I32, ci.data(), nc, err, sizeof err) < 0) {
            std::fprintf(stderr, "ERROR : Failed to read in coupling (%s)\n", err);
            return false;
        }
        offset_in_indices.assign(N + 1, 0);
        indices.assign(std::max<int64_t>(nc, 1), 0);
        const int64_t k = nnx_csr_subsample(nr - 1, rp.data(), ci.data(), N, offset_in_indices.data(),
                                            indices.data(), &dropped);
        if (k < 0) {
            std::fprintf(stderr, "Not enough points in file %s: N = %d, %lld available\n", fl_name.c_str(), N,
                         (long long)(nr - 1));
            return false;
        }
        indices.resize(k);
        return true;
    }
};

// Neuron positions and the spherical electrode regions built from them.
struct Morphology {
    std::vector<float> xyz;  // 3 per neuron, interleaved

    Morphology() = default;
    explicit Morphology(std::vector<float> xyz_) : xyz(std::move(xyz_)) {}
    int64_t size() const { return (int64_t)xyz.size() / 3; }

    // region n = the k neurons nearest to neuron electrode_inds[n]
    void GenSphericalRegions(std::vector<int>& h_regions, std::vector<int>& h_region_inds,
                             const std::vector<int>& electrode_inds, int const k) const {
        const int64_t nr = (int64_t)electrode_inds.size();
        h_regions.assign(nr * k, 0);
        h_region_inds.assign(nr + 1, 0);
        if (nnx_spherical_regions(size(), xyz.data(), nr, electrode_inds.data(), nullptr, k, h_region_inds.data(),
                                  h_regions.data()) != 0)
            throw Error(NNX_ERR_ARG, "GenSphericalRegions: k or an electrode index out of range");
    }
    // region n = the k neurons nearest to the point electrode_positions[n]
    void GenSphericalRegions(std::vector<int>& h_regions, std::vector<int>& h_region_inds,
                             const std::vector<std::array<float, 3>>& electrode_positions, int const k) const {
        const int64_t nr = (int64_t)electrode_positions.size();
        std::vector<float> c(3 * nr);
        for (int64_t r = 0; r < nr; ++r)
            for (int d = 0; d < 3; ++d) c[3 * r + d] = electrode_positions[r][d];
        h_regions.assign(nr * k, 0);
        h_region_inds.assign(nr + 1, 0);
        if (nnx_spherical_regions(size(), xyz.data(), nr, nullptr, c.data(), k, h_region_inds.data(),
                                  h_regions.data()) != 0)
            throw Error(NNX_ERR_ARG, "GenSphericalRegions: k out of range");
    }
};

// Positions (and region ids) from an HDF5 file holding /x, /y, /z (and
// /region_id), sampled to N neurons and mapped to [-1, 1]^3 as
// FileMorphology::GenerateFromHDF does.
struct FileMorphology : Morphology {
    std::string fl_name, format;
    bool read_regions;
    std::vector<int> regions, unique_regions;

    explicit FileMorphology(const std::string& fl_name_, bool read_regions_ = false)
        : fl_name(fl_name_), format(fl_name_.substr(fl_name_.find_last_of('.') + 1)), read_regions(read_regions_) {}
    bool Generate(int const N) {
        if (format != "h5" && format != "H5") return false;
        const char* ids[4] = {"/x", "/y", "/z", "/region_id"};
        std::vector<float> ax[4];
        char err[512];
        for (int d = 0; d < (read_regions ? 4 : 3); ++d) {
            const int64_t n = nnx_h5_count(fl_name.c_str(), ids[d], err, sizeof err);
            if (n < 1) return false;
            ax[d].resize(n);
            if (nnx_h5_read(fl_name.c_str(), ids[d], NNX_H5_F32, ax[d].data(), n, err, sizeof err) < 0) return false;
        }
        const int64_t np = (int64_t)ax[0].size();
        if (ax[1].size() != (size_t)np || ax[2].size() != (size_t)np) return false;
        xyz.assign(3 * (size_t)N, 0.0f);
        if (nnx_morphology_from_axes(np, ax[0].data(), ax[1].data(), ax[2].data(), N, xyz.data()) != 0) {
            std::fprintf(stderr, "Not enough points in file %s\n", fl_name.c_str());
            return false;
        }
        if (read_regions) {
            const int64_t stride = np / N;
            regions.resize(N);
            for (int i = 0; i < N; ++i) regions[i] = (int)ax[3][stride * i];
            unique_regions = regions;
            std::sort(unique_regions.begin(), unique_regions.end());
            unique_regions.erase(std::unique(unique_regions.begin(), unique_regions.end()), unique_regions.end());
        }
        return true;
    }
};

}  // namespace nnx_compat

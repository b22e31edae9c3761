// compat_demo.cpp - the reference's simulation loop (src/main.cu:112-181:
// coupling -> paracrine -> ephys per increment, coupling through the
// reference's own CouplingStruct over a ChunkedAdjacencyList,
// src/main.cu:22-33,264-265) written against nnx_compat.hpp, with inputs and
// outputs as raw files so that a test can check the run against the oracle.
// The ephys step draws the reference's white noise (mean 3, std 1.5) unless
// the environment sets COMPAT_DEMO_NOISE_STD (e.g. 0 for the deterministic
// mode).
//
//   compat_demo DIR nincs nx ny nz h D decay par_dt timescale coup_strength
//               [print_incs render_N k]
//
// DIR/off.i32 (N+1), DIR/tgt.i32 (nnz), DIR/sv.f32 (17 x N, SoA in the
// SIM_PARAMS order), DIR/pos.f32 (3 x N).  Writes DIR/sv_out.f32,
// DIR/raster.i32 ((step, neuron) pairs of the post-step firing),
// DIR/nfiring.i32 (per-increment return of the coupling), DIR/g_par.f32,
// DIR/density.f32.  With print_incs, render_N electrodes at neurons
// (N / render_N) * e, each the mean of its k nearest neurons
// (src/main.cu:279-285), are read at every print increment before its
// coupling (src/main.cu:133-141): DIR/electrodes.f32 holds one row per print
// increment, the render_N means then the network mean; DIR/regions.i32 the
// region indices.  If DIR/w.f32 (nnz weights) exists, Hebbian learning
// (window 20, heb_inc 0.05) updates them after every increment's ephys, on
// the firing set of that increment's coupling (src_wholebrain/main.cu:
// 158-168); the result is DIR/w_out.f32.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "nnx_compat.hpp"

using namespace nnx_compat;

template <class T> static std::vector<T> load(const std::string& path) {
    std::vector<T> v;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    v.resize(bytes / sizeof(T));
    if (!v.empty() && std::fread(v.data(), sizeof(T), v.size(), f) != v.size())
        throw std::runtime_error("short read " + path);
    std::fclose(f);
    return v;
}

template <class T> static void save(const std::string& path, const std::vector<T>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

// src/main.cu:22-33, unchanged
struct CouplingStruct 
{
    ChunkedAdjacencyList& list;
    AtomicCoupling& coupling;
    Float* downstream_inputs;
    Float* neuron_outputs;
    
    // Performs coupling and stores results in downstream_inputs. 
    // Returns number of neurons firing.
    int operator() () 
    {
        return coupling(list, downstream_inputs, neuron_outputs);
    }
};

// the network of the input files (the reference generates one on the host,
// e.g. UniformRandomNetwork, src/network_factory.cu:96-131)
class CsrFileNetwork : public Network {
   public:
    CsrFileNetwork(const std::vector<int>& off, const std::vector<int>& tgt) {
        offset_in_indices = off;
        indices = tgt;
    }
    bool Generate(int const) override { return true; }
};

static void record(const AtomicCoupling& c, int n_firing, int step, std::vector<int32_t>& raster) {
    if (step < 0 || n_firing == 0) return;
    std::vector<int32_t> l = c.GetFiringNeurons().host();
    for (int q = 0; q < n_firing; ++q) {
        raster.push_back(step);
        raster.push_back(l[q]);
    }
}

int main(int argc, char** argv) {
    if (argc != 12 && argc != 15) {
        std::fprintf(stderr,
                     "usage: %s DIR nincs nx ny nz h D decay par_dt timescale coup_strength "
                     "[print_incs render_N k]\n",
                     argv[0]);
        return 2;
    }
    try {
        const std::string dir = argv[1];
        const int nincs = std::atoi(argv[2]);
        const int nx = std::atoi(argv[3]), ny = std::atoi(argv[4]), nz = std::atoi(argv[5]);
        const Float h = std::atof(argv[6]), D = std::atof(argv[7]), decay = std::atof(argv[8]);
        const Float par_dt = std::atof(argv[9]);
        const int timescale = std::atoi(argv[10]);
        const Float coup = std::atof(argv[11]);

        const std::vector<int32_t> off = load<int32_t>(dir + "/off.i32");
        const std::vector<int32_t> tgt = load<int32_t>(dir + "/tgt.i32");
        const std::vector<Float> sv = load<Float>(dir + "/sv.f32");
        const std::vector<Float> pos = load<Float>(dir + "/pos.f32");
        const int N = (int)off.size() - 1;
        if (N <= 0 || sv.size() != (size_t)NUM_SIM * N || pos.size() != 3 * (size_t)N)
            throw std::runtime_error("input sizes do not match");

        std::vector<DeviceVector<Float>> vars;
        std::vector<Float*> ptrs;
        for (int k = 0; k < NUM_SIM; ++k) {
            vars.emplace_back(std::vector<Float>(sv.begin() + (size_t)k * N, sv.begin() + (size_t)(k + 1) * N));
            ptrs.push_back(vars.back().data());
        }
        DeviceVector<Float*> sim_vars(ptrs);
        // connectivity and in-degree as src/main.cu:264-265 builds them
        CsrFileNetwork network(off, tgt);
        ChunkedAdjacencyList chunked_list = network.ToChunkedAdjacencyList(ptrs[SIM_LOCAL_COUNT]);
        AdjacencyList& list = chunked_list.list;  // the unchunked list, for Hebbian learning
        std::unique_ptr<Hebbian> heb;
        {
            FILE* wf = std::fopen((dir + "/w.f32").c_str(), "rb");
            if (wf) {
                std::fclose(wf);
                list.weights = DeviceVector<Float>(load<Float>(dir + "/w.f32"));
                heb = std::make_unique<Hebbian>(20, 0.05f, (int)tgt.size());
            }
        }
        DeviceVector<Float> px(std::vector<Float>(pos.begin(), pos.begin() + N));
        DeviceVector<Float> py(std::vector<Float>(pos.begin() + N, pos.begin() + 2 * N));
        DeviceVector<Float> pz(std::vector<Float>(pos.begin() + 2 * N, pos.end()));
        DeviceVector<Float> g_par_vec(N);
        Float* g_par = g_par_vec.data();

        HHEphys ephys(N);
        if (const char* ns = std::getenv("COMPAT_DEMO_NOISE_STD")) ephys.params().noise_std = std::atof(ns);
        AtomicCoupling coupling(N);
        std::vector<CouplingStruct> coupling_structs{{chunked_list, coupling, ptrs[SIM_LOCAL_IN], ptrs[SIM_LOCAL_OUTPUT]}};
        ParacrineModule paracrine(g_par, N, nx, ny, nz, par_dt, timescale, coup, h, D, decay,
                                  px.data(), py.data(), pz.data());

        const int print_incs = argc == 15 ? std::atoi(argv[12]) : 0;
        std::unique_ptr<ElectrodeNet> electrodes;
        int render_N = 0;
        std::vector<Float> erows;
        if (print_incs > 0) {
            render_N = std::atoi(argv[13]);
            const int k = std::atoi(argv[14]);
            std::vector<float> xyz(3 * (size_t)N);
            for (int i = 0; i < N; ++i)
                for (int d = 0; d < 3; ++d) xyz[3 * (size_t)i + d] = pos[(size_t)d * N + i];
            Morphology morphology(xyz);
            std::vector<int> electrode_inds(render_N), h_regions, h_regions_inds;
            for (int e = 0; e < render_N; ++e) electrode_inds[e] = (N / render_N) * e;
            morphology.GenSphericalRegions(h_regions, h_regions_inds, electrode_inds, k);
            electrodes = std::make_unique<ElectrodeNet>(h_regions, h_regions_inds);
            save(dir + "/regions.i32", h_regions);
        }

        std::vector<int32_t> raster, nfiring;
        for (int i = 0; i < nincs; ++i) {
            if (electrodes && i % print_incs == 0) {
                electrodes->DoElectrodeMeasurements(ptrs[SIM_V]);
                const Float* avg = electrodes->ElectrodeOutputHandle();
                erows.insert(erows.end(), avg, avg + render_N);
                erows.push_back(ElectrodeNet::MeasureAverageEntireNetwork(ptrs[SIM_V], N));
            }
            // Perform coupling. The number of neurons firing is returned by each of these calls.
            int n_firing;
            for(auto& coup_struct: coupling_structs)
                n_firing = coup_struct();
            nfiring.push_back(n_firing);
            record(coupling, n_firing, i - 1, raster);
            if (i % paracrine.timescale == 0)
                paracrine.update_on_own_timescale();
            else
                paracrine.update_on_fine_timescale(coupling);
            ephys.SimulateEphys(sim_vars.data());
            if (heb) heb->UpdateWeights(list, coupling.is_firing, coupling.firing_neurons);
        }
        std::vector<Float> out;
        for (int k = 0; k < NUM_SIM; ++k) {
            std::vector<Float> a = vars[k].host();
            out.insert(out.end(), a.begin(), a.end());
        }
        // the firing of the last increment (the coupling input it writes is
        // not part of sv_out)
        record(coupling, coupling_structs[0](), nincs - 1, raster);
        if (electrodes) save(dir + "/electrodes.f32", erows);
        if (heb) save(dir + "/w_out.f32", list.weights.host());
        save(dir + "/sv_out.f32", out);
        save(dir + "/raster.i32", raster);
        save(dir + "/nfiring.i32", nfiring);
        save(dir + "/g_par.f32", g_par_vec.host());
        std::vector<Float> dens((size_t)nx * ny * nz);
        check(nnx_memcpy_d2h(ctx(), dens.data(), paracrine.get_grid().density(), dens.size() * sizeof(Float)));
        save(dir + "/density.f32", dens);
        std::printf("compat_demo: N=%d nincs=%d spikes=%zu\n", N, nincs, raster.size() / 2);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "compat_demo: %s\n", e.what());
        return 1;
    }
    return 0;
}

/*
 * nnx.h - C ABI of the MI355X (gfx950) spiking-network timestep.
 *
 * One library (libnnx.so, built with hipcc for gfx950 only) implements the
 * per-timestep hot path of the neuronephys simulator:
 *   ephys     - Hodgkin-Huxley neuron update (reference leapfrog or RK4),
 *   coupling  - spike detection, ascending spike list, propagation over CSR/CSC,
 *   paracrine - release deposit and the implicit-diffusion CG solve.
 *
 * Conventions
 *   - Every pointer named d_* or documented as "device" is device memory.
 *   - All calls are asynchronous on the context's stream unless documented.
 *   - No call exits the process: errors are returned as nnx_status and the
 *     message is kept in the context (nnx_last_error).
 *   - A context is not thread-safe: one per host thread / stream.
 *   - Spike counts stay in device memory; nothing on the step path reads
 *     them back unless the caller asks (nnx_read_count).
 *
 * The reference interfaces each entry point replaces are cited as
 * file:line of the neuronephys_clean sources.
 */
#ifndef NNX_H
#define NNX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nnx_ctx nnx_ctx;

typedef enum nnx_status {
    NNX_OK = 0,
    NNX_ERR_ARG = 1,         /* invalid argument (shape, pointer, enum)           */
    NNX_ERR_HIP = 2,         /* HIP runtime error                                  */
    NNX_ERR_NONFINITE = 3,   /* a NaN/Inf was found where the caller asked to check */
    NNX_ERR_NOCONV = 4,      /* iterative solver hit maxit before rtol            */
    NNX_ERR_UNSUPPORTED = 5, /* valid but unimplemented combination               */
    NNX_ERR_STATE = 6        /* call order / context state error                  */
} nnx_status;

/* element types */
enum { NNX_F32 = 0, NNX_F64 = 1 };

/* HH integrators.  LEAPFROG is the reference's semi-implicit update with the
 * three kinetic synapses (src/ephys.cu:28-96); RK4 is the classical 4-stage
 * Runge-Kutta with one exponential synapse (no reference counterpart).      */
enum { NNX_INT_LEAPFROG = 0, NNX_INT_RK4 = 1 };

/* spike rules.  LEVEL: output > out_thresh (src/coupling.cu:103-109).
 * CROSSING: V_prev < v_thresh <= V_new (upward crossing).                  */
enum { NNX_SPIKE_LEVEL = 0, NNX_SPIKE_CROSSING = 1 };

/* coupling modes (what one presynaptic spike contributes to each target).
 *   WEIGHTED_PULL  : g_in[t] = sum of w over spiking sources, summed in
 *                    ascending source order, dense gather over the CSC.
 *   WEIGHTED_EVENT : same sum, bit-identical to WEIGHTED_PULL, but reads only
 *                    the out-edges of spiking sources (event-driven).
 *   COUNT          : +1 per spiking source (src/coupling.cu:34-48).
 *   SIGNAL         : +output[source] per spiking source
 *                    (src_wholebrain/coupling.cu:12-33).
 *   WEIGHTED_FUSED : (nnx_run_steps only; fp32 RK4 with crossing spikes)
 *                    the WEIGHTED_EVENT sum, bit-identical, computed inside
 *                    one persistent launch per step that also performs the
 *                    RK4 update and pushes the new spikes to their targets. */
enum { NNX_COUPLE_WEIGHTED_PULL = 0, NNX_COUPLE_WEIGHTED_EVENT = 1,
       NNX_COUPLE_COUNT = 2, NNX_COUPLE_SIGNAL = 3, NNX_COUPLE_WEIGHTED_FUSED = 4 };

/* stencils / time schemes for the paracrine grid */
enum { NNX_STENCIL_7 = 7, NNX_STENCIL_27 = 27 };
enum { NNX_PRECOND_NONE = 0, NNX_PRECOND_JACOBI = 1, NNX_PRECOND_LEARNED = 2, NNX_PRECOND_MG = 3,
       NNX_PRECOND_MG_LEARNED = 4 };
/* multigrid smoothers */
enum { NNX_MG_CHEBYSHEV = 0, NNX_MG_JACOBI = 1 };

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */

/* Creates a context on `device`.  A non-NULL `hip_stream` (hipStream_t) is
 * borrowed, not owned; with NULL the context creates a blocking stream of
 * its own (default priority) and destroys it in nnx_destroy.  Every call
 * enqueues on that stream.                                                  */
nnx_status nnx_create(nnx_ctx** out, int device, void* hip_stream);
nnx_status nnx_destroy(nnx_ctx* ctx);
const char* nnx_last_error(const nnx_ctx* ctx);
nnx_status nnx_set_stream(nnx_ctx* ctx, void* hip_stream);
nnx_status nnx_synchronize(nnx_ctx* ctx);
/* library version string, e.g. "nnx 0.1 gfx950" */
const char* nnx_version(void);

/* Device-memory helpers so that host code compiled without a HIP toolchain
 * (the compat classes) can own buffers.  The reference uses cudaMalloc /
 * thrust::device_vector for the same buffers (src/main.cu:327-372).        */
nnx_status nnx_malloc(nnx_ctx* ctx, void** d_ptr, size_t bytes);
nnx_status nnx_free(nnx_ctx* ctx, void* d_ptr);
nnx_status nnx_memcpy_h2d(nnx_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
nnx_status nnx_memcpy_d2h(nnx_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
nnx_status nnx_memcpy_d2d(nnx_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
nnx_status nnx_memset(nnx_ctx* ctx, void* d_dst, int value, size_t bytes);

/* Stream-ordered timing on the context's stream (hipEvent pairs), for
 * callers that time kernels without a HIP toolchain of their own.          */
nnx_status nnx_event_create(nnx_ctx* ctx, void** ev);
nnx_status nnx_event_destroy(nnx_ctx* ctx, void* ev);
nnx_status nnx_event_record(nnx_ctx* ctx, void* ev);
/* milliseconds between two recorded events (synchronises on `end`)        */
nnx_status nnx_event_elapsed(nnx_ctx* ctx, void* start, void* end, float* ms);
/* the context's hipStream_t                                               */
void* nnx_get_stream(const nnx_ctx* ctx);

/* ------------------------------------------------------------------------ */
/* ephys                                                                     */
/* ------------------------------------------------------------------------ */

/* Structure-of-arrays neuron state.  Each pointer is a device array of n
 * elements of `dtype`.  The names follow the reference's SoA indices
 * (src/defines.cuh:57-78, src_wholebrain/ephys.cuh:62-78).  Fields a mode
 * does not use may be NULL.                                                 */
typedef struct nnx_hh_state {
    int32_t dtype;            /* NNX_F32 | NNX_F64 */
    int32_t pad_;
    int64_t n;
    void* v;                  /* membrane potential, mV        (SIM_PARAMS_V) */
    void* m;                  /* Na activation                 (SIM_PARAMS_M) */
    void* h;                  /* Na inactivation               (SIM_PARAMS_H) */
    void* ng;                 /* K activation "n"              (SIM_PARAMS_N) */
    /* RK4 / exponential synapse */
    void* s;                  /* synaptic conductance, mS/cm^2 */
    const void* i_ext;        /* static drive, uA/cm^2 (RK4) */
    /* leapfrog / kinetic synapses */
    void* y_local;            /* SIM_PARAMS_LOCAL_Y */
    void* y_i;                /* SIM_PARAMS_IY */
    void* y_e;                /* SIM_PARAMS_EY */
    const void* in_local;     /* SIM_PARAMS_LOCAL_IN */
    const void* count_local;  /* SIM_PARAMS_LOCAL_COUNT (in-degree) */
    const void* in_i;         /* SIM_PARAMS_IIN */
    const void* count_i;      /* SIM_PARAMS_ICOUNT */
    const void* in_e;         /* SIM_PARAMS_EIN */
    const void* count_e;      /* SIM_PARAMS_ECOUNT */
    const void* appcur;       /* SIM_PARAMS_APPCUR */
    void* out_local;          /* SIM_PARAMS_LOCAL_OUTPUT */
    void* out_i;              /* SIM_PARAMS_IOUTPUT */
    void* out_e;              /* SIM_PARAMS_EOUTPUT */
} nnx_hh_state;

/* Model parameters (all in the reference's units: mV, ms, mS/cm^2, uA/cm^2).
 * The six rate functions are the reference's (src/ephys.cu:57-63), written
 * for a -60 mV rest, evaluated at V + v_shift; v_shift = 5 gives the
 * classic -65 mV squid axon.                                               */
typedef struct nnx_hh_params {
    double C, gna, gk, gl, ena, ek, el;
    double v_shift;
    /* RK4: exponential synapse, I_syn = s * (V - e_syn), s decays with tau_syn */
    double tau_syn, e_syn;
    /* leapfrog: kinetic synapses + sigmoid output (src/defines.cuh:94-101) */
    double tmax, vt, kp, ar, ad;
    double g_local, e_local, g_i, e_i, g_e, e_e;
    double noise_mean, noise_std;   /* white noise added to the drive (leapfrog); std 0 = deterministic */
    uint64_t noise_seed;
    /* spike rules */
    double v_thresh;                /* CROSSING threshold, mV */
    double out_thresh;              /* LEVEL threshold on the output */
} nnx_hh_params;

/* Fills `p` with the reference HH constants (src/defines.cuh:87-102):
 * leapfrog, v_shift 0, ENa 55, EK -72, EL -50, noise mean 3 / std 1.5.    */
void nnx_hh_params_reference(nnx_hh_params* p);
/* Fills `p` with the squid constants of the RK4 configuration: C 1,
 * gNa 120, gK 36, gL 0.3, ENa 50, EK -77, EL -54.4, v_shift 5, tau 2 ms,
 * E_syn 0 mV, threshold 0 mV.                                             */
void nnx_hh_params_squid(nnx_hh_params* p);

/* One timestep of every neuron.  Replaces HHEphys::SimulateEphys
 * (src/ephys.cu:98-103, src_wholebrain/ephys.cu:219-224) and its kernel
 * Leapfrog_HH (src/ephys.cu:28-96).
 *   g_in       RK4 only: per-neuron conductance increment of this step
 *              (coupling output), dtype of the state; NULL = none.
 *   spike_bits out: bit i of word i/64 = neuron i spiked at this step under
 *              `spike_rule`; ceil(n/64) uint64 words; NULL = not wanted.
 *   step       global step index (keys the counter-based white noise).
 * White noise (noise_std != 0): the leapfrog step adds
 * noise_std * z + noise_mean to the drive, z = nnx_noise_normal's draw of
 * (noise_seed, step, neuron) (the reference's curand_normal,
 * src/ephys.cu:82); RK4 returns NNX_ERR_UNSUPPORTED for it (the benchmark
 * model's drive is static).                                               */
nnx_status nnx_hh_step(nnx_ctx* ctx, const nnx_hh_state* st, const nnx_hh_params* p,
                       int integrator, int spike_rule, double dt, uint64_t step,
                       const void* g_in, uint64_t* spike_bits);

/* Diekman suprachiasmatic-nucleus neuron (DiekmanEphys, src/ephys.cuh,
 * Leapfrog_diekman, src/ephys.cu:250-347): Hodgkin-Huxley-type Na, K,
 * L-type and non-L-type Ca, Ca-activated K, persistent Na, two Ca pools and
 * a kinetic GABA synapse, all gates by the trapezoid rule, the submembrane
 * Ca by the root of the trapezoid rule's quadratic, V semi-implicitly.
 * Arrays are device arrays of n elements of dtype, in the reference's
 * DIEKMAN_VAR order (src/defines.cuh:153-173).                             */
typedef struct nnx_diekman_state {
    int32_t dtype;
    int32_t pad_;
    int64_t n;
    void *v, *m, *h, *ng, *rl, *rnl, *fnl, *s, *p, *y, *cas, *cac;  /* updated */
    void* iin;                /* DIEKMAN_VAR_IIN (coupling input, written by the coupling) */
    const void* icount;       /* DIEKMAN_VAR_ICOUNT (in-degree) */
    void* ioutput;            /* DIEKMAN_VAR_IOUTPUT, written after the update */
    const void *egaba, *gkca, *gkleak;  /* per-neuron parameters */
} nnx_diekman_state;

/* The constants of diekman_params (src/defines.cuh), and the ones
 * Leapfrog_diekman fixes in its body (gnaP 2.85, tau_Rl = tau_Rnl = 3.1,
 * tau_P 100 ms).  The reference stores them as float.                      */
typedef struct nnx_diekman_params {
    double gna, gnaleak, gk, gcal, gcanl, ena, ek, eca, k1, k2, kcas, kcac, taucas, taucac, bcas, bcac;
    double ar, ad, tmax, vt, kp, g_gaba, appcur, gnap, tau_rl, tau_rnl, tau_p;
    double out_thresh;        /* level rule on IOUTPUT (0.1, src/coupling.cu:103-109) */
} nnx_diekman_params;
void nnx_diekman_params_reference(nnx_diekman_params* p);

/* One step of every neuron (DiekmanEphys::SimulateEphys).  spike_bits
 * (optional): bit i = IOUTPUT of neuron i after the step > out_thresh.
 * fp32 follows the reference's mix of float and double operations; the
 * exponentials are this library's (see the HH step), powf(x, 0.25) is
 * sqrt(sqrt(x)), powf(x, 1.5) is x sqrt(x) and powf(x, 4) is (x^2)^2.    */
nnx_status nnx_diekman_step(nnx_ctx* ctx, const nnx_diekman_state* st, const nnx_diekman_params* p, double dt,
                            uint64_t* spike_bits);

/* Slow-wave "AN" neuron with excitatory and inhibitory synapses (SANEphys,
 * src/ephys.cuh:34-39: Euler_an_ie_model with dxdt_an_ie_model,
 * src/ephys.cu:105-141,203-245): V, K activation nK, Ca, AMPA (Ye) and GABA
 * (Yi) gates, one forward-Euler step; the white-noise drive is its mean
 * (the reference draws it from cuRAND).  Arrays in the AN_IE_VAR order
 * (src/defines.cuh:107-117).                                               */
typedef struct nnx_san_state {
    int32_t dtype;
    int32_t pad_;
    int64_t n;
    void *v, *nk, *ca, *ye, *yi;             /* updated */
    const void *ein, *ecount, *iin, *icount; /* synaptic inputs and in-degrees */
    void *eoutput, *ioutput;                 /* written after the update (both the same sigmoid) */
} nnx_san_state;

/* an_ie_params (src/defines.cuh:118-149); stored as float there.           */
typedef struct nnx_san_params {
    double c, a, v_l, v_na, v_k, v_ca, kd, v_ampa, v_gaba, alpha_ca, tau_ampa, tau_gaba;
    double g_l, g_k, g_nap, g_ca, g_kca, g_ampa, g_gaba, tau_ca, noise_mean, tmax, vt, kp;
    double out_thresh;                       /* level rule on the outputs (0.1) */
} nnx_san_params;
void nnx_san_params_reference(nnx_san_params* p);

/* One Euler step of every neuron (SANEphys::SimulateEphys).  spike_bits
 * (optional): bit i = output of neuron i after the step > out_thresh.
 * powf(x, 3.5) is x^3 sqrt(x) and powf(x, 4) is (x^2)^2.                  */
nnx_status nnx_san_step(nnx_ctx* ctx, const nnx_san_state* st, const nnx_san_params* p, double dt,
                        uint64_t* spike_bits);

/* Standard normal draws of the leapfrog step's white noise: d_out[i] (FP64)
 * = the draw of neuron first + i at `step` with `seed`: Philox4x64-10 at
 * counter {step, neuron, 0x4e4f495345, 0} with key {seed, 0}, words 0 and 1
 * as u1 = ((w0 >> 11) + 1) 2^-53 and u2 = (w1 >> 11) 2^-53, then
 * sqrt(-2 ln u1) cos(2 pi u2) with this library's ln and cos and a correctly
 * rounded sqrt.  The step draws noise_std * z + noise_mean (in the state's
 * type) per neuron.  Replaces init_random + curand_normal
 * (src/ephys.cu:4-10,82): a stateless counter-based stream instead of 48
 * bytes of XORWOW state per neuron.                                        */
nnx_status nnx_noise_normal(nnx_ctx* ctx, uint64_t seed, uint64_t step, int64_t first, int64_t n,
                            double* d_out);

/* Returns NNX_ERR_NONFINITE if any V is NaN/Inf (synchronous).             */
nnx_status nnx_check_finite(nnx_ctx* ctx, int dtype, int64_t n, const void* d_x);

/* ------------------------------------------------------------------------ */
/* spikes and coupling                                                      */
/* ------------------------------------------------------------------------ */

/* Level rule from outputs: bits[i] = out[i] > thresh.  Replaces the thrust
 * transform with CheckFiringOperator (src/coupling.cu:103-109,118).        */
nnx_status nnx_spike_threshold(nnx_ctx* ctx, int dtype, int64_t n, const void* d_out,
                               double thresh, uint64_t* d_bits);

/* Ascending list of set bits (stable compaction) and its length, both left
 * in device memory.  Replaces thrust::copy_if over a counting iterator
 * (src/coupling.cu:121-127, src_wholebrain/coupling.cu:52-57).             */
nnx_status nnx_spike_compact(nnx_ctx* ctx, int64_t n, const uint64_t* d_bits,
                             int32_t* d_list, int32_t* d_count);

/* Synchronous read of a device int32 (the count), through pinned memory.   */
nnx_status nnx_read_count(nnx_ctx* ctx, const int32_t* d_count, int32_t* h_count);

/* Connectivity of one network, both directions, all device arrays.
 * out_*: CSR of out-edges exactly as the reference's AdjacencyList
 *        (indices / offset_in_indices, src/coupling.cuh:4-8).
 * in_* : its transpose (CSC), in-edges of each target sorted by source.
 * out_to_in[e]: position in the CSC of CSR edge e (WEIGHTED_EVENT only).   */
typedef struct nnx_graph {
    int64_t n;
    int64_t nnz;
    const int32_t* out_offsets;   /* n+1 */
    const int32_t* out_targets;   /* nnz */
    const float*   out_weights;   /* nnz or NULL */
    const int32_t* in_offsets;    /* n+1 */
    const int32_t* in_sources;    /* nnz */
    const float*   in_weights;    /* nnz or NULL */
    const int32_t* out_to_in;     /* nnz or NULL */
} nnx_graph;

/* Builds the CSC (in_offsets, in_sources, in_weights) and out_to_in of a CSR
 * on the device.  Deterministic: in-edges end up sorted by source, ties
 * (duplicate edges) by CSR position.  Replaces the host loop that counts
 * in-degree in Network::ToAdjacencyList (src/network_factory.cu:5-27) and
 * cusparseCsr2cscEx2 (src/paracrine.cu:403-432).  in_degree (optional,
 * dtype) receives the in-degree as the reference's COUNT array.            */
nnx_status nnx_graph_transpose(nnx_ctx* ctx, int64_t n, int64_t nnz,
                               const int32_t* d_out_offsets, const int32_t* d_out_targets,
                               const float* d_out_weights,
                               int32_t* d_in_offsets, int32_t* d_in_sources, float* d_in_weights,
                               int32_t* d_out_to_in, int in_degree_dtype, void* d_in_degree);

/* Spike propagation.  Replaces AtomicCoupling::operator()
 * (src/coupling.cu:111-178) and AtomicCoupling::PerformCoupling
 * (src_wholebrain/coupling.cu:43-69).  Inputs are the spike bitmap of the
 * previous step (and, for COUNT/SIGNAL/EVENT, its ascending list + device
 * count from nnx_spike_compact).  The result overwrites d_g_in (n elements
 * of `dtype`); every mode is deterministic.                                */
nnx_status nnx_couple(nnx_ctx* ctx, const nnx_graph* g, int mode, int dtype,
                      const uint64_t* d_bits, const int32_t* d_list, const int32_t* d_count,
                      const void* d_outputs, void* d_g_in);

/* ------------------------------------------------------------------------ */
/* paracrine                                                                */
/* ------------------------------------------------------------------------ */

/* A box grid, x fastest: index = i + nx*(j + ny*k) (src/paracrine.cu:172-178).
 * Operators (Dirichlet 0 outside the box):
 *   STENCIL_7 : A = diag*I - c*sum(6 neighbours), diag = 1 + 6c, i.e.
 *               (I - c*L7) of a backward-Euler diffusion step, c = dt*D/h^2;
 *               with kx/ky/kz set, c varies per face (c is then unused):
 *               kx has (nx+1)*ny*nz entries, kx[i + (nx+1)*(j + ny*k)] the
 *               face between voxels i-1 and i of row (j,k) (i = 0 and i = nx
 *               are the boundary faces); ky has nx*(ny+1)*nz entries indexed
 *               i + nx*(j + (ny+1)*k); kz has nx*ny*(nz+1) indexed
 *               i + nx*(j + ny*k).  Each entry is dt*D_face/h^2 and the
 *               diagonal is 1 + the sum of the voxel's six faces.
 *   STENCIL_27: A = (1 + decay_dt/2) I - (c/2) L27, L27 the reference's
 *               high-order compact Laplacian {1,3,14,-128}/30 (src/paracrine.cu:40-43,80-81),
 *               c = dt*D/h^2: the Crank-Nicolson operator.
 * block > 0 (constant STENCIL_7 only; nx, ny, nz multiples of it): the box
 * is a lattice of independent block^3 grids, each with Dirichlet 0 on its
 * own faces - the operator has no coupling across a block face, so one
 * solve on the box solves every block (the per-network grids of an
 * ensemble laid side by side).                                              */
typedef struct nnx_grid {
    int32_t nx, ny, nz;
    int32_t stencil;
    double c;
    double decay_dt;
    const float* kx;   /* variable face coefficients or NULL (STENCIL_7 only) */
    const float* ky;
    const float* kz;
    int32_t block;     /* 0: one grid; > 0: a lattice of block^3 grids (above) */
    int32_t pad_;
} nnx_grid;

typedef struct nnx_cg_opts {
    double rtol;          /* stop when ||r||/||b|| <= rtol (recursive residual) */
    int32_t maxit;
    int32_t precond;      /* NNX_PRECOND_* */
    int32_t check_every;  /* iterations between host convergence checks (>=1) */
    int32_t use_x0;       /* 1: start from x; 0: start from 0 */
    int32_t stall_iters;  /* flexible CG: fall back to Jacobi after this many non-improving iterations */
    int32_t smooth;       /* learned preconditioner: damped-Jacobi sweeps after the network (0: 1) */
    int32_t mg_nu;        /* multigrid: smoothing sweeps before and after each level (0: 3) */
    int32_t mg_smoother;  /* NNX_MG_CHEBYSHEV (0) or NNX_MG_JACOBI */
    int32_t mg_cycle;     /* multigrid: visits of each coarser level, 1 V-cycle, 2 (or 0) W-cycle */
    int32_t reserved_;
} nnx_cg_opts;

typedef struct nnx_cg_result {
    int32_t iters;        /* inner iterations performed */
    int32_t outer;        /* refinement steps (nnx_cg_solve_refine) */
    int32_t converged;
    int32_t precond_used; /* precond of the final sweep (fallback reports JACOBI) */
    double relres;        /* final recursive residual ||r||/||b|| */
    double true_relres;   /* ||b - A x||/||b|| recomputed in FP64 */
    double ms;            /* wall time of the solve (host clock, includes syncs) */
    double learned_weight; /* learned preconditioners: mean weight theta given to the
                              network's correction (nnx_fcg_solve); 0 otherwise */
} nnx_cg_result;

/* y = A x for the grid's operator (fp32).                                   */
nnx_status nnx_stencil_apply(nnx_ctx* ctx, const nnx_grid* g, const float* d_x, float* d_y);

/* Preconditioned CG on A x = b, fp32 vectors, FP64 dots/norms, scalars kept
 * on the device; the host reads convergence every check_every iterations.
 * Replaces Density_on_grid_3d::update_density's CG loop
 * (src/paracrine.cu:106-129).                                              */
nnx_status nnx_cg_solve(nnx_ctx* ctx, const nnx_grid* g, const float* d_b, float* d_x,
                        const nnx_cg_opts* o, nnx_cg_result* res);

/* The same Jacobi (or unpreconditioned) CG with no host read and no host
 * synchronisation, so that it can be captured into a hipGraph: the init and
 * exactly o->maxit iterations are enqueued; the iterations after
 * convergence are launched but return at once (the stop test runs on the
 * device).  d_result (device, 4 doubles) receives the iterations performed,
 * the final relative recursive residual and 1.0 if it met o->rtol (else
 * 0.0).  check_every, stall_iters, smooth and the mg_* fields are unused.
 * No counterpart in the reference (its CG reads every scalar back to the
 * host, src/paracrine.cu:106-129).                                           */
nnx_status nnx_cg_solve_async(nnx_ctx* ctx, const nnx_grid* g, const float* d_b, float* d_x,
                              const nnx_cg_opts* o, double* d_result);

/* nnx_cg_solve_async without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, hosts that forbid hipMalloc).
 * The CG working set lives in a fixed 48 MiB array of the library, one per
 * GPU, held until the process ends: grids up to 2.5M voxels with variable
 * coefficients and Jacobi (e.g. 128^3), 3.1M otherwise; larger ones return
 * NNX_ERR_ARG before anything is enqueued.  Same
 * kernels and results as nnx_cg_solve_async, bit for bit.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams
 * or threads) race.                                                         */
nnx_status nnx_cg_solve_async_noalloc(nnx_ctx* ctx, const nnx_grid* g, const float* d_b, float* d_x,
                                      const nnx_cg_opts* o, double* d_result);

/* Mixed-precision iterative refinement: FP64 b and x, the outer residual
 * r = b - A x recomputed in FP64, inner fp32 CG on A d = r, x += d, until
 * ||r||/||b|| <= o->rtol.  Each inner solve stops at the relative residual
 * max(inner_rtol, 0.5 * o->rtol / current relres): inner_rtol is the
 * tightest inner tolerance (about 1e-6 is the useful limit of fp32).       */
nnx_status nnx_cg_solve_refine(nnx_ctx* ctx, const nnx_grid* g, const double* d_b, double* d_x,
                               const nnx_cg_opts* o, double inner_rtol, int max_outer,
                               nnx_cg_result* res);

/* One reference paracrine update (Crank-Nicolson 27-point, decay, explicit-
 * Euler initial guess, at most max_iter CG iterations, stop on the absolute
 * squared residual < error_bound), restating Density_on_grid_3d::update_density
 * (src/paracrine.cu:63-134) including its use of h where dt is meant in b
 * (src/paracrine.cu:105) and its early return (src/paracrine.cu:114-117).
 * `data` is updated in place.  res->iters/relres report the loop.          */
nnx_status nnx_update_density_reference(nnx_ctx* ctx, int nx, int ny, int nz, double h,
                                        double diffusion, double decay, double timestep,
                                        double error_bound, int max_iter, float* d_data,
                                        nnx_cg_result* res);

/* Release deposit: voxel_counts[voxel_of[i]] += 1 for each spiking neuron i
 * of the list (integer atomics: order-independent).                        */
nnx_status nnx_deposit_count(nnx_ctx* ctx, const int32_t* d_voxel_of, const int32_t* d_list,
                             const int32_t* d_count, int32_t* d_voxel_counts);

/* release[list[i]] += amount for the count neurons of the ascending spike
 * list (each neuron at most once: no conflicts).  The per-step release
 * accumulation of ParacrineModule::update_on_fine_timescale
 * (src/paracrine.cuh:246-253), applied to the firing neurons.               */
nnx_status nnx_release_add(nnx_ctx* ctx, const int32_t* d_list, const int32_t* d_count,
                           double amount, float* d_release);

/* b = c_prev + amount * voxel_counts, then voxel_counts = 0.               */
nnx_status nnx_rhs_from_counts(nnx_ctx* ctx, int64_t nvox, const float* d_c_prev,
                               int32_t* d_voxel_counts, double amount, float* d_b);

/* Trilinear neuron<->grid transfer (Scatter_grid_converter_3d,
 * src/paracrine.cu:136-296), both as CSR SpMV over device CSR matrices the
 * host builder produced (nnx_host.h).  interp: y[n] = sum_j W[n,j] g[j];
 * scatter: g[j] += sum_n W^T[j,n] x[n] (rows of the transpose, neurons in
 * ascending order: deterministic).                                         */
nnx_status nnx_csr_spmv(nnx_ctx* ctx, int64_t rows, const int32_t* d_offsets, const int32_t* d_cols,
                        const float* d_vals, const float* d_x, float* d_y, int accumulate);

/* The paracrine field solved on its own cadence inside the timestep, as
 * ParacrineModule::update_on_own_timescale does every `timescale` steps
 * (src/main.cu:171-177, src/paracrine.cuh:255-265): before each step s with
 * s % every == 0 the releases counted since the last solve become the
 * source,
 *     b = conc + release * voxel_counts,   voxel_counts = 0,
 * and conc is replaced by the solution of A conc = b (A the grid's
 * operator), warm-started from the previous field.  The solve is Jacobi-CG
 * with FP64 iterative refinement: the residual b - A conc in FP64, an inner
 * fp32 CG (FP64 dots) on it, conc += correction, until the FP64 relative
 * residual ||b - A conc|| / ||b|| <= rtol.  Nothing is read back: max_outer
 * refinement steps of max_inner CG iterations each are enqueued (captured
 * with the steps), and every kernel after convergence returns at once.
 * Releases: every spike of the steps before s (the reference's
 * update_on_fine_timescale skips the spikes its coupling reads on a solve
 * step; here none is skipped).                                             */
typedef struct nnx_paracrine {
    int64_t every;            /* the reference's timescale in steps (<= 0: off) */
    nnx_grid grid;            /* nx*ny*nz = the number of voxel_counts entries */
    double release;           /* source added per spike at its voxel (0.01) */
    double rtol;              /* FP64 relative residual of each solve (1e-8) */
    double inner_rtol;        /* tightest inner fp32 CG tolerance (0: 1e-6) */
    int32_t max_inner;        /* CG iterations enqueued for the first refinement step, half as
                               * many for each later one (their tolerance is looser) (0: 64) */
    int32_t max_outer;        /* refinement steps enqueued (0: 3) */
    double* conc;             /* device, nvox FP64: the field (in: previous, out: new) */
    float* at_neurons;        /* optional device n floats: conc at each neuron's voxel
                               * after each solve (the write-back into d_g_paracrine,
                               * src/paracrine.cuh:261-264, at the neuron's voxel) */
    double* stats;            /* optional device 8 doubles, accumulated (the caller
                               * zeroes them): solves, inner iterations, refinement
                               * steps, max relres, solves above rtol, and the last
                               * solve's relres, inner iterations, refinement steps */
} nnx_paracrine;

/* ------------------------------------------------------------------------ */
/* ensembles of independent networks (no counterpart in the reference,      */
/* which simulates one network, src/main.cu:112-182)                         */
/* ------------------------------------------------------------------------ */

/* nnet networks in one index space: network i owns neurons
 * [net_begin[i], net_begin[i] + net_size[i]) (net_begin[i] a multiple of 64,
 * networks in ascending order), edges only within a network.  The step is
 * the RK4 model of nnx_run_steps (exponential synapses, 0 mV crossing,
 * event-driven weighted coupling in ascending source order), fp32.          */
typedef struct nnx_ens_plan {
    nnx_hh_state state;       /* fp32; n = the padded total (a multiple of 64) */
    nnx_hh_params params;
    nnx_graph graph;          /* block-diagonal CSR, CSC and out_to_in over the same index space */
    int32_t nnet, pad_;
    const int64_t* net_begin; /* host arrays [nnet] */
    const int64_t* net_size;
    double dt;
    uint64_t* bits[2];        /* spike bitmaps by step parity (n / 64 words each) */
    const int32_t* voxel_of;  /* optional: +1 at voxel_counts[voxel_of[i]] per spike of i */
    int32_t* voxel_counts;
    int32_t* net_spikes;      /* optional [nnet]: spikes per network, accumulated */
    int32_t* raster;          /* optional (step, neuron) pairs; within a step unordered */
    int64_t* raster_len;
    int64_t raster_cap;
    nnx_paracrine paracrine;  /* optional solve on its cadence (every = 0: off), as in
                               * nnx_step_plan; with a block lattice grid (nnx_grid.block)
                               * every network's own grid is solved at once */
} nnx_ens_plan;

/* launch shape of the fused step: threads per workgroup (64, 128, 256, 512),
 * neurons per tile (a multiple of 64), persistent workgroups; and the
 * flagless-event threshold `bulk`: when a step fires more than `bulk`
 * neurons, the next step pushes its events without marking their targets,
 * and every target then scans its event bits (< 0: never; 0: the library's
 * rule - for one network, more events than one per 64 targets; for an
 * ensemble, never).  None changes any result.                           */
typedef struct nnx_ens_config {
    int32_t block, tile, grid, bulk;
} nnx_ens_config;

/* n_steps steps from first_step (first_step == 0 clears the pending
 * coupling and the spike bitmaps) with one launch per step, optionally
 * replayed from a captured hipGraph; cfg NULL = {256, 256, 8192}.  The
 * result does not depend on cfg.  With plan->paracrine.every > 0 the solve
 * runs before every step s with s % every == 0, as in nnx_run_steps.       */
nnx_status nnx_ensemble_run(nnx_ctx* ctx, const nnx_ens_plan* plan, const nnx_ens_config* cfg,
                            uint64_t first_step, int n_steps, int use_graph);
/* Coordinate search over block, tile and grid (from the default
 * {256, 256, 8192}; a candidate replaces the current best only when faster).
 * Every candidate times the same window of steps_per_candidate steps from
 * first_step, from the same state: the state the steps change (neuron
 * state, pending events, spike bitmaps, step counters) is saved in a scratch
 * buffer and restored before each timed run and after the search, and the
 * candidates run without the plan's accumulating outputs (voxel deposit,
 * spike counts, raster) and without its paracrine solve.  So the simulation does not advance (*steps_run =
 * 0) and no output changes; choose a window that holds the bursts the run
 * will see.  Each candidate: one untimed run, then the faster of two timed
 * ones.  Synchronizes the stream; not under capture.                       */
nnx_status nnx_ensemble_autotune(nnx_ctx* ctx, const nnx_ens_plan* plan, uint64_t first_step,
                                 int steps_per_candidate, nnx_ens_config* best, double* best_us_per_step,
                                 int* steps_run);

/* ------------------------------------------------------------------------ */
/* learned preconditioner (no counterpart in the reference: its paracrine    */
/* solve is unpreconditioned CG, src/paracrine.cu:106-129)                   */
/* ------------------------------------------------------------------------ */

/* A 3-level 3-D convolutional U-Net M(r) ~ A^-1 r, BF16 weights and
 * activations with FP32 accumulation on matrix cores:
 *   e0 = act(conv3(x; 2->16)), e1 = act(conv3(pool(e0); 16->32)),
 *   e2 = act(conv3(pool(e1); 32->32)), d1 = act(conv3(e1 + up(e2); 32->16)),
 *   d0 = act(conv3(e0 + up(d1); 16->16)), z = rms(r) (w_h . d0 + b_h),
 * x = (q, q * feature) per voxel with q = r / rms(r), conv3 = 3x3x3 with
 * zero padding,
 * act(v) = v for v > 0, 0.1 v otherwise (leaky ReLU),
 * pool = 2x2x2 mean, up = nearest x2.  Parameters (fp32, in this order):
 * per layer W[cout][27][cin] (tap = dx + 3 dy + 9 dz fastest-x) then b[cout];
 * then the head's 16 weights and bias: nnx_unet_param_count() floats.
 * Training holds the biases at zero (the seeded init sets them to zero), so
 * a trained network maps r = 0 to 0 and scales with r.
 * Grids: nx % 64 == 0, ny % 32 == 0, nz % 32 == 0.                          */
typedef struct nnx_unet nnx_unet;
int64_t nnx_unet_param_count(void);
/* h_params (host, param_count floats) or NULL for a seeded He-uniform init */
nnx_status nnx_unet_create(nnx_ctx* ctx, const float* h_params, uint64_t seed, nnx_unet** out);
nnx_status nnx_unet_destroy(nnx_ctx* ctx, nnx_unet* net);
nnx_status nnx_unet_get_params(nnx_ctx* ctx, const nnx_unet* net, float* h_params);
/* Two-grid mode (coarse != 0): the network runs on the 2x coarser grid;
 * its input is the mean over 2x2x2 blocks of (q, q * feature) and its
 * output is prolonged back by cell-centred trilinear interpolation (weights
 * 3/4 and 1/4 per axis, zero outside the box).  Grids then need
 * nx % 128 == 0 and ny, nz % 64 == 0.  The mode is not part of the
 * parameters; training and use must agree.                                 */
nnx_status nnx_unet_set_coarse(nnx_ctx* ctx, nnx_unet* net, int coarse);
/* gradient of the last nnx_unet_train_step (same layout as the parameters) */
nnx_status nnx_unet_get_grad(nnx_ctx* ctx, const nnx_unet* net, float* h_grad);
/* z = M(r) for `batch` grids stored one after another (fp32; d_feat may be
 * NULL for a zero feature channel).                                         */
nnx_status nnx_unet_apply(nnx_ctx* ctx, nnx_unet* net, int nx, int ny, int nz, int batch,
                          const float* d_r, const float* d_feat, float* d_z);
/* One Adam step (lr; lr = 0 only computes the gradient) on the loss
 * (1/B) sum_b |d_b|^2 / |t_b|^2 with t_b = e_b / s_b, d_b = N(r_b / s_b) - t_b,
 * s_b = rms(r_b), N the network without the output scale; with face
 * coefficients (d_kx/ky/kz: each sample's nnx_grid kx/ky/kz arrays, one
 * after another) the A-norm loss (1/B) sum_b d_b.A_b d_b / t_b.A_b t_b
 * instead, and with smooth = m > 0 (at most 4) the smoothed A-norm loss
 * (1/B) sum_b |S_b^m d_b|_A^2 / t_b.A_b t_b, S_b = I - (2/3) D_b^-1 A_b (the
 * error left after m damped-Jacobi post-smoothing sweeps).  *h_loss (if
 * not NULL) receives the loss before the step (the call then synchronises). */
nnx_status nnx_unet_train_step(nnx_ctx* ctx, nnx_unet* net, int nx, int ny, int nz, int batch,
                               const float* d_r, const float* d_feat, const float* d_e,
                               const float* d_kx, const float* d_ky, const float* d_kz, int smooth,
                               double lr, double* h_loss);

/* Single convolution layers (BF16 channels-last tensors as uint16_t, weights
 * fp32 [cout][27 cin]); (cin, cout) in {(2,16), (16,16), (16,32), (32,16),
 * (32,32)}.  out = act(conv3(in + coarse_scale * up(coarse)) + bias) with
 * act the leaky ReLU when relu != 0 (identity otherwise), pool =
 * its 2x2x2 mean; wgrad: dw[co][tap cin + ci] = sum_v g[v][co] in'[v + tap][ci]
 * and db[co] = sum_v g[v][co], in' = in + coarse_scale * up(coarse).
 * Dimensions: nx a multiple of 16, and ny a multiple of 16 (8 for cin 32)
 * for the z-marching kernel, else of 8 (4 for cin 32) with nz a multiple
 * of 4 for the tile kernel; nx a multiple of 32 and ny, nz even with coarse
 * or pool.  wgrad: nx a multiple of 16, ny and nz of 4 (8 with coarse).    */
nnx_status nnx_conv3_bf16(nnx_ctx* ctx, int cin, int cout, int nx, int ny, int nz, int batch,
                          const uint16_t* d_in, const uint16_t* d_coarse, double coarse_scale,
                          const float* d_w, const float* d_bias, int relu, uint16_t* d_out,
                          uint16_t* d_pool);
nnx_status nnx_conv3_wgrad_bf16(nnx_ctx* ctx, int cin, int cout, int nx, int ny, int nz, int batch,
                                const uint16_t* d_in, const uint16_t* d_coarse, double coarse_scale,
                                const uint16_t* d_g, float* d_dw, float* d_db);

/* z = M(r): one multigrid cycle from a zero initial guess for the
 * variable-coefficient 7-point operator of g (kx, ky, kz set).  Cell-centred
 * 2x2x2 coarsening while every dimension is even and >= 8; a coarse face
 * carries 1/16 of the sum of the 4 fine faces it covers; block-mean
 * restriction of the residual; trilinear prolongation (zero beyond the
 * boundary); nu smoothing sweeps before and after each level (nu = 0: 3):
 * smoother NNX_MG_CHEBYSHEV, a Chebyshev polynomial in D^-1 A on [1/8, 2],
 * or NNX_MG_JACOBI, damped Jacobi (w = 0.8); 16 damped-Jacobi sweeps on the
 * coarsest level.  cycle 1: V-cycle; 2 (or 0): W-cycle (each coarser level
 * of more than 4096 cells visited twice per visit of its parent, the second
 * time on the residual the first visit leaves; smaller levels once).  No
 * counterpart in the reference.                                             */
nnx_status nnx_mg_apply(nnx_ctx* ctx, const nnx_grid* g, int nu, int smoother, int cycle, const float* d_r,
                        float* d_z);

/* Flexible CG (Notay, one stored direction): fp32 vectors, FP64 reductions,
 * precond LEARNED: a two-level preconditioner with the network N as the
 * coarse correction between damped-Jacobi smoothings (w = 2/3, D the
 * operator's diagonal): z0 = w D^-1 r, u = N(r - A z0),
 * z1 = z0 + theta u with theta = u.(r - A z0) / u.(A u) (the weight that
 * minimises the energy norm of the error left, so the network's correction
 * never makes z1 worse than z0 in that norm; res->learned_weight reports
 * its mean), then o->smooth sweeps z <- z + w D^-1 (r - A z) (at least
 * one); d_feat is N's feature input.
 * precond MG: one multigrid cycle (nnx_mg_apply with o->mg_nu, mg_smoother, mg_cycle).
 * precond MG_LEARNED: the cycle in place of z0 above: z0 = V(r), then u,
 * theta and z1 as above and o->smooth damped-Jacobi sweeps.
 * precond JACOBI: M = D^-1;
 * after o->stall_iters (default 10) iterations without a new best
 * residual a learned preconditioner is replaced by its base (Jacobi or the
 * V-cycle; res reports the preconditioner of the final iterations).
 * MG and MG_LEARNED need the variable-coefficient grid (kx, ky, kz).        */
nnx_status nnx_fcg_solve(nnx_ctx* ctx, const nnx_grid* g, nnx_unet* net, const float* d_feat,
                         const float* d_b, float* d_x, const nnx_cg_opts* o, nnx_cg_result* res);
/* FP64 iterative refinement around nnx_fcg_solve (as nnx_cg_solve_refine).  */
nnx_status nnx_fcg_solve_refine(nnx_ctx* ctx, const nnx_grid* g, nnx_unet* net, const float* d_feat,
                                const double* d_b, double* d_x, const nnx_cg_opts* o,
                                double inner_rtol, int max_outer, nnx_cg_result* res);

/* ------------------------------------------------------------------------ */
/* electrode readout                                                         */
/* ------------------------------------------------------------------------ */

/* Regions of an electrode net, an adjacency list as the reference's
 * (ElectrodeNet's d_regions, src/electrode_net.cuh:6-22): region r holds the
 * neurons idx[offsets[r] .. offsets[r+1]), in summation order.  nnz =
 * offsets[n_regions] (a host copy; it bounds every idx read).  Entries
 * outside [0, n) make the region's mean NaN and count as not firing;
 * offsets are clamped to [0, nnz] (a region with end <= start is empty).   */
typedef struct nnx_regions {
    int64_t n_regions;
    int64_t nnz;
    const int32_t* offsets;   /* n_regions+1, device */
    const int32_t* idx;       /* nnz, device */
} nnx_regions;

/* Mean V of each region and the number of its neurons whose spike bit is
 * set.  Replaces ElectrodeNet::DoElectrodeMeasurements and its
 * electrodes_measure_kernel (src/electrode_net.cu:3-36,
 * src_wholebrain/electrode_net.cu:3-49).  avg[r] (dtype) = the sum of V over
 * the region in idx order, one add at a time in dtype, divided by the
 * region's size: bit-identical to the reference's sequential loop, NaN for
 * an empty region as there.  d_bits: spike bitmap (ceil(n/64) words; the
 * reference's is_firing flags), NULL = no counts (d_firing unused).  Both
 * outputs stay on the device (the reference copies them to the host).     */
nnx_status nnx_electrode_measure(nnx_ctx* ctx, const nnx_regions* r, int dtype, int64_t n,
                                 const void* d_v, const uint64_t* d_bits, void* d_avg,
                                 int32_t* d_firing);

/* Mean of V over all n neurons into *d_mean (dtype, device).  Replaces
 * ElectrodeNet::MeasureAverageEntireNetwork (src/electrode_net.cu:44-47),
 * whose fp32 thrust::reduce has an unspecified order; here the sum is taken
 * in FP64 in a fixed order (n < 2^31), divided by n and rounded once.       */
nnx_status nnx_network_mean(nnx_ctx* ctx, int dtype, int64_t n, const void* d_v, void* d_mean);

/* Synchrony statistics of one region (SynchronyNet,
 * src_wholebrain/electrode_net.cu:54-126): per call (one timestep),
 *   vi_sum[j] += V[idx[j]],  vi2_sum[j] += V[idx[j]]*V[idx[j]]  (product in
 *   dtype, sums in FP64), j < n_idx;
 *   for each sub-population s (the first sub_counts[s] entries of idx):
 *   Vsum = their sum (FP64 in a fixed order, rounded to dtype; the reference
 *   uses an fp32 thrust::reduce), Vavg = Vsum / count in dtype, widened to
 *   FP64; vavg_sum[s] += Vavg, vavg2_sum[s] += Vavg*Vavg.
 * Sums are caller-owned device FP64 arrays (n_idx, n_idx, n_sub, n_sub),
 * zeroed by the caller before the first step; 1 <= sub_counts[s] <= n_idx
 * (device int32 array) is checked on the device (a count out of range
 * leaves that sub-population's sums NaN).                                   */
nnx_status nnx_synchrony_accumulate(nnx_ctx* ctx, int dtype, int64_t n, const void* d_v, int64_t n_idx,
                                    const int32_t* d_idx, int32_t n_sub, const int32_t* d_sub_counts,
                                    double* d_vi_sum, double* d_vi2_sum, double* d_vavg_sum,
                                    double* d_vavg2_sum);

/* chi^2 of each sub-population after n_steps accumulated timesteps
 * (SynchronyNet::GetSynchronyMeasures, src_wholebrain/electrode_net.cu:
 * 99-126), the reference's host arithmetic in the same order:
 *   chi[s] = count * var(Vavg) / sum_{j<count} var(V_j),
 *   var(Vavg) = (vavg2 - vavg^2 / T) / T, var(V_j) = vi2/T - (vi/T)^2 with
 *   T = n_steps (the per-neuron division skipped when T = 0, as there).
 * d_chi: n_sub doubles on the device.                                       */
nnx_status nnx_synchrony_measures(nnx_ctx* ctx, int64_t n_idx, int32_t n_sub, const int32_t* d_sub_counts,
                                  const double* d_vi_sum, const double* d_vi2_sum,
                                  const double* d_vavg_sum, const double* d_vavg2_sum, int64_t n_steps,
                                  double* d_chi);

/* ------------------------------------------------------------------------ */
/* Hebbian plasticity                                                        */
/* ------------------------------------------------------------------------ */

/* Hebbian learning of per-synapse weights (Hebbian, src_wholebrain/
 * hebbian.cuh:3-27, hebbian.cu:3-69).  The reference keeps a countdown per
 * synapse, restarts it at `window` whenever the source fires, and on every
 * call decrements every running countdown: at zero the weight goes down by
 * heb_inc * exp(-|w - 1|); if the target fires first it goes up by as much
 * and the countdown stops.  The same rule is applied here event by event: a
 * call touches the out-synapses of the sources firing now, the in-synapses
 * of the targets firing now and the out-synapses of the sources whose
 * window ends now, instead of every synapse; weights and which synapses are
 * counting down are identical to the reference's after every call.
 * State (device arrays the caller owns; nnx_hebbian_init clears it):       */
typedef struct nnx_hebbian {
    int32_t window;          /* the reference's window_duration (>= 1); 0 in a step plan = off */
    float heb_inc;
    float* weights;          /* nnz, CSR order (AdjacencyList::weights) */
    uint8_t* armed;          /* nnz: 1 while the synapse's countdown runs */
    int32_t* last_fire;      /* n: step of each neuron's last firing seen */
    uint64_t* ring;          /* window * ceil(n/64) words: the last `window` firing bitmaps */
    int32_t* in_to_out;      /* nnz: CSR position of each CSC entry (built by init) */
} nnx_hebbian;

/* Clears armed, last_fire and ring, and builds in_to_out from the graph's
 * out_to_in.  Weights are left as they are.                                 */
nnx_status nnx_hebbian_init(nnx_ctx* ctx, const nnx_graph* g, const nnx_hebbian* h);

/* One call of Hebbian::UpdateWeights with the firing set d_firing_bits
 * (the bitmap the coupling read: the reference's is_firing and
 * firing_inds).  `step` numbers the calls: consecutive calls must pass
 * consecutive steps (the countdowns advance by one per call).               */
nnx_status nnx_hebbian_update(nnx_ctx* ctx, const nnx_graph* g, const nnx_hebbian* h, int64_t step,
                              const uint64_t* d_firing_bits);

/* ------------------------------------------------------------------------ */
/* composite timestep                                                        */
/* ------------------------------------------------------------------------ */


/* Everything one timestep touches, in the reference's order
 * coupling -> paracrine -> ephys (src/main.cu:166-181,
 * src_wholebrain/mouse_cortex.cu:104-120).                                 */
typedef struct nnx_step_plan {
    nnx_hh_state state;
    nnx_hh_params params;
    nnx_graph graph;
    int32_t integrator, spike_rule, couple_mode, record;  /* record: append rasters */
    double dt;
    void* g_in;              /* n, state dtype: coupling output / ephys input */
    uint64_t* bits[2];       /* ping-pong spike bitmaps, ceil(n/64) words each */
    int32_t* list;           /* n: ascending spike list of the last step */
    int32_t* count;          /* 1: its length */
    /* paracrine release accumulation (optional; voxel_of NULL = off) */
    const int32_t* voxel_of; /* n */
    int32_t* voxel_counts;   /* nvox */
    /* raster recording (optional): (step, neuron) pairs appended while the
     * total stays below raster_cap; raster_len is a device int64.          */
    int32_t* raster;         /* 2*raster_cap */
    int64_t* raster_len;
    int64_t raster_cap;
    /* WEIGHTED_FUSED launch shape (0 = default): workgroup size (64, 128,
     * 256 or 512), tile (neurons per work item, a multiple of 64), grid
     * (workgroups); and the burst threshold: when a step fires more than
     * `fused_burst` neurons (<= 0: never) the next step pulls every target's
     * in-edges instead of reading its pending events (pushing events is the
     * cheaper of the two until a large part of the network fires at once);
     * and `fused_bulk`, the flagless-event threshold of nnx_ens_config.bulk.
     * None of these changes any result.                                     */
    int32_t fused_block, fused_tile, fused_grid, fused_burst;
    int32_t fused_bulk, fused_pad_;
    /* electrode readout (optional; readout_every <= 0 = off), as the
     * reference's print step (src/main.cu:133-141) and ElectrodesExample
     * (src_wholebrain/main.cu:36-47): before each step s with
     * s % readout_every == 0, row s / readout_every (rows >= readout_cap are
     * dropped) receives nnx_electrode_measure of the state entering step s -
     * n_regions means in readout_avg (state dtype), the counts over the
     * spike bitmap the previous step's coupling read (the reference's
     * is_firing) in readout_firing (NULL = not kept) - and
     * nnx_network_mean in readout_mean (one value per row, NULL = not
     * kept).  The rows are written on the device; nothing synchronises.     */
    nnx_regions regions;
    int64_t readout_every, readout_cap;
    void* readout_avg;         /* readout_cap * n_regions */
    int32_t* readout_firing;   /* readout_cap * n_regions */
    void* readout_mean;        /* readout_cap */
    /* Hebbian learning (optional; hebbian.window = 0 = off): after the
     * coupling of each step s, nnx_hebbian_update(step s) with the spike
     * bitmap that coupling read, as the reference's HebbianLearningExample
     * calls UpdateWeights after each Simulate() (src_wholebrain/main.cu:
     * 158-168).  The learned weights are not the coupling's weights (the
     * reference's coupling does not read them either).                    */
    nnx_hebbian hebbian;
    /* paracrine solve on its cadence (optional; paracrine.every <= 0 = off;
     * needs voxel_of and voxel_counts)                                     */
    nnx_paracrine paracrine;
} nnx_step_plan;

/* Runs n_steps timesteps starting at global step `first_step`.  With
 * use_graph = 1 the step sequence is captured once into a hipGraph per
 * (plan contents, n_steps, phase) and replayed; plans must not move between
 * calls.  The context keeps at most NNX_GRAPH_CACHE executable graphs
 * (least recently used ones are destroyed first).
 * WEIGHTED_FUSED: g_in is not used; each call starts from the previous
 * step's list/count and leaves the last step's list/count, so calls with
 * other coupling modes can be interleaved; raster entries of one step are
 * appended in unspecified order (steps in order).                          */
nnx_status nnx_run_steps(nnx_ctx* ctx, nnx_step_plan* plan, uint64_t first_step, int n_steps,
                         int use_graph);

/* Executable graphs kept per context by nnx_run_steps and
 * nnx_ensemble_run.                                                         */
#define NNX_GRAPH_CACHE 32
/* Destroys every cached graph (after the context's stream has drained);
 * a caller frees a plan's buffers after this or after nnx_destroy.  The
 * reference has no graphs (every launch is synchronised,
 * src/main.cu:112-182).                                                     */
nnx_status nnx_graph_cache_clear(nnx_ctx* ctx);
/* Number of cached executable graphs and the number instantiated since the
 * context was created (either pointer may be NULL).                        */
nnx_status nnx_graph_cache_stats(const nnx_ctx* ctx, int* cached, int64_t* instantiated);

#ifdef __cplusplus
}
#endif
#endif /* NNX_H */

"""ctypes binding of the C ABI (include/nnx.h, include/nnx_host.h).

libnnx.so holds the gfx950 kernels; it is built in-tree by
``__graft_entry__.build()`` (make -C neuronephys_clean_amd).  There is no CPU
fallback: if the library is missing, every device entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnnx.so")
HOST_LIB_PATH = os.path.join(_HERE, "libnnx_host.so")

F32, F64 = 0, 1
INT_LEAPFROG, INT_RK4 = 0, 1
SPIKE_LEVEL, SPIKE_CROSSING = 0, 1
COUPLE_WEIGHTED_PULL, COUPLE_WEIGHTED_EVENT, COUPLE_COUNT, COUPLE_SIGNAL, COUPLE_WEIGHTED_FUSED = 0, 1, 2, 3, 4
STENCIL_7, STENCIL_27 = 7, 27
PRECOND_NONE, PRECOND_JACOBI, PRECOND_LEARNED, PRECOND_MG, PRECOND_MG_LEARNED = 0, 1, 2, 3, 4
MG_CHEBYSHEV, MG_JACOBI = 0, 1

STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_NONFINITE", 4: "ERR_NOCONV",
          5: "ERR_UNSUPPORTED", 6: "ERR_STATE"}

vp = C.c_void_p
i32p = C.POINTER(C.c_int32)


class NNXError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class HHState(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("pad_", C.c_int32), ("n", C.c_int64)] + [
        (name, vp) for name in (
            "v", "m", "h", "ng", "s", "i_ext", "y_local", "y_i", "y_e", "in_local",
            "count_local", "in_i", "count_i", "in_e", "count_e", "appcur", "out_local",
            "out_i", "out_e")]


PARAM_FIELDS = ("C", "gna", "gk", "gl", "ena", "ek", "el", "v_shift", "tau_syn", "e_syn",
                "tmax", "vt", "kp", "ar", "ad", "g_local", "e_local", "g_i", "e_i", "g_e",
                "e_e", "noise_mean", "noise_std", "noise_seed", "v_thresh", "out_thresh")


class HHParams(C.Structure):
    _fields_ = [(f, C.c_uint64 if f == "noise_seed" else C.c_double) for f in PARAM_FIELDS]

    def as_dict(self):
        return {f: getattr(self, f) for f in PARAM_FIELDS}


class Graph(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("out_offsets", vp), ("out_targets", vp),
                ("out_weights", vp), ("in_offsets", vp), ("in_sources", vp), ("in_weights", vp),
                ("out_to_in", vp)]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("stencil", C.c_int32),
                ("c", C.c_double), ("decay_dt", C.c_double), ("kx", vp), ("ky", vp), ("kz", vp),
                ("block", C.c_int32), ("pad_", C.c_int32)]


class CgOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("maxit", C.c_int32), ("precond", C.c_int32),
                ("check_every", C.c_int32), ("use_x0", C.c_int32), ("stall_iters", C.c_int32),
                ("smooth", C.c_int32), ("mg_nu", C.c_int32), ("mg_smoother", C.c_int32),
                ("mg_cycle", C.c_int32), ("reserved_", C.c_int32)]


class CgResult(C.Structure):
    _fields_ = [("iters", C.c_int32), ("outer", C.c_int32), ("converged", C.c_int32),
                ("precond_used", C.c_int32), ("relres", C.c_double), ("true_relres", C.c_double),
                ("ms", C.c_double), ("learned_weight", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


DIEKMAN_PARAM_FIELDS = ("gna", "gnaleak", "gk", "gcal", "gcanl", "ena", "ek", "eca", "k1", "k2", "kcas", "kcac",
                        "taucas", "taucac", "bcas", "bcac", "ar", "ad", "tmax", "vt", "kp", "g_gaba", "appcur",
                        "gnap", "tau_rl", "tau_rnl", "tau_p", "out_thresh")
DIEKMAN_STATE = ("v", "m", "h", "ng", "rl", "rnl", "fnl", "s", "p", "y", "cas", "cac")


class DiekmanParams(C.Structure):
    _fields_ = [(f, C.c_double) for f in DIEKMAN_PARAM_FIELDS]

    def consts(self):
        """The 27 model constants (everything but out_thresh), in order."""
        return [getattr(self, f) for f in DIEKMAN_PARAM_FIELDS[:27]]


class DiekmanState(C.Structure):
    _fields_ = ([("dtype", C.c_int32), ("pad_", C.c_int32), ("n", C.c_int64)] + [(f, vp) for f in DIEKMAN_STATE] +
                [("iin", vp), ("icount", vp), ("ioutput", vp), ("egaba", vp), ("gkca", vp), ("gkleak", vp)])


SAN_PARAM_FIELDS = ("c", "a", "v_l", "v_na", "v_k", "v_ca", "kd", "v_ampa", "v_gaba", "alpha_ca", "tau_ampa",
                    "tau_gaba", "g_l", "g_k", "g_nap", "g_ca", "g_kca", "g_ampa", "g_gaba", "tau_ca", "noise_mean",
                    "tmax", "vt", "kp", "out_thresh")


class SanParams(C.Structure):
    _fields_ = [(f, C.c_double) for f in SAN_PARAM_FIELDS]

    def consts(self):
        """The 24 model constants (everything but out_thresh), in order."""
        return [getattr(self, f) for f in SAN_PARAM_FIELDS[:24]]


class SanState(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("pad_", C.c_int32), ("n", C.c_int64)] + [
        (f, vp) for f in ("v", "nk", "ca", "ye", "yi", "ein", "ecount", "iin", "icount", "eoutput", "ioutput")]


class Regions(C.Structure):
    _fields_ = [("n_regions", C.c_int64), ("nnz", C.c_int64), ("offsets", vp), ("idx", vp)]


class Hebbian(C.Structure):
    _fields_ = [("window", C.c_int32), ("heb_inc", C.c_float), ("weights", vp), ("armed", vp),
                ("last_fire", vp), ("ring", vp), ("in_to_out", vp)]


class Paracrine(C.Structure):
    _fields_ = [("every", C.c_int64), ("grid", Grid), ("release", C.c_double), ("rtol", C.c_double),
                ("inner_rtol", C.c_double), ("max_inner", C.c_int32), ("max_outer", C.c_int32),
                ("conc", vp), ("at_neurons", vp), ("stats", vp)]


PARA_STATS = ("solves", "inner_iters", "outer_steps", "max_relres", "unconverged", "last_relres",
              "last_inner_iters", "last_outer_steps")


class StepPlan(C.Structure):
    _fields_ = [("state", HHState), ("params", HHParams), ("graph", Graph),
                ("integrator", C.c_int32), ("spike_rule", C.c_int32), ("couple_mode", C.c_int32),
                ("record", C.c_int32), ("dt", C.c_double), ("g_in", vp), ("bits", vp * 2),
                ("list", vp), ("count", vp), ("voxel_of", vp), ("voxel_counts", vp),
                ("raster", vp), ("raster_len", vp), ("raster_cap", C.c_int64),
                ("fused_block", C.c_int32), ("fused_tile", C.c_int32), ("fused_grid", C.c_int32),
                ("fused_burst", C.c_int32), ("fused_bulk", C.c_int32), ("fused_pad_", C.c_int32),
                ("regions", Regions), ("readout_every", C.c_int64),
                ("readout_cap", C.c_int64), ("readout_avg", vp), ("readout_firing", vp),
                ("readout_mean", vp), ("hebbian", Hebbian), ("paracrine", Paracrine)]


class EnsPlan(C.Structure):
    _fields_ = [("state", HHState), ("params", HHParams), ("graph", Graph), ("nnet", C.c_int32),
                ("pad_", C.c_int32), ("net_begin", vp), ("net_size", vp), ("dt", C.c_double),
                ("bits", vp * 2), ("voxel_of", vp), ("voxel_counts", vp), ("net_spikes", vp),
                ("raster", vp), ("raster_len", vp), ("raster_cap", C.c_int64), ("paracrine", Paracrine)]


class EnsConfig(C.Structure):
    _fields_ = [("block", C.c_int32), ("tile", C.c_int32), ("grid", C.c_int32), ("bulk", C.c_int32)]

    def as_dict(self):
        return {"block": self.block, "tile": self.tile, "grid": self.grid, "bulk": self.bulk}


_P = C.POINTER
_SIGS = {
    "nnx_create": ([_P(vp), C.c_int, vp], C.c_int),
    "nnx_destroy": ([vp], C.c_int),
    "nnx_last_error": ([vp], C.c_char_p),
    "nnx_set_stream": ([vp, vp], C.c_int),
    "nnx_synchronize": ([vp], C.c_int),
    "nnx_version": ([], C.c_char_p),
    "nnx_malloc": ([vp, _P(vp), C.c_size_t], C.c_int),
    "nnx_free": ([vp, vp], C.c_int),
    "nnx_memcpy_h2d": ([vp, vp, vp, C.c_size_t], C.c_int),
    "nnx_memcpy_d2h": ([vp, vp, vp, C.c_size_t], C.c_int),
    "nnx_memcpy_d2d": ([vp, vp, vp, C.c_size_t], C.c_int),
    "nnx_memset": ([vp, vp, C.c_int, C.c_size_t], C.c_int),
    "nnx_event_create": ([vp, _P(vp)], C.c_int),
    "nnx_event_destroy": ([vp, vp], C.c_int),
    "nnx_event_record": ([vp, vp], C.c_int),
    "nnx_event_elapsed": ([vp, vp, vp, _P(C.c_float)], C.c_int),
    "nnx_get_stream": ([vp], vp),
    "nnx_hh_params_reference": ([_P(HHParams)], None),
    "nnx_hh_params_squid": ([_P(HHParams)], None),
    "nnx_hh_step": ([vp, _P(HHState), _P(HHParams), C.c_int, C.c_int, C.c_double, C.c_uint64,
                     vp, vp], C.c_int),
    "nnx_check_finite": ([vp, C.c_int, C.c_int64, vp], C.c_int),
    "nnx_noise_normal": ([vp, C.c_uint64, C.c_uint64, C.c_int64, C.c_int64, vp], C.c_int),
    "nnx_spike_threshold": ([vp, C.c_int, C.c_int64, vp, C.c_double, vp], C.c_int),
    "nnx_spike_compact": ([vp, C.c_int64, vp, vp, vp], C.c_int),
    "nnx_read_count": ([vp, vp, _P(C.c_int32)], C.c_int),
    "nnx_graph_transpose": ([vp, C.c_int64, C.c_int64, vp, vp, vp, vp, vp, vp, vp, C.c_int, vp],
                            C.c_int),
    "nnx_couple": ([vp, _P(Graph), C.c_int, C.c_int, vp, vp, vp, vp, vp], C.c_int),
    "nnx_stencil_apply": ([vp, _P(Grid), vp, vp], C.c_int),
    "nnx_cg_solve": ([vp, _P(Grid), vp, vp, _P(CgOpts), _P(CgResult)], C.c_int),
    "nnx_cg_solve_refine": ([vp, _P(Grid), vp, vp, _P(CgOpts), C.c_double, C.c_int,
                             _P(CgResult)], C.c_int),
    "nnx_cg_solve_async": ([vp, _P(Grid), vp, vp, _P(CgOpts), vp], C.c_int),
    "nnx_cg_solve_async_noalloc": ([vp, _P(Grid), vp, vp, _P(CgOpts), vp], C.c_int),
    "nnx_graph_cache_clear": ([vp], C.c_int),
    "nnx_graph_cache_stats": ([vp, _P(C.c_int), _P(C.c_int64)], C.c_int),
    "nnx_update_density_reference": ([vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_int, vp,
                                      _P(CgResult)], C.c_int),
    "nnx_deposit_count": ([vp, vp, vp, vp, vp], C.c_int),
    "nnx_release_add": ([vp, vp, vp, C.c_double, vp], C.c_int),
    "nnx_rhs_from_counts": ([vp, C.c_int64, vp, vp, C.c_double, vp], C.c_int),
    "nnx_csr_spmv": ([vp, C.c_int64, vp, vp, vp, vp, vp, C.c_int], C.c_int),
    "nnx_run_steps": ([vp, _P(StepPlan), C.c_uint64, C.c_int, C.c_int], C.c_int),
    "nnx_electrode_measure": ([vp, _P(Regions), C.c_int, C.c_int64, vp, vp, vp, vp], C.c_int),
    "nnx_network_mean": ([vp, C.c_int, C.c_int64, vp, vp], C.c_int),
    "nnx_synchrony_accumulate": ([vp, C.c_int, C.c_int64, vp, C.c_int64, vp, C.c_int32, vp, vp, vp, vp,
                                  vp], C.c_int),
    "nnx_synchrony_measures": ([vp, C.c_int64, C.c_int32, vp, vp, vp, vp, vp, C.c_int64, vp], C.c_int),
    "nnx_hebbian_init": ([vp, _P(Graph), _P(Hebbian)], C.c_int),
    "nnx_diekman_params_reference": ([_P(DiekmanParams)], None),
    "nnx_diekman_step": ([vp, _P(DiekmanState), _P(DiekmanParams), C.c_double, vp], C.c_int),
    "nnx_san_params_reference": ([_P(SanParams)], None),
    "nnx_san_step": ([vp, _P(SanState), _P(SanParams), C.c_double, vp], C.c_int),
    "nnx_hebbian_update": ([vp, _P(Graph), _P(Hebbian), C.c_int64, vp], C.c_int),
    "nnx_unet_param_count": ([], C.c_int64),
    "nnx_unet_create": ([vp, vp, C.c_uint64, _P(vp)], C.c_int),
    "nnx_unet_destroy": ([vp, vp], C.c_int),
    "nnx_unet_get_params": ([vp, vp, vp], C.c_int),
    "nnx_unet_set_coarse": ([vp, vp, C.c_int], C.c_int),
    "nnx_unet_get_grad": ([vp, vp, vp], C.c_int),
    "nnx_unet_apply": ([vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp], C.c_int),
    "nnx_unet_train_step": ([vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp,
                             C.c_int, C.c_double, _P(C.c_double)], C.c_int),
    "nnx_conv3_bf16": ([vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp,
                        C.c_double, vp, vp, C.c_int, vp, vp], C.c_int),
    "nnx_conv3_wgrad_bf16": ([vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp,
                              C.c_double, vp, vp, vp], C.c_int),
    "nnx_ensemble_run": ([vp, _P(EnsPlan), _P(EnsConfig), C.c_uint64, C.c_int, C.c_int], C.c_int),
    "nnx_ensemble_autotune": ([vp, _P(EnsPlan), C.c_uint64, C.c_int, _P(EnsConfig), _P(C.c_double),
                               _P(C.c_int)], C.c_int),
    "nnx_fcg_solve": ([vp, _P(Grid), vp, vp, vp, vp, _P(CgOpts), _P(CgResult)], C.c_int),
    "nnx_mg_apply": ([vp, _P(Grid), C.c_int, C.c_int, C.c_int, vp, vp], C.c_int),
    "nnx_fcg_solve_refine": ([vp, _P(Grid), vp, vp, vp, vp, _P(CgOpts), C.c_double, C.c_int,
                              _P(CgResult)], C.c_int),
}

_HOST_SIGS = {
    "nnx_philox4x64_10": ([_P(C.c_uint64), _P(C.c_uint64), _P(C.c_uint64)], None),
    "nnx_gen_uniform": ([C.c_uint64, C.c_uint64, C.c_int64, C.c_double, C.c_double, vp], None),
    "nnx_gen_normal": ([C.c_uint64, C.c_uint64, C.c_int64, C.c_double, C.c_double, vp], None),
    "nnx_gen_lognormal": ([C.c_uint64, C.c_uint64, C.c_int64, C.c_double, C.c_double, vp], None),
    "nnx_gen_er_fixed": ([C.c_uint64, C.c_int64, C.c_int32, C.c_int, vp, vp], C.c_int),
    "nnx_gen_uniform_targets": ([C.c_uint64, C.c_int64, vp, C.c_int, vp], C.c_int),
    "nnx_gen_outdeg_lognormal": ([C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_int32, vp],
                                 C.c_int64),
    "nnx_gen_outdeg_poisson": ([C.c_uint64, C.c_int64, C.c_double, vp], C.c_int64),
    "nnx_gen_spatial_targets": ([C.c_uint64, C.c_int64, vp, C.c_double, C.c_double, C.c_double,
                                 vp, vp], C.c_int64),
    "nnx_csr_to_csc": ([C.c_int64, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp], None),
    "nnx_hh_steady_state": ([C.c_int64, vp, C.c_double, vp, vp, vp], None),
    "nnx_voxel_of": ([C.c_int64, vp, C.c_double, C.c_int32, C.c_int32, C.c_int32, vp], None),
    "nnx_trilinear_matrix": ([C.c_int64, vp, vp, vp, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                              vp, vp, vp], C.c_int64),
    "nnx_csr_transpose": ([C.c_int64, C.c_int64, vp, vp, vp, vp, vp, vp], None),
    "nnx_gen_grf": ([C.c_uint64, C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_double, vp],
                    None),
    "nnx_face_coefficients": ([C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp], None),
    "nnx_spherical_regions": ([C.c_int64, vp, C.c_int64, vp, vp, C.c_int32, vp, vp], C.c_int),
    "nnx_h5_count": ([C.c_char_p, C.c_char_p, vp, C.c_size_t], C.c_int64),
    "nnx_h5_read": ([C.c_char_p, C.c_char_p, C.c_int, vp, C.c_int64, vp, C.c_size_t], C.c_int64),
    "nnx_csr_subsample": ([C.c_int64, vp, vp, C.c_int64, vp, vp, _P(C.c_int64)], C.c_int64),
    "nnx_morphology_from_axes": ([C.c_int64, vp, vp, vp, C.c_int64, vp], C.c_int),
    "nnx_read_connection_file": ([C.c_char_p, _P(C.c_int64), _P(C.c_int64), vp, vp, vp, vp, C.c_size_t],
                                 C.c_int64),
}

_lib = None
_host = None


def _bind(lib, sigs):
    for name, (args, res) in sigs.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res


def lib():
    """libnnx.so (device library); raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NNXError(2, f"{LIB_PATH} is missing: run __graft_entry__.build() first "
                              "(there is no CPU fallback)")
        _lib = C.CDLL(LIB_PATH)
        _bind(_lib, _SIGS)
    return _lib


def host():
    """libnnx_host.so (host builders)."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise NNXError(2, f"{HOST_LIB_PATH} is missing: run __graft_entry__.build() first")
        _host = C.CDLL(HOST_LIB_PATH)
        _bind(_host, _HOST_SIGS)
    return _host


def exported_symbols():
    """Names the C headers declare (used by the load/export test)."""
    return list(_SIGS), list(_HOST_SIGS)


class Context:
    """Owns one nnx_ctx (device, stream, scratch, captured graphs)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self._l = lib()
        h = vp()
        st = self._l.nnx_create(C.byref(h), device, vp(stream) if stream else None)
        if st != 0:
            raise NNXError(st, f"nnx_create(device={device}) failed")
        self.h = h
        self.device = device

    def check(self, status):
        if status != 0:
            raise NNXError(status, self._l.nnx_last_error(self.h).decode())
        return status

    def call(self, name, *args):
        return self.check(getattr(self._l, name)(self.h, *args))

    def synchronize(self):
        self.call("nnx_synchronize")

    def graph_cache(self):
        """(cached executable graphs, graphs instantiated so far)."""
        n, tot = C.c_int(), C.c_int64()
        self.call("nnx_graph_cache_stats", C.byref(n), C.byref(tot))
        return n.value, tot.value

    def clear_graphs(self):
        self.call("nnx_graph_cache_clear")

    def close(self):
        if getattr(self, "h", None):
            self._l.nnx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx = {}


def context(device: int = 0) -> Context:
    """Process-wide context per device."""
    if device not in _ctx:
        _ctx[device] = Context(device)
    return _ctx[device]


def params_reference() -> HHParams:
    p = HHParams()
    lib().nnx_hh_params_reference(C.byref(p))
    return p


def params_diekman() -> DiekmanParams:
    p = DiekmanParams()
    lib().nnx_diekman_params_reference(C.byref(p))
    return p


def params_san() -> SanParams:
    p = SanParams()
    lib().nnx_san_params_reference(C.byref(p))
    return p


def params_squid() -> HHParams:
    p = HHParams()
    lib().nnx_hh_params_squid(C.byref(p))
    return p

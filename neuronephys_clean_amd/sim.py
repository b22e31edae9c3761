"""Device simulation driver over nnx_run_steps.

Mirrors the reference's step loop (_RunAndPrintSimulation,
src/main.cu:112-182; MouseCortex::Simulate, src_wholebrain/mouse_cortex.cu:104-120):
coupling on the previous step's spikes -> paracrine release -> ephys, with
the spike list, raster and release counts produced on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import HHParams, StepPlan
from .configs import Network
from .devmem import DeviceArray

COUPLE = {"pull": _lib.COUPLE_WEIGHTED_PULL, "event": _lib.COUPLE_WEIGHTED_EVENT,
          "count": _lib.COUPLE_COUNT, "signal": _lib.COUPLE_SIGNAL, "fused": _lib.COUPLE_WEIGHTED_FUSED}


def upload_graph(ctx, net: Network, device_csc: bool = False):
    """CSR (+ weights) and its CSC/out_to_in on the device.  With device_csc
    the transpose is built by nnx_graph_transpose instead of the host."""
    d = {"off": DeviceArray.from_numpy(ctx, net.off), "tgt": DeviceArray.from_numpy(ctx, net.tgt),
         "w": DeviceArray.from_numpy(ctx, net.w)}
    n, nnz = net.n, net.nnz
    if device_csc:
        d["in_off"] = DeviceArray(ctx, n + 1, np.int32)
        d["in_src"] = DeviceArray(ctx, max(nnz, 1), np.int32)
        d["in_w"] = DeviceArray(ctx, max(nnz, 1), np.float32)
        d["o2i"] = DeviceArray(ctx, max(nnz, 1), np.int32)
        d["deg"] = DeviceArray(ctx, n, np.float32)
        ctx.call("nnx_graph_transpose", n, nnz, d["off"].ptr, d["tgt"].ptr, d["w"].ptr,
                 d["in_off"].ptr, d["in_src"].ptr, d["in_w"].ptr, d["o2i"].ptr, _lib.F32,
                 d["deg"].ptr)
    else:
        in_off, in_src, in_w, o2i, deg = net.ensure_csc()
        for k, a in (("in_off", in_off), ("in_src", in_src), ("in_w", in_w), ("o2i", o2i),
                     ("deg", deg)):
            d[k] = DeviceArray.from_numpy(ctx, a)
    g = _lib.Graph(n=n, nnz=nnz, out_offsets=d["off"].ptr, out_targets=d["tgt"].ptr,
                   out_weights=d["w"].ptr, in_offsets=d["in_off"].ptr,
                   in_sources=d["in_src"].ptr, in_weights=d["in_w"].ptr, out_to_in=d["o2i"].ptr)
    return g, d


def _drop_graphs(sim):
    """Drops the context's captured step graphs when a simulation goes
    away: they hold its buffers' addresses (the cache would otherwise keep
    them until it evicts them)."""
    ctx = getattr(sim, "ctx", None)
    if ctx is not None and getattr(ctx, "h", None):
        try:
            ctx.clear_graphs()
        except Exception:
            pass


class Simulation:
    """RK4 squid network with exponential synapses and weighted coupling."""

    def __del__(self):
        _drop_graphs(self)

    def __init__(self, net: Network, dtype=np.float32, couple: str = "event",
                 record_cap: int = 0, deposit: bool = False, params: HHParams | None = None,
                 ctx: _lib.Context | None = None, device_csc: bool = False, fused=None):
        """couple "fused": one persistent launch per step; `fused` =
        (block, tile, grid, burst) of that launch (0 = default)."""
        self.ctx = ctx or _lib.context()
        self.net = net
        self.T = np.dtype(dtype)
        ctx = self.ctx
        T = self.T
        n = net.n
        self.params = params or _lib.params_squid()
        self.a = {}
        a = self.a
        for k, v in (("v", net.V0), ("m", net.m0), ("h", net.h0), ("ng", net.n0),
                     ("s", np.zeros(n)), ("i_ext", net.iext)):
            a[k] = DeviceArray.from_numpy(ctx, np.asarray(v, T))
        self.graph, self.g = upload_graph(ctx, net, device_csc)
        nwords = (n + 63) // 64
        a["g_in"] = DeviceArray.zeros(ctx, n, T)
        a["bits0"] = DeviceArray.zeros(ctx, max(nwords, 1), np.uint64)
        a["bits1"] = DeviceArray.zeros(ctx, max(nwords, 1), np.uint64)
        a["list"] = DeviceArray.zeros(ctx, max(n, 1), np.int32)
        a["count"] = DeviceArray.zeros(ctx, 1, np.int32)
        a["raster_len"] = DeviceArray.zeros(ctx, 1, np.int64)
        a["raster"] = DeviceArray.zeros(ctx, max(2 * record_cap, 2), np.int32)
        self.record_cap = record_cap
        if deposit:
            a["voxel_of"] = DeviceArray.from_numpy(ctx, net.ensure_voxel())
            a["voxel_counts"] = DeviceArray.zeros(ctx, int(np.prod(net.grid)), np.int32)
        st = _lib.HHState(dtype=_lib.F32 if T == np.float32 else _lib.F64, n=n,
                          v=a["v"].ptr, m=a["m"].ptr, h=a["h"].ptr, ng=a["ng"].ptr, s=a["s"].ptr,
                          i_ext=a["i_ext"].ptr)
        p = StepPlan()
        p.state = st
        p.params = self.params
        p.graph = self.graph
        p.integrator = _lib.INT_RK4
        p.spike_rule = _lib.SPIKE_CROSSING
        p.couple_mode = COUPLE[couple]
        p.record = 1 if record_cap > 0 else 0
        p.dt = net.dt
        p.g_in = a["g_in"].ptr
        p.bits[0] = a["bits0"].ptr
        p.bits[1] = a["bits1"].ptr
        p.list = a["list"].ptr
        p.count = a["count"].ptr
        if deposit:
            p.voxel_of = a["voxel_of"].ptr
            p.voxel_counts = a["voxel_counts"].ptr
        p.raster = a["raster"].ptr
        p.raster_len = a["raster_len"].ptr
        p.raster_cap = record_cap
        if fused:  # (block, tile, grid, burst[, bulk])
            f = tuple(fused) + (0,) * (5 - len(fused))
            p.fused_block, p.fused_tile, p.fused_grid, p.fused_burst, p.fused_bulk = f
        self.plan = p
        self.step = 0

    def run(self, n_steps: int, use_graph: bool = False, chunk: int | None = None):
        if not use_graph or not chunk:
            self.ctx.call("nnx_run_steps", C.byref(self.plan), self.step, n_steps, int(use_graph))
            self.step += n_steps
            return
        done = 0
        while done < n_steps:
            k = min(chunk, n_steps - done)
            self.ctx.call("nnx_run_steps", C.byref(self.plan), self.step, k, 1)
            self.step += k
            done += k

    def synchronize(self):
        self.ctx.synchronize()

    def set_readout(self, offsets, idx, every: int, cap: int, firing: bool = True, mean: bool = True):
        """Electrode readout every `every` steps into device rows (readout.attach)."""
        from . import readout
        readout.attach(self, offsets, idx, every, cap, firing, mean)

    def readout(self) -> dict:
        from . import readout
        return readout.rows(self)

    def set_paracrine(self, every: int, rtol: float = 1e-8, inner_rtol: float = 1e-6, max_inner: int = 0,
                      max_outer: int = 0, at_neurons: bool = False, conc0=None, grid: _lib.Grid | None = None):
        """Solves the network's paracrine field inside the step every `every`
        steps (nnx_paracrine; needs deposit=True): b = conc + amount * counts,
        conc = A^-1 b to an FP64 relative residual rtol.  `every` <= 0 turns
        it off."""
        net, a, ctx = self.net, self.a, self.ctx
        q = self.plan.paracrine
        if every <= 0:
            q.every = 0
            return
        if "voxel_counts" not in a:
            raise ValueError("the paracrine solve needs the release deposit (deposit=True)")
        nvox = int(np.prod(net.grid))
        c0 = np.zeros(nvox) if conc0 is None else np.asarray(conc0, np.float64)
        a["conc"] = DeviceArray.from_numpy(ctx, c0)
        a["para_stats"] = DeviceArray.zeros(ctx, 8, np.float64)
        if at_neurons:
            a["at_neurons"] = DeviceArray.zeros(ctx, net.n, np.float32)
        g = grid or _lib.Grid(net.grid[0], net.grid[1], net.grid[2], _lib.STENCIL_7, net.c, 0.0, None, None, None)
        self._para_grid = g
        q.every = every
        q.grid = g
        q.release = net.amount
        q.rtol = rtol
        q.inner_rtol = inner_rtol
        q.max_inner = max_inner
        q.max_outer = max_outer
        q.conc = a["conc"].ptr
        q.at_neurons = a["at_neurons"].ptr if at_neurons else None
        q.stats = a["para_stats"].ptr

    def paracrine_stats(self) -> dict:
        v = self.a["para_stats"].numpy()
        return dict(zip(_lib.PARA_STATS, (float(x) for x in v)))

    def set_fused(self, block=0, tile=0, grid=0, burst=None, bulk=None):
        """Launch shape of the fused step and its burst and flagless-event
        thresholds (results do not depend on them)."""
        p = self.plan
        p.fused_block, p.fused_tile, p.fused_grid = block, tile, grid
        if burst is not None:
            p.fused_burst = burst
        if bulk is not None:
            p.fused_bulk = bulk

    def raster(self) -> np.ndarray:
        """(step, neuron) rows, ordered by step then neuron."""
        ln = int(self.a["raster_len"].numpy()[0])
        r = self.a["raster"].numpy().reshape(-1, 2)[:min(ln, self.record_cap)]
        if self.plan.couple_mode == _lib.COUPLE_WEIGHTED_FUSED and len(r):
            r = r[np.lexsort((r[:, 1], r[:, 0]))]
        return r, ln

    def state(self) -> dict:
        return {k: self.a[k].numpy() for k in ("v", "m", "h", "ng", "s")}

    def spike_count_last(self) -> int:
        return int(self.a["count"].numpy()[0])


SOA = ("V", "M", "N", "H", "LOCAL_Y", "IY", "EY", "LOCAL_OUTPUT", "LOCAL_IN", "LOCAL_COUNT",
       "EOUTPUT", "EIN", "ECOUNT", "IOUTPUT", "IIN", "ICOUNT", "APPCUR")


class ReferenceSimulation:
    """The reference's HH network: leapfrog with kinetic synapses, LEVEL
    firing (output > 0.1), COUNT (src/) or SIGNAL (src_wholebrain/) coupling
    into LOCAL_IN with LOCAL_COUNT = in-degree; the 17 SoA arrays of
    src/defines.cuh:57-78."""

    def __del__(self):
        _drop_graphs(self)

    def __init__(self, net: Network, sv0: dict, dtype=np.float32, couple: str = "count",
                 record_cap: int = 0, params: HHParams | None = None, dt: float = 0.4,
                 ctx: _lib.Context | None = None, noise: bool = False):
        """noise: the reference's white noise (std 1.5, mean 3) drawn per
        (seed, step, neuron); otherwise its mean (the deterministic mode)."""
        self.ctx = ctx or _lib.context()
        ctx = self.ctx
        self.T = np.dtype(dtype)
        n = net.n
        self.params = params or _lib.params_reference()
        if not noise:
            self.params.noise_std = 0.0
        self.sv = {k: DeviceArray.from_numpy(ctx, np.asarray(sv0.get(k, np.zeros(n)), self.T))
                   for k in SOA}
        self.graph, self.g = upload_graph(ctx, net)
        nwords = (n + 63) // 64
        a = self.a = {}
        a["bits0"] = DeviceArray.zeros(ctx, max(nwords, 1), np.uint64)
        a["bits1"] = DeviceArray.zeros(ctx, max(nwords, 1), np.uint64)
        a["list"] = DeviceArray.zeros(ctx, max(n, 1), np.int32)
        a["count"] = DeviceArray.zeros(ctx, 1, np.int32)
        a["raster_len"] = DeviceArray.zeros(ctx, 1, np.int64)
        a["raster"] = DeviceArray.zeros(ctx, max(2 * record_cap, 2), np.int32)
        self.record_cap = record_cap
        # the step's first coupling reads the firing of the initial outputs
        # (src/coupling.cu:118): seed the "previous" bitmap from them
        ctx.call("nnx_spike_threshold", _lib.F32 if self.T == np.float32 else _lib.F64, n,
                 self.sv["LOCAL_OUTPUT"].ptr, self.params.out_thresh, a["bits1"].ptr)
        ctx.call("nnx_spike_compact", n, a["bits1"].ptr, a["list"].ptr, a["count"].ptr)
        sv = self.sv
        st = _lib.HHState(dtype=_lib.F32 if self.T == np.float32 else _lib.F64, n=n,
                          v=sv["V"].ptr, m=sv["M"].ptr, h=sv["H"].ptr, ng=sv["N"].ptr,
                          y_local=sv["LOCAL_Y"].ptr, y_i=sv["IY"].ptr, y_e=sv["EY"].ptr,
                          in_local=sv["LOCAL_IN"].ptr, count_local=sv["LOCAL_COUNT"].ptr,
                          in_i=sv["IIN"].ptr, count_i=sv["ICOUNT"].ptr, in_e=sv["EIN"].ptr,
                          count_e=sv["ECOUNT"].ptr, appcur=sv["APPCUR"].ptr,
                          out_local=sv["LOCAL_OUTPUT"].ptr, out_i=sv["IOUTPUT"].ptr,
                          out_e=sv["EOUTPUT"].ptr)
        p = StepPlan()
        p.state = st
        p.params = self.params
        p.graph = self.graph
        p.integrator = _lib.INT_LEAPFROG
        p.spike_rule = _lib.SPIKE_LEVEL
        p.couple_mode = COUPLE[couple]
        p.record = 1 if record_cap > 0 else 0
        p.dt = dt
        p.bits[0] = a["bits0"].ptr
        p.bits[1] = a["bits1"].ptr
        p.list = a["list"].ptr
        p.count = a["count"].ptr
        p.raster = a["raster"].ptr
        p.raster_len = a["raster_len"].ptr
        p.raster_cap = record_cap
        self.plan = p
        self.step = 0

    def run(self, n_steps: int, use_graph: bool = False, chunk: int | None = None):
        done = 0
        while done < n_steps:
            k = n_steps - done if not chunk else min(chunk, n_steps - done)
            self.ctx.call("nnx_run_steps", C.byref(self.plan), self.step, k, int(use_graph))
            self.step += k
            done += k

    def set_readout(self, offsets, idx, every: int, cap: int, firing: bool = True, mean: bool = True):
        """Electrode readout every `every` steps into device rows (readout.attach)."""
        from . import readout
        readout.attach(self, offsets, idx, every, cap, firing, mean)

    def readout(self) -> dict:
        from . import readout
        return readout.rows(self)

    def raster(self):
        ln = int(self.a["raster_len"].numpy()[0])
        return self.a["raster"].numpy().reshape(-1, 2)[:min(ln, self.record_cap)], ln

    def state(self) -> dict:
        return {k: self.sv[k].numpy() for k in SOA}


class EnsembleSim:
    """An ensemble (configs.Ensemble) stepped by nnx_ensemble_run."""

    def __del__(self):
        _drop_graphs(self)

    def __init__(self, ens, ctx: _lib.Context | None = None, record_cap: int = 0, deposit: bool = False,
                 params: HHParams | None = None):
        self.ctx = ctx or _lib.context()
        self.ens = ens
        ctx = self.ctx
        n = ens.n
        self.params = params or _lib.params_squid()
        a = self.a = {}
        for k, v in (("v", ens.V0), ("m", ens.m0), ("h", ens.h0), ("ng", ens.n0), ("s", np.zeros(n)),
                     ("i_ext", ens.iext)):
            a[k] = DeviceArray.from_numpy(ctx, np.asarray(v, np.float32))
        d = {"off": DeviceArray.from_numpy(ctx, ens.off), "tgt": DeviceArray.from_numpy(ctx, ens.tgt),
             "w": DeviceArray.from_numpy(ctx, ens.w)}
        nnz = ens.nnz
        for k, dt in (("in_off", np.int32), ("in_src", np.int32), ("in_w", np.float32), ("o2i", np.int32)):
            d[k] = DeviceArray(ctx, max(nnz, 1) if k != "in_off" else n + 1, dt)
        ctx.call("nnx_graph_transpose", n, nnz, d["off"].ptr, d["tgt"].ptr, d["w"].ptr, d["in_off"].ptr,
                 d["in_src"].ptr, d["in_w"].ptr, d["o2i"].ptr, _lib.F32, None)
        self.g = d
        a["bits0"] = DeviceArray.zeros(ctx, n // 64, np.uint64)
        a["bits1"] = DeviceArray.zeros(ctx, n // 64, np.uint64)
        a["net_spikes"] = DeviceArray.zeros(ctx, len(ens.size), np.int32)
        a["raster"] = DeviceArray.zeros(ctx, max(2 * record_cap, 2), np.int32)
        a["raster_len"] = DeviceArray.zeros(ctx, 1, np.int64)
        if deposit:
            a["voxel_counts"] = DeviceArray.zeros(ctx, int(np.prod(ens.grid)), np.int32)
            a["voxel_of"] = DeviceArray.from_numpy(ctx, ens.voxel_of)
        self._begin = np.ascontiguousarray(ens.begin, np.int64)
        self._size = np.ascontiguousarray(ens.size, np.int64)
        p = _lib.EnsPlan()
        p.state = _lib.HHState(dtype=_lib.F32, n=n, v=a["v"].ptr, m=a["m"].ptr, h=a["h"].ptr, ng=a["ng"].ptr,
                               s=a["s"].ptr, i_ext=a["i_ext"].ptr)
        p.params = self.params
        p.graph = _lib.Graph(n=n, nnz=nnz, out_offsets=d["off"].ptr, out_targets=d["tgt"].ptr,
                             out_weights=d["w"].ptr, in_offsets=d["in_off"].ptr, in_sources=d["in_src"].ptr,
                             in_weights=d["in_w"].ptr, out_to_in=d["o2i"].ptr)
        p.nnet = len(ens.size)
        p.net_begin = self._begin.ctypes.data
        p.net_size = self._size.ctypes.data
        p.dt = ens.dt
        p.bits[0] = a["bits0"].ptr
        p.bits[1] = a["bits1"].ptr
        p.net_spikes = a["net_spikes"].ptr
        if record_cap:
            p.raster = a["raster"].ptr
            p.raster_len = a["raster_len"].ptr
            p.raster_cap = record_cap
        if deposit:
            p.voxel_of = a["voxel_of"].ptr
            p.voxel_counts = a["voxel_counts"].ptr
        self.plan = p
        self.record_cap = record_cap
        self.step = 0
        self.cfg = None

    def run(self, n_steps: int, cfg=None, use_graph: bool = False, chunk: int = 100):
        cfg = cfg or self.cfg
        pc = C.byref(cfg) if cfg is not None else None
        if not use_graph:
            self.ctx.call("nnx_ensemble_run", C.byref(self.plan), pc, self.step, n_steps, 0)
            self.step += n_steps
            return
        done = 0
        while done < n_steps:
            k = min(chunk, n_steps - done)
            self.ctx.call("nnx_ensemble_run", C.byref(self.plan), pc, self.step, k, 1)
            self.step += k
            done += k

    def set_paracrine(self, every: int, rtol: float = 1e-8, inner_rtol: float = 1e-6, max_inner: int = 0,
                      max_outer: int = 0, release: float = 0.01):
        """Solves every network's 64^3 paracrine grid inside the step every
        `every` steps: one solve of the release box with Dirichlet faces
        between the networks' grids (nnx_grid.block = 64); needs
        deposit=True.  `every` <= 0 turns it off."""
        ens, a, ctx = self.ens, self.a, self.ctx
        q = self.plan.paracrine
        if every <= 0:
            q.every = 0
            return
        if "voxel_counts" not in a:
            raise ValueError("the paracrine solve needs the release deposit (deposit=True)")
        nx, ny, nz = ens.grid
        a["conc"] = DeviceArray.zeros(ctx, nx * ny * nz, np.float64)
        a["para_stats"] = DeviceArray.zeros(ctx, 8, np.float64)
        q.every = every
        q.grid = _lib.Grid(nx, ny, nz, _lib.STENCIL_7, ens.c, 0.0, None, None, None, 64)
        q.release = release
        q.rtol = rtol
        q.inner_rtol = inner_rtol
        q.max_inner = max_inner
        q.max_outer = max_outer
        q.conc = a["conc"].ptr
        q.at_neurons = None
        q.stats = a["para_stats"].ptr

    def paracrine_stats(self) -> dict:
        v = self.a["para_stats"].numpy()
        return dict(zip(_lib.PARA_STATS, (float(x) for x in v)))

    def autotune(self, steps_per_candidate: int = 20):
        best, us, ran = _lib.EnsConfig(), C.c_double(), C.c_int()
        self.ctx.call("nnx_ensemble_autotune", C.byref(self.plan), self.step, steps_per_candidate,
                      C.byref(best), C.byref(us), C.byref(ran))
        self.step += ran.value
        self.cfg = best
        return best, us.value, ran.value

    def synchronize(self):
        self.ctx.synchronize()

    def state(self) -> dict:
        return {k: self.a[k].numpy() for k in ("v", "m", "h", "ng", "s")}

    def raster(self):
        ln = int(self.a["raster_len"].numpy()[0])
        r = self.a["raster"].numpy().reshape(-1, 2)[:min(ln, self.record_cap)]
        return r[np.lexsort((r[:, 1], r[:, 0]))], ln

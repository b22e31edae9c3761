"""Electrode readout over the C ABI (nnx_electrode_measure, nnx_network_mean,
nnx_synchrony_*), in the shape of the reference's ElectrodeNet and
SynchronyNet (src/electrode_net.cuh:6-38, src_wholebrain/electrode_net.cuh:
6-92), plus the step-graph readout of nnx_run_steps.

The device computes; these classes only hold device buffers and copy results
back when asked (the reference copies after every measurement).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .devmem import DeviceArray


def regions_struct(off: DeviceArray, idx: DeviceArray, n_regions: int, nnz: int) -> _lib.Regions:
    return _lib.Regions(n_regions=n_regions, nnz=nnz, offsets=off.ptr, idx=idx.ptr if nnz else None)


def _dt(dtype) -> int:
    return _lib.F32 if np.dtype(dtype) == np.float32 else _lib.F64


class ElectrodeNet:
    """Regions stored as an adjacency list (regions_inds, regions_offsets),
    the reference constructor's argument order."""

    def __init__(self, ctx: _lib.Context, regions_inds, regions_offsets):
        self.ctx = ctx
        off = np.ascontiguousarray(regions_offsets, np.int32)
        idx = np.ascontiguousarray(regions_inds, np.int32)
        self.n_regions = len(off) - 1
        self.nnz = len(idx)
        self.d_off = DeviceArray.from_numpy(ctx, off)
        self.d_idx = DeviceArray.from_numpy(ctx, idx if len(idx) else np.zeros(1, np.int32))
        self.r = regions_struct(self.d_off, self.d_idx, self.n_regions, self.nnz)
        self.d_firing = DeviceArray.zeros(ctx, max(self.n_regions, 1), np.int32)
        self.d_avg = None

    def measure(self, v: DeviceArray, bits: DeviceArray | None = None):
        """DoElectrodeMeasurements: (mean V per region, firing count per region or None)."""
        if self.d_avg is None or self.d_avg.dtype != v.dtype:
            self.d_avg = DeviceArray.zeros(self.ctx, max(self.n_regions, 1), v.dtype)
        self.ctx.call("nnx_electrode_measure", C.byref(self.r), _dt(v.dtype), v.size, v.ptr,
                      bits.ptr if bits is not None else None, self.d_avg.ptr,
                      self.d_firing.ptr if bits is not None else None)
        avg = self.d_avg.numpy()[:self.n_regions]
        return avg, (self.d_firing.numpy()[:self.n_regions] if bits is not None else None)


def network_mean(ctx: _lib.Context, v: DeviceArray):
    """MeasureAverageEntireNetwork."""
    out = DeviceArray(ctx, 1, v.dtype)
    ctx.call("nnx_network_mean", _dt(v.dtype), v.size, v.ptr, out.ptr)
    return out.numpy()[0]


class SynchronyNet:
    """chi^2 synchrony of the sub-populations formed by the first
    sub_counts[s] neurons of region_inds."""

    def __init__(self, ctx: _lib.Context, region_inds, sub_counts):
        self.ctx = ctx
        idx = np.ascontiguousarray(region_inds, np.int32)
        cnt = np.ascontiguousarray(sub_counts, np.int32)
        self.n_idx, self.n_sub = len(idx), len(cnt)
        self.d_idx = DeviceArray.from_numpy(ctx, idx if len(idx) else np.zeros(1, np.int32))
        self.d_cnt = DeviceArray.from_numpy(ctx, cnt if len(cnt) else np.zeros(1, np.int32))
        self.sums = [DeviceArray.zeros(ctx, max(k, 1), np.float64)
                     for k in (self.n_idx, self.n_idx, self.n_sub, self.n_sub)]
        self.d_chi = DeviceArray.zeros(ctx, max(self.n_sub, 1), np.float64)
        self.steps = 0

    def measure_variances(self, v: DeviceArray):
        self.ctx.call("nnx_synchrony_accumulate", _dt(v.dtype), v.size, v.ptr, self.n_idx, self.d_idx.ptr,
                      self.n_sub, self.d_cnt.ptr, *(s.ptr for s in self.sums))
        self.steps += 1

    def measures(self) -> np.ndarray:
        self.ctx.call("nnx_synchrony_measures", self.n_idx, self.n_sub, self.d_cnt.ptr,
                      *(s.ptr for s in self.sums), self.steps, self.d_chi.ptr)
        return self.d_chi.numpy()[:self.n_sub]


def attach(sim, offsets, idx, every: int, cap: int, firing: bool = True, mean: bool = True):
    """Adds the print-step readout to a simulation's step plan: every `every`
    steps, one row of region means (and firing counts, network mean) is
    written on the device; rows >= cap are dropped."""
    ctx, a, p = sim.ctx, sim.a, sim.plan
    off = np.ascontiguousarray(offsets, np.int32)
    ix = np.ascontiguousarray(idx, np.int32)
    nreg = len(off) - 1
    a["ro_off"] = DeviceArray.from_numpy(ctx, off)
    a["ro_idx"] = DeviceArray.from_numpy(ctx, ix if len(ix) else np.zeros(1, np.int32))
    a["ro_avg"] = DeviceArray.zeros(ctx, max(cap * nreg, 1), sim.T)
    a["ro_firing"] = DeviceArray.zeros(ctx, max(cap * nreg, 1), np.int32)
    a["ro_mean"] = DeviceArray.zeros(ctx, max(cap, 1), sim.T)
    p.regions = regions_struct(a["ro_off"], a["ro_idx"], nreg, len(ix))
    p.readout_every = int(every)
    p.readout_cap = int(cap)
    p.readout_avg = a["ro_avg"].ptr
    p.readout_firing = a["ro_firing"].ptr if firing else None
    p.readout_mean = a["ro_mean"].ptr if mean else None
    sim.readout_shape = (cap, nreg)


def rows(sim) -> dict:
    """The readout rows written so far: avg (cap, n_regions), firing, mean."""
    cap, nreg = sim.readout_shape
    a, p = sim.a, sim.plan
    out = {"avg": a["ro_avg"].numpy()[:cap * nreg].reshape(cap, nreg)}
    out["firing"] = a["ro_firing"].numpy()[:cap * nreg].reshape(cap, nreg) if p.readout_firing else None
    out["mean"] = a["ro_mean"].numpy()[:cap] if p.readout_mean else None
    return out

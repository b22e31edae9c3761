"""Hebbian learning over the C ABI (nnx_hebbian_init / nnx_hebbian_update),
in the shape of the reference's Hebbian class (src_wholebrain/hebbian.cuh:
3-27): device state for one network, updated once per timestep with the
spike bitmap the coupling read."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .devmem import DeviceArray


class HebbianState:
    def __init__(self, ctx: _lib.Context, graph: _lib.Graph, window: int, heb_inc: float, weights):
        self.ctx, self.graph = ctx, graph
        n, nnz = graph.n, graph.nnz
        w = np.ascontiguousarray(weights, np.float32)
        assert len(w) == nnz
        self.d = {"w": DeviceArray.from_numpy(ctx, w if nnz else np.zeros(1, np.float32)),
                  "armed": DeviceArray(ctx, max(nnz, 1), np.uint8),
                  "last": DeviceArray(ctx, max(n, 1), np.int32),
                  "ring": DeviceArray(ctx, max(window * ((n + 63) // 64), 1), np.uint64),
                  "i2o": DeviceArray(ctx, max(nnz, 1), np.int32)}
        d = self.d
        self.h = _lib.Hebbian(window=window, heb_inc=heb_inc, weights=d["w"].ptr, armed=d["armed"].ptr,
                              last_fire=d["last"].ptr, ring=d["ring"].ptr, in_to_out=d["i2o"].ptr)
        ctx.call("nnx_hebbian_init", C.byref(graph), C.byref(self.h))
        self.nnz = nnz

    def update(self, step: int, bits: DeviceArray):
        self.ctx.call("nnx_hebbian_update", C.byref(self.graph), C.byref(self.h), int(step), bits.ptr)

    def weights(self) -> np.ndarray:
        return self.d["w"].numpy()[:self.nnz]

    def armed(self) -> np.ndarray:
        return self.d["armed"].numpy()[:self.nnz]


def attach(sim, window: int, heb_inc: float, weights) -> HebbianState:
    """Hebbian learning inside a simulation's step plan (after each step's
    coupling, on the spikes it read)."""
    st = HebbianState(sim.ctx, sim.graph, window, heb_inc, weights)
    sim.plan.hebbian = st.h
    sim.hebbian = st
    return st

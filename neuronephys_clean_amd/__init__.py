"""MI355X-native per-timestep hot path of the neuronephys spiking-network
simulator: Hodgkin-Huxley update, deterministic spike propagation over CSR,
paracrine implicit-diffusion CG.

The compute lives in libnnx.so (hand-written gfx950 HIP kernels behind the C
ABI of include/nnx.h); this package is the host-side mirror used by the
tests and the benchmark.  Nothing here computes on the CPU in place of the
device: if libnnx.so is missing every device call raises.
"""
from . import _lib  # noqa: F401
from ._lib import (COUPLE_COUNT, COUPLE_SIGNAL, COUPLE_WEIGHTED_EVENT,  # noqa: F401
                   COUPLE_WEIGHTED_PULL, F32, F64, INT_LEAPFROG, INT_RK4, NNXError,
                   PRECOND_JACOBI, PRECOND_LEARNED, PRECOND_NONE, SPIKE_CROSSING, SPIKE_LEVEL,
                   STENCIL_7, STENCIL_27, context, params_reference, params_squid)

__all__ = ["context", "params_reference", "params_squid", "NNXError"]

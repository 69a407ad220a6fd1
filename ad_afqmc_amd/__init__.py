"""MI355X-native phaseless-AFQMC walker engine.

A drop-in for the reference's propagator / trial-wavefunction calls
(propagate_constrained, calc_overlap, calc_force_bias, calc_energy on RHF/UHF
trials, and the sampler block loop) whose arithmetic runs in hand-written HIP
kernels for gfx950 (libafq.so, C ABI in include/afq.h).  PyTorch-ROCm tensors
are the device buffers.
"""
from . import _native  # noqa: F401

__version__ = "0.1.0"

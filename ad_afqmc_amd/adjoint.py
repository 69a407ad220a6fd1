"""Derivatives of the phaseless segment energy with respect to the Hamiltonian.

The reference obtains one- and two-body observables as derivatives of a block
energy with respect to h1 (and the Cholesky vectors) by running jax.jvp / jax.vjp
through a checkpointed phaseless segment (sampling.py:300-341,
driver.py:824-904).  Here the reverse pass is hand-written on the device
(libafq.so `afq_segment_energy_vjp`): the segment is checkpointed once per
block and each block is recomputed and pulled back in reverse.

`segment_energy` is a torch.autograd.Function, so the value composes with
torch autograd; `segment_rdm1` is the reference's reverse mode with
observable h1 (dF/dh1, the mixed one-body estimator of the segment).
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

from .engine import EngineJob, as_c128, as_f64


def _job(h0, h1, chol, mo, n_walkers, dt, n_exp_terms) -> EngineJob:
    h1 = as_f64(h1)
    chol = as_f64(chol)
    mo = as_f64(mo)
    N, O = mo.shape
    G = chol.shape[0]
    if chol.dim() != 2 or chol.shape[1] != N * N:
        raise ValueError(f"chol must be (G, {N * N}), got {tuple(chol.shape)}")
    if h1.shape != (2, N, N):
        raise ValueError(f"h1 must be (2, {N}, {N}), got {tuple(h1.shape)}")
    job = EngineJob(N, (O, O), G, n_walkers, "restricted", "rhf", n_exp_terms, dt, "fp64")
    d = mo @ mo.T
    job.set_hamiltonian(float(h0), h1, chol)
    job.set_trial(mo, None, torch.stack([d, d]))
    job.build_measurement()
    job.build_propagation()
    return job


class SegmentEnergy(torch.autograd.Function):
    """F(h1, chol) for fixed walkers, weights, e_estimate and fields."""

    @staticmethod
    def forward(ctx, h1, chol, h0, mo, walkers, weights, e_estimate, fields, dt, n_exp_terms):
        need = bool(ctx.needs_input_grad[0] or ctx.needs_input_grad[1])
        job = _job(h0, h1.detach(), chol.detach(), mo, walkers.shape[0], dt, n_exp_terms)
        try:
            F, gh1, gL = job.segment_energy_vjp(walkers, weights, e_estimate, fields, want_grad=need)
        finally:
            job.close()
        if need:
            ctx.save_for_backward(gh1, gL)
        return F.reshape(())

    @staticmethod
    def backward(ctx, gF):
        gh1, gL = ctx.saved_tensors
        return gF * gh1, gF * gL, None, None, None, None, None, None, None, None


def segment_energy(h1: torch.Tensor, chol: torch.Tensor, mo, walkers, weights, e_estimate: float, fields, dt: float,
                   h0: float = 0.0, n_exp_terms: int = 6) -> torch.Tensor:
    """Differentiable segment energy (RHF trial, restricted walkers, FP64; a
    UHF trial with unrestricted walkers goes through
    EngineJob.segment_energy_vjp with walkers_dn).

    h1 (2,N,N), chol (G,N*N), mo (N,O), walkers (W,N,O) c128, weights (W,),
    fields (n_blocks, n_steps, W, G)."""
    return SegmentEnergy.apply(h1, chol, h0, as_f64(mo), as_c128(walkers), as_f64(weights), float(e_estimate),
                               as_f64(fields), float(dt), int(n_exp_terms))


def segment_rdm1(h1, chol, mo, walkers, weights, e_estimate: float, fields, dt: float, h0: float = 0.0,
                 n_exp_terms: int = 6) -> Tuple[float, torch.Tensor, torch.Tensor]:
    """(F, dF/dh1 (2,N,N), dF/dchol (G,N*N)) in one reverse sweep -- the
    reference's "reverse" AD mode (driver.py:865-876)."""
    job = _job(h0, h1, chol, mo, as_c128(walkers).shape[0], dt, n_exp_terms)
    try:
        F, gh1, gL = job.segment_energy_vjp(walkers, weights, e_estimate, fields, want_grad=True)
    finally:
        job.close()
    return float(F.item()), gh1, gL

"""Phaseless propagator whose steps run in libafq.so (ad_afqmc/propagation.py:111-441).

`propagate_constrained` keeps the reference's functional contract (returns an
updated prop_data; the inputs are not modified).  The sampler uses the engine
in place instead.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import engine as E
from .binding import engine_for
from .walkers import RHFWalkers, UHFWalkers


@dataclass
class propagator_afqmc:
    """dt, n_walkers, n_exp_terms (Taylor terms of exp(VHS)), walker_type as
    in the reference; `precision` selects the arithmetic of the VHS / Taylor
    / exchange contractions ("fp64"; the reference's vhs_mixed_precision
    corresponds to a reduced-precision mode)."""

    dt: float = 0.01
    n_walkers: int = 50
    n_exp_terms: int = 6
    n_chunks: int = 1
    walker_type: str = "restricted"
    precision: str = "fp64"

    def _job(self, trial, ham_data, wave_data, n_walkers=None):
        return engine_for(trial, wave_data, ham_data, self.walker_type, n_walkers or self.n_walkers, self)

    def init_prop_data(self, trial, wave_data: dict, ham_data: dict, seed: int, init_walkers=None) -> dict:
        """propagation.py:123-152 (the JAX PRNG key becomes a Philox seed and counter)."""
        pd = {}
        pd["weights"] = torch.ones(self.n_walkers, dtype=torch.float64, device=E.device())
        pd["seed"] = int(seed)
        pd["rng_counter"] = 0
        pd["zeta_rng"] = np.random.default_rng([int(seed), 0x5EED])
        pd["walkers"] = init_walkers if init_walkers is not None else trial.get_init_walkers(
            wave_data, self.n_walkers, self.walker_type
        )
        if "e_estimate" in ham_data:
            pd["e_estimate"] = E.as_f64(ham_data["e_estimate"]).reshape(())
        else:
            e = trial.calc_energy(pd["walkers"], ham_data, wave_data).real
            pd["e_estimate"] = (e.sum() / self.n_walkers).reshape(())
        pd["pop_control_ene_shift"] = pd["e_estimate"].clone()
        pd["overlaps"] = trial.calc_overlap(pd["walkers"], wave_data)
        pd["normed_overlaps"] = pd["overlaps"]
        pd["norms"] = torch.ones(self.n_walkers, dtype=torch.complex128, device=E.device())
        pd["n_killed_walkers"] = 0
        return pd

    def apply_trotprop(self, trial, ham_data: dict, walkers, shifted_fields, wave_data: dict):
        """_apply_trotprop (propagation.py:154-272): returns propagated walkers."""
        job = self._job(trial, ham_data, wave_data, walkers.n_walkers)
        up, dn = (t.clone() if t is not None else None for t in walkers.tensors())
        job.trotprop(up, dn, E.as_c128(shifted_fields))
        return RHFWalkers(up) if dn is None else UHFWalkers([up, dn])

    def propagate_constrained(self, trial, ham_data: dict, prop_data: dict, fields, wave_data: dict,
                              return_force_bias: bool = False) -> dict:
        """One phaseless step (propagation.py:274-336)."""
        walkers = prop_data["walkers"]
        job = self._job(trial, ham_data, wave_data, walkers.n_walkers)
        up, dn = (t.clone() if t is not None else None for t in walkers.tensors())
        weights = E.as_f64(prop_data["weights"]).clone()
        overlaps = E.as_c128(prop_data["overlaps"]).clone()
        scalars = torch.stack(
            [E.as_f64(prop_data["e_estimate"]).reshape(()), E.as_f64(prop_data["pop_control_ene_shift"]).reshape(())]
        ).contiguous()
        fb = None
        if return_force_bias:
            fb = torch.empty((walkers.n_walkers, job.nchol), dtype=torch.complex128, device=up.device)
        job.propagate(up, dn, E.as_f64(fields), weights, overlaps, scalars, fb)
        out = dict(prop_data)
        out["walkers"] = RHFWalkers(up) if dn is None else UHFWalkers([up, dn])
        out["weights"] = weights
        out["overlaps"] = overlaps
        out["pop_control_ene_shift"] = scalars[1]
        if fb is not None:
            out["force_bias"] = fb
        return out

    def __hash__(self) -> int:
        return hash(tuple(self.__dict__.values()))

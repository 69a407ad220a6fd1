"""Sampler whose blocks run on the device (ad_afqmc/sampling.py:15-168, 343-365).

One `propagate_phaseless` call runs n_sr_blocks x n_ene_blocks device blocks
(afq_block: n_prop_steps phaseless steps with on-device Philox fields, QR,
overlap and energy measurement, clipping and the population-control update)
and a comb resampling after each SR block, with no host synchronisation.  The
fields come from the engine's counter-based generator instead of JAX's
threefry stream, so results agree with the reference statistically, not
bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

import torch

from . import engine as E
from .binding import engine_for
from .walkers import RHFWalkers, UHFWalkers


@dataclass
class sampler:
    n_prop_steps: int = 50
    n_ene_blocks: int = 50
    n_sr_blocks: int = 1
    n_blocks: int = 50

    def propagate_phaseless(self, ham, ham_data: dict, prop, prop_data: dict, trial, wave_data: dict
                            ) -> Tuple[torch.Tensor, dict]:
        """Returns (weighted mean block energy as a device scalar, new prop_data)."""
        walkers = prop_data["walkers"]
        W = walkers.n_walkers
        job = engine_for(trial, wave_data, ham_data, prop.walker_type, W, prop)
        up, dn = (t.clone() if t is not None else None for t in walkers.tensors())
        weights = E.as_f64(prop_data["weights"]).clone()
        overlaps = job.overlap(up, dn)  # sampling.py:355
        e_est = E.as_f64(prop_data["e_estimate"]).reshape(())
        scalars = torch.stack([e_est, e_est.clone()]).contiguous()  # shift <- e_estimate, :357
        counter = int(prop_data.get("rng_counter", 0))
        seed = int(prop_data.get("seed", 0))
        zrng = prop_data.get("zeta_rng")
        outs = []
        for _ in range(self.n_sr_blocks):
            for _ in range(self.n_ene_blocks):
                out, counter = job.block(up, dn, weights, overlaps, scalars, self.n_prop_steps, seed, counter)
                outs.append(out)
            zeta = float(zrng.random()) if zrng is not None else 0.5
            job.stochastic_reconfiguration(up, dn, weights, torch.tensor([zeta], dtype=torch.float64, device=up.device))
            overlaps = job.overlap(up, dn)  # :167
        blocks = torch.stack(outs)
        energy = (blocks[:, 0] * blocks[:, 1]).sum() / blocks[:, 1].sum()
        new = dict(prop_data)
        new["walkers"] = RHFWalkers(up) if dn is None else UHFWalkers([up, dn])
        new["weights"] = weights
        new["overlaps"] = overlaps
        new["pop_control_ene_shift"] = scalars[1]
        new["rng_counter"] = counter
        new["block_energies"] = blocks
        return energy, new

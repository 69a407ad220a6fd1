"""Hamiltonian intermediates built on the device (ad_afqmc/hamiltonian.py:12-84).

Call order follows the reference driver (ad_afqmc/driver.py:68-71): measurement
intermediates first (they symmetrise h1 in ham_data), then propagation ones.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _native as nat
from .binding import engine_for, record


@dataclass
class hamiltonian:
    norb: int

    def build_measurement_intermediates(self, ham_data: dict, trial, wave_data: dict) -> dict:
        """h1 <- (h1 + h1^T)/2 per spin; rot_h1 = T^H h1; rot_chol_g = T^H L_g
        (rhf: wavefunctions.py:1215-1232, uhf: :1472-1496)."""
        for k in ("rot_h1", "rot_chol"):
            ham_data.pop(k, None)
        wt = "restricted" if trial.kind == "rhf" else "unrestricted"
        job = engine_for(trial, wave_data, ham_data, wt, 1)
        job.build_measurement()
        ham_data["h1"] = job.get(nat.AFQ_INT_H1)
        record(wave_data, "h1", ham_data["h1"])
        if trial.kind == "rhf":
            ham_data["rot_h1"] = job.get(nat.AFQ_INT_ROT_H1, 0)
            ham_data["rot_chol"] = job.get(nat.AFQ_INT_ROT_CHOL, 0)
        else:
            ham_data["rot_h1"] = [job.get(nat.AFQ_INT_ROT_H1, s) for s in range(2)]
            ham_data["rot_chol"] = [job.get(nat.AFQ_INT_ROT_CHOL, s) for s in range(2)]
        record(wave_data, "rot_h1", ham_data["rot_h1"])
        record(wave_data, "rot_chol", ham_data["rot_chol"])
        return ham_data

    def build_propagation_intermediates(self, ham_data: dict, prop, trial, wave_data: dict) -> dict:
        """mf_shifts, h0_prop, exp_h1 = expm(-dt/2 h1_mod) (propagation.py:378-438)."""
        for k in ("mf_shifts", "h0_prop", "exp_h1"):
            ham_data.pop(k, None)
        job = engine_for(trial, wave_data, ham_data, prop.walker_type, prop.n_walkers, prop)
        job.build_propagation()
        ham_data["mf_shifts"] = job.get(nat.AFQ_INT_MF_SHIFTS)
        ham_data["h0_prop"] = job.get(nat.AFQ_INT_H0_PROP)
        if prop.walker_type == "restricted":
            ham_data["exp_h1"] = job.get(nat.AFQ_INT_EXP_H1, 0)
        else:
            ham_data["exp_h1"] = torch.stack([job.get(nat.AFQ_INT_EXP_H1, s) for s in range(2)])
        for k in ("mf_shifts", "h0_prop", "exp_h1"):
            record(wave_data, k, ham_data[k])
        return ham_data

    def __hash__(self) -> int:
        return hash(self.norb)

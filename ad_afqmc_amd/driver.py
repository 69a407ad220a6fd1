"""Single-GPU energy driver (ad_afqmc/driver.py:24-192, 653-715, 1100-1146).

The reference's MPI gathers collapse to one rank: block energies and weights
are read back once per driver block (as float32, as the reference gathers
them), SR runs on the device, and the final estimate goes through outlier
rejection and the weighted blocking analysis.
"""
from __future__ import annotations

import time
from typing import Optional, Tuple

import numpy as np
import torch

from . import stat_utils
from .walkers import RHFWalkers, UHFWalkers

DEFAULT_OPTIONS = dict(
    # ad_afqmc/launch_script.py:115-124
    dt=0.01,
    n_walkers=50,
    n_prop_steps=50,
    n_ene_blocks=50,
    n_sr_blocks=1,
    n_blocks=50,
    n_ene_blocks_eql=5,
    n_sr_blocks_eql=10,
    n_eql=1,
    seed=0,
)


def _orthonormalize_and_sr(prop_data: dict, zeta: float):
    w = prop_data["walkers"].orthonormalize()
    w, weights = w.stochastic_reconfiguration_local(prop_data["weights"], zeta)
    prop_data["walkers"], prop_data["weights"] = w, weights
    return prop_data


def _run_equilibration(ham, ham_data, propagator, prop_data, trial, wave_data, sampler_eq, verbose):
    for n in range(1, sampler_eq.n_blocks + 1):
        be, prop_data = sampler_eq.propagate_phaseless(ham, ham_data, propagator, prop_data, trial, wave_data)
        be32 = np.float32(be.item())
        bw32 = np.float32(prop_data["weights"].sum().item())
        zeta = float(prop_data["zeta_rng"].random())
        prop_data = _orthonormalize_and_sr(prop_data, zeta)
        e_est = float(prop_data["e_estimate"]) if not isinstance(prop_data["e_estimate"], torch.Tensor) else prop_data[
            "e_estimate"
        ].item()
        prop_data["e_estimate"] = torch.tensor(np.float64(0.9 * e_est + 0.1 * np.float64(be32)), dtype=torch.float64,
                                               device=prop_data["weights"].device)
        if verbose and n % max(sampler_eq.n_blocks // 5, 1) == 0:
            print(f"# {n:>10}      {bw32:<20.9e} {be32:<20.9e}", flush=True)
    return prop_data


def afqmc_energy(ham_data: dict, ham, propagator, trial, wave_data: dict, sampler, options: dict,
                 init_walkers=None, verbose: bool = True) -> Tuple[float, float]:
    """Equilibration, sampling blocks, blocking analysis; returns (energy, error)."""
    opts = dict(DEFAULT_OPTIONS)
    opts.update(options)
    t0 = time.time()
    if "rdm1" not in wave_data:
        wave_data["rdm1"] = trial.get_rdm1(wave_data)
    ham_data = ham.build_measurement_intermediates(ham_data, trial, wave_data)
    ham_data = ham.build_propagation_intermediates(ham_data, propagator, trial, wave_data)
    prop_data = propagator.init_prop_data(trial, wave_data, ham_data, opts["seed"], init_walkers)
    if abs(prop_data["overlaps"].sum().item()) < 1.0e-6:
        raise ValueError("Initial overlaps are zero. Pass walkers with non-zero overlap.")
    sampler_eq = type(sampler)(
        n_prop_steps=50, n_ene_blocks=opts["n_ene_blocks_eql"], n_sr_blocks=opts["n_sr_blocks_eql"],
        n_blocks=opts["n_eql"]
    )
    if verbose:
        print("# Equilibration sweeps:", flush=True)
    prop_data = _run_equilibration(ham, ham_data, propagator, prop_data, trial, wave_data, sampler_eq, verbose)
    if verbose:
        print("#\n# Sampling sweeps:", flush=True)
    nb = sampler.n_blocks
    weights_b = np.zeros(nb)
    energies_b = np.zeros(nb)
    for n in range(nb):
        be, prop_data = sampler.propagate_phaseless(ham, ham_data, propagator, prop_data, trial, wave_data)
        be32 = np.float32(be.item())
        bw32 = np.float32(prop_data["weights"].sum().item())
        weights_b[n], energies_b[n] = bw32, be32
        zeta = float(prop_data["zeta_rng"].random())
        prop_data = _orthonormalize_and_sr(prop_data, zeta)
        e_est = prop_data["e_estimate"].item()
        prop_data["e_estimate"] = torch.tensor(0.9 * e_est + 0.1 * np.float64(be32), dtype=torch.float64,
                                               device=prop_data["weights"].device)
        if verbose and n % max(nb // 10, 1) == 0:
            m = np.sum(weights_b[: n + 1] * energies_b[: n + 1]) / np.sum(weights_b[: n + 1])
            print(f" {n:5d}      {m:.9e}      {time.time() - t0:.2e}", flush=True)
    samples = np.stack((weights_b, energies_b)).T
    clean, _ = stat_utils.reject_outliers(samples, 1)
    e, err = stat_utils.blocking_analysis(clean[:, 0], clean[:, 1], neql=0, printQ=verbose)
    if err is None:
        err = 0.0
    if verbose:
        print(f"AFQMC energy: {e:.8f} +/- {err:.8f}", flush=True)
    return float(e), float(err)

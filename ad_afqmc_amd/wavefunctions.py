"""RHF / UHF trial wave functions whose measurements run in libafq.so.

Mirrors the reference's trial interface (ad_afqmc/wavefunctions.py:19-494 for
the dispatch, :1038-1236 for rhf, :1238-1500 for uhf): calc_overlap,
calc_force_bias, calc_energy on RHFWalkers / UHFWalkers batches, get_rdm1,
get_init_walkers.  Restricted walkers with a UHF trial use the reference's
default mapping walker[:, :nelec[s]] (wavefunctions.py:181-186).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np
import torch

from . import engine as E
from .binding import engine_for
from .walkers import RHFWalkers, UHFWalkers, walker_type_of


def _mo_np(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


@dataclass
class wave_function:
    norb: int
    nelec: Tuple[int, int]
    n_chunks: int = 1
    projector: Optional[str] = None
    kind: str = field(default="", init=False)

    def _check(self, walkers):
        if self.projector is not None:
            raise NotImplementedError("symmetry projectors are not supported by the engine")
        return walker_type_of(walkers)

    def _job(self, walkers, wave_data, ham_data=None):
        wt = self._check(walkers)
        return engine_for(self, wave_data, ham_data, wt, walkers.n_walkers)

    # -- measurements (wavefunctions.py:93-395)
    def calc_overlap(self, walkers, wave_data: dict) -> torch.Tensor:
        """< psi_T | walker > per walker."""
        up, dn = walkers.tensors()
        return self._job(walkers, wave_data).overlap(up, dn)

    def calc_green(self, walkers, wave_data: dict):
        """Half-rotated Green's functions G_s = (W_s (T_s^H W_s)^-1)^T, (W, nocc_s, N)."""
        up, dn = walkers.tensors()
        gu, gd = self._job(walkers, wave_data).green(up, dn)
        if isinstance(walkers, RHFWalkers) and self.kind == "rhf":
            return gu
        return gu, gd

    def calc_force_bias(self, walkers, ham_data: dict, wave_data: dict) -> torch.Tensor:
        """< psi_T | L_g | walker > / < psi_T | walker >, (W, G)."""
        up, dn = walkers.tensors()
        return self._job(walkers, wave_data, ham_data).force_bias(up, dn)

    def calc_energy(self, walkers, ham_data: dict, wave_data: dict) -> torch.Tensor:
        """Local energy < psi_T | H | walker > / < psi_T | walker >, (W,)."""
        up, dn = walkers.tensors()
        return self._job(walkers, wave_data, ham_data).energy(up, dn)

    # -- trial data (wavefunctions.py:396-466)
    def get_rdm1(self, wave_data: dict):
        if "rdm1" in wave_data:
            return wave_data["rdm1"]
        return self._calc_rdm1(wave_data)

    def _calc_rdm1(self, wave_data: dict):
        raise NotImplementedError

    def get_init_walkers(self, wave_data: dict, n_walkers: int, walker_type: str):
        """Natural orbitals of the trial 1-RDM, replicated (wavefunctions.py:427-466)."""
        rdm1 = E.as_f64(self.get_rdm1(wave_data))
        if walker_type == "restricted":
            nat = torch.linalg.eigh(rdm1[0] + rdm1[1])[1].flip(-1)[:, : self.nelec[0]]
            return RHFWalkers(nat.to(torch.complex128).unsqueeze(0).repeat(n_walkers, 1, 1))
        if walker_type == "unrestricted":
            up = torch.linalg.eigh(rdm1[0])[1].flip(-1)[:, : self.nelec[0]]
            dn = torch.linalg.eigh(rdm1[1])[1].flip(-1)[:, : self.nelec[1]]
            return UHFWalkers(
                [
                    up.to(torch.complex128).unsqueeze(0).repeat(n_walkers, 1, 1),
                    dn.to(torch.complex128).unsqueeze(0).repeat(n_walkers, 1, 1),
                ]
            )
        raise Exception("Unknown walker_type.")


@dataclass
class rhf(wave_function):
    """Restricted Hartree-Fock trial; wave_data["mo_coeff"] is (norb, nelec)."""

    n_opt_iter: int = 30

    def __post_init__(self):
        assert self.nelec[0] == self.nelec[1], "RHF requires equal number of up and down electrons."
        self.kind = "rhf"

    def _calc_rdm1(self, wave_data: dict):
        mo = E.as_f64(wave_data["mo_coeff"])
        d = mo @ mo.T
        return torch.stack([d, d])


@dataclass
class uhf(wave_function):
    """Unrestricted Hartree-Fock trial; wave_data["mo_coeff"] is [(norb, n_up), (norb, n_dn)]."""

    n_opt_iter: int = 30

    def __post_init__(self):
        self.kind = "uhf"

    def _calc_rdm1(self, wave_data: dict):
        up, dn = E.as_f64(wave_data["mo_coeff"][0]), E.as_f64(wave_data["mo_coeff"][1])
        return torch.stack([up @ up.T, dn @ dn.T])

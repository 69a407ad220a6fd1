"""Synthetic Hamiltonians / trials for the benchmark and parity workloads.

  h1   = diag(linspace(-2, 1, N)) + 0.05 sym(X),  X_ij ~ N(0,1)
  L_g  = sym(A_g),  (A_g)_ij ~ N(0, s_g^2),  s_g = 0.5 exp(-g / (0.2 G))
  RHF trial: lowest O eigenvectors of h1
  UHF trial: lowest O_a / O_b eigenvectors of h1 +/- 0.1 diag(xi), xi ~ N(0,1)
  h0 = 0
with sym(A) = (A + A^T)/2 and numpy.random.default_rng(seed).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np


def _sym(a):
    return (a + np.swapaxes(a, -1, -2)) / 2.0


@dataclass
class System:
    norb: int
    nelec: tuple
    nchol: int
    h0: float
    h1: np.ndarray  # (2, N, N)
    chol: np.ndarray  # (G, N*N)
    mo_up: np.ndarray
    mo_dn: Optional[np.ndarray]  # None for RHF

    @property
    def rhf(self) -> bool:
        return self.mo_dn is None

    def rdm1(self) -> np.ndarray:
        dn = self.mo_up if self.mo_dn is None else self.mo_dn
        return np.array([self.mo_up @ self.mo_up.T, dn @ dn.T])

    def wave_data(self) -> dict:
        mo = self.mo_up if self.rhf else [self.mo_up, self.mo_dn]
        return {"mo_coeff": mo, "rdm1": self.rdm1()}

    def ham_data(self) -> dict:
        return {"h0": self.h0, "h1": self.h1.copy(), "chol": self.chol, "ene0": 0.0}


def make_system(norb: int, nocc_up: int, nchol: int, nocc_dn: Optional[int] = None, uhf: bool = False,
                seed: int = 0, l_scale: float = 1.0) -> System:
    """l_scale multiplies every s_g.  l_scale = 1 is the generator as stated;
    its two-body scale grows like sqrt(N) (|sqrt(dt) sum_g x_g L_g| ~ 7 at
    N=200, G=1200, dt=0.005), so walkers are killed within a few steps.
    l_scale = 1/sqrt(N) gives O(1) Cholesky-vector norms, as in molecules."""
    rng = np.random.default_rng(seed)
    N, G = norb, nchol
    h1 = np.diag(np.linspace(-2.0, 1.0, N)) + 0.05 * _sym(rng.standard_normal((N, N)))
    sig = l_scale * 0.5 * np.exp(-np.arange(G) / (0.2 * G))
    chol = np.empty((G, N * N))
    for g in range(G):  # one gamma at a time keeps peak memory at the size of chol
        a = rng.standard_normal((N, N)) * sig[g]
        chol[g] = _sym(a).reshape(-1)
    if not uhf:
        mo = np.linalg.eigh(h1)[1][:, :nocc_up]
        return System(N, (nocc_up, nocc_up), G, 0.0, np.array([h1, h1]), chol, mo, None)
    nocc_dn = nocc_up if nocc_dn is None else nocc_dn
    xi = rng.standard_normal(N)
    up = np.linalg.eigh(h1 + 0.1 * np.diag(xi))[1][:, :nocc_up]
    dn = np.linalg.eigh(h1 - 0.1 * np.diag(xi))[1][:, :nocc_dn]
    return System(N, (nocc_up, nocc_dn), G, 0.0, np.array([h1, h1]), chol, up, dn)


def perturbed_walkers(mo: np.ndarray, n_walkers: int, scale: float = 0.1, seed: int = 1) -> np.ndarray:
    """Complex walkers near the trial orbitals (W, N, O): mo + scale*(X + iY)."""
    rng = np.random.default_rng(seed)
    N, O = mo.shape
    noise = rng.standard_normal((n_walkers, N, O)) + 1j * rng.standard_normal((n_walkers, N, O))
    return mo[None].astype(np.complex128) + scale * noise


# Benchmark / parity workloads (sizes of the published configurations)
CONFIGS = {
    "rhf_n200_o40_g1200_w512": dict(norb=200, nocc_up=40, nchol=1200, uhf=False, n_walkers=512, dt=0.005,
                                    n_prop_steps=10),
    "uhf_n300_o60_58_g2000_w256": dict(norb=300, nocc_up=60, nocc_dn=58, nchol=2000, uhf=True, n_walkers=256,
                                       dt=0.005, n_prop_steps=10),
    "rhf_n100_o20_g500_w256": dict(norb=100, nocc_up=20, nchol=500, uhf=False, n_walkers=256, dt=0.005,
                                   n_prop_steps=10),
}

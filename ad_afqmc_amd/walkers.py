"""Walker batches on the device (ad_afqmc/walkers.py:85-284).

RHFWalkers hold one (W, N, nelec_up) complex128 tensor; UHFWalkers hold
[(W, N, nelec_up), (W, N, nelec_dn)].  Re-orthonormalisation and comb
resampling run in libafq.so; the reference's chunked vmap (`n_chunks`) has no
counterpart because every kernel is batched over all walkers.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple, Union

import torch

from . import engine as E


def _zeta_tensor(zeta) -> torch.Tensor:
    if isinstance(zeta, torch.Tensor):
        return zeta.to(device=E.device(), dtype=torch.float64).reshape(1).contiguous()
    return torch.tensor([float(zeta)], dtype=torch.float64, device=E.device())


@dataclass
class RHFWalkers:
    """Restricted walkers (walkers.py:85-177)."""

    data: torch.Tensor

    def __post_init__(self):
        self.data = E.as_c128(self.data)

    @property
    def n_walkers(self) -> int:
        return self.data.shape[0]

    def orthonormalize(self) -> "RHFWalkers":
        """QR re-orthonormalisation, R discarded (walkers.py:155-157)."""
        d = self.data.clone()
        E.walkers_orthonormalize(d)
        return self.__class__(d)

    def stochastic_reconfiguration_local(self, weights, zeta) -> Tuple["RHFWalkers", torch.Tensor]:
        """Comb resampling (walkers.py:139-145, sr.py:24-34)."""
        d = self.data.clone()
        w = E.as_f64(weights).clone()
        E.walkers_sr(d, None, w, _zeta_tensor(zeta))
        return self.__class__(d), w

    def multiply_constants(self, constants) -> "RHFWalkers":
        c = E.as_c128(constants).reshape(-1, 1, 1)
        return self.__class__(self.data * c)

    def tensors(self):
        return self.data, None


@dataclass
class UHFWalkers:
    """Unrestricted walkers (walkers.py:180-284)."""

    data: List[torch.Tensor]

    def __post_init__(self):
        self.data = [E.as_c128(self.data[0]), E.as_c128(self.data[1])]

    @property
    def n_walkers(self) -> int:
        return self.data[0].shape[0]

    def orthonormalize(self) -> "UHFWalkers":
        """Per-spin QR re-orthonormalisation (walkers.py:260-262).  The
        reference multiplies the first column by sign(prod diag R); the
        engine's Gram-Schmidt R has a positive diagonal, so the overlap phase
        is preserved the same way."""
        up, dn = self.data[0].clone(), self.data[1].clone()
        E.walkers_orthonormalize(up)
        E.walkers_orthonormalize(dn)
        return self.__class__([up, dn])

    def stochastic_reconfiguration_local(self, weights, zeta) -> Tuple["UHFWalkers", torch.Tensor]:
        up, dn = self.data[0].clone(), self.data[1].clone()
        w = E.as_f64(weights).clone()
        E.walkers_sr(up, dn, w, _zeta_tensor(zeta))
        return self.__class__([up, dn]), w

    def multiply_constants(self, constants) -> "UHFWalkers":
        c = E.as_c128(constants)
        return self.__class__([self.data[0] * c[0].reshape(-1, 1, 1), self.data[1] * c[1].reshape(-1, 1, 1)])

    def tensors(self):
        return self.data[0], self.data[1]


walker_batch = Union[RHFWalkers, UHFWalkers]


def walker_type_of(walkers) -> str:
    if isinstance(walkers, RHFWalkers):
        return "restricted"
    if isinstance(walkers, UHFWalkers):
        return "unrestricted"
    raise NotImplementedError(f"Walker type {type(walkers)} not supported")

"""Molecular integrals for hydrogen chains without pyscf.

The reference prepares its Hamiltonian with pyscf (pyscf_interface.py:52-117:
RHF, integrals in the RHF orbital basis, modified Cholesky of the electron
repulsion integrals at `chol_cut`, FCIDUMP_chol).  For molecules made of
hydrogen atoms in minimal STO-nG bases every basis function is a contracted
s-type Gaussian, and all integrals have closed forms (Boys function F0):

  S_ab  = (pi/p)^{3/2} exp(-mu R_AB^2)                 p = a+b, mu = ab/p
  T_ab  = mu (3 - 2 mu R_AB^2) S_ab
  V_ab  = -2 pi/p Z exp(-mu R_AB^2) F0(p R_PC^2)
  (ab|cd) = 2 pi^{5/2} / (p q sqrt(p+q)) exp(-mu R_AB^2 - nu R_CD^2) F0(pq/(p+q) R_PQ^2)

(Szabo & Ostlund, Modern Quantum Chemistry, appendix A).  This module builds
them, runs closed-shell SCF (DIIS), decomposes the AO integrals by pivoted
(modified) Cholesky to `chol_cut` and rotates to the RHF orbitals -- the same
steps and conventions as the reference's preparation, so that its CPU-runnable
H10 case runs here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np
from scipy.special import erf

# Hydrogen 1s contractions, Slater exponent 1.24 (Hehre, Stewart, Pople 1969;
# the values distributed as the STO-3G / STO-6G basis sets).
STO_NG_H = {
    "sto-3g": (np.array([3.42525091, 0.62391373, 0.16885540]), np.array([0.15432897, 0.53532814, 0.44463454])),
    "sto-6g": (
        np.array([35.52322122, 6.513143725, 1.822142904, 0.625955266, 0.243076747, 0.100112428]),
        np.array([0.00916359628, 0.04936149294, 0.16853830490, 0.37056279970, 0.41649152980, 0.13033408410]),
    ),
}


def boys0(t):
    t = np.asarray(t, dtype=np.float64)
    out = np.ones_like(t)
    big = t > 1e-12
    st = np.sqrt(t[big])
    out[big] = 0.5 * math.sqrt(math.pi) * erf(st) / st
    small = ~big
    out[small] = 1.0 - t[small] / 3.0
    return out


@dataclass
class Basis:
    centers: np.ndarray  # (n, 3)
    alphas: np.ndarray  # (nprim,)
    coefs: np.ndarray  # (nprim,) normalised contraction coefficients (primitive norms included)


def hydrogen_basis(coords: np.ndarray, basis: str = "sto-6g") -> Basis:
    a, d = STO_NG_H[basis.lower()]
    c = d * (2.0 * a / math.pi) ** 0.75
    # normalise the contraction exactly
    p = a[:, None] + a[None, :]
    s = np.sum(c[:, None] * c[None, :] * (math.pi / p) ** 1.5)
    return Basis(np.asarray(coords, dtype=np.float64), a, c / math.sqrt(s))


def one_electron(b: Basis, charges: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Overlap, kinetic and nuclear attraction matrices (n x n)."""
    R = b.centers
    n = R.shape[0]
    a, c = b.alphas, b.coefs
    S = np.zeros((n, n))
    T = np.zeros((n, n))
    V = np.zeros((n, n))
    p = a[:, None] + a[None, :]
    mu = a[:, None] * a[None, :] / p
    cc = c[:, None] * c[None, :]
    for i in range(n):
        for j in range(n):
            r2 = float(np.sum((R[i] - R[j]) ** 2))
            s = (math.pi / p) ** 1.5 * np.exp(-mu * r2)
            S[i, j] = np.sum(cc * s)
            T[i, j] = np.sum(cc * mu * (3.0 - 2.0 * mu * r2) * s)
            P = (a[:, None, None] * R[i] + a[None, :, None] * R[j]) / p[:, :, None]
            v = 0.0
            for C, Z in zip(R, charges):
                rpc2 = np.sum((P - C) ** 2, axis=2)
                v = v - np.sum(cc * 2.0 * math.pi / p * Z * np.exp(-mu * r2) * boys0(p * rpc2))
            V[i, j] = v
    return S, T, V


def two_electron(b: Basis) -> np.ndarray:
    """(ij|kl) in chemists' notation, (n, n, n, n)."""
    R = b.centers
    n = R.shape[0]
    a, c = b.alphas, b.coefs
    npr = a.shape[0]
    # primitive pair data for every basis pair (i, j)
    p = a[:, None] + a[None, :]
    mu = a[:, None] * a[None, :] / p
    cc = c[:, None] * c[None, :]
    pairs = []
    for i in range(n):
        for j in range(n):
            r2 = float(np.sum((R[i] - R[j]) ** 2))
            K = cc * np.exp(-mu * r2)
            P = (a[:, None, None] * R[i] + a[None, :, None] * R[j]) / p[:, :, None]
            pairs.append((K.reshape(-1), p.reshape(-1), P.reshape(-1, 3)))
    eri = np.zeros((n * n, n * n))
    for x in range(n * n):
        K1, p1, P1 = pairs[x]
        for y in range(x, n * n):
            K2, p2, P2 = pairs[y]
            pq = p1[:, None] * p2[None, :]
            sp = p1[:, None] + p2[None, :]
            r2 = np.sum((P1[:, None, :] - P2[None, :, :]) ** 2, axis=2)
            val = np.sum(
                K1[:, None] * K2[None, :] * 2.0 * math.pi**2.5 / (pq * np.sqrt(sp)) * boys0(pq / sp * r2)
            )
            eri[x, y] = eri[y, x] = val
    del npr
    return eri.reshape(n, n, n, n)


def nuclear_repulsion(coords: np.ndarray, charges: np.ndarray) -> float:
    e = 0.0
    for i in range(len(charges)):
        for j in range(i):
            e += charges[i] * charges[j] / float(np.linalg.norm(coords[i] - coords[j]))
    return e


def rhf(S, H, eri, nocc: int, max_iter: int = 200, tol: float = 1e-10) -> Tuple[float, np.ndarray, np.ndarray]:
    """Closed-shell Roothaan SCF with DIIS; returns (electronic energy, C, orbital energies)."""
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s**-0.5) @ U.T
    e, Cp = np.linalg.eigh(X.T @ H @ X)
    C = X @ Cp
    focks: List[np.ndarray] = []
    errs: List[np.ndarray] = []
    E_old = 0.0
    for it in range(max_iter):
        D = C[:, :nocc] @ C[:, :nocc].T
        J = np.einsum("ijkl,kl->ij", eri, D)
        K = np.einsum("ikjl,kl->ij", eri, D)
        F = H + 2.0 * J - K
        E = float(np.sum(D * (H + F)))
        err = X.T @ (F @ D @ S - S @ D @ F) @ X
        if it > 0 and abs(E - E_old) < tol and np.max(np.abs(err)) < 1e-7:
            e, Cp = np.linalg.eigh(X.T @ F @ X)
            return E, X @ Cp, e
        E_old = E
        focks.append(F)
        errs.append(err)
        if len(focks) > 8:
            focks.pop(0)
            errs.pop(0)
        m = len(focks)
        if m > 1:
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for i in range(m):
                for j in range(m):
                    B[i, j] = np.sum(errs[i] * errs[j])
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            w = np.linalg.lstsq(B, rhs, rcond=None)[0][:m]
            F = sum(wi * Fi for wi, Fi in zip(w, focks))
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
    raise RuntimeError("SCF did not converge")


def modified_cholesky(M: np.ndarray, max_error: float) -> np.ndarray:
    """Pivoted Cholesky of a positive semidefinite (n x n) matrix: vectors L_g
    with M ~ sum_g L_g L_g^T, stopping when the largest residual diagonal is
    below max_error (pyscf_interface.py:287-391 does the same on the AO
    integrals)."""
    n = M.shape[0]
    diag = np.array(np.diagonal(M), dtype=np.float64)
    vecs = []
    while True:
        nu = int(np.argmax(diag))
        dmax = diag[nu]
        if dmax <= max_error or len(vecs) >= n:
            break
        col = M[:, nu].copy()
        for v in vecs:
            col -= v * v[nu]
        v = col / math.sqrt(dmax)
        vecs.append(v)
        diag = diag - v * v
    return np.array(vecs)


@dataclass
class MolecularHamiltonian:
    h0: float  # nuclear repulsion
    h1: np.ndarray  # (N, N) in the RHF orbital basis
    chol: np.ndarray  # (G, N*N) in the RHF orbital basis
    nelec: Tuple[int, int]
    e_rhf: float
    mo_coeff: np.ndarray  # AO -> MO

    @property
    def norb(self) -> int:
        return self.h1.shape[0]

    def trial_mo(self) -> np.ndarray:
        """RHF trial in its own orbital basis: the first nocc unit vectors."""
        return np.eye(self.norb)[:, : self.nelec[0]]

    def ham_data(self) -> dict:
        """The reference's ham_data dict (launch_script.py:213-243)."""
        return {"h0": self.h0, "h1": np.array([self.h1, self.h1]), "chol": self.chol.copy(), "ene0": 0.0}

    def wave_data(self) -> dict:
        mo = self.trial_mo()
        d = mo @ mo.T
        return {"mo_coeff": mo, "rdm1": np.array([d, d])}


def hydrogen_chain(n_atoms: int, spacing: float, basis: str = "sto-6g", chol_cut: float = 1e-5) -> MolecularHamiltonian:
    """Linear H_n chain along z (bohr), closed shell (n even)."""
    if n_atoms % 2:
        raise ValueError("closed-shell chains need an even number of atoms")
    coords = np.zeros((n_atoms, 3))
    coords[:, 2] = spacing * np.arange(n_atoms)
    Z = np.ones(n_atoms)
    b = hydrogen_basis(coords, basis)
    S, T, V = one_electron(b, Z)
    eri = two_electron(b)
    H = T + V
    nocc = n_atoms // 2
    e_el, C, _ = rhf(S, H, eri, nocc)
    enuc = nuclear_repulsion(coords, Z)
    n = S.shape[0]
    L_ao = modified_cholesky(eri.reshape(n * n, n * n), chol_cut).reshape(-1, n, n)
    L_mo = np.einsum("pi,gpq,qj->gij", C, L_ao, C, optimize=True)
    h1 = C.T @ H @ C
    return MolecularHamiltonian(enuc, h1, L_mo.reshape(-1, n * n), (nocc, nocc), e_el + enuc, C)

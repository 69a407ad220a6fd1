"""Brute-force second quantisation on a few spatial orbitals (test helper).

Spin orbitals 0..N-1 are alpha, N..2N-1 beta; Fock states are bitmasks.
H = h0 + sum_{pq,s} h_pq a+_ps a_qs
      + 1/2 sum_g sum_{pqrs,st} L_g,pq L_g,rs a+_ps a+_rt a_st a_qs
built as 1/2 sum_g (V_g^2 - sum_s (L_g L_g)_ps a+_ps a_ss) with V_g = sum L_g,pq E_pq.
"""
import itertools

import numpy as np


def _ops(nso):
    dim = 1 << nso
    ann = []
    for j in range(nso):
        m = np.zeros((dim, dim))
        for s in range(dim):
            if s >> j & 1:
                sign = (-1) ** bin(s & ((1 << j) - 1)).count("1")
                m[s ^ (1 << j), s] = sign
        ann.append(m)
    return ann


def hamiltonian_matrix(h0, h1, L):
    N = h1.shape[0]
    a = _ops(2 * N)
    cre = [x.T for x in a]
    E = np.zeros((N, N, 1 << 2 * N, 1 << 2 * N))
    for p, q in itertools.product(range(N), range(N)):
        E[p, q] = cre[p] @ a[q] + cre[N + p] @ a[N + q]
    H = h0 * np.eye(1 << 2 * N) + np.einsum("pq,pqij->ij", h1, E)
    for g in range(L.shape[0]):
        V = np.einsum("pq,pqij->ij", L[g], E)
        LL = L[g] @ L[g]
        H = H + 0.5 * (V @ V - np.einsum("pq,pqij->ij", LL, E))
    return H, cre


def determinant(C_up, C_dn, cre):
    N = C_up.shape[0]
    dim = cre[0].shape[0]
    v = np.zeros(dim, dtype=np.complex128)
    v[0] = 1.0
    for i in range(C_up.shape[1]):
        v = sum(C_up[p, i] * (cre[p] @ v) for p in range(N))
    for i in range(C_dn.shape[1]):
        v = sum(C_dn[p, i] * (cre[N + p] @ v) for p in range(N))
    return v


def local_energy(h0, h1, L, T_up, T_dn, W_up, W_dn):
    H, cre = hamiltonian_matrix(h0, h1, L)
    bra = determinant(T_up, T_dn, cre).conj()
    ket = determinant(W_up, W_dn, cre)
    return (bra @ H @ ket) / (bra @ ket), bra @ ket

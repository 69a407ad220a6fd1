"""Input formats and molecular setup (CPU): the FCIDUMP_chol reader against the
reference's own fixture, and the hydrogen integrals/SCF/Cholesky against
textbook values and exact identities."""
import math
import os

import numpy as np
import pytest

from ad_afqmc_amd import fcidump
from ad_afqmc_amd import integrals as I

HERE = os.path.dirname(os.path.abspath(__file__))


def test_hdf5_reader_matches_h5py_conversion():
    """tests/golden/h6_FCIDUMP_chol is the reference's driver-test file
    (tests/driver_test_files/FCIDUMP_chol); h6_fcidump_chol.npz was written from
    it with h5py."""
    h = fcidump.load(os.path.join(HERE, "golden", "h6_FCIDUMP_chol"))
    z = np.load(os.path.join(HERE, "golden", "h6_fcidump_chol.npz"))
    nelec, norb, ms, nchol = (int(x) for x in z["header"])
    assert h.norb == norb == 6 and h.nelec_sp == ((nelec + abs(ms)) // 2, (nelec - abs(ms)) // 2)
    assert h.h0 == float(z["energy_core"])
    assert np.array_equal(h.h1.reshape(-1), z["hcore"])
    assert np.array_equal(h.chol.reshape(-1), z["chol"]) and h.chol.shape == (nchol, norb, norb)
    assert np.array_equal(h.mo_coeffs[0], z["mo_coeff"][0]) and np.array_equal(h.mo_coeffs[1], z["mo_coeff"][1])


def test_npz_round_trip(tmp_path):
    h = fcidump.load(os.path.join(HERE, "golden", "h6_FCIDUMP_chol"))
    p = str(tmp_path / "ham.npz")
    fcidump.save_npz(p, h.h0, h.h1, h.chol, h.nelec_sp)
    g = fcidump.load(p)
    assert g.h0 == h.h0 and g.nelec_sp == h.nelec_sp
    assert np.array_equal(g.h1, h.h1) and np.array_equal(g.chol, h.chol)
    # hcore_mod = h1 - v0 as the reference writes it (pyscf_interface.py:99-100)
    assert np.allclose(g.h1_mod, h.h1_mod, atol=1e-12)


def test_hdf5_reader_rejects_other_files(tmp_path):
    p = tmp_path / "x.h5"
    p.write_bytes(b"not hdf5 at all")
    with pytest.raises(fcidump.HDF5FormatError):
        fcidump.read_hdf5_datasets(str(p))


def test_h2_sto3g_matches_szabo_ostlund():
    """H2, R = 1.4 bohr, STO-3G (Szabo & Ostlund section 3.5.2)."""
    b = I.hydrogen_basis(np.array([[0.0, 0, 0], [0, 0, 1.4]]), "sto-3g")
    S, T, V = I.one_electron(b, np.ones(2))
    eri = I.two_electron(b)
    H = T + V
    assert abs(S[0, 1] - 0.6593) < 1e-4 and abs(T[0, 0] - 0.7600) < 1e-4 and abs(T[0, 1] - 0.2365) < 1e-4
    assert abs(H[0, 0] + 1.1204) < 1e-4 and abs(H[0, 1] + 0.9584) < 1e-4
    assert abs(eri[0, 0, 0, 0] - 0.7746) < 1e-4 and abs(eri[0, 0, 1, 1] - 0.5697) < 1e-4
    assert abs(eri[1, 0, 0, 0] - 0.4441) < 1e-4 and abs(eri[1, 0, 1, 0] - 0.2970) < 1e-4
    # 8-fold permutational symmetry
    assert np.allclose(eri, eri.transpose(1, 0, 2, 3)) and np.allclose(eri, eri.transpose(2, 3, 0, 1))
    E, C, e = I.rhf(S, H, eri, 1)
    assert abs(E + 1.8310) < 1e-4 and abs(E + 1 / 1.4 + 1.1167) < 1e-4
    assert abs(e[0] + 0.578) < 1e-3 and abs(e[1] - 0.670) < 1e-3


@pytest.mark.parametrize("basis,energy", [("sto-3g", -0.46658), ("sto-6g", -0.47104)])
def test_hydrogen_atom_energy(basis, energy):
    """<1s|T+V|1s> of the contraction; a Slater 1s with exponent 1.24 gives
    1.24^2/2 - 1.24 = -0.47120, which STO-nG approaches from above."""
    b = I.hydrogen_basis(np.zeros((1, 3)), basis)
    S, T, V = I.one_electron(b, np.ones(1))
    assert abs(S[0, 0] - 1.0) < 1e-12
    e = T[0, 0] + V[0, 0]
    assert abs(e - energy) < 2e-5 and e > 1.24**2 / 2 - 1.24


def test_boys_function_limits():
    t = np.array([0.0, 1e-14, 1e-3, 1.0, 30.0])
    f = I.boys0(t)
    assert f[0] == 1.0 and abs(f[1] - 1.0) < 1e-13
    assert abs(f[2] - (1 - 1e-3 / 3 + 1e-6 / 10 - 1e-9 / 42)) < 1e-13  # series of F0
    assert abs(f[3] - 0.7468241328124271) < 1e-15  # 0.5 sqrt(pi) erf(1)
    assert abs(f[4] - 0.5 * math.sqrt(math.pi / 30.0)) < 1e-12


def test_modified_cholesky_reconstructs_to_threshold():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((30, 12))
    M = A @ A.T
    L = I.modified_cholesky(M, 1e-8)
    assert L.shape[0] <= 12
    assert np.max(np.abs(L.T @ L - M)) < 1e-7


def test_h10_chain_hamiltonian():
    h = I.hydrogen_chain(10, 1.6, "sto-6g", 1e-5)
    N, O = 10, 5
    assert h.norb == N and h.nelec == (5, 5) and abs(h.h0 - sum(1.0 / (1.6 * abs(i - j)) for i in range(10) for j in range(i))) < 1e-12
    L = h.chol.reshape(-1, N, N)
    # the Cholesky Hamiltonian in the RHF orbital basis reproduces the SCF energy at the threshold
    c = np.einsum("gii->g", L[:, :O, :O])
    e = h.h0 + 2 * np.trace(h.h1[:O, :O]) + 2 * np.sum(c * c) - np.einsum("gij,gji->", L[:, :O, :O], L[:, :O, :O])
    assert abs(e - h.e_rhf) < 1e-5
    # Brillouin: the Fock matrix of the Cholesky Hamiltonian has no occupied-virtual block
    J = np.einsum("gpq,g->pq", L, 2 * c)
    K = np.einsum("gpi,giq->pq", L[:, :, :O], L[:, :O, :])
    F = h.h1 + J - K
    assert np.max(np.abs(F[:O, O:])) < 1e-5
    # vectors are symmetric and reconstruct the MO integrals to ~ the threshold
    assert np.allclose(L, L.transpose(0, 2, 1), atol=1e-12)
    assert h.e_rhf < -5.2 and h.e_rhf > -5.3

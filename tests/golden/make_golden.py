"""Writes tests/golden/golden_wavefunctions.npz.

The inputs are replayed from NumPy's legacy MT19937 stream exactly in the order
the reference test module draws them (tests/test_wavefunctions.py:12-29 for the
RHF inputs, :56-73 for the UHF inputs; no RNG is consumed in between, the
determinant file read at :42-44 is a parser).  The expected scalars are the
reference's asserted values, tests/test_wavefunctions.py:128-184.

Nothing from the reference is imported or executed; only its literal test
constants are recorded here.

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np


def replay():
    np.random.seed(102)  # test_wavefunctions.py:12-13
    norb, nelec, nchol = 10, (5, 5), 5
    mo = np.eye(norb)[:, : nelec[0]]  # :17
    walker = np.random.rand(norb, nelec[0]) + 1.0j * np.random.rand(norb, nelec[0])  # :19-21
    h0 = np.random.rand(1)[0]  # :24-26
    rot_h1 = np.random.rand(nelec[0], norb)  # :26
    rot_chol = np.random.rand(nchol, nelec[0], norb)  # :27
    h1 = np.random.rand(norb, norb)  # :28 (replicated for both spins)
    chol = np.random.rand(nchol, norb, norb)  # :29
    nelec_sp = (3, 2)  # :56
    mo_u_up = np.random.rand(norb, nelec_sp[0])  # :60
    mo_u_dn = np.random.rand(norb, nelec_sp[1])  # :61
    walker_up = np.random.rand(norb, nelec_sp[0]) + 1.0j * np.random.rand(norb, nelec_sp[0])  # :69-71
    walker_dn = np.random.rand(norb, nelec_sp[1]) + 1.0j * np.random.rand(norb, nelec_sp[1])  # :71-73
    return dict(
        norb=norb,
        nchol=nchol,
        mo=mo,
        walker=walker,
        h0=h0,
        rot_h1=rot_h1,
        rot_chol=rot_chol,
        h1=h1,
        chol=chol,
        mo_u_up=mo_u_up,
        mo_u_dn=mo_u_dn,
        walker_up=walker_up,
        walker_dn=walker_dn,
        # reference asserted values (real parts), tests/test_wavefunctions.py:128-184
        rhf_overlap=-0.10794844182417201,
        rhf_green_sum=12.181348093111438,
        rhf_force_bias_sum=66.13455680423321,
        rhf_energy=217.79874063608622,
        uhf_overlap=-0.4029825074695857,
        uhf_green_sum=3.6324117896217394,
        uhf_force_bias_sum=10.441272099672341,
        uhf_energy=-1.7203463308366032,
    )


if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_wavefunctions.npz")
    np.savez(out, **replay())
    print("wrote", out)

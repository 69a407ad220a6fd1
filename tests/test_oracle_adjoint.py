"""CPU pins of the hand-written reverse pass of a phaseless segment
(oracle/afqmc_adjoint.py) against central finite differences of the forward
segment energy, which itself is the oracle's block loop."""
import math

import numpy as np
import pytest

from oracle import afqmc_adjoint as A
from oracle import afqmc_oracle as O


def small_system(N=6, Oc=2, G=8, W=5, seed=3, lsc=0.25):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((N, N))
    h1 = np.diag(np.linspace(-2.0, 1.0, N)) + 0.05 * (x + x.T) / 2
    sig = lsc * 0.5 * np.exp(-np.arange(G) / (0.2 * G))
    a = rng.standard_normal((G, N, N)) * sig[:, None, None]
    L = (a + a.transpose(0, 2, 1)) / 2
    mo = np.linalg.eigh(h1)[1][:, :Oc]
    walkers = mo[None] + 0.1 * (rng.standard_normal((W, N, Oc)) + 1j * rng.standard_normal((W, N, Oc)))
    weights = rng.uniform(0.5, 1.5, W)
    h1_2 = np.array([h1, h1])
    return h1_2, L.reshape(G, N * N), mo, walkers, weights


def _case(n_blocks=2, n_steps=3, dt=0.01, seed=3):
    h1, chol, mo, walkers, weights = small_system(seed=seed)
    rng = np.random.default_rng(seed + 100)
    G = chol.shape[0]
    fields = rng.standard_normal((n_blocks, n_steps, walkers.shape[0], G))
    t = O.RHFTrial(mo.shape[0], (mo.shape[1], mo.shape[1]), mo)
    _, rot_h1, rot_chol = O.rhf_measurement_intermediates(h1, chol, mo)
    e_est = float(np.mean(np.real(t.calc_energy(walkers, 0.3, rot_h1, rot_chol))))
    return dict(h0=0.3, h1=h1, chol=chol, mo=mo, walkers=walkers, weights=weights, e_estimate=e_est, fields=fields,
                dt=dt)


def _F(c, h1=None, chol=None):
    return A.segment_energy(c["h0"], c["h1"] if h1 is None else h1, c["chol"] if chol is None else chol, c["mo"],
                            c["walkers"], c["weights"], c["e_estimate"], c["fields"], c["dt"])


def test_vjp_forward_value_matches_block_loop():
    c = _case()
    F, _, _ = A.segment_energy_vjp(c["h0"], c["h1"], c["chol"], c["mo"], c["walkers"], c["weights"],
                                   c["e_estimate"], c["fields"], c["dt"])
    assert abs(F - _F(c)) < 1e-12 * abs(F)


@pytest.mark.parametrize("seed,dt,n_blocks,n_steps", [(3, 0.01, 2, 3), (7, 0.01, 2, 3), (3, 0.05, 3, 4)])
def test_vjp_matches_central_differences(seed, dt, n_blocks, n_steps):
    # agreement is ~1e-10 absolute; dropping the population-control-shift
    # feedback alone moves the directional derivative by ~5e-6 (dt=0.01)
    c = _case(seed=seed, dt=dt, n_blocks=n_blocks, n_steps=n_steps)
    F, gh1, gL = A.segment_energy_vjp(c["h0"], c["h1"], c["chol"], c["mo"], c["walkers"], c["weights"],
                                      c["e_estimate"], c["fields"], c["dt"])
    rng = np.random.default_rng(seed + 5)
    h = 1e-5
    # h1: a general (non-symmetric, spin-dependent) direction
    D = rng.standard_normal(c["h1"].shape)
    fd = (_F(c, h1=c["h1"] + h * D) - _F(c, h1=c["h1"] - h * D)) / (2 * h)
    an = float(np.sum(gh1 * D))
    assert abs(fd - an) <= 1e-7 * max(1.0, abs(an)), (fd, an)
    # chol: a symmetric direction per vector, and a general one
    N = c["mo"].shape[0]
    for symmetric in (True, False):
        E = rng.standard_normal((c["chol"].shape[0], N, N))
        if symmetric:
            E = (E + E.transpose(0, 2, 1)) / 2
        E = E.reshape(c["chol"].shape)
        fd = (_F(c, chol=c["chol"] + h * E) - _F(c, chol=c["chol"] - h * E)) / (2 * h)
        an = float(np.sum(gL * E))
        assert abs(fd - an) <= 1e-7 * max(1.0, abs(an)), (symmetric, fd, an)


def test_vjp_h1_gradient_is_symmetric_and_spin_equal():
    c = _case()
    _, gh1, _ = A.segment_energy_vjp(c["h0"], c["h1"], c["chol"], c["mo"], c["walkers"], c["weights"],
                                     c["e_estimate"], c["fields"], c["dt"])
    assert np.allclose(gh1[0], gh1[0].T, atol=1e-14) and np.allclose(gh1[0], gh1[1], atol=1e-14)


def test_zero_fields_trial_walkers_gradient_is_trial_rdm():
    """With W = T, zero fields and dt -> 0 the segment energy is the trial's
    variational energy, whose h1 gradient is the RHF 1-RDM per spin: T T^T
    (the quantity the reference's reverse AD mode returns, driver.py:865-876)."""
    c = _case(n_blocks=1, n_steps=1, dt=1e-7)
    W = c["walkers"].shape[0]
    c["walkers"] = np.array([c["mo"] + 0j] * W)
    c["fields"] = np.zeros_like(c["fields"])
    _, gh1, _ = A.segment_energy_vjp(c["h0"], c["h1"], c["chol"], c["mo"], c["walkers"], c["weights"],
                                     c["e_estimate"], c["fields"], c["dt"])
    d = c["mo"] @ c["mo"].T
    assert np.allclose(gh1[0], d, atol=1e-5) and np.allclose(gh1[1], d, atol=1e-5)

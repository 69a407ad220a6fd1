"""CPU checks that pin the FP64 oracle independently of its own formulas:
the reference's golden scalars, brute-force second quantisation, invariants
and closed forms.  Each check is chosen so that a dropped term, a wrong sign
or index, or a transposed operand fails it."""
import math
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import afqmc_oracle as O
from tests import fock

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "golden_wavefunctions.npz")


def _golden():
    return np.load(GOLD)


# --------------------------------------------------------------- golden scalars
def test_golden_rhf():
    d = _golden()
    t = O.RHFTrial(10, (5, 5), d["mo"])
    w = d["walker"][None]
    assert np.isclose(t.calc_overlap(w).real[0], d["rhf_overlap"], rtol=1e-12, atol=0)
    assert np.isclose(t.calc_green(w)[0].sum().real, d["rhf_green_sum"], rtol=1e-12, atol=0)
    assert np.isclose(t.calc_force_bias(w, d["rot_chol"]).sum().real, d["rhf_force_bias_sum"], rtol=1e-12, atol=0)
    assert np.isclose(t.calc_energy(w, d["h0"], d["rot_h1"], d["rot_chol"]).real[0], d["rhf_energy"], rtol=1e-12,
                      atol=0)


def test_golden_uhf():
    d = _golden()
    u = O.UHFTrial(10, (3, 2), d["mo_u_up"], d["mo_u_dn"])
    w = [d["walker_up"][None], d["walker_dn"][None]]
    rh = [d["rot_h1"][:3], d["rot_h1"][:2]]
    rc = [d["rot_chol"][:, :3], d["rot_chol"][:, :2]]
    assert np.isclose(u.calc_overlap(w).real[0], d["uhf_overlap"], rtol=1e-12, atol=0)
    g = u.calc_green(w)
    assert np.isclose(g[0].sum().real + g[1].sum().real, d["uhf_green_sum"], rtol=1e-12, atol=0)
    assert np.isclose(u.calc_force_bias(w, rc).sum().real, d["uhf_force_bias_sum"], rtol=1e-12, atol=0)
    assert np.isclose(u.calc_energy(w, d["h0"], rh, rc).real[0], d["uhf_energy"], rtol=1e-12, atol=0)


# --------------------------------------------------------------- brute force
def _small(N=3, G=4, seed=3):
    rng = np.random.default_rng(seed)
    h1 = rng.standard_normal((N, N))
    h1 = (h1 + h1.T) / 2
    L = rng.standard_normal((G, N, N)) * 0.5
    L = (L + L.transpose(0, 2, 1)) / 2
    return rng, h1, L


def test_energy_bruteforce_rhf():
    N, O_ = 3, 2
    rng, h1, L = _small(N)
    T = np.linalg.qr(rng.standard_normal((N, O_)))[0]
    W = rng.standard_normal((N, O_)) + 1j * rng.standard_normal((N, O_))
    h0 = 0.7
    _, rot_h1, rot_chol = O.rhf_measurement_intermediates(np.array([h1, h1]), L.reshape(G := L.shape[0], -1), T)
    t = O.RHFTrial(N, (O_, O_), T)
    e = t.calc_energy(W[None], h0, rot_h1, rot_chol)[0]
    ov = t.calc_overlap(W[None])[0]
    e_ref, ov_ref = fock.local_energy(h0, h1, L, T, T, W, W)
    assert abs(e - e_ref) < 1e-11 * max(1, abs(e_ref))
    assert abs(ov - ov_ref) < 1e-12 * max(1, abs(ov_ref))


def test_energy_bruteforce_uhf():
    N = 3
    rng, h1, L = _small(N, seed=5)
    Tu = rng.standard_normal((N, 2))
    Td = rng.standard_normal((N, 1))
    Wu = rng.standard_normal((N, 2)) + 1j * rng.standard_normal((N, 2))
    Wd = rng.standard_normal((N, 1)) + 1j * rng.standard_normal((N, 1))
    h0 = -0.3
    _, rot_h1, rot_chol = O.uhf_measurement_intermediates(np.array([h1, h1]), L.reshape(L.shape[0], -1), Tu, Td)
    u = O.UHFTrial(N, (2, 1), Tu, Td)
    e = u.calc_energy([Wu[None], Wd[None]], h0, rot_h1, rot_chol)[0]
    e_ref, _ = fock.local_energy(h0, h1, L, Tu, Td, Wu, Wd)
    assert abs(e - e_ref) < 1e-11 * max(1, abs(e_ref))


def test_force_bias_bruteforce():
    # fb_g = <T| V_g |W>/<T|W> with V_g = sum_pq L_g,pq E_pq
    N = 3
    rng, h1, L = _small(N, seed=7)
    T = np.linalg.qr(rng.standard_normal((N, 2)))[0]
    W = rng.standard_normal((N, 2)) + 1j * rng.standard_normal((N, 2))
    _, _, rot_chol = O.rhf_measurement_intermediates(np.array([h1, h1]), L.reshape(L.shape[0], -1), T)
    fb = O.RHFTrial(N, (2, 2), T).calc_force_bias(W[None], rot_chol)[0]
    for g in range(L.shape[0]):
        e_ref, _ = fock.local_energy(0.0, L[g], np.zeros((0, N, N)), T, T, W, W)
        assert abs(fb[g] - e_ref) < 1e-12 * max(1, abs(e_ref))


# --------------------------------------------------------------- invariants
def test_right_multiplication_invariance():
    rng = np.random.default_rng(11)
    N, Oo, G = 6, 3, 5
    T = np.linalg.qr(rng.standard_normal((N, Oo)))[0]
    L = rng.standard_normal((G, N, N))
    L = (L + L.transpose(0, 2, 1)) / 2
    h1 = rng.standard_normal((N, N))
    _, rh, rc = O.rhf_measurement_intermediates(np.array([h1, h1]), L.reshape(G, -1), T)
    t = O.RHFTrial(N, (Oo, Oo), T)
    W = rng.standard_normal((2, N, Oo)) + 1j * rng.standard_normal((2, N, Oo))
    R = rng.standard_normal((Oo, Oo)) + 1j * rng.standard_normal((Oo, Oo))
    WR = W @ R
    assert np.allclose(t.calc_overlap(WR), t.calc_overlap(W) * np.linalg.det(R) ** 2, rtol=1e-12)
    assert np.allclose(t.calc_force_bias(WR, rc), t.calc_force_bias(W, rc), rtol=1e-11)
    assert np.allclose(t.calc_energy(WR, 0.1, rh, rc), t.calc_energy(W, 0.1, rh, rc), rtol=1e-11)


def test_fb_equals_minus_i_mf_at_trial():
    rng = np.random.default_rng(2)
    N, Oo, G = 7, 3, 6
    T = np.linalg.qr(rng.standard_normal((N, Oo)))[0]
    L = rng.standard_normal((G, N, N))
    L = (L + L.transpose(0, 2, 1)) / 2
    h1 = rng.standard_normal((N, N))
    h1s, rh, rc = O.rhf_measurement_intermediates(np.array([h1, h1]), L.reshape(G, -1), T)
    rdm1 = O.rdm1_from_mo(T)
    pi = O.propagation_intermediates(0.0, h1s, L.reshape(G, -1), rdm1, 0.01)
    fb = O.RHFTrial(N, (Oo, Oo), T).calc_force_bias(T[None] + 0j, rc)[0]
    assert np.allclose(fb, -1j * pi["mf_shifts"], atol=1e-12)
    shifts = -math.sqrt(0.01) * (1j * fb - pi["mf_shifts"])
    assert np.allclose(shifts, 0, atol=1e-12)


def test_energy_at_trial_closed_form():
    rng = np.random.default_rng(4)
    N, Oo, G = 6, 2, 4
    T = np.linalg.qr(rng.standard_normal((N, N)))[0]
    occ = T[:, :Oo]
    L = rng.standard_normal((G, N, N))
    L = (L + L.transpose(0, 2, 1)) / 2
    h1 = rng.standard_normal((N, N))
    h1 = (h1 + h1.T) / 2
    _, rh, rc = O.rhf_measurement_intermediates(np.array([h1, h1]), L.reshape(G, -1), occ)
    e = O.RHFTrial(N, (Oo, Oo), occ).calc_energy(occ[None] + 0j, 0.2, rh, rc)[0]
    hm = occ.T @ h1 @ occ
    Lm = np.einsum("pi,gpq,qj->gij", occ, L, occ)
    e_cf = 0.2 + 2 * np.trace(hm) + 2 * np.sum(np.einsum("gii->g", Lm) ** 2) - np.einsum("gij,gji->", Lm, Lm)
    assert abs(e - e_cf) < 1e-12


def test_propagation_intermediates_closed_forms():
    rng = np.random.default_rng(8)
    N, G = 5, 4
    L = rng.standard_normal((G, N, N))
    L = (L + L.transpose(0, 2, 1)) / 2
    h1 = rng.standard_normal((N, N))
    h1 = (h1 + h1.T) / 2
    T = np.linalg.qr(rng.standard_normal((N, 2)))[0]
    rdm1 = O.rdm1_from_mo(T)
    dt = 0.05
    pi = O.propagation_intermediates(0.4, np.array([h1, h1]), L.reshape(G, -1), rdm1, dt)
    v0 = 0.5 * sum(L[g] @ L[g] for g in range(G))
    m = np.array([np.sum(L[g] * (rdm1[0] + rdm1[1])) for g in range(G)])
    v1 = -sum(m[g] * L[g] for g in range(G))
    assert np.allclose(pi["v0"], v0, atol=1e-13)
    assert np.allclose(pi["v1"], v1, atol=1e-13)
    assert np.isclose(pi["h0_prop"], -0.4 + 0.5 * np.sum(m**2), atol=1e-13)
    hm = h1 - v0 - v1
    ev, U = np.linalg.eigh(hm)
    assert np.allclose(pi["exp_h1"], U @ np.diag(np.exp(-dt * ev / 2)) @ U.T, atol=1e-13)


def test_taylor_within_remainder_bound():
    rng = np.random.default_rng(9)
    N, Oo, G = 6, 2, 5
    L = rng.standard_normal((G, N * N)) * 0.3
    x = rng.standard_normal((1, G)) + 0.2j * rng.standard_normal((1, G))
    dt = 0.01
    vhs = O.vhs_from_fields(x, L, dt)[0]
    W = rng.standard_normal((N, Oo)) + 0j
    I = np.eye(N)
    out = O.apply_exp_vhs(W, vhs, I, 6)
    exact = scipy.linalg.expm(vhs) @ W
    nrm = np.linalg.norm(vhs, 2)
    bound = nrm**6 / math.factorial(6) * math.exp(nrm) * np.linalg.norm(W, 2)
    err6 = np.linalg.norm(out - exact, 2)
    assert err6 <= bound * 1.01 + 1e-15
    # the error is the first omitted term, VHS^6 W / 6!, to leading order
    lead6 = np.linalg.norm(np.linalg.matrix_power(vhs, 6) @ W, 2) / math.factorial(6)
    assert 0.5 * lead6 < err6 < 1.5 * lead6
    # a dropped term (5 instead of 6) leaves VHS^5 W / 5! instead
    out5 = O.apply_exp_vhs(W, vhs, I, 5)
    lead5 = np.linalg.norm(np.linalg.matrix_power(vhs, 5) @ W, 2) / math.factorial(5)
    err5 = np.linalg.norm(out5 - exact, 2)
    assert 0.5 * lead5 < err5 < 1.5 * lead5 and err5 > 3 * err6


def test_propagate_dt_to_zero_keeps_weights():
    rng = np.random.default_rng(10)
    N, Oo, G, Wn = 6, 2, 5, 4
    L = rng.standard_normal((G, N, N)) * 0.3
    L = (L + L.transpose(0, 2, 1)) / 2
    h1 = rng.standard_normal((N, N))
    T = np.linalg.qr(rng.standard_normal((N, Oo)))[0]
    h1s, rh, rc = O.rhf_measurement_intermediates(np.array([h1, h1]), L.reshape(G, -1), T)
    dt = 1e-12
    pi = O.propagation_intermediates(0.0, h1s, L.reshape(G, -1), O.rdm1_from_mo(T), dt)
    trial = O.RHFTrial(N, (Oo, Oo), T)
    W = O.init_walkers(O.rdm1_from_mo(T), (Oo, Oo), Wn, "restricted")
    W = W + 0.1 * rng.standard_normal(W.shape)
    state = dict(walkers=W, weights=np.ones(Wn), overlaps=trial.calc_overlap(W), e_estimate=0.0,
                 pop_control_ene_shift=0.0)
    ham = dict(chol=L.reshape(G, -1), rot_chol=rc, mf_shifts=pi["mf_shifts"], h0_prop=pi["h0_prop"],
               exp_h1=pi["exp_h1"])
    out = O.propagate_constrained(trial, state, rng.standard_normal((Wn, G)), ham, dt)
    assert np.allclose(out["weights"], 1.0, atol=1e-5)
    assert np.allclose(out["walkers"], W, atol=1e-5)


def test_qr_invariants():
    rng = np.random.default_rng(12)
    W = rng.standard_normal((3, 8, 3)) + 1j * rng.standard_normal((3, 8, 3))
    T = np.linalg.qr(rng.standard_normal((8, 3)))[0]
    q, _ = O.qr_restricted(W)
    assert np.allclose(np.einsum("wpi,wpj->wij", q.conj(), q), np.eye(3), atol=1e-13)
    # span preserved -> Green's function unchanged
    assert np.allclose(O.green(q, T), O.green(W, T), atol=1e-11)
    qu, _ = O.qr_unrestricted([W, W[:, :, :2]])
    # UHF sign fix keeps the overlap phase: det(T^H Q) / det(T^H W) is real positive
    r = O.det_overlap(qu[0], T) / O.det_overlap(W, T)
    assert np.allclose(r.imag, 0, atol=1e-12) and np.all(r.real > 0)


def test_sr_known_answers():
    idx, w = O.sr_indices(np.ones(6), 0.3)
    assert list(idx) == list(range(6)) and np.allclose(w, 1.0)
    weights = np.array([0.0, 3.0, 1.0, 0.0, 2.0])
    idx, w = O.sr_indices(weights, 0.5)
    # brute-force comb: tooth k at total*(k+zeta)/W, pick first cumulative >= tooth
    cum = np.cumsum(weights)
    teeth = cum[-1] * (np.arange(5) + 0.5) / 5
    brute = [next(i for i, c in enumerate(cum) if c >= t) for t in teeth]
    assert list(idx) == brute and np.allclose(w, 6.0 / 5)
    assert 0 not in idx and 3 not in idx


def test_blocking_and_outliers():
    rng = np.random.default_rng(13)
    n = 4000
    e = rng.standard_normal(n) * 0.5 + 1.0
    w = np.ones(n)
    mean, err = O.blocking_analysis(w, e)
    assert abs(mean - 1.0) < 5 * 0.5 / math.sqrt(n)
    assert 0.5 * 0.5 / math.sqrt(n) < err < 2.5 * 0.5 / math.sqrt(n)
    data = np.stack([w[:100], e[:100]]).T
    data[17, 1] = 1e3
    clean, mask = O.reject_outliers(data, 1)
    assert clean.shape[0] == 99 and not mask[17]


@pytest.mark.parametrize("walker_type", ["restricted", "unrestricted"])
def test_init_walkers_span_the_trial_space(walker_type):
    """get_init_walkers (wavefunctions.py:427-466): the natural orbitals of the
    trial 1-RDM span the trial's occupied space, so the walker projector equals
    the trial projector T T^T per spin and every walker overlaps the trial."""
    rng = np.random.default_rng(21)
    N, nup, ndn, Wn = 9, 4, 3, 3
    Tu = np.linalg.qr(rng.standard_normal((N, nup)))[0]
    Td = Tu[:, :ndn] if walker_type == "restricted" else np.linalg.qr(rng.standard_normal((N, ndn)))[0]
    rdm1 = np.array([Tu @ Tu.T, Td @ Td.T])
    nelec = (nup, nup) if walker_type == "restricted" else (nup, ndn)
    if walker_type == "restricted":
        rdm1 = np.array([Tu @ Tu.T, Tu @ Tu.T])
        spins = [(O.init_walkers(rdm1, nelec, Wn, walker_type), Tu)]
    else:
        up, dn = O.init_walkers(rdm1, nelec, Wn, walker_type)
        spins = [(up, Tu), (dn, Td)]
    for W, T in spins:
        assert W.shape == (Wn, N, T.shape[1]) and np.iscomplexobj(W)
        for w in W:
            assert np.allclose(w, W[0])
            assert np.allclose(np.real(w) @ np.real(w).T, T @ T.T, atol=1e-12)
            assert abs(abs(np.linalg.det(T.T @ w)) - 1.0) < 1e-12

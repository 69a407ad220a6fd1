"""CPU pins of the oracle's phaseless importance function, population-control
constant and mean-field split (propagation.py:296-334, 386-437), each from what
the mathematics fixes rather than from the oracle's own formulas:

* operator identity on the whole Fock space (brute-force second quantisation):
  H = -h0_prop + sum h1_mod E + 1/2 sum_g (V_g - m_g)^2, with m_g = -i mf_g;
* closed form at W = T with zero fields: the field shift vanishes and
  I = exp(dt (E_shift + h0_prop)) det(T^T exp(-dt h1_mod) T)^{2 | per spin};
* the Girsanov / Hubbard-Stratonovich identity: the Gaussian mean of the
  unclipped complex importance function is
  <T| exp(-dt (H - E_shift)) |W> / <T|W> = 1 - dt (E_L(W) - E_shift) + O(dt^2),
  with E_L the golden- and brute-force-pinned local energy.  The mean is taken
  by tensor Gauss-Hermite quadrature (exact to rounding for this integrand) and
  the O(dt^2) terms are removed by Richardson extrapolation over three dt.

`test_mutations_fail_the_pins` plants a sign error in each term of the
oracle's importance function (fb_term, shift_term, h0_prop, E_shift) and
shows that a pin above fails for every one of them.
"""
import math

import numpy as np
import pytest
from numpy.polynomial.hermite_e import hermegauss

from oracle import afqmc_oracle as O
from tests import fock


def _system(N=4, G=2, seed=21, uhf=False, nocc=(2, 2)):
    rng = np.random.default_rng(seed)
    h1 = rng.standard_normal((N, N))
    h1 = (h1 + h1.T) / 2
    L = rng.standard_normal((G, N, N)) * 0.6
    L = (L + L.transpose(0, 2, 1)) / 2
    if uhf:
        Tu = np.linalg.qr(rng.standard_normal((N, nocc[0])))[0]
        Td = np.linalg.qr(rng.standard_normal((N, nocc[1])))[0]
        mos = (Tu, Td)
    else:
        T = np.linalg.qr(rng.standard_normal((N, nocc[0])))[0]
        mos = (T, T)
    Wu = mos[0] + 0.25 * (rng.standard_normal((N, nocc[0])) + 1j * rng.standard_normal((N, nocc[0])))
    Wd = mos[1] + 0.25 * (rng.standard_normal((N, nocc[1])) + 1j * rng.standard_normal((N, nocc[1])))
    return dict(N=N, G=G, h0=0.3, h1=h1, L=L, mos=mos, W=(Wu, Wd), uhf=uhf, nocc=nocc)


def _oracle_case(s, dt):
    """Oracle trial, walker batch builder and propagation dict at time step dt."""
    N, G, L = s["N"], s["G"], s["L"]
    chol = L.reshape(G, -1)
    h1x2 = np.array([s["h1"], s["h1"]])
    Tu, Td = s["mos"]
    if s["uhf"]:
        h1s, rh, rc = O.uhf_measurement_intermediates(h1x2, chol, Tu, Td)
        trial = O.UHFTrial(N, s["nocc"], Tu, Td)
        rdm1 = O.rdm1_from_mo(Tu, Td)
        wtype = "unrestricted"
    else:
        h1s, rh, rc = O.rhf_measurement_intermediates(h1x2, chol, Tu)
        trial = O.RHFTrial(N, s["nocc"], Tu)
        rdm1 = O.rdm1_from_mo(Tu)
        wtype = "restricted"
    pi = O.propagation_intermediates(s["h0"], h1s, chol, rdm1, dt, wtype)
    ham = dict(h0=s["h0"], chol=chol, rot_h1=rh, rot_chol=rc, mf_shifts=pi["mf_shifts"], h0_prop=pi["h0_prop"],
               exp_h1=pi["exp_h1"])

    def batch(Wu, Wd, M):
        if s["uhf"]:
            return [np.repeat(Wu[None], M, 0), np.repeat(Wd[None], M, 0)]
        return np.repeat(Wu[None], M, 0)

    return trial, batch, ham, pi


def _gauss_hermite_grid(G, n):
    x, w = hermegauss(n)
    w = w / math.sqrt(2.0 * math.pi)
    grids = np.meshgrid(*([x] * G), indexing="ij")
    wts = np.meshgrid(*([w] * G), indexing="ij")
    pts = np.stack([g.reshape(-1) for g in grids], axis=1)
    wt = np.prod(np.stack([g.reshape(-1) for g in wts], axis=1), axis=1)
    return pts, wt


def _mean_importance(s, dt, e_shift, n_nodes=20):
    """Gaussian mean of the oracle's unclipped importance function at walker W."""
    trial, batch, ham, _ = _oracle_case(s, dt)
    pts, wt = _gauss_hermite_grid(s["G"], n_nodes)
    M = pts.shape[0]
    walkers = batch(*s["W"], M)
    state = dict(walkers=walkers, weights=np.ones(M), overlaps=trial.calc_overlap(walkers), e_estimate=e_shift,
                 pop_control_ene_shift=e_shift)
    out = O.propagate_constrained(trial, state, pts, ham, dt)
    return np.sum(wt * out["importance"])


def _girsanov_residual(s, e_shift=0.7, h=2e-3):
    """|extrapolated d/d(dt) of E_x[I] at dt=0 + (E_L - E_shift)| and the scale of E_L."""
    r = [(_mean_importance(s, k * h, e_shift) - 1.0) / (k * h) for k in (1, 2, 4)]
    slope = (8.0 * r[0] - 6.0 * r[1] + r[2]) / 3.0
    trial, batch, ham, _ = _oracle_case(s, h)
    e_loc = trial.calc_energy(batch(*s["W"], 1), ham["h0"], ham["rot_h1"], ham["rot_chol"])[0]
    return abs(slope + (e_loc - e_shift)), abs(e_loc)


CASES = {
    "rhf-restricted": dict(uhf=False, nocc=(2, 2), seed=21),
    "uhf-unrestricted": dict(uhf=True, nocc=(2, 1), seed=22),
}
GIRSANOV_TOL = 2e-6


@pytest.mark.parametrize("case", list(CASES))
def test_girsanov_mean_of_importance_function(case):
    s = _system(**CASES[case])
    res, scale = _girsanov_residual(s)
    assert res < GIRSANOV_TOL * max(1.0, scale), (res, scale)


def _fock_matrices(s, dt):
    """Full-Fock-space H and its mean-field decomposition built from the
    oracle's h0_prop, h1_mod and mf_shifts (restricted walkers: one h1_mod)."""
    N, G, L = s["N"], s["G"], s["L"]
    _, _, ham, pi = _oracle_case(s, dt)
    H, _ = fock.hamiltonian_matrix(s["h0"], s["h1"], L)
    h1_mod = pi["h1_mod"] if pi["h1_mod"].ndim == 2 else pi["h1_mod"][0]
    assert abs(np.imag(pi["h0_prop"])) == 0.0 and np.all(np.real(pi["mf_shifts"]) == 0.0)
    m = np.imag(pi["mf_shifts"])  # mf_g = i m_g
    H_dec, _ = fock.hamiltonian_matrix(-np.real(pi["h0_prop"]), h1_mod, np.zeros((0, N, N)))
    dim = H.shape[0]
    for g in range(G):
        V, _ = fock.hamiltonian_matrix(0.0, L[g], np.zeros((0, N, N)))
        D = V - m[g] * np.eye(dim)
        H_dec = H_dec + 0.5 * D @ D
    return H, H_dec, m


@pytest.mark.parametrize("case", list(CASES))
def test_mean_field_decomposition_is_the_hamiltonian(case):
    s = _system(N=3, **{**CASES[case], "nocc": (1, 1)})
    H, H_dec, m = _fock_matrices(s, 0.01)
    assert np.max(np.abs(H - H_dec)) < 1e-12 * max(1.0, np.max(np.abs(H)))
    # m_g is the trial's expectation of V_g: 2 tr(P L_g) per restricted trial
    Tu, Td = s["mos"]
    m_ref = np.array([np.trace(Tu.T @ L @ Tu) + np.trace(Td.T @ L @ Td) for L in s["L"]])
    assert np.allclose(m, m_ref, atol=1e-13)


@pytest.mark.parametrize("case", list(CASES))
def test_trial_walker_zero_field_closed_form(case):
    s = _system(**CASES[case])
    dt, e_shift = 0.01, 0.45
    trial, batch, ham, pi = _oracle_case(s, dt)
    Tu, Td = s["mos"]
    walkers = batch(Tu + 0j, Td + 0j, 1)
    state = dict(walkers=walkers, weights=np.ones(1), overlaps=trial.calc_overlap(walkers), e_estimate=e_shift,
                 pop_control_ene_shift=e_shift)
    out = O.propagate_constrained(trial, state, np.zeros((1, s["G"])), ham, dt)
    # the force bias at the trial is -i mf, so no shift, no shift_term, no fb_term
    assert np.allclose(out["shifted_fields"], 0.0, atol=1e-14)
    hm = pi["h1_mod"]
    hms = (hm, hm) if hm.ndim == 2 else (hm[0], hm[1])

    def det_exp(T, h):
        ev, U = np.linalg.eigh(h)
        return np.linalg.det(T.T @ U @ np.diag(np.exp(-dt * ev)) @ U.T @ T)

    ratio = det_exp(Tu, hms[0]) * det_exp(Td, hms[1])
    expected = np.exp(dt * (e_shift + pi["h0_prop"])) * ratio
    assert abs(out["importance"][0] - expected) < 1e-13 * abs(expected)
    assert abs(out["phaseless_factor"][0] - abs(expected)) < 1e-13 * abs(expected)


def _mutant(name):
    """A copy of one oracle function with the sign of one term flipped."""
    orig_terms, orig_imp, orig_pi = O.step_terms, O.importance_function, O.propagation_intermediates
    if name == "fb_term":
        def f(*a, **k):
            fs, sh, st, ft = orig_terms(*a, **k)
            return fs, sh, st, -ft
        return "step_terms", f
    if name == "shift_term":
        def f(*a, **k):
            fs, sh, st, ft = orig_terms(*a, **k)
            return fs, sh, -st, ft
        return "step_terms", f
    if name == "e_shift":
        def f(shift_term, fb_term, ov_new, ov_old, e_shift, h0_prop, dt):
            return orig_imp(shift_term, fb_term, ov_new, ov_old, -e_shift, h0_prop, dt)
        return "importance_function", f
    if name == "h0_prop":
        def f(*a, **k):
            d = orig_pi(*a, **k)
            d = dict(d)
            d["h0_prop"] = -d["h0_prop"]
            return d
        return "propagation_intermediates", f
    raise ValueError(name)


@pytest.mark.parametrize("term", ["fb_term", "shift_term", "h0_prop", "e_shift"])
@pytest.mark.parametrize("case", list(CASES))
def test_mutations_fail_the_pins(monkeypatch, term, case):
    s = _system(**CASES[case])
    attr, fn = _mutant(term)
    monkeypatch.setattr(O, attr, fn)
    res, scale = _girsanov_residual(s)
    # the planted error is caught with a wide margin
    assert res > 1e3 * GIRSANOV_TOL * max(1.0, scale), (term, res)


def test_pins_table_names_existing_tests():
    """Every test the oracle's PINS table cites exists under tests/."""
    import os
    import re

    here = os.path.dirname(os.path.abspath(__file__))
    defined = set()
    for f in os.listdir(here):
        if f.endswith(".py"):
            defined |= set(re.findall(r"^def (test_\w+)", open(os.path.join(here, f)).read(), re.M))
    cited = set()
    for v in O.PINS.values():
        cited |= {t.rstrip("*") for t in re.findall(r"test_\w+", v) if not t.startswith("test_wavefunctions")
                  and not t.startswith("test_driver") and not t.startswith("test_hamiltonian")}
    cited -= {os.path.splitext(f)[0] for f in os.listdir(here)}  # module names
    # prefix citations (test_golden_*, test_energy_bruteforce_*) match at least one test
    missing = [c for c in cited if not any(d == c or (c.endswith("_") and d.startswith(c)) for d in defined)]
    assert not missing, missing

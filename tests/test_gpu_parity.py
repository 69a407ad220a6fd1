"""HIP engine (libafq.so through the reference-mirroring facade) vs the FP64
oracle on identical inputs.  Tolerance: 1e-10 relative (max-abs error over
max-abs value of each quantity), FP64 throughout."""
import math
import os

import numpy as np
import pytest
import torch

from ad_afqmc_amd import engine as E
from ad_afqmc_amd import hamiltonian as hm
from ad_afqmc_amd import propagation as pr
from ad_afqmc_amd import synthetic as syn
from ad_afqmc_amd import walkers as wk
from ad_afqmc_amd import wavefunctions as wf
from oracle import afqmc_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def npy(t):
    if isinstance(t, (list, tuple)):
        return [npy(x) for x in t]
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


# ------------------------------------------------------------------ golden vectors
def test_golden_rhf_engine():
    d = np.load(os.path.join(HERE, "golden", "golden_wavefunctions.npz"))
    trial = wf.rhf(10, (5, 5))
    mo = d["mo"]
    wave = {"mo_coeff": mo, "rdm1": np.array([mo @ mo.T] * 2)}
    ham = {"h0": float(d["h0"]), "rot_h1": d["rot_h1"], "rot_chol": d["rot_chol"], "h1": np.array([d["h1"]] * 2),
           "chol": d["chol"].reshape(5, -1), "ene0": 0.0}
    w = wk.RHFWalkers(d["walker"][None])
    ov = npy(trial.calc_overlap(w, wave))[0]
    g = npy(trial.calc_green(w, wave))[0]
    fb = npy(trial.calc_force_bias(w, ham, wave))[0]
    e = npy(trial.calc_energy(w, ham, wave))[0]
    assert math.isclose(ov.real, d["rhf_overlap"], rel_tol=1e-12)
    assert math.isclose(g.sum().real, d["rhf_green_sum"], rel_tol=1e-12)
    assert math.isclose(fb.sum().real, d["rhf_force_bias_sum"], rel_tol=1e-12)
    assert math.isclose(e.real, d["rhf_energy"], rel_tol=1e-12)
    t = O.RHFTrial(10, (5, 5), mo)
    assert rel(fb, t.calc_force_bias(d["walker"][None], d["rot_chol"])[0]) < TOL
    assert rel(e, t.calc_energy(d["walker"][None], d["h0"], d["rot_h1"], d["rot_chol"])[0]) < TOL


def test_golden_uhf_engine():
    d = np.load(os.path.join(HERE, "golden", "golden_wavefunctions.npz"))
    trial = wf.uhf(10, (3, 2))
    mu, md = d["mo_u_up"], d["mo_u_dn"]
    wave = {"mo_coeff": [mu, md], "rdm1": np.array([mu @ mu.T, md @ md.T])}
    rh = [d["rot_h1"][:3], d["rot_h1"][:2]]
    rc = [np.ascontiguousarray(d["rot_chol"][:, :3]), np.ascontiguousarray(d["rot_chol"][:, :2])]
    ham = {"h0": float(d["h0"]), "rot_h1": rh, "rot_chol": rc, "h1": np.array([d["h1"]] * 2),
           "chol": d["chol"].reshape(5, -1), "ene0": 0.0}
    w = wk.UHFWalkers([d["walker_up"][None], d["walker_dn"][None]])
    ov = npy(trial.calc_overlap(w, wave))[0]
    gu, gd = npy(trial.calc_green(w, wave))
    fb = npy(trial.calc_force_bias(w, ham, wave))[0]
    e = npy(trial.calc_energy(w, ham, wave))[0]
    assert math.isclose(ov.real, d["uhf_overlap"], rel_tol=1e-11)
    assert math.isclose(gu.sum().real + gd.sum().real, d["uhf_green_sum"], rel_tol=1e-11)
    assert math.isclose(fb.sum().real, d["uhf_force_bias_sum"], rel_tol=1e-11)
    assert math.isclose(e.real, d["uhf_energy"], rel_tol=1e-11)


# ------------------------------------------------------------------ random systems
CASES = [
    # (trial, walker_type, N, nocc_up, nocc_dn, G, W)
    ("rhf", "restricted", 12, 4, 4, 10, 6),
    ("uhf", "unrestricted", 12, 4, 3, 10, 6),
    ("rhf", "unrestricted", 10, 3, 3, 7, 5),
    ("uhf", "restricted", 11, 4, 2, 9, 5),
    ("rhf", "restricted", 70, 17, 17, 130, 9),
    ("uhf", "unrestricted", 67, 21, 18, 97, 7),
    # 40 walker columns: the register-panel Taylor and exp(-dt/2 h1) products
    ("rhf", "restricted", 96, 40, 40, 150, 8),
    # 48 < O <= 64: the two-wave FP64 exchange
    ("uhf", "unrestricted", 72, 60, 52, 40, 4),
    # nocc > 64: Gauss-Jordan and Cholesky in global memory, the stored-f exchange
    ("uhf", "unrestricted", 100, 80, 70, 60, 4),
    ("rhf", "restricted", 210, 200, 200, 16, 2),
]


def _setup(trial_kind, wtype, N, ou, od, G, W, seed=0, dt=0.01, chol_asym=0.0):
    # Cholesky scale ~ 1/sqrt(N) keeps the walker population alive at this dt
    sysm = syn.make_system(N, ou, G, nocc_dn=od, uhf=(trial_kind == "uhf"), seed=seed,
                           l_scale=min(1.0, 2.0 / math.sqrt(N)))
    if chol_asym:
        # an antisymmetric part above the engine's symmetry threshold (1e-12 of max |L|)
        a = np.random.default_rng(seed + 9).standard_normal((G, N, N))
        sysm.chol += chol_asym * (a - a.transpose(0, 2, 1)).reshape(G, N * N)
    wave = sysm.wave_data()
    ham = sysm.ham_data()
    trial = wf.rhf(N, (ou, ou)) if trial_kind == "rhf" else wf.uhf(N, (ou, od))
    prop = pr.propagator_afqmc(dt=dt, n_walkers=W, walker_type=wtype)
    H = hm.hamiltonian(N)
    ham = H.build_measurement_intermediates(ham, trial, wave)
    ham = H.build_propagation_intermediates(ham, prop, trial, wave)
    # oracle side
    if trial_kind == "rhf":
        h1s, orh, orc = O.rhf_measurement_intermediates(sysm.h1, sysm.chol, sysm.mo_up)
        otrial = O.RHFTrial(N, (ou, ou), sysm.mo_up)
    else:
        h1s, orh, orc = O.uhf_measurement_intermediates(sysm.h1, sysm.chol, sysm.mo_up, sysm.mo_dn)
        otrial = O.UHFTrial(N, (ou, od), sysm.mo_up, sysm.mo_dn)
    opi = O.propagation_intermediates(sysm.h0, h1s, sysm.chol, sysm.rdm1(), dt, wtype)
    oham = dict(h0=sysm.h0, chol=sysm.chol, rot_h1=orh, rot_chol=orc, **{k: opi[k] for k in ("mf_shifts", "h0_prop", "exp_h1")})
    # perturbation ~ 0.3/sqrt(N) per element keeps the walkers' energies near the trial's
    amp = 0.3 / math.sqrt(N)
    if wtype == "restricted":
        wn = syn.perturbed_walkers(sysm.mo_up, W, amp, seed + 1)
        walkers, owalk = wk.RHFWalkers(wn), wn
    else:
        dn_mo = sysm.mo_up if sysm.mo_dn is None else sysm.mo_dn
        wu = syn.perturbed_walkers(sysm.mo_up, W, amp, seed + 1)
        wd = syn.perturbed_walkers(dn_mo, W, amp, seed + 2)
        walkers, owalk = wk.UHFWalkers([wu, wd]), [wu, wd]
    return sysm, trial, prop, ham, wave, walkers, otrial, oham, owalk, h1s


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1]}-N{c[2]}" for c in CASES])
def test_intermediates_and_measurements(case):
    sysm, trial, prop, ham, wave, walkers, otrial, oham, owalk, h1s = _setup(*case)
    assert rel(npy(ham["h1"]), h1s) < TOL
    if case[0] == "rhf":
        assert rel(npy(ham["rot_h1"]), oham["rot_h1"]) < TOL
        assert rel(npy(ham["rot_chol"]), oham["rot_chol"]) < TOL
    else:
        for s in range(2):
            assert rel(npy(ham["rot_h1"][s]), oham["rot_h1"][s]) < TOL
            assert rel(npy(ham["rot_chol"][s]), oham["rot_chol"][s]) < TOL
    assert rel(npy(ham["mf_shifts"]), oham["mf_shifts"]) < TOL
    assert rel(npy(ham["h0_prop"]), oham["h0_prop"]) < TOL
    assert rel(npy(ham["exp_h1"]), oham["exp_h1"]) < 1e-13 * 10
    assert rel(npy(trial.calc_overlap(walkers, wave)), otrial.calc_overlap(owalk)) < TOL
    assert rel(npy(trial.calc_force_bias(walkers, ham, wave)), otrial.calc_force_bias(owalk, oham["rot_chol"])) < TOL
    e = npy(trial.calc_energy(walkers, ham, wave))
    eo = otrial.calc_energy(owalk, oham["h0"], oham["rot_h1"], oham["rot_chol"])
    assert rel(e, eo) < TOL


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1]}-N{c[2]}" for c in CASES])
def test_propagate_constrained_steps(case):
    sysm, trial, prop, ham, wave, walkers, otrial, oham, owalk, _ = _setup(*case)
    W, G = case[6], case[5]
    rng = np.random.default_rng(42)
    pd = prop.init_prop_data(trial, wave, ham, seed=3, init_walkers=walkers)
    e_est = float(pd["e_estimate"].item())
    st = dict(walkers=owalk, weights=np.ones(W), overlaps=otrial.calc_overlap(owalk), e_estimate=e_est,
              pop_control_ene_shift=e_est)
    assert rel(npy(pd["overlaps"]), st["overlaps"]) < TOL
    for step in range(3):
        fields = rng.standard_normal((W, G))
        pd = prop.propagate_constrained(trial, ham, pd, fields, wave, return_force_bias=True)
        st = O.propagate_constrained(otrial, st, fields, oham, prop.dt, prop.n_exp_terms)
        assert rel(npy(pd["force_bias"]), st["force_bias"]) < TOL, step
        if case[1] == "restricted":
            assert rel(npy(pd["walkers"].data), st["walkers"]) < TOL, step
        else:
            assert rel(npy(pd["walkers"].data[0]), st["walkers"][0]) < TOL, step
            assert rel(npy(pd["walkers"].data[1]), st["walkers"][1]) < TOL, step
        assert rel(npy(pd["overlaps"]), st["overlaps"]) < TOL, step
        assert rel(npy(pd["weights"]), st["weights"]) < TOL, step
        assert np.count_nonzero(st["weights"]) >= W // 2, step  # the step is not trivially clipped
        assert abs(pd["pop_control_ene_shift"].item() - st["pop_control_ene_shift"]) < TOL * max(1, abs(e_est))


# N = 70: the VHS product runs over the packed lower block triangle when L_g is
# symmetric (exactly so here) and over all N^2 columns when it is not.
@pytest.mark.parametrize("asym", [0.0, 1e-9], ids=["symmetric-L", "asymmetric-L"])
@pytest.mark.parametrize("case", [("rhf", "restricted", 70, 17, 17, 130, 9), ("uhf", "unrestricted", 70, 21, 18, 97, 7)],
                         ids=["rhf-N70", "uhf-N70"])
def test_vhs_symmetric_and_general_cholesky(case, asym):
    sysm, trial, prop, ham, wave, walkers, otrial, oham, owalk, _ = _setup(*case, chol_asym=asym)
    W, G = case[6], case[5]
    rng = np.random.default_rng(5)
    pd = prop.init_prop_data(trial, wave, ham, seed=3, init_walkers=walkers)
    e_est = float(pd["e_estimate"].item())
    st = dict(walkers=owalk, weights=np.ones(W), overlaps=otrial.calc_overlap(owalk), e_estimate=e_est,
              pop_control_ene_shift=e_est)
    for step in range(2):
        fields = rng.standard_normal((W, G))
        pd = prop.propagate_constrained(trial, ham, pd, fields, wave)
        st = O.propagate_constrained(otrial, st, fields, oham, prop.dt, prop.n_exp_terms)
        wa = npy(pd["walkers"].data)
        for s in range(1 if case[1] == "restricted" else 2):
            got = wa if case[1] == "restricted" else wa[s]
            ref = st["walkers"] if case[1] == "restricted" else st["walkers"][s]
            assert rel(got, ref) < TOL, (step, s)
        assert rel(npy(pd["weights"]), st["weights"]) < TOL, step


@pytest.mark.parametrize("asym", [0.0, 1e-9], ids=["symmetric-L", "asymmetric-L"])
@pytest.mark.parametrize("case", [("rhf", "restricted", 70, 17, 17, 130, 9), ("uhf", "unrestricted", 67, 21, 18, 97, 7)],
                         ids=["rhf-N70", "uhf-N67"])
def test_bf16x3_propagator_symmetric_and_general_cholesky(case, asym):
    """The fused split-BF16 propagator on the planes split from a packed
    symmetric VHS (mirrored) and from a general N x N one: one step of each
    against the FP64 oracle."""
    kind, wtype, N, ou, od, G, W = case
    sysm, trial, prop, ham, wave, walkers, otrial, oham, owalk, _ = _setup(*case, chol_asym=asym)
    wave_lp, ham_lp = sysm.wave_data(), sysm.ham_data()
    prop_lp = pr.propagator_afqmc(dt=prop.dt, n_walkers=W, walker_type=wtype, precision="bf16x3")
    H = hm.hamiltonian(N)
    ham_lp = H.build_measurement_intermediates(ham_lp, trial, wave_lp)
    ham_lp = H.build_propagation_intermediates(ham_lp, prop_lp, trial, wave_lp)
    pd = prop_lp.init_prop_data(trial, wave_lp, ham_lp, seed=3, init_walkers=walkers)
    e_est = float(pd["e_estimate"].item())
    st = dict(walkers=owalk, weights=np.ones(W), overlaps=otrial.calc_overlap(owalk), e_estimate=e_est,
              pop_control_ene_shift=e_est)
    fields = np.random.default_rng(11).standard_normal((W, G))
    pd = prop_lp.propagate_constrained(trial, ham_lp, pd, fields, wave_lp)
    st = O.propagate_constrained(otrial, st, fields, oham, prop.dt, prop.n_exp_terms)
    wa = npy(pd["walkers"].data)
    for s in range(1 if wtype == "restricted" else 2):
        got = wa if wtype == "restricted" else wa[s]
        ref = st["walkers"] if wtype == "restricted" else st["walkers"][s]
        assert rel(got, ref) < LP_TOL["bf16x3"], s
    assert rel(npy(pd["weights"]), st["weights"]) < 10 * LP_TOL["bf16x3"]


def test_trotprop_matches_oracle():
    sysm, trial, prop, ham, wave, walkers, otrial, oham, owalk, _ = _setup("uhf", "unrestricted", 20, 5, 4, 15, 4)
    rng = np.random.default_rng(5)
    xs = rng.standard_normal((4, 15)) + 0.3j * rng.standard_normal((4, 15))
    out = prop.apply_trotprop(trial, ham, walkers, xs, wave)
    ref = O.apply_trotprop(owalk, xs, oham["chol"], oham["exp_h1"], prop.dt, prop.n_exp_terms)
    assert rel(npy(out.data[0]), ref[0]) < TOL
    assert rel(npy(out.data[1]), ref[1]) < TOL


def test_phaseless_clipping_known_answers():
    # a walker whose overlap ratio flips sign gets cos(theta) < 0 -> weight 0;
    # weights above 100 are zeroed.
    sysm, trial, prop, ham, wave, walkers, otrial, oham, owalk, _ = _setup("rhf", "restricted", 12, 4, 4, 10, 6)
    pd = prop.init_prop_data(trial, wave, ham, seed=3, init_walkers=walkers)
    ov = pd["overlaps"].clone()
    ov[1] = -ov[1]
    w = pd["weights"].clone()
    w[2] = 150.0
    pd["overlaps"], pd["weights"] = ov, w
    fields = np.zeros((6, 10))
    out = prop.propagate_constrained(trial, ham, pd, fields, wave)
    st = dict(walkers=owalk, weights=npy(w), overlaps=npy(ov), e_estimate=pd["e_estimate"].item(),
              pop_control_ene_shift=pd["e_estimate"].item())
    ref = O.propagate_constrained(otrial, st, fields, oham, prop.dt)
    wout = npy(out["weights"])
    assert wout[1] == 0.0 and ref["weights"][1] == 0.0
    assert wout[2] == 0.0 and ref["weights"][2] == 0.0
    assert rel(wout, ref["weights"]) < TOL


def test_orthonormalize_matches_oracle_up_to_phases():
    sysm, trial, prop, ham, wave, walkers, otrial, oham, owalk, _ = _setup("uhf", "unrestricted", 30, 7, 5, 12, 5)
    q = walkers.orthonormalize()
    qo = O.orthonormalize(owalk)
    for s in range(2):
        qg = npy(q.data[s])
        # columns agree up to a sign (LAPACK's R diagonal is real with either sign)
        m = np.einsum("wpi,wpj->wij", qg.conj(), qo[s])
        assert np.allclose(np.abs(m), np.eye(qg.shape[2])[None], atol=1e-10)
        assert np.allclose(np.einsum("wpi,wpj->wij", qg.conj(), qg), np.eye(qg.shape[2])[None], atol=1e-12)
    # overlaps, Green's functions and energies are identical after QR
    assert rel(npy(trial.calc_overlap(q, wave)), otrial.calc_overlap(qo)) < TOL
    e = npy(trial.calc_energy(q, ham, wave))
    assert rel(e, otrial.calc_energy(qo, oham["h0"], oham["rot_h1"], oham["rot_chol"])) < TOL


@pytest.mark.parametrize("case", [("uhf", "unrestricted", 30, 7, 5, 12, 5), ("rhf", "restricted", 70, 17, 17, 130, 9),
                                  ("rhf", "restricted", 100, 80, 80, 60, 4)])
def test_block_qr_equals_gram_schmidt_and_oracle(case):
    """The block loop's QR (Cholesky QR, twice) against Gram-Schmidt with
    re-orthogonalisation (both give R with a positive real diagonal, so the
    same Q) and against the oracle's Householder QR up to column signs, on
    walkers after three propagation steps."""
    from ad_afqmc_amd.binding import engine_for

    kind, wtype, N, ou, od, G, W = case
    sysm, trial, prop, ham, wave, walkers, otrial, oham, owalk, _ = _setup(*case)
    job = engine_for(trial, wave, ham, wtype, W, prop)
    up, dn = (t.clone() if t is not None else None for t in walkers.tensors())
    weights = torch.ones(W, dtype=torch.float64, device="cuda")
    ov = job.overlap(up, dn)
    sc = torch.zeros(2, dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(3)
    for _ in range(3):
        job.propagate(up, dn, torch.as_tensor(rng.standard_normal((W, G)), device="cuda"), weights, ov, sc)
    ref = [t.clone() for t in (up, dn) if t is not None]
    gs = [wk.RHFWalkers(r).orthonormalize().data for r in ref] if wtype == "restricted" else \
        wk.UHFWalkers(ref).orthonormalize().data
    job.orthonormalize(up, dn)
    for q, g, r in zip([t for t in (up, dn) if t is not None], gs, ref):
        qn, gn, rn = npy(q), npy(g), npy(r)
        assert np.allclose(np.einsum("wpi,wpj->wij", qn.conj(), qn), np.eye(qn.shape[2])[None], atol=1e-12)
        assert rel(qn, gn) < 1e-10
        qo = O.qr_restricted(rn)[0]
        m = np.einsum("wpi,wpj->wij", qn.conj(), qo)
        assert np.allclose(np.abs(m), np.eye(qn.shape[2])[None], atol=1e-10)


def test_stochastic_reconfiguration_matches_oracle():
    rng = np.random.default_rng(7)
    W, N, Oo = 37, 8, 3
    data = rng.standard_normal((W, N, Oo)) + 1j * rng.standard_normal((W, N, Oo))
    weights = rng.random(W) * 2
    weights[[3, 9, 20]] = 0.0
    zeta = 0.3141
    w2, wt2 = wk.RHFWalkers(data).stochastic_reconfiguration_local(weights, zeta)
    idx, wref = O.sr_indices(weights, zeta)
    assert np.array_equal(npy(w2.data), data[idx])
    assert rel(npy(wt2), wref) < 1e-14


def test_normal_fill_moments_and_determinism():
    x = npy(E.normal_fill(1_000_001, seed=11, counter=5))
    y = npy(E.normal_fill(1_000_001, seed=11, counter=5))
    z = npy(E.normal_fill(1000, seed=11, counter=6))
    assert np.array_equal(x, y)
    assert not np.allclose(x[:1000], z)
    assert abs(x.mean()) < 5e-3 and abs(x.var() - 1) < 5e-3
    assert abs(np.mean(x**3)) < 1e-2 and abs(np.mean(x**4) - 3) < 3e-2


def test_h6_driver_energy_statistical():
    """Reference H6 fixture (UHF trial, unrestricted walkers, 50 walkers,
    tests/test_driver.py:74-86: -3.238196747261496).  The field stream differs
    from JAX's, so agreement is statistical."""
    from ad_afqmc_amd import driver, sampling

    d = np.load(os.path.join(HERE, "golden", "h6_fcidump_chol.npz"))
    nelec, norb, ms, nchol = (int(x) for x in d["header"])
    nup, ndn = (nelec + abs(ms)) // 2, (nelec - abs(ms)) // 2
    mo = d["mo_coeff"]
    trial = wf.uhf(norb, (nup, ndn))
    wave = {"mo_coeff": [mo[0][:, :nup], mo[1][:, :ndn]]}
    wave["rdm1"] = np.array([wave["mo_coeff"][0] @ wave["mo_coeff"][0].T, wave["mo_coeff"][1] @ wave["mo_coeff"][1].T])
    h1 = d["hcore"].reshape(norb, norb)
    ham = {"h0": float(d["energy_core"]), "h1": np.array([h1, h1]), "chol": d["chol"].reshape(nchol, -1), "ene0": 0.0}
    prop = pr.propagator_afqmc(dt=0.01, n_walkers=50, walker_type="unrestricted")
    smp = sampling.sampler(n_prop_steps=50, n_ene_blocks=1, n_sr_blocks=10, n_blocks=40)
    e, err = driver.afqmc_energy(ham, hm.hamiltonian(norb), prop, trial, wave, smp,
                                 dict(n_eql=1, seed=98), verbose=False)
    assert abs(e - (-3.238196747261496)) < 5 * err + 2e-3, (e, err)


# ------------------------------------------------------------------ reduced precision
LP_CASES = [
    ("rhf", "restricted", 12, 4, 4, 10, 6),
    ("uhf", "unrestricted", 67, 21, 18, 97, 7),
    ("rhf", "restricted", 70, 17, 17, 130, 9),
    ("rhf", "restricted", 96, 40, 40, 150, 8),
    # 112 walker columns (both spins) on the bf16x3 register-panel Taylor kernel; O > 48 exchanges
    ("uhf", "unrestricted", 72, 60, 52, 40, 4),
    # nocc > 64 (the stored-f exchange in FP64 arithmetic in every mode)
    ("rhf", "restricted", 100, 80, 80, 60, 4),
]
# relative tolerances vs the FP64 oracle for one measurement / one step
LP_TOL = {"fp32": 2e-5, "bf16x3": 5e-4}


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("case", LP_CASES, ids=[f"{c[0]}-{c[1]}-N{c[2]}" for c in LP_CASES])
def test_reduced_precision_against_oracle(case, precision):
    kind, wtype, N, ou, od, G, W = case
    sysm, trial, prop, ham, wave, walkers, otrial, oham, owalk, _ = _setup(*case)
    # same system, reduced-precision job (new wave_data so a new job is bound)
    wave_lp = sysm.wave_data()
    ham_lp = sysm.ham_data()
    prop_lp = pr.propagator_afqmc(dt=prop.dt, n_walkers=W, walker_type=wtype, precision=precision)
    H = hm.hamiltonian(N)
    ham_lp = H.build_measurement_intermediates(ham_lp, trial, wave_lp)
    ham_lp = H.build_propagation_intermediates(ham_lp, prop_lp, trial, wave_lp)
    tol = LP_TOL[precision]
    e = npy(trial.calc_energy(walkers, ham_lp, wave_lp))
    eo = otrial.calc_energy(owalk, oham["h0"], oham["rot_h1"], oham["rot_chol"])
    assert rel(e, eo) < tol
    fb = npy(trial.calc_force_bias(walkers, ham_lp, wave_lp))
    assert rel(fb, otrial.calc_force_bias(owalk, oham["rot_chol"])) < tol
    pd = prop_lp.init_prop_data(trial, wave_lp, ham_lp, seed=3, init_walkers=walkers)
    e_est = float(pd["e_estimate"].item())
    st = dict(walkers=owalk, weights=np.ones(W), overlaps=otrial.calc_overlap(owalk), e_estimate=e_est,
              pop_control_ene_shift=e_est)
    fields = np.random.default_rng(42).standard_normal((W, G))
    pd = prop_lp.propagate_constrained(trial, ham_lp, pd, fields, wave_lp)
    st = O.propagate_constrained(otrial, st, fields, oham, prop.dt, prop.n_exp_terms)
    if wtype == "restricted":
        assert rel(npy(pd["walkers"].data), st["walkers"]) < tol
    else:
        assert rel(npy(pd["walkers"].data[0]), st["walkers"][0]) < tol
        assert rel(npy(pd["walkers"].data[1]), st["walkers"][1]) < tol
    assert rel(npy(pd["overlaps"]), st["overlaps"]) < 10 * tol
    assert rel(npy(pd["weights"]), st["weights"]) < 10 * tol
    # the FP64-kept parts are untouched: overlap of the same walkers is exact
    assert rel(npy(trial.calc_overlap(walkers, wave_lp)), otrial.calc_overlap(owalk)) < TOL


@pytest.mark.parametrize("kind", ["rhf", "uhf"])
def test_init_walkers_span_the_trial_space(kind):
    """The façade's get_init_walkers (wavefunctions.py:427-466) on the device:
    the walkers' projector per spin is the trial's T T^T, as for the oracle's."""
    print(f"[init walkers] {kind}", flush=True)
    rng = np.random.default_rng(22)
    N, nup, ndn, Wn = 10, 4, 3, 5
    Tu = np.linalg.qr(rng.standard_normal((N, nup)))[0]
    if kind == "rhf":
        trial, wtype, nelec, Ts = wf.rhf(N, (nup, nup)), "restricted", (nup, nup), [Tu]
        wave = {"mo_coeff": Tu, "rdm1": np.array([Tu @ Tu.T, Tu @ Tu.T])}
    else:
        Td = np.linalg.qr(rng.standard_normal((N, ndn)))[0]
        trial, wtype, nelec, Ts = wf.uhf(N, (nup, ndn)), "unrestricted", (nup, ndn), [Tu, Td]
        wave = {"mo_coeff": [Tu, Td], "rdm1": np.array([Tu @ Tu.T, Td @ Td.T])}
    tensors = [t for t in trial.get_init_walkers(wave, Wn, wtype).tensors() if t is not None]
    assert len(tensors) == len(Ts)
    for t, T in zip(tensors, Ts):
        W = t.cpu().numpy()
        assert W.shape == (Wn, N, T.shape[1])
        for w in W:
            assert np.allclose(np.real(w) @ np.real(w).T, T @ T.T, atol=1e-12)
            assert np.allclose(np.imag(w), 0.0)

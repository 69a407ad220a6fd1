"""Parity at the benchmark workloads' full sizes:

* RHF N=200, O=40, G=1200, 512 walkers (restricted walkers);
* UHF N=300, O=60/58, G=2000, 256 walkers (unrestricted walkers);

with the synthetic generator at Cholesky scale 0.3/sqrt(N) (so that the
population lives; the stated scale clips every walker within a few steps).
The whole walker batch runs through afq_propagate_constrained for two steps
and through afq_calc_energy; a sample of walkers spread over the batch is
checked against the FP64 oracle, and every walker for finiteness and weight
range.  FP64: the oracle's own two-step trajectory, 1e-10 relative.  FP32 and
bf16x3: each step from the engine's previous state, at the reduced-precision
tolerances of one step (2e-5 / 5e-4 relative for walkers, force biases and
energies; 10x for overlaps and weights)."""
import math

import numpy as np
import pytest
import torch

from ad_afqmc_amd import hamiltonian as hm
from ad_afqmc_amd import propagation as pr
from ad_afqmc_amd import synthetic as syn
from ad_afqmc_amd import walkers as wk
from ad_afqmc_amd import wavefunctions as wf
from oracle import afqmc_oracle as O
from tests.test_gpu_parity import LP_TOL, TOL, npy, rel

pytestmark = pytest.mark.gpu

WORKLOADS = ["rhf_n200_o40_g1200_w512", "uhf_n300_o60_58_g2000_w256"]
_CACHE = {}


def _workload(name):
    """System, oracle intermediates, perturbed walkers and fields (cached per module)."""
    if name in _CACHE:
        return _CACHE[name]
    _CACHE.clear()
    cfg = syn.CONFIGS[name]
    N, G, W = cfg["norb"], cfg["nchol"], cfg["n_walkers"]
    Oa = cfg["nocc_up"]
    Ob = cfg.get("nocc_dn", Oa)
    uhf = cfg["uhf"]
    s = syn.make_system(N, Oa, G, nocc_dn=Ob, uhf=uhf, seed=0, l_scale=0.3 / math.sqrt(N))
    dt = cfg["dt"]
    wtype = "unrestricted" if uhf else "restricted"
    if uhf:
        h1s, orh, orc = O.uhf_measurement_intermediates(s.h1, s.chol, s.mo_up, s.mo_dn)
        otrial = O.UHFTrial(N, (Oa, Ob), s.mo_up, s.mo_dn)
    else:
        h1s, orh, orc = O.rhf_measurement_intermediates(s.h1, s.chol, s.mo_up)
        otrial = O.RHFTrial(N, (Oa, Oa), s.mo_up)
    opi = O.propagation_intermediates(s.h0, h1s, s.chol, s.rdm1(), dt, wtype)
    oham = dict(h0=s.h0, chol=s.chol, rot_h1=orh, rot_chol=orc, mf_shifts=opi["mf_shifts"], h0_prop=opi["h0_prop"],
                exp_h1=opi["exp_h1"])
    amp = 0.3 / math.sqrt(N)
    if uhf:
        walkers = [syn.perturbed_walkers(s.mo_up, W, amp, 1), syn.perturbed_walkers(s.mo_dn, W, amp, 2)]
    else:
        walkers = syn.perturbed_walkers(s.mo_up, W, amp, 1)
    rng = np.random.default_rng(42)
    fields = [rng.standard_normal((W, G)) for _ in range(2)]
    sample = np.array(sorted({0, 1, W // 3, W // 2, W - 2, W - 1}))
    _CACHE[name] = dict(cfg=cfg, s=s, otrial=otrial, oham=oham, walkers=walkers, fields=fields, sample=sample,
                        wtype=wtype, uhf=uhf, dt=dt, traj=None, energies=None)
    return _CACHE[name]


def _sub(walkers, idx):
    return [walkers[0][idx], walkers[1][idx]] if isinstance(walkers, list) else walkers[idx]


def _engine(wl, precision):
    cfg, s = wl["cfg"], wl["s"]
    N, W = cfg["norb"], cfg["n_walkers"]
    trial = wf.uhf(N, s.nelec) if wl["uhf"] else wf.rhf(N, s.nelec)
    prop = pr.propagator_afqmc(dt=wl["dt"], n_walkers=W, walker_type=wl["wtype"], precision=precision)
    wave, ham = s.wave_data(), s.ham_data()
    H = hm.hamiltonian(N)
    ham = H.build_measurement_intermediates(ham, trial, wave)
    ham = H.build_propagation_intermediates(ham, prop, trial, wave)
    w0 = wl["walkers"]
    walkers = wk.UHFWalkers(list(w0)) if wl["uhf"] else wk.RHFWalkers(w0)
    pd = prop.init_prop_data(trial, wave, ham, seed=0, init_walkers=walkers)
    return trial, prop, wave, ham, pd


def _walkers_np(pd, uhf):
    d = pd["walkers"].data
    return [npy(d[0]), npy(d[1])] if uhf else npy(d)


def _oracle_state(wl, walkers, weights, overlaps, e_est, shift):
    idx = wl["sample"]
    return dict(walkers=_sub(walkers, idx), weights=np.asarray(weights)[idx], overlaps=np.asarray(overlaps)[idx],
                e_estimate=e_est, pop_control_ene_shift=shift)


def _check_batch(wl, pd):
    w = npy(pd["weights"])
    W = wl["cfg"]["n_walkers"]
    assert np.all(np.isfinite(w)) and np.all(w >= 0.0) and np.all(w <= 100.0)
    assert np.count_nonzero(w) >= 0.9 * W
    ov = npy(pd["overlaps"])
    assert np.all(np.isfinite(ov))
    for x in (pd["walkers"].data if wl["uhf"] else [pd["walkers"].data]):
        assert bool(torch.isfinite(torch.view_as_real(x)).all())


def _compare_step(wl, pd, st, tol, fb=None):
    idx = wl["sample"]
    got = _walkers_np(pd, wl["uhf"])
    if wl["uhf"]:
        for s in range(2):
            assert rel(got[s][idx], st["walkers"][s]) < tol, s
    else:
        assert rel(got[idx], st["walkers"]) < tol
    assert rel(npy(pd["overlaps"])[idx], st["overlaps"]) < 10 * tol
    assert rel(npy(pd["weights"])[idx], st["weights"]) < 10 * tol
    if fb is not None:
        assert rel(npy(fb)[idx], st["force_bias"]) < tol


def _shift_of(wl, e_est, weights):
    """pop_control_ene_shift after a step from the whole batch's weights (propagation.py:332-334)."""
    W = wl["cfg"]["n_walkers"]
    return e_est - 0.1 * math.log(float(np.sum(weights)) / W) / wl["dt"]


# FP64 first for each workload: the reduced modes reuse its trajectory's walkers
ORDER = [(n, p) for n in WORKLOADS for p in ("fp64", "fp32", "bf16x3")]


@pytest.mark.parametrize("name,precision", ORDER, ids=[f"{n}-{p}" for n, p in ORDER])
def test_full_size_steps_and_energy(name, precision):
    wl = _workload(name)
    trial, prop, wave, ham, pd = _engine(wl, precision)
    e_est = float(pd["e_estimate"].item())
    W = wl["cfg"]["n_walkers"]
    otrial, oham = wl["otrial"], wl["oham"]
    if precision == "fp64":
        # the oracle's own trajectory on the sample; the shift it sees is the
        # one the whole batch produces (checked against the batch's weights)
        st = _oracle_state(wl, wl["walkers"], np.ones(W), npy(pd["overlaps"]), e_est, e_est)
        assert rel(st["overlaps"], otrial.calc_overlap(st["walkers"])) < TOL
        traj = []
        for step in range(2):
            pd = prop.propagate_constrained(trial, ham, pd, wl["fields"][step], wave, return_force_bias=True)
            _check_batch(wl, pd)
            st = O.propagate_constrained(otrial, st, wl["fields"][step][wl["sample"]], oham, wl["dt"])
            _compare_step(wl, pd, st, TOL, pd["force_bias"])
            shift = float(pd["pop_control_ene_shift"].item())
            assert abs(shift - _shift_of(wl, e_est, npy(pd["weights"]))) < TOL * max(1.0, abs(e_est))
            st["pop_control_ene_shift"] = shift
            traj.append((_walkers_np(pd, wl["uhf"]), npy(pd["weights"]), npy(pd["overlaps"]), shift))
        wl["traj"] = traj
        e = npy(trial.calc_energy(pd["walkers"], ham, wave))
        assert np.all(np.isfinite(e))
        eo = otrial.calc_energy(st["walkers"], oham["h0"], oham["rot_h1"], oham["rot_chol"])
        wl["energies"] = (_walkers_np(pd, wl["uhf"]), eo)
        assert rel(e[wl["sample"]], eo) < TOL
        return
    if wl["traj"] is None:
        pytest.skip("needs the FP64 case of this workload first")
    tol = LP_TOL[precision]
    prev = (wl["walkers"], np.ones(W), npy(pd["overlaps"]), e_est)
    for step in range(2):
        pd = prop.propagate_constrained(trial, ham, pd, wl["fields"][step], wave, return_force_bias=True)
        _check_batch(wl, pd)
        st = _oracle_state(wl, prev[0], prev[1], prev[2], e_est, prev[3])
        st = O.propagate_constrained(otrial, st, wl["fields"][step][wl["sample"]], oham, wl["dt"])
        _compare_step(wl, pd, st, tol, pd["force_bias"])
        shift = float(pd["pop_control_ene_shift"].item())
        assert abs(shift - _shift_of(wl, e_est, npy(pd["weights"]))) < 1e-10 * max(1.0, abs(e_est))
        prev = (_walkers_np(pd, wl["uhf"]), npy(pd["weights"]), npy(pd["overlaps"]), shift)
    # measurement on the FP64 trajectory's walkers, against the oracle's energies of them
    w64, eo = wl["energies"]
    batch = wk.UHFWalkers([w64[0], w64[1]]) if wl["uhf"] else wk.RHFWalkers(w64)
    e = npy(trial.calc_energy(batch, ham, wave))
    assert np.all(np.isfinite(e))
    assert rel(e[wl["sample"]], eo) < tol

"""Device reverse pass of the phaseless segment (afq_segment_energy_vjp):
against the oracle's reverse pass (itself pinned by central differences) on a
small system, and against central differences of the device forward at
N=100, O=20, G=500, W=256 over 50 steps checkpointed every 10; for a UHF
trial with unrestricted walkers, against the RHF pass with equal spins and
against central differences of the device forward."""
import math

import numpy as np
import pytest
import torch

from ad_afqmc_amd import _native as nat
from ad_afqmc_amd import synthetic as syn
from oracle import afqmc_adjoint as OA
from tests.test_oracle_adjoint import _case

pytestmark = pytest.mark.gpu


def _dev(x, dt=torch.float64):
    return torch.as_tensor(np.asarray(x)).to("cuda", dt)


def _gpu_vjp(c, want_grad=True):
    from ad_afqmc_amd.adjoint import _job

    job = _job(c["h0"], c["h1"], c["chol"], c["mo"], c["walkers"].shape[0], c["dt"], 6)
    try:
        F, gh1, gL = job.segment_energy_vjp(_dev(c["walkers"], torch.complex128), _dev(c["weights"]), c["e_estimate"],
                                            _dev(c["fields"]), want_grad=want_grad)
        torch.cuda.synchronize()
        return (float(F.item()), None if gh1 is None else gh1.cpu().numpy(), None if gL is None else gL.cpu().numpy())
    finally:
        job.close()


@pytest.mark.parametrize("seed,dt,n_blocks,n_steps", [(3, 0.01, 2, 3), (3, 0.05, 3, 4)])
def test_vjp_matches_oracle_small(seed, dt, n_blocks, n_steps):
    print(f"[adjoint] small seed={seed} dt={dt} blocks={n_blocks} steps={n_steps}", flush=True)
    c = _case(seed=seed, dt=dt, n_blocks=n_blocks, n_steps=n_steps)
    F, gh1, gL = _gpu_vjp(c)
    Fo, gh1o, gLo = OA.segment_energy_vjp(c["h0"], c["h1"], c["chol"], c["mo"], c["walkers"], c["weights"],
                                          c["e_estimate"], c["fields"], c["dt"])
    assert abs(F - Fo) <= 1e-12 * abs(Fo)
    assert np.max(np.abs(gh1 - gh1o)) <= 1e-10 * np.max(np.abs(gh1o))
    assert np.max(np.abs(gL - gLo)) <= 1e-10 * np.max(np.abs(gLo))


def test_recorded_and_recomputed_tapes_agree(monkeypatch):
    """Every block's tapes recorded in the forward sweep, and each block
    recomputed from its checkpoint in the reverse sweep (tape memory limit 0):
    the same value and gradients, bit for bit."""
    print("[adjoint] recorded vs recomputed tapes, 3 blocks x 4 steps", flush=True)
    c = _case(seed=3, dt=0.05, n_blocks=3, n_steps=4)
    F1, g1, L1 = _gpu_vjp(c)
    monkeypatch.setenv("AFQ_SEGMENT_TAPE_LIMIT", "0")
    F2, g2, L2 = _gpu_vjp(c)
    assert F1 == F2
    assert np.array_equal(g1, g2) and np.array_equal(L1, L2)


def _uhf_job(h0, h1, chol, mo_up, mo_dn, n_walkers, dt, n_exp_terms=6):
    from ad_afqmc_amd.engine import EngineJob

    N = mo_up.shape[0]
    job = EngineJob(N, (mo_up.shape[1], mo_dn.shape[1]), chol.shape[0], n_walkers, "unrestricted", "uhf", n_exp_terms,
                    dt, "fp64")
    job.set_hamiltonian(float(h0), h1, chol)
    job.set_trial(mo_up, mo_dn, np.array([mo_up @ mo_up.T, mo_dn @ mo_dn.T]))
    job.build_measurement()
    job.build_propagation()
    return job


def _gpu_vjp_uhf(c, want_grad=True):
    job = _uhf_job(c["h0"], c["h1"], c["chol"], c["mo_up"], c["mo_dn"], c["up"].shape[0], c["dt"])
    try:
        F, gh1, gL = job.segment_energy_vjp(_dev(c["up"], torch.complex128), _dev(c["weights"]), c["e_estimate"],
                                            _dev(c["fields"]), want_grad=want_grad,
                                            walkers_dn=_dev(c["dn"], torch.complex128))
        torch.cuda.synchronize()
        return (float(F.item()), None if gh1 is None else gh1.cpu().numpy(), None if gL is None else gL.cpu().numpy())
    finally:
        job.close()


def test_uhf_vjp_with_equal_spins_matches_rhf():
    """A UHF trial with both spins equal to an RHF trial, and unrestricted
    walkers with equal spin panels, is the RHF segment: the unrestricted
    reverse pass (spin loops, per-spin factors, the Coulomb trace coupling the
    spins) reproduces the restricted one, which the oracle pins."""
    print("[adjoint] UHF with equal spins against RHF, 3 blocks x 4 steps", flush=True)
    c = _case(seed=3, dt=0.05, n_blocks=3, n_steps=4)
    F, gh1, gL = _gpu_vjp(c)
    cu = dict(c, mo_up=c["mo"], mo_dn=c["mo"], up=c["walkers"], dn=c["walkers"])
    Fu, gh1u, gLu = _gpu_vjp_uhf(cu)
    assert abs(Fu - F) <= 1e-12 * abs(F)
    assert np.max(np.abs(gh1u - gh1)) <= 1e-10 * np.max(np.abs(gh1))
    assert np.max(np.abs(gLu - gL)) <= 1e-10 * np.max(np.abs(gL))


def test_uhf_vjp_matches_central_differences():
    """UHF trial (O = 4 / 3, distinct spin orbitals), unrestricted walkers:
    the gradient against central differences of the device forward in random
    per-spin h1 and symmetric Cholesky directions."""
    print("[adjoint] UHF N=12 O=4/3 G=10 W=8, 2 blocks x 3 steps", flush=True)
    N, G, W = 12, 10, 8
    sysm = syn.make_system(N, 4, G, nocc_dn=3, uhf=True, seed=2, l_scale=0.3 / math.sqrt(N))
    up = syn.perturbed_walkers(sysm.mo_up, W, 0.05, seed=3)
    dn = syn.perturbed_walkers(sysm.mo_dn, W, 0.05, seed=4)
    rng = np.random.default_rng(9)
    c = dict(h0=sysm.h0, h1=np.array(sysm.h1), chol=np.array(sysm.chol).reshape(G, N * N), mo_up=sysm.mo_up,
             mo_dn=sysm.mo_dn, up=up, dn=dn, weights=rng.uniform(0.5, 1.5, W), fields=rng.standard_normal((2, 3, W, G)),
             dt=0.01)
    job = _uhf_job(c["h0"], c["h1"], c["chol"], c["mo_up"], c["mo_dn"], W, c["dt"])
    try:
        e = job.energy(_dev(up, torch.complex128), _dev(dn, torch.complex128))
        c["e_estimate"] = float(torch.mean(e.real))
    finally:
        job.close()
    F, gh1, gL = _gpu_vjp_uhf(c)
    assert math.isfinite(F)
    h = 1e-5
    for trial in range(2):
        D = rng.standard_normal(c["h1"].shape)
        cp, cm = dict(c), dict(c)
        cp["h1"], cm["h1"] = c["h1"] + h * D, c["h1"] - h * D
        fd = (_gpu_vjp_uhf(cp, False)[0] - _gpu_vjp_uhf(cm, False)[0]) / (2 * h)
        an = float(np.sum(gh1 * D))
        print(f"[adjoint] UHF h1 direction {trial}: fd={fd:.12e} an={an:.12e}", flush=True)
        assert abs(an) > 1e-3  # a derivative the check can resolve
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
        E = rng.standard_normal((G, N, N))
        E = ((E + E.transpose(0, 2, 1)) / 2).reshape(G, N * N) * 0.01
        cp, cm = dict(c), dict(c)
        cp["chol"], cm["chol"] = c["chol"] + h * E, c["chol"] - h * E
        fd = (_gpu_vjp_uhf(cp, False)[0] - _gpu_vjp_uhf(cm, False)[0]) / (2 * h)
        an = float(np.sum(gL * E))
        print(f"[adjoint] UHF chol direction {trial}: fd={fd:.12e} an={an:.12e}", flush=True)
        assert abs(an) > 1e-5
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))


def _cfg5():
    N, Oc, G, W = 100, 20, 500, 256
    s = syn.make_system(N, Oc, G, seed=0, l_scale=0.3 / math.sqrt(N))
    walkers = syn.perturbed_walkers(s.mo_up, W, 0.3 / math.sqrt(N), seed=1)
    rng = np.random.default_rng(5)
    fields = rng.standard_normal((5, 10, W, G))
    c = dict(h0=s.h0, h1=np.array(s.h1), chol=np.array(s.chol).reshape(G, N * N), mo=s.mo_up, walkers=walkers,
             weights=np.ones(W), fields=fields, dt=0.005)
    from tests.test_oracle_adjoint import O as Oo

    t = Oo.RHFTrial(N, (Oc, Oc), s.mo_up)
    _, rh, rc = Oo.rhf_measurement_intermediates(c["h1"], c["chol"], s.mo_up)
    c["e_estimate"] = float(np.mean(np.real(t.calc_energy(walkers[:16], s.h0, rh, rc))))
    return c


def test_vjp_matches_central_differences_n100_g500_w256():
    print("[adjoint] N=100 O=20 G=500 W=256, 5 blocks x 10 steps", flush=True)
    c = _cfg5()
    F, gh1, gL = _gpu_vjp(c)
    assert math.isfinite(F)
    rng = np.random.default_rng(11)
    h = 1e-5
    for trial in range(2):
        D = rng.standard_normal(c["h1"].shape)
        cp, cm = dict(c), dict(c)
        cp["h1"], cm["h1"] = c["h1"] + h * D, c["h1"] - h * D
        fd = (_gpu_vjp(cp, False)[0] - _gpu_vjp(cm, False)[0]) / (2 * h)
        an = float(np.sum(gh1 * D))
        print(f"[adjoint] h1 direction {trial}: fd={fd:.12e} an={an:.12e}", flush=True)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
        E = rng.standard_normal((c["chol"].shape[0], 100, 100))
        E = ((E + E.transpose(0, 2, 1)) / 2).reshape(c["chol"].shape) * 0.01
        cp, cm = dict(c), dict(c)
        cp["chol"], cm["chol"] = c["chol"] + h * E, c["chol"] - h * E
        fd = (_gpu_vjp(cp, False)[0] - _gpu_vjp(cm, False)[0]) / (2 * h)
        an = float(np.sum(gL * E))
        print(f"[adjoint] chol direction {trial}: fd={fd:.12e} an={an:.12e}", flush=True)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))


def test_autograd_function_and_rdm1_agree():
    from ad_afqmc_amd.adjoint import segment_energy, segment_rdm1

    c = _case(seed=7)
    h1 = _dev(c["h1"]).requires_grad_()
    chol = _dev(c["chol"]).requires_grad_()
    F = segment_energy(h1, chol, c["mo"], c["walkers"], c["weights"], c["e_estimate"], c["fields"], c["dt"], h0=c["h0"])
    (3.0 * F).backward()
    F2, gh1, gL = segment_rdm1(c["h1"], c["chol"], c["mo"], c["walkers"], c["weights"], c["e_estimate"], c["fields"],
                               c["dt"], h0=c["h0"])
    assert abs(float(F.detach()) - F2) <= 1e-14 * abs(F2)
    assert torch.allclose(h1.grad, 3.0 * gh1, rtol=0, atol=1e-14)
    assert torch.allclose(chol.grad, 3.0 * gL, rtol=0, atol=1e-14)


def test_vjp_rejects_reduced_precision_job():
    from ad_afqmc_amd.engine import EngineJob

    c = _case()
    N, Oc = c["mo"].shape
    job = EngineJob(N, (Oc, Oc), c["chol"].shape[0], 5, "restricted", "rhf", 6, c["dt"], "fp32")
    try:
        d = c["mo"] @ c["mo"].T
        job.set_hamiltonian(c["h0"], c["h1"], c["chol"])
        job.set_trial(c["mo"], None, np.array([d, d]))
        job.build_measurement()
        job.build_propagation()
        with pytest.raises(nat.AfqError) as e:
            job.segment_energy_vjp(_dev(c["walkers"], torch.complex128), _dev(c["weights"]), 0.0, _dev(c["fields"]))
        assert e.value.code == -5
    finally:
        job.close()

"""Device reverse pass of the phaseless segment (afq_segment_energy_vjp):
against the oracle's reverse pass (itself pinned by central differences) on a
small system, and against central differences of the device forward at
N=100, O=20, G=500, W=256 over 50 steps checkpointed every 10."""
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

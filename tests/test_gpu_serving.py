"""Serving on the GPU: one served job's block energies against the oracle's
block loop on the same fields and comb offsets, and the 64-job Poisson stream
(N ~ U{30..250}, O = N/5, G = 5N, W ~ U{64..512}, 300 steps, 4 jobs/s,
bf16x3, 200 GB admission cap) with every job finite and sampled jobs bitwise
equal alone and in the mix."""
import numpy as np
import pytest
import torch

from ad_afqmc_amd import engine as E
from ad_afqmc_amd import serving
from ad_afqmc_amd.serving import GpuExecutor, JobSpec, Scheduler, poisson_stream, summarize
from oracle import afqmc_oracle as O
from tests.test_gpu_parity import TOL, npy, rel

pytestmark = pytest.mark.gpu


class RecordingExecutor(GpuExecutor):
    """Keeps each job's initial walkers and scalars for the oracle replay."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.initial = {}

    def start(self, spec):
        h = super().start(spec)
        h["stream"].synchronize()
        self.initial[spec.job_id] = dict(up=npy(h["up"]), overlaps=npy(h["overlaps"]), scalars=npy(h["scalars"]))
        return h


def test_served_job_matches_oracle_block_loop():
    nsteps = 10
    spec = JobSpec(job_id=7, norb=30, nocc=6, nchol=150, n_walkers=64, n_steps=60, seed=31, precision="fp64")
    ex = RecordingExecutor(n_prop_steps=nsteps)
    res = Scheduler(ex, hbm_cap_bytes=50e9).run([spec], realtime=False)[spec.job_id]
    assert res.error is None
    blocks = res.block_energies
    assert blocks.shape == (6, 2)
    # the same system, drawn again from the job's seed
    sysd = {k: npy(v) for k, v in serving._device_system(spec).items()}
    N, Oc, G, W = spec.norb, spec.nocc, spec.nchol, spec.n_walkers
    h1s, rh, rc = O.rhf_measurement_intermediates(sysd["h1"], sysd["chol"], sysd["mo"])
    pi = O.propagation_intermediates(0.0, h1s, sysd["chol"], sysd["rdm1"], spec.dt)
    ham = dict(h0=0.0, chol=sysd["chol"], rot_h1=rh, rot_chol=rc, mf_shifts=pi["mf_shifts"], h0_prop=pi["h0_prop"],
               exp_h1=pi["exp_h1"])
    trial = O.RHFTrial(N, (Oc, Oc), sysd["mo"])
    init = ex.initial[spec.job_id]
    e0 = float(init["scalars"][0])
    st = dict(walkers=init["up"], weights=np.ones(W), overlaps=trial.calc_overlap(init["up"]), e_estimate=e0,
              pop_control_ene_shift=e0)
    assert rel(init["overlaps"], st["overlaps"]) < TOL
    e_init = trial.calc_energy(init["up"], 0.0, rh, rc).real
    assert abs(np.mean(e_init) - e0) < TOL * max(1.0, abs(e0))
    for b in range(blocks.shape[0]):
        fields = npy(E.normal_fill(nsteps * W * G, spec.seed, b)).reshape(nsteps, W, G)
        st, be, bw = O.block_scan(trial, st, fields, ham, spec.dt)
        assert abs(blocks[b, 0] - be) < TOL * max(1.0, abs(be)), (b, blocks[b, 0], be)
        assert abs(blocks[b, 1] - bw) < TOL * max(1.0, abs(bw)), (b, blocks[b, 1], bw)
        zeta = float(npy(E.sr_zeta(spec.seed, b))[0])
        st["walkers"], st["weights"] = O.stochastic_reconfiguration(st["walkers"], st["weights"], zeta)
        st["overlaps"] = trial.calc_overlap(st["walkers"])
        # the served job's energy estimate follows its block energies (driver.py:168)
        st["e_estimate"] = 0.9 * st["e_estimate"] + 0.1 * be
    assert abs(res.energy - np.sum(blocks[:, 0] * blocks[:, 1]) / np.sum(blocks[:, 1])) < 1e-12 * abs(res.energy)


def test_poisson_stream_of_64_jobs():
    specs = poisson_stream(64, rate=4.0, seed=0, precision="bf16x3")
    sch = Scheduler(GpuExecutor(n_prop_steps=10), hbm_cap_bytes=200e9, max_active=64)
    torch.cuda.synchronize()
    import time

    t0 = time.perf_counter()
    res = sch.run(specs, realtime=True)
    torch.cuda.synchronize()
    summ = summarize(res, time.perf_counter() - t0)
    print(f"serving stream: {summ}, max concurrent {sch.max_concurrent}, peak admitted {sch.peak_used / 1e9:.2f} GB")
    assert summ["completed"] == 64 and summ["refused"] == 0
    assert sch.peak_used <= 200e9 and sch.max_concurrent >= 2
    for r in res.values():
        assert r.error is None and r.admitted >= r.arrival - 1e-9 and r.finished > r.admitted
        assert r.block_energies.shape == (30, 2) and np.all(np.isfinite(r.block_energies))
        assert np.all(r.block_energies[:, 1] > 0)
    # three jobs (smallest, median and largest walker-steps x N^2) alone: bitwise equal
    size = sorted(specs, key=lambda s: s.n_walkers * s.norb**2)
    for s in (size[0], size[len(size) // 2], size[-1]):
        alone = Scheduler(GpuExecutor(n_prop_steps=10), hbm_cap_bytes=200e9).run([s], realtime=False)[s.job_id]
        assert np.array_equal(alone.block_energies, res[s.job_id].block_energies), s.job_id

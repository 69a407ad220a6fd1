"""Serving loop: admission / continuous batching on the CPU (fake executor),
and on the GPU that every job's result is identical alone and in the mix."""
import math

import numpy as np
import pytest

from ad_afqmc_amd.serving import JobSpec, Scheduler, footprint_bytes, poisson_stream, summarize


class FakeExecutor:
    """Each job needs `blocks` issue() calls; a block finishes on the next poll."""

    def __init__(self, blocks=3):
        self.blocks = blocks
        self.log = []
        self.live = set()

    def start(self, spec):
        self.live.add(spec.job_id)
        self.log.append(("start", spec.job_id, len(self.live)))
        return {"spec": spec, "issued": 0}

    def ready(self, h):
        return h["issued"] < self.blocks

    def issue(self, h):
        h["issued"] += 1

    def done(self, h):
        return h["issued"] >= self.blocks

    def finish(self, h):
        return float(h["spec"].job_id), np.zeros((self.blocks, 2)), 10

    def release(self, h):
        self.live.discard(h["spec"].job_id)
        self.log.append(("release", h["spec"].job_id, len(self.live)))


def fake_fp(spec, nsteps):
    return spec.norb * 1_000_000  # 1 MB per orbital


def test_admission_respects_cap_and_refuses_oversize():
    specs = [JobSpec(job_id=i, norb=n, nocc=2, nchol=10, n_walkers=8, arrival=0.0) for i, n in enumerate([30, 40, 50, 500])]
    ex = FakeExecutor()
    sch = Scheduler(ex, hbm_cap_bytes=100e6, max_active=8, footprint=fake_fp)
    res = sch.run(specs, realtime=False)
    assert res[3].error is not None and "exceeds the HBM cap" in res[3].error
    assert all(res[i].error is None and res[i].energy == float(i) for i in range(3))
    assert sch.peak_used <= 100e6
    # 30+40 fit together, 50 must wait for a release
    assert sch.max_concurrent == 2


def test_continuous_batching_admits_as_slots_free():
    specs = [JobSpec(job_id=i, norb=10, nocc=2, nchol=10, n_walkers=8, arrival=0.0) for i in range(6)]
    ex = FakeExecutor(blocks=2)
    sch = Scheduler(ex, hbm_cap_bytes=1e12, max_active=2, footprint=fake_fp)
    res = sch.run(specs, realtime=False)
    assert all(r.error is None for r in res.values())
    live = [n for (what, _, n) in ex.log if what == "start"]
    assert max(live) == 2
    # every job after the first two starts right after some job released its slot
    order = [(w, j) for (w, j, _) in ex.log]
    assert order.index(("start", 2)) > min(order.index(("release", 0)), order.index(("release", 1)))
    s = summarize(res, wall=1.0)
    assert s["completed"] == 6 and s["refused"] == 0


def test_realtime_arrivals_and_latency_summary():
    specs = [JobSpec(job_id=i, norb=10, nocc=2, nchol=10, n_walkers=8, arrival=0.01 * i) for i in range(4)]
    ex = FakeExecutor(blocks=1)
    res = Scheduler(ex, hbm_cap_bytes=1e12, footprint=fake_fp).run(specs, realtime=True)
    for r in res.values():
        assert r.admitted >= r.arrival - 1e-9 and r.finished >= r.admitted
    s = summarize(res, wall=0.05)
    assert s["p99_time_to_energy_s"] >= s["p50_time_to_energy_s"] >= 0


def test_footprint_uses_library_sizes():
    a = footprint_bytes(JobSpec(job_id=0, norb=100, nocc=20, nchol=500, n_walkers=64), 10)
    b = footprint_bytes(JobSpec(job_id=0, norb=100, nocc=20, nchol=500, n_walkers=128), 10)
    assert 0 < a < b
    with pytest.raises(ValueError):
        footprint_bytes(JobSpec(job_id=0, norb=10, nocc=20, nchol=5, n_walkers=4), 10)


def test_poisson_stream_shapes():
    s = poisson_stream(20, rate=4.0, seed=3)
    assert len(s) == 20 and all(30 <= j.norb <= 250 and j.nocc == j.norb // 5 and j.nchol == 5 * j.norb for j in s)
    t = [j.arrival for j in s]
    assert t == sorted(t) and 1.0 < t[-1] < 20.0


@pytest.mark.gpu
def test_results_identical_alone_and_in_mix():
    from ad_afqmc_amd.serving import GpuExecutor

    specs = [
        JobSpec(job_id=0, norb=40, nocc=8, nchol=200, n_walkers=64, n_steps=40, seed=11),
        JobSpec(job_id=1, norb=64, nocc=12, nchol=320, n_walkers=96, n_steps=40, seed=12, precision="fp32"),
        JobSpec(job_id=2, norb=30, nocc=6, nchol=150, n_walkers=32, n_steps=40, seed=13, precision="bf16x3"),
    ]
    alone = {}
    for s in specs:
        r = Scheduler(GpuExecutor(n_prop_steps=10), hbm_cap_bytes=50e9).run([s], realtime=False)
        alone[s.job_id] = r[s.job_id]
    mix = Scheduler(GpuExecutor(n_prop_steps=10), hbm_cap_bytes=50e9).run(specs, realtime=False)
    for s in specs:
        a, m = alone[s.job_id], mix[s.job_id]
        assert a.error is None and m.error is None
        assert np.all(np.isfinite(a.block_energies))
        # same kernels on the same inputs in the same order: bitwise equal
        assert np.array_equal(a.block_energies, m.block_energies), s.job_id
        assert a.energy == m.energy

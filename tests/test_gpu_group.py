"""Grouped blocks (afq_group_sr_block): several bf16x3 jobs stepped by one
launch per kernel, against each job's own afq_sr_block (the single-job block,
itself checked against the oracle's block loop), and the grouped serving
executor: every job's block energies identical alone and in the mix."""
import numpy as np
import pytest
import torch

from ad_afqmc_amd import serving
from ad_afqmc_amd.engine import JobGroup
from ad_afqmc_amd.serving import GpuExecutor, GroupedExecutor, JobSpec, Scheduler

pytestmark = pytest.mark.gpu

SPECS = [
    JobSpec(job_id=0, norb=30, nocc=6, nchol=150, n_walkers=64, n_steps=30, seed=11, precision="bf16x3"),
    JobSpec(job_id=1, norb=52, nocc=10, nchol=260, n_walkers=96, n_steps=30, seed=12, precision="bf16x3"),
    JobSpec(job_id=2, norb=75, nocc=17, nchol=375, n_walkers=130, n_steps=30, seed=13, precision="bf16x3"),
    # three column blocks per propagator slab and two row blocks per wave
    JobSpec(job_id=3, norb=140, nocc=40, nchol=700, n_walkers=64, n_steps=30, seed=14, precision="bf16x3"),
]


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def test_group_matches_single_job_blocks():
    nsteps, nblocks = 5, 3
    single = GpuExecutor(n_prop_steps=nsteps, use_graphs=False)
    grouped = GpuExecutor(n_prop_steps=nsteps, use_graphs=False)
    hs = [single.start(s) for s in SPECS]
    hg = [grouped.start(s) for s in SPECS]
    torch.cuda.synchronize()
    for a, b in zip(hs, hg):
        assert torch.equal(a["up"], b["up"]) and torch.equal(a["overlaps"], b["overlaps"])
    group = JobGroup([(h["job"], h["up"], h["weights"], h["overlaps"], h["scalars"], h["counter"], h["out"],
                       h["spec"].seed) for h in hg], nsteps)
    for blk in range(nblocks):
        ref = []
        for h in hs:
            single._block(h)  # afq_sr_block, then e_estimate <- 0.9 e + 0.1 E_block
            ref.append(h["out"].clone())
        group.sr_block()
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(hs, hg)):
            assert _rel(b["out"].cpu(), ref[i].cpu()) < 1e-10, (blk, i)
            assert _rel(b["weights"].cpu(), a["weights"].cpu()) < 1e-10, (blk, i)
            assert _rel(b["overlaps"].cpu(), a["overlaps"].cpu()) < 1e-10, (blk, i)
            assert _rel(b["up"].cpu(), a["up"].cpu()) < 1e-10, (blk, i)
            assert _rel(b["scalars"].cpu(), a["scalars"].cpu()) < 1e-10, (blk, i)
            assert int(b["counter"].item()) == blk + 1
        assert np.all([np.isfinite(h["out"].cpu().numpy()).all() for h in hg])
    group.close()
    for h in hs:
        single.release(h)
    for h in hg:
        grouped.release(h)


def test_group_rejects_other_precisions():
    ex = GpuExecutor(n_prop_steps=5, use_graphs=False)
    spec = JobSpec(job_id=5, norb=30, nocc=6, nchol=150, n_walkers=64, n_steps=10, seed=3, precision="fp64")
    h = ex.start(spec)
    with pytest.raises(RuntimeError, match="split-BF16"):
        JobGroup([(h["job"], h["up"], h["weights"], h["overlaps"], h["scalars"], h["counter"], h["out"], 3)], 5)
    ex.release(h)


@pytest.mark.parametrize("solo_above", [None, 100], ids=["all-grouped", "large-solo"])
def test_grouped_serving_alone_and_in_mix(solo_above):
    specs = [JobSpec(job_id=i, norb=n, nocc=max(1, n // 5), nchol=5 * n, n_walkers=w, n_steps=st, seed=100 + i,
                     precision="bf16x3")
             for i, (n, w, st) in enumerate([(30, 64, 60), (64, 100, 40), (96, 160, 50), (48, 80, 30), (130, 96, 30)])]
    # two size classes (N <= 100 and N > 100), up to two rounds in flight per
    # class; or the N = 130 job on its own stream
    ex = GroupedExecutor(n_prop_steps=10, solo_above=solo_above)
    sch = Scheduler(ex, hbm_cap_bytes=50e9, max_active=8)
    res = sch.run(specs, realtime=False)
    assert ex.rounds >= 6 and sch.max_concurrent >= 2
    for s in specs:
        r = res[s.job_id]
        assert r.error is None and r.block_energies.shape == (s.n_steps // 10, 2)
        assert np.all(np.isfinite(r.block_energies)) and np.all(r.block_energies[:, 1] > 0)
        alone = Scheduler(GroupedExecutor(n_prop_steps=10, solo_above=solo_above), hbm_cap_bytes=50e9).run(
            [s], realtime=False)[s.job_id]
        assert np.array_equal(alone.block_energies, r.block_energies), s.job_id

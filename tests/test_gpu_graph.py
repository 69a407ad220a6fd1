"""Graph capture of the device block loop: afq_sr_block recorded once into a
HIP graph and replayed gives the same walkers, weights and block energies, bit
for bit, as the same calls made eagerly -- with eager work and host
synchronisation between the replays."""
import math

import numpy as np
import pytest
import torch

from ad_afqmc_amd import hamiltonian as hm
from ad_afqmc_amd import propagation as pr
from ad_afqmc_amd import synthetic as syn
from ad_afqmc_amd import wavefunctions as wf
from ad_afqmc_amd.binding import engine_for

pytestmark = pytest.mark.gpu


def _state(precision):
    N, O, G, W = 30, 6, 150, 32
    s = syn.make_system(N, O, G, seed=4, l_scale=1.0 / math.sqrt(N))
    wave, ham = s.wave_data(), s.ham_data()
    trial = wf.rhf(N, (O, O))
    prop = pr.propagator_afqmc(dt=0.005, n_walkers=W, walker_type="restricted", precision=precision)
    H = hm.hamiltonian(N)
    ham = H.build_measurement_intermediates(ham, trial, wave)
    ham = H.build_propagation_intermediates(ham, prop, trial, wave)
    pd = prop.init_prop_data(trial, wave, ham, seed=0)
    job = engine_for(trial, wave, ham, "restricted", W, prop)
    up = pd["walkers"].data.clone()
    weights = pd["weights"].clone()
    ov = job.overlap(up, None)
    e0 = pd["e_estimate"].reshape(())
    sc = torch.stack([e0, e0.clone()]).contiguous()
    return job, wave, up, weights, ov, sc


def _run(precision, use_graph, nsteps=5, nblocks=5):
    job, wave, up, weights, ov, sc = _state(precision)
    job.reserve_fields(nsteps)
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    outs = []
    call = lambda: job.sr_block(up, None, weights, ov, sc, nsteps, 21, counter, out)
    call()
    outs.append(out.clone())
    if use_graph:
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.graph(g, stream=side):
            call()
        torch.cuda.current_stream().wait_stream(side)
        for _ in range(nblocks - 1):
            g.replay()
            outs.append(out.clone())
            _ = torch.ones(1000, device="cuda").cumsum(0)  # unrelated eager work between replays
            torch.cuda.synchronize()
    else:
        for _ in range(nblocks - 1):
            call()
            outs.append(out.clone())
    torch.cuda.synchronize()
    res = (torch.stack(outs).cpu(), up.cpu(), weights.cpu(), ov.cpu(), int(counter.item()))
    job.close()
    return res


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_sr_block_graph_replays_match_eager(precision):
    print(f"[graph] {precision}", flush=True)
    a = _run(precision, False)
    b = _run(precision, True)
    assert a[4] == b[4] == 5
    assert torch.all(torch.isfinite(a[0]))
    for x, y in zip(a[:4], b[:4]):
        assert torch.equal(x, y)


def test_sr_block_matches_block_plus_sr_with_the_same_stream():
    """afq_sr_block's fields are afq_block's stream for the same counter."""
    job, wave, up, weights, ov, sc = _state("fp64")
    job.reserve_fields(4)
    up2, w2, ov2, sc2 = up.clone(), weights.clone(), ov.clone(), sc.clone()
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    job.sr_block(up, None, weights, ov, sc, 4, 33, counter, out)
    o2, c2 = job.block(up2, None, w2, ov2, sc2, 4, 33, 0)
    torch.cuda.synchronize()
    assert c2 == 1 and int(counter.item()) == 1
    assert torch.equal(out, o2)  # same fields, same block
    job.close()

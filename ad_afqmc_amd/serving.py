"""Continuous-batching service for many AFQMC jobs on one GPU.

Requests (jobs of different N, O, G and walker counts) arrive over time.  A
job is admitted as soon as a slot and its HBM footprint (as libafq.so would
allocate it, `afq_job_footprint`) fit under the cap; admitted jobs keep their
Hamiltonian, rotated Cholesky tensor and walkers resident and each runs its
device blocks (afq_sr_block: phaseless steps, QR, energy, then comb
resampling, replayed from a captured HIP graph) on its own HIP stream, so work
of different jobs overlaps on the device.  A
job holds at most one block in flight; the host issues the next block of every
job whose previous block has finished, admits newly arrived jobs in between,
and retires finished jobs (freeing their memory and slot) as soon as their
last block completes.  The reference runs many-job workloads (LNO fragments)
one after another through separate processes (ad_afqmc/lnoutils.py:175-177).

Each job's arithmetic depends only on its own inputs and seed, so its energy is
the same whether it runs alone or in the mix.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np


@dataclass
class JobSpec:
    job_id: int
    norb: int
    nocc: int
    nchol: int
    n_walkers: int
    n_steps: int = 300
    dt: float = 0.005
    seed: int = 0
    precision: str = "fp64"
    arrival: float = 0.0  # seconds after the start of the stream
    l_scale: Optional[float] = None  # Cholesky scale (default 0.3/sqrt(norb): a live walker population)


@dataclass
class JobResult:
    job_id: int
    arrival: float
    admitted: float = math.nan
    finished: float = math.nan
    energy: float = math.nan
    block_energies: Optional[np.ndarray] = None
    walker_steps: int = 0
    error: Optional[str] = None

    @property
    def time_to_energy(self) -> float:
        return self.finished - self.arrival


def footprint_bytes(spec: JobSpec, n_prop_steps: int) -> int:
    """Device bytes a running job holds: libafq.so's own buffers (Hamiltonian,
    trial, intermediates, scratch and the field buffer) plus the caller-owned
    walkers, weights, overlaps and scalars.  The executor keeps nothing else
    per job once it is set up."""
    from . import _native as nat

    d = nat.JobDesc(
        norb=spec.norb, nelec_up=spec.nocc, nelec_dn=spec.nocc, nchol=spec.nchol, n_walkers=spec.n_walkers,
        walker_type=nat.AFQ_WALKERS_RESTRICTED, trial_type=nat.AFQ_TRIAL_RHF, n_exp_terms=6,
        precision=nat.PRECISIONS[spec.precision], reserved=0, dt=spec.dt,
    )
    b = nat.lib().afq_job_footprint(C.byref(d), int(n_prop_steps))
    if b < 0:
        raise ValueError(f"job {spec.job_id}: invalid sizes ({nat.lib().afq_last_error().decode()})")
    walkers = spec.n_walkers * spec.norb * spec.nocc * 16
    return int(b + walkers + spec.n_walkers * 24 + 4096)


class Scheduler:
    """Admission and issue logic, independent of the device.

    executor must provide start(spec) -> handle, ready(handle) -> bool (the
    previous block finished), issue(handle) -> None (enqueue the next block),
    done(handle) -> bool (all blocks issued and finished), finish(handle) ->
    (energy, block_energies, walker_steps), release(handle)."""

    def __init__(self, executor, hbm_cap_bytes: float = 200e9, max_active: int = 64, n_prop_steps: int = 10,
                 footprint: Callable[[JobSpec, int], int] = footprint_bytes, clock=time.perf_counter):
        self.ex = executor
        self.cap = float(hbm_cap_bytes)
        self.max_active = int(max_active)
        self.n_prop_steps = int(n_prop_steps)
        self.footprint = footprint
        self.clock = clock
        self.used = 0
        self.peak_used = 0
        self.max_concurrent = 0

    def run(self, specs: List[JobSpec], realtime: bool = True) -> Dict[int, JobResult]:
        pending = deque(sorted(specs, key=lambda s: s.arrival))
        waiting: deque = deque()
        active = {}  # job_id -> (spec, handle, bytes)
        # without realtime every job is there at t = 0, and its latency counts from then
        results: Dict[int, JobResult] = {s.job_id: JobResult(s.job_id, s.arrival if realtime else 0.0) for s in specs}
        t0 = self.clock()
        now = lambda: (self.clock() - t0)
        while pending or waiting or active:
            t = now()
            while pending and (pending[0].arrival <= t or not realtime):
                waiting.append(pending.popleft())
            # admission, first come first served; a job that can never fit is refused
            while waiting:
                spec = waiting[0]
                try:
                    need = self.footprint(spec, self.n_prop_steps)
                except ValueError as e:
                    results[spec.job_id].error = str(e)
                    waiting.popleft()
                    continue
                if need > self.cap:
                    results[spec.job_id].error = f"footprint {need / 1e9:.2f} GB exceeds the HBM cap {self.cap / 1e9:.2f} GB"
                    waiting.popleft()
                    continue
                if self.used + need > self.cap or len(active) >= self.max_active:
                    break
                waiting.popleft()
                handle = self.ex.start(spec)
                self.used += need
                self.peak_used = max(self.peak_used, self.used)
                active[spec.job_id] = (spec, handle, need)
                results[spec.job_id].admitted = now()
            self.max_concurrent = max(self.max_concurrent, len(active))
            # issue the next block of every job whose previous one finished
            progressed = False
            for jid, (spec, h, need) in list(active.items()):
                if self.ex.done(h):
                    e, blocks, ws = self.ex.finish(h)
                    r = results[jid]
                    r.finished, r.energy, r.block_energies, r.walker_steps = now(), e, blocks, ws
                    self.ex.release(h)
                    self.used -= need
                    del active[jid]
                    progressed = True
                elif self.ex.ready(h):
                    self.ex.issue(h)
                    progressed = True
            if not progressed:
                if not active and pending and realtime:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
                else:
                    time.sleep(0.0002)
        return results


class GpuExecutor:
    """Runs each admitted job on its own HIP stream through libafq.so.

    A job's blocks are `afq_sr_block` calls (block + comb resampling, stream
    counter in device memory).  With use_graphs, the first block runs eagerly
    (it also warms up the library) and the second is captured into a HIP graph,
    which every later block replays: one graph launch per block instead of a
    few hundred kernel launches from the host."""

    def __init__(self, n_prop_steps: int = 10, use_graphs: bool = True):
        import torch

        self.torch = torch
        self.n_prop_steps = int(n_prop_steps)
        self.use_graphs = bool(use_graphs)
        self._setup = None

    def start(self, spec: JobSpec):
        """Builds the job on a setup stream shared by all jobs (so the
        transient generator and intermediate tensors return to one pool of
        torch's caching allocator and are reused by the next job's setup) and
        keeps only what the job's blocks use: the engine job (with its own
        resident copies of the Hamiltonian, trial and intermediates), the
        walkers and the per-job scalars.  Its blocks run on a stream of its
        own that waits for the setup."""
        torch = self.torch
        from . import hamiltonian as hm
        from . import propagation as pr
        from . import wavefunctions as wf
        from .binding import engine_for

        if self._setup is None:
            self._setup = torch.cuda.Stream()
        setup = self._setup
        stream = torch.cuda.Stream()
        with torch.cuda.stream(setup):
            h = _device_system(spec)
            trial = wf.rhf(spec.norb, (spec.nocc, spec.nocc))
            prop = pr.propagator_afqmc(dt=spec.dt, n_walkers=spec.n_walkers, walker_type="restricted",
                                       precision=spec.precision)
            wave = {"mo_coeff": h["mo"], "rdm1": h["rdm1"]}
            ham = {"h0": 0.0, "h1": h["h1"], "chol": h["chol"], "ene0": 0.0}
            H = hm.hamiltonian(spec.norb)
            ham = H.build_measurement_intermediates(ham, trial, wave)
            ham = H.build_propagation_intermediates(ham, prop, trial, wave)
            pd = prop.init_prop_data(trial, wave, ham, seed=spec.seed)
            job = engine_for(trial, wave, ham, "restricted", spec.n_walkers, prop)
            job.reserve_fields(self.n_prop_steps)
            up = pd["walkers"].data.clone()
            weights = pd["weights"].clone()
            overlaps = job.overlap(up, None)
            e0 = pd["e_estimate"].reshape(())
            scalars = torch.stack([e0, e0.clone()]).contiguous()
            counter = torch.zeros(1, dtype=torch.int64, device="cuda")
            out = torch.empty(2, dtype=torch.float64, device="cuda")
            ready = torch.cuda.Event()
            ready.record(setup)
        # the binding in `wave` refers to every uploaded array (chol included):
        # dropping wave, ham, pd and h here returns them to the setup pool
        del h, wave, ham, pd
        stream.wait_event(ready)
        for t in (up, weights, overlaps, scalars, counter, out):
            t.record_stream(stream)
        nblocks = max(1, -(-spec.n_steps // self.n_prop_steps))
        ev = torch.cuda.Event()
        ev.record(stream)
        return dict(spec=spec, stream=stream, job=job, up=up, weights=weights, overlaps=overlaps, scalars=scalars,
                    counter=counter, out=out, graph=None, issued=0, nblocks=nblocks, outs=[], event=ev)

    def ready(self, h) -> bool:
        return h["issued"] < h["nblocks"] and h["event"].query()

    def _block(self, h):
        h["job"].sr_block(h["up"], None, h["weights"], h["overlaps"], h["scalars"], self.n_prop_steps,
                          h["spec"].seed, h["counter"], h["out"])
        # the driver's energy estimate follows the block energy (driver.py:168):
        # e_estimate <- 0.9 e_estimate + 0.1 E_block, on the device
        h["scalars"][:1].mul_(0.9).add_(h["out"][:1], alpha=0.1)

    def issue(self, h):
        torch = self.torch
        stream = h["stream"]
        with torch.cuda.stream(stream):
            if h["graph"] is not None:
                h["graph"].replay()
            elif self.use_graphs and h["issued"] == 1:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=stream):
                    self._block(h)  # recorded, not run
                h["graph"] = g
                g.replay()
            else:
                self._block(h)
            h["outs"].append(h["out"].clone())
            h["event"] = torch.cuda.Event()
            h["event"].record(stream)
        h["issued"] += 1

    def done(self, h) -> bool:
        return h["issued"] >= h["nblocks"] and h["event"].query()

    def finish(self, h):
        blocks = self.torch.stack(h["outs"]).cpu().numpy()
        e = float(np.sum(blocks[:, 0] * blocks[:, 1]) / np.sum(blocks[:, 1]))
        return e, blocks, h["spec"].n_walkers * self.n_prop_steps * h["nblocks"]

    def release(self, h):
        h["stream"].synchronize()
        h["graph"] = None
        h["job"].close()
        h.clear()


class _Lane:
    """One size class of GroupedExecutor: its stream, live jobs, current group
    (and graph), the events of its rounds in flight and retired groups waiting
    for their last round."""

    def __init__(self, stream):
        self.stream = stream
        self.live = []
        self.members = ()
        self.group = None
        self.graph = None
        self.rounds_same = 0
        self.inflight = deque()  # events of issued rounds, oldest first
        self.retired = []  # (group, graph, event of its last round)

    def busy(self) -> int:
        while self.inflight and self.inflight[0].query():
            self.inflight.popleft()
        return len(self.inflight)

    def reap(self):
        """Closes the retired groups whose last round has finished."""
        keep = []
        for grp, graph, ev in self.retired:
            if ev.query():
                grp.close()
            else:
                keep.append((grp, graph, ev))
        self.retired = keep


class GroupedExecutor(GpuExecutor):
    """Runs admitted jobs in groups: every round issues one grouped block
    (afq_group_sr_block: each kernel of the block launched once over all
    member jobs) for every member with blocks left.

    Jobs are sorted into size classes by N (`lanes`: the class bounds), one
    group and one stream per class, so that a class's rounds last about as
    long as its members' blocks and the classes' rounds overlap on the device.
    A class keeps up to `depth` rounds in flight: the next round is queued
    behind the current one instead of waiting for it.  A job admitted while
    rounds are in flight joins the next round; a job whose blocks are all
    issued leaves.  When the member set changes the class's group (the device
    tables of the members' launch parameters) is rebuilt; the old one is
    closed once its last round has finished.  With use_graphs a member set
    that lasts more than one round is captured into a HIP graph on its second
    round and replayed afterwards.  Jobs must be restricted-walker RHF jobs in
    bf16x3 (the grouped block's scope).

    Jobs above `solo_above` orbitals (None: none) are not grouped: they run as
    GpuExecutor runs every job, on their own stream with their own graph; such
    a job's kernels fill the device alone, where grouping the small jobs saves
    their launches and fills the device between them."""

    def __init__(self, n_prop_steps: int = 10, use_graphs: bool = True, lanes=(100, 175), depth: int = 2,
                 solo_above=None):
        super().__init__(n_prop_steps, use_graphs)
        self.bounds = tuple(int(b) for b in lanes)
        self.depth = max(1, int(depth))
        self.solo_above = None if solo_above is None else int(solo_above)
        self._lanes = None
        self.rounds = 0

    def _lane_list(self):
        if self._lanes is None:
            self._lanes = [_Lane(self.torch.cuda.Stream()) for _ in range(len(self.bounds) + 1)]
        return self._lanes

    def _lane(self, spec: JobSpec) -> _Lane:
        k = sum(1 for b in self.bounds if spec.norb > b)
        return self._lane_list()[k]

    def start(self, spec: JobSpec):
        if spec.precision != "bf16x3":
            raise ValueError("grouped serving runs bf16x3 jobs")
        h = super().start(spec)
        if self.solo_above is not None and spec.norb > self.solo_above:
            h["lane"] = None
            return h
        h["lane"] = self._lane(spec)
        h["lane"].live.append(h)
        return h

    def ready(self, h) -> bool:
        if h["lane"] is None:
            return super().ready(h)
        if h["issued"] >= h["nblocks"]:
            return False
        if h["issued"] == 0 and not h["event"].query():
            return False  # setup still running
        return h["lane"].busy() < self.depth

    def _reap(self):
        """Closes retired groups once no class has a round in flight (freeing
        device memory waits for the whole device), or when many have piled up."""
        lanes = self._lane_list()
        if all(lane.busy() == 0 for lane in lanes) or sum(len(lane.retired) for lane in lanes) > 32:
            for lane in lanes:
                lane.reap()

    def _retire(self, lane: _Lane):
        if lane.group is not None:
            ev = self.torch.cuda.Event()
            ev.record(lane.stream)
            lane.retired.append((lane.group, lane.graph, ev))
        lane.group, lane.graph, lane.members, lane.rounds_same = None, None, (), 0

    def issue(self, h):
        """Queues one grouped block for every job of h's class with blocks left
        whose setup has finished (h among them); the other members' issue()
        calls find them already issued."""
        torch = self.torch
        from .engine import JobGroup

        lane = h["lane"]
        self._reap()
        if lane is None:
            super().issue(h)
            return
        if lane.busy() >= self.depth:
            return
        live = [x for x in lane.live if x["issued"] < x["nblocks"] and (x["issued"] > 0 or x["event"].query())]
        if not live:
            return
        members = tuple(id(x) for x in live)
        stream = lane.stream
        with torch.cuda.stream(stream):
            if members != lane.members:
                self._retire(lane)
                for x in live:
                    if x["issued"] == 0:
                        stream.wait_stream(x["stream"])
                lane.group = JobGroup([(x["job"], x["up"], x["weights"], x["overlaps"], x["scalars"], x["counter"],
                                        x["out"], x["spec"].seed) for x in live], self.n_prop_steps)
                lane.members = members
            if lane.graph is not None:
                lane.graph.replay()
            elif self.use_graphs and lane.rounds_same == 1:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=stream):
                    lane.group.sr_block()  # recorded, not run
                lane.graph = g
                g.replay()
            else:
                lane.group.sr_block()
            lane.rounds_same += 1
            for x in live:
                x["outs"].append(x["out"].clone())
                x["issued"] += 1
            ev = torch.cuda.Event()
            ev.record(stream)
            lane.inflight.append(ev)
            for x in live:
                x["event"] = ev
        self.rounds += 1

    def done(self, h) -> bool:
        return h["issued"] >= h["nblocks"] and h["event"].query()

    def release(self, h):
        lane = h["lane"]
        if lane is not None:
            if h in lane.live:
                lane.live.remove(h)
            if id(h) in lane.members:
                # the group refers to this job's buffers: it must not be replayed again
                self._retire(lane)
        self._reap()
        super().release(h)


def _device_system(spec: JobSpec):
    """The synthetic generator of ad_afqmc_amd.synthetic, drawn on the device
    from torch's generator seeded with the job seed (fast for large jobs), with
    the Cholesky-vector scale s_g multiplied by spec.l_scale (0.3/sqrt(N) by
    default; at the unscaled generator every walker is clipped within a few
    steps)."""
    import torch

    N, O, G = spec.norb, spec.nocc, spec.nchol
    gen = torch.Generator(device="cuda").manual_seed(int(spec.seed))
    dev = torch.device("cuda")
    x = torch.randn((N, N), generator=gen, dtype=torch.float64, device=dev)
    h1 = torch.diag(torch.linspace(-2.0, 1.0, N, dtype=torch.float64, device=dev)) + 0.05 * (x + x.T) / 2
    ls = spec.l_scale if spec.l_scale is not None else 0.3 / math.sqrt(N)
    sig = ls * 0.5 * torch.exp(-torch.arange(G, dtype=torch.float64, device=dev) / (0.2 * G))
    a = torch.randn((G, N, N), generator=gen, dtype=torch.float64, device=dev) * sig.view(G, 1, 1)
    chol = ((a + a.transpose(1, 2)) / 2).reshape(G, N * N).contiguous()
    mo = torch.linalg.eigh(h1)[1][:, :O].contiguous()
    d = mo @ mo.T
    return dict(h1=torch.stack([h1, h1]), chol=chol, mo=mo, rdm1=torch.stack([d, d]))


def poisson_stream(n_jobs: int, rate: float, seed: int = 0, nmin: int = 30, nmax: int = 250, wmin: int = 64,
                   wmax: int = 512, n_steps: int = 300, precision: str = "fp64") -> List[JobSpec]:
    """Jobs with N ~ U{nmin..nmax}, O = N/5, G ~ 5N, W ~ U{wmin..wmax}, Poisson arrivals at `rate` per second."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_jobs))
    specs = []
    for i in range(n_jobs):
        N = int(rng.integers(nmin, nmax + 1))
        specs.append(JobSpec(job_id=i, norb=N, nocc=max(1, N // 5), nchol=5 * N, n_walkers=int(rng.integers(wmin, wmax + 1)),
                             n_steps=n_steps, seed=1000 + i, precision=precision, arrival=float(t[i])))
    return specs


def summarize(results: Dict[int, JobResult], wall: float) -> dict:
    ok = [r for r in results.values() if r.error is None and not math.isnan(r.finished)]
    tte = np.array([r.time_to_energy for r in ok]) if ok else np.array([math.nan])
    ws = sum(r.walker_steps for r in ok)
    return {
        "jobs": len(results),
        "completed": len(ok),
        "refused": sum(1 for r in results.values() if r.error is not None),
        "wall_s": wall,
        "walker_steps_per_s": ws / wall if wall > 0 else math.nan,
        "p50_time_to_energy_s": float(np.percentile(tte, 50)),
        "p99_time_to_energy_s": float(np.percentile(tte, 99)),
    }

#!/usr/bin/env python
"""Walker-step throughput of the phaseless-AFQMC engine on one MI355X.

One bench step = one device sampler block (afq_sr_block) on the synthetic UHF
workload where the mixed-precision target is stated (N=300, O=60/58, G=2000,
256 walkers, dt=0.005), in split-BF16 x3 by default: 10 phaseless steps with
on-device fields, QR re-orthonormalisation, overlap and local-energy
measurement, the population-control update, then comb resampling and the
overlaps of the new population.
value = walkers * 10 * steps / wall time of the timed region (walker-steps/s).
The same workload is timed in FP64 over the same number of blocks
(speedup_vs_fp64) and in FP32; the RHF N=200 workload, the H10 chain and the
64-job serving stream are reported beside it.

The generator is the stated one with the Cholesky-vector scale s_g multiplied
by 0.3/sqrt(N), so that the walker population lives and block energies are
defined (at the stated scale every walker is clipped within a few steps; the
dense work per block is the same, which `stated_generator` shows).

Prints one JSON line.  --gpus N > 1 is refused: this engine runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAKS = {  # MI355X published dense peaks
    "fp64": 78.6e12,
    "fp32": 157.3e12,
    "bf16": 2.5e15,
    "hbm": 8.0e12,
}


def flops_per_walker_step(N, Oa, Ob, G, m, one_spin, K=5):
    """Algorithmic real flops per walker-step (complex x real MAC = 4, complex x
    complex MAC = 8), with the measurement (QR + energy) amortised over m steps."""
    occ = [Oa] if one_spin else [Oa, Ob]
    sumO = sum(occ)
    f = {}
    f["vhs"] = 4.0 * G * N * N
    f["taylor"] = 8.0 * K * N * N * sumO
    f["exp_h1"] = 8.0 * N * N * sumO
    green = sum(4.0 * o * o * N + 8.0 * o**3 + 8.0 * N * o * o for o in occ)
    f["green_fb"] = green
    f["force_bias"] = 4.0 * G * N * sumO
    f["overlap"] = sum(4.0 * o * o * N + 8.0 / 3.0 * o**3 for o in occ)
    f["energy"] = (sum(4.0 * G * o * o * N + 8.0 * G * o * o for o in occ) + green) / m
    f["qr"] = sum(8.0 * (2.0 * N * o * o - 2.0 / 3.0 * o**3) for o in occ) / m
    return sum(f.values()), f


def whole_step_roofline(N, Oa, Ob, G, m, one_spin, precision, value):
    """Whole-step roofline: achieved = F_total x walker-steps/s; attainable =
    sum_k F_k / sum_k (F_k / P_k) with P_k the peak of the precision contraction
    k runs in (bf16x3 counted at a third of the bf16 peak).  In the reduced modes
    the VHS, Taylor, force-bias and energy contractions run reduced; exp(-dt/2
    h1) runs on FP32 (on FP64 at 33..48 walker columns); the Green's function,
    overlap and QR stay FP64."""
    f_total, f = flops_per_walker_step(N, Oa, Ob, G, m, one_spin)
    red = {"fp64": PEAKS["fp64"], "fp32": PEAKS["fp32"], "bf16x3": PEAKS["bf16"] / 3.0}[precision]
    cols = Oa if one_spin else Oa  # exp(-dt/2 h1) runs per spin buffer
    peaks = {"vhs": red, "taylor": red, "force_bias": red, "energy": red,
             "exp_h1": PEAKS["fp64"] if precision == "fp64" or 32 < cols <= 48 else PEAKS["fp32"],
             "green_fb": PEAKS["fp64"], "overlap": PEAKS["fp64"], "qr": PEAKS["fp64"]}
    attainable = f_total / sum(fk / peaks[k] for k, fk in f.items())
    achieved = f_total * value
    return {"achieved": achieved / 1e12, "attainable": attainable / 1e12, "unit": "TFLOP/s",
            "frac": achieved / attainable, "flops_per_walker_step": f_total, "measurement_interval": m}


def sym_packed_cols(N):
    """Columns of the VHS product when the Cholesky vectors are symmetric: the
    16 x 16 orbital blocks on and below the diagonal."""
    nb = (N + 15) // 16
    return nb * (nb + 1) // 2 * 256


def vhs_columns(chol, N):
    """Columns the engine's VHS product covers for these Cholesky vectors (it
    packs the lower block triangle when max |L - L^T| <= 1e-12 max |L| and that
    saves at least 15% of the N^2 columns)."""
    import numpy as np

    npk = sym_packed_cols(N)
    if 20 * npk > 17 * N * N:
        return N * N
    dmax = amax = 0.0
    for g0 in range(0, chol.shape[0], 64):
        L = chol[g0 : g0 + 64].reshape(-1, N, N)
        dmax = max(dmax, float(np.max(np.abs(L - L.transpose(0, 2, 1)))))
        amax = max(amax, float(np.max(np.abs(L))))
    return npk if dmax <= 1e-12 * amax else N * N


def launch_flops(cls, N, Oa, Ob, G, W, unrestricted=False, vhs_cols=None):
    """Flops of one launch of a kernel class, as the kernel computes them,
    averaged over the launches of a block: the VHS product covers `vhs_cols`
    columns (N^2 by default); a Taylor launch covers the walker panel (both
    spins' columns side by side for unrestricted walkers); the one-body,
    exchange, force-bias, Green's-function and overlap launches alternate
    between the spins of an unrestricted or UHF case."""
    occ = (Oa, Ob) if (unrestricted or Oa != Ob) else (Oa,)
    mean = lambda f: sum(f(o) for o in occ) / len(occ)
    if cls == "vhs":
        return 4.0 * G * (N * N if vhs_cols is None else vhs_cols) * W
    if cls == "taylor":
        return 8.0 * N * N * ((Oa + Ob) if unrestricted else Oa) * W
    if cls == "exp_h1":
        return mean(lambda o: 4.0 * N * N * o * W) if unrestricted else 4.0 * N * N * Oa * W
    if cls == "exchange":
        return mean(lambda o: 4.0 * G * o * o * N * W)
    if cls == "force_bias":
        return mean(lambda o: 4.0 * G * o * N * W)
    if cls == "green":
        return mean(lambda o: 8.0 * o * o * N * W)
    if cls == "overlap_gemm":
        return mean(lambda o: 4.0 * o * o * N * W)
    return None


def _oracle_setup(name, W, seed, l_coef):
    """FP64 oracle on workload `name` with W walkers near the trial: returns
    (step(), measure(), n_prop_steps).  step() runs one phaseless step of all
    W walkers; measure() the block's QR, overlaps, energies and bookkeeping."""
    import numpy as np

    from ad_afqmc_amd import synthetic as syn
    from oracle import afqmc_oracle as O

    cfg = syn.CONFIGS[name]
    N, G = cfg["norb"], cfg["nchol"]
    Oa = cfg["nocc_up"]
    Ob = cfg.get("nocc_dn", Oa)
    uhf = cfg["uhf"]
    s = syn.make_system(N, Oa, G, nocc_dn=Ob, uhf=uhf, seed=0, l_scale=l_coef / math.sqrt(N))
    dt = cfg["dt"]
    wtype = "unrestricted" if uhf else "restricted"
    if uhf:
        h1s, rh, rc = O.uhf_measurement_intermediates(s.h1, s.chol, s.mo_up, s.mo_dn)
        trial = O.UHFTrial(N, (Oa, Ob), s.mo_up, s.mo_dn)
    else:
        h1s, rh, rc = O.rhf_measurement_intermediates(s.h1, s.chol, s.mo_up)
        trial = O.RHFTrial(N, (Oa, Oa), s.mo_up)
    pi = O.propagation_intermediates(s.h0, h1s, s.chol, s.rdm1(), dt, wtype)
    ham = dict(h0=s.h0, chol=s.chol, rot_h1=rh, rot_chol=rc, mf_shifts=pi["mf_shifts"], h0_prop=pi["h0_prop"],
               exp_h1=pi["exp_h1"])
    w0 = O.init_walkers(s.rdm1(), s.nelec, W, wtype)
    e0 = float(np.mean(trial.calc_energy(w0, s.h0, rh, rc).real))
    st = {"s": dict(walkers=w0, weights=np.ones(W), overlaps=trial.calc_overlap(w0), e_estimate=e0,
                    pop_control_ene_shift=e0)}
    rng = np.random.default_rng(seed)

    def step():
        st["s"] = O.propagate_constrained(trial, st["s"], rng.standard_normal((W, G)), ham, dt)

    def measure():
        x = st["s"]
        x["walkers"] = O.orthonormalize(x["walkers"])
        x["overlaps"] = trial.calc_overlap(x["walkers"])
        e = np.real(trial.calc_energy(x["walkers"], ham["h0"], ham["rot_h1"], ham["rot_chol"]))
        return float(np.sum(e * x["weights"]) / np.sum(x["weights"]))

    return step, measure, cfg["n_prop_steps"]


def _one_thread():
    """Limits BLAS to one thread; numpy is imported first so that its BLAS is
    loaded when the limit is applied (a fresh worker process has not loaded it)."""
    try:
        import numpy  # noqa: F401
        from threadpoolctl import threadpool_limits

        return threadpool_limits(limits=1)
    except Exception:  # pragma: no cover
        return None


def _time_sample(name, W, sample_steps, seed, l_coef):
    """Seconds for `sample_steps` oracle steps and for one measurement."""
    step, measure, m = _oracle_setup(name, W, seed, l_coef)
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        step()
    t_steps = time.perf_counter() - t0
    t0 = time.perf_counter()
    e = measure()
    t_meas = time.perf_counter() - t0
    return t_steps, t_meas, m, e


def cpu_baseline(name, sample_walkers, sample_steps, l_coef=0.3, seed=0):
    """The FP64 oracle on one host core, on a bounded sample of the workload:
    `sample_walkers` walkers through `sample_steps` phaseless steps and one
    measurement (QR + overlaps + energies), scaled to walker-steps/s of a block
    of m steps with one measurement: W m / (m t_step + t_meas)."""
    limits = _one_thread()
    try:
        t_steps, t_meas, m, e = _time_sample(name, sample_walkers, sample_steps, seed, l_coef)
    finally:
        if limits is not None:
            limits.restore_original_limits()
    t_step = t_steps / sample_steps
    W = sample_walkers
    return {
        "value": W * m / (m * t_step + t_meas),
        "unit": "walker-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{W} walkers x {sample_steps} steps ({t_steps:.1f} s) + one QR/overlap/energy measurement "
        f"({t_meas:.1f} s) of {name} (Cholesky scale {l_coef}/sqrt(N), live population, mean energy {e:.4f}), "
        f"scaled to blocks of {m} steps + 1 measurement; FP64 NumPy oracle, 1 BLAS thread",
    }


def _cpu_worker(name, W, sample_steps, seed, l_coef, barrier, queue):
    _limits = _one_thread()
    step, measure, m = _oracle_setup(name, W, seed, l_coef)
    barrier.wait(timeout=900)
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        step()
    t1 = time.perf_counter()
    measure()
    queue.put((W, sample_steps, m, t0, t1, time.perf_counter()))


def cpu_baseline_parallel(name, sample_walkers, sample_steps, procs=16, l_coef=0.3):
    """The same oracle sample in `procs` walker-sharded processes of one BLAS
    thread each (the reference's MPI walker parallelism with seed + rank):
    aggregate walker-steps/s over the span from the first start to the last end,
    scaled like cpu_baseline.  Workers are spawned fresh, before the caller
    touches the GPU."""
    import multiprocessing as mp

    ctx = mp.get_context("spawn")
    barrier, queue = ctx.Barrier(procs), ctx.Queue()
    ps = [ctx.Process(target=_cpu_worker, args=(name, sample_walkers, sample_steps, 1 + r, l_coef, barrier, queue))
          for r in range(procs)]
    for p_ in ps:
        p_.start()
    res = []
    deadline = time.time() + 1200
    try:
        for _ in ps:
            res.append(queue.get(timeout=max(1.0, deadline - time.time())))
    except Exception:  # a worker failed: report no parallel baseline rather than wait
        pass
    finally:
        for p_ in ps:
            p_.join(timeout=10)
            if p_.is_alive():
                p_.terminate()
    if len(res) < procs:
        return None
    m = res[0][2]
    start = min(r[3] for r in res)
    span_steps = max(r[4] for r in res) - start
    span_all = max(r[5] for r in res) - start
    t_meas = span_all - span_steps
    t_step = span_steps / sample_steps
    work = sum(r[0] for r in res)
    return {
        "value": work * m / (m * t_step + t_meas),
        "unit": "walker-steps/s",
        "cores": procs,
        "kind": "port",
        "sample": f"{procs} processes x {sample_walkers} walkers x {sample_steps} steps + one measurement of {name} "
        f"(live population), {span_all:.1f} s span, scaled to blocks of {m} steps; FP64 NumPy oracle, "
        f"1 BLAS thread each",
    }


def host_description():
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_count": os.cpu_count(), "cpu_model": model}


def _new_job(system, name, precision, nsteps):
    """Façade objects and an engine job of workload `name` in `precision`, with
    the device field buffer reserved for blocks of nsteps steps."""
    import torch

    from ad_afqmc_amd import hamiltonian as hm
    from ad_afqmc_amd import propagation as pr
    from ad_afqmc_amd import wavefunctions as wf
    from ad_afqmc_amd.binding import engine_for

    cfg = syn_config(name)
    N, W = cfg["norb"], cfg["n_walkers"]
    uhf = cfg["uhf"]
    Oa = cfg["nocc_up"]
    Ob = cfg.get("nocc_dn", Oa)
    wave, ham = system.wave_data(), system.ham_data()
    trial = wf.uhf(N, (Oa, Ob)) if uhf else wf.rhf(N, (Oa, Oa))
    wtype = "unrestricted" if uhf else "restricted"
    prop = pr.propagator_afqmc(dt=cfg["dt"], n_walkers=W, walker_type=wtype, precision=precision)
    H = hm.hamiltonian(N)
    ham = H.build_measurement_intermediates(ham, trial, wave)
    ham = H.build_propagation_intermediates(ham, prop, trial, wave)
    pd = prop.init_prop_data(trial, wave, ham, seed=0)
    job = engine_for(trial, wave, ham, wtype, W, prop)
    job.reserve_fields(nsteps)
    up, dn = (t.clone() if t is not None else None for t in pd["walkers"].tensors())
    state = dict(up=up, dn=dn, weights=pd["weights"].clone(), overlaps=job.overlap(up, dn),
                 scalars=torch.stack([pd["e_estimate"], pd["e_estimate"].clone()]).contiguous(),
                 counter=torch.zeros(1, dtype=torch.int64, device=up.device),
                 out=torch.empty(2, dtype=torch.float64, device=up.device))
    return job, state, (wave, ham, pd, trial, prop)


def _sr_block(job, st, nsteps, seed):
    """One device block + comb resampling + overlaps of the new population
    (afq_sr_block), then the driver's e_estimate <- 0.9 e_estimate + 0.1 E_block
    (driver.py:168) on the device; returns a copy of {block_energy, block_weight}."""
    job.sr_block(st["up"], st["dn"], st["weights"], st["overlaps"], st["scalars"], nsteps, seed, st["counter"],
                 st["out"])
    st["scalars"][:1].mul_(0.9).add_(st["out"][:1], alpha=0.1)
    return st["out"].clone()


def _clone_state(st):
    return {k: (v.clone() if v is not None else None) for k, v in st.items()}


def precision_study(cfg_name: str, n_blocks: int, n_eql: int, l_coef: float = 0.3):
    """Energy of each precision mode against FP64 on identical Philox fields and
    comb offsets (afq_sr_block draws both from the block counter): the
    trajectories stay correlated, so the paired block-energy differences
    measure the systematic error with a small error bar."""
    import numpy as np
    import torch

    from ad_afqmc_amd import stat_utils

    cfg = syn_config(cfg_name)
    W, nsteps = cfg["n_walkers"], cfg["n_prop_steps"]
    s = make_system(cfg_name, l_coef)
    out = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        job, st, keep = _new_job(s, cfg_name, prec, nsteps)
        blocks = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(n_eql + n_blocks):
            blocks.append(_sr_block(job, st, nsteps, 99))
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        b = torch.stack(blocks).cpu().numpy()[n_eql:]
        mean, err = stat_utils.blocking_analysis(b[:, 1], b[:, 0])
        out[prec] = dict(blocks=b, energy=float(mean), error=float(err) if err is not None else None,
                         walker_steps_per_s=W * nsteps * (n_eql + n_blocks) / t, finite=bool(np.all(np.isfinite(b))))
        job.close()
        del job, st, keep
        torch.cuda.empty_cache()
    ref = out["fp64"]["blocks"]
    summary = {"workload": cfg_name, "l_scale": f"{l_coef}/sqrt(N)",
               "blocks": n_blocks, "equilibration_blocks": n_eql, "steps_per_block": nsteps,
               "block": "afq_sr_block (steps, QR, energy, comb resampling, overlaps of the new population)"}
    for prec, r in out.items():
        d = r["blocks"][:, 0] - ref[:, 0]
        summary[prec] = {
            "energy_ha": r["energy"],
            "stat_error_ha": r["error"],
            "delta_vs_fp64_mha": 1e3 * float(np.mean(d)),
            "delta_stderr_mha": 1e3 * float(np.std(d, ddof=1) / math.sqrt(len(d))) if len(d) > 1 else None,
            "max_abs_block_delta_mha": 1e3 * float(np.max(np.abs(d))),
            "walker_steps_per_s": r["walker_steps_per_s"],
            "finite": r["finite"],
        }
    summary.update(_restart_paired(s, cfg_name, n_blocks, n_eql))
    return summary


def _restart_paired(s, cfg_name, n_blocks, n_eql):
    """Per-block error without trajectory divergence: along the FP64 trajectory,
    each block is re-run from the same FP64 state (walkers, weights, overlaps,
    shift, block counter) in each reduced-precision mode, and the local energies
    of the same FP64 walkers are measured in each mode.  The means of these
    differences are the systematic errors per block; their spread is small
    because both sides start from identical states."""
    import numpy as np

    cfg = syn_config(cfg_name)
    nsteps = cfg["n_prop_steps"]
    jobs, states, keep = {}, {}, {}
    for prec in ("fp64", "fp32", "bf16x3"):
        jobs[prec], states[prec], keep[prec] = _new_job(s, cfg_name, prec, nsteps)
    J, st = jobs["fp64"], states["fp64"]
    blk = {p: [] for p in jobs}
    ene = {p: [] for p in jobs}
    for i in range(n_eql + n_blocks):
        snap = _clone_state(st)
        if i >= n_eql:
            for p, job in jobs.items():
                e = job.energy(st["up"], st["dn"]).real
                ene[p].append(float((st["weights"] * e).sum() / st["weights"].sum()))
        o = _sr_block(J, st, nsteps, 99)
        if i >= n_eql:
            blk["fp64"].append(float(o[0]))
            for p in ("fp32", "bf16x3"):
                o2 = _sr_block(jobs[p], _clone_state(snap), nsteps, 99)
                blk[p].append(float(o2[0]))
    out = {}
    for p in ("fp32", "bf16x3"):
        res = {}
        for name, series in (("restart_paired_block", blk), ("energy_same_walkers", ene)):
            d = 1e3 * (np.array(series[p]) - np.array(series["fp64"]))
            res[name] = {"delta_mha": float(np.mean(d)), "stderr_mha": float(np.std(d, ddof=1) / math.sqrt(len(d))),
                         "max_abs_mha": float(np.max(np.abs(d)))}
        out[p + "_per_block"] = res
    for j in jobs.values():
        j.close()
    return out


def syn_config(name):
    from ad_afqmc_amd import synthetic as syn

    return syn.CONFIGS[name]


def make_system(name, l_coef=0.3):
    """The synthetic generator for workload `name` (seed 0) with the Cholesky
    scale s_g multiplied by l_coef/sqrt(N); l_coef=None gives the stated scale."""
    from ad_afqmc_amd import synthetic as syn

    cfg = syn.CONFIGS[name]
    Oa = cfg["nocc_up"]
    ls = 1.0 if l_coef is None else l_coef / math.sqrt(cfg["norb"])
    return syn.make_system(cfg["norb"], Oa, cfg["nchol"], nocc_dn=cfg.get("nocc_dn", Oa), uhf=cfg["uhf"], seed=0,
                           l_scale=ls)


def _precision_errors(name):
    """Measured systematic energy errors of each mode on `name` (per block, restart-paired)."""
    f = os.path.join(ROOT, "profiles", f"precision_study_{name}.json")
    out = {}
    if not os.path.exists(f):
        return out
    try:
        d = json.load(open(f))
    except Exception:
        return out
    for p in ("fp32", "bf16x3"):
        rp = d.get(p + "_per_block", {}).get("restart_paired_block")
        if rp:
            out[p] = {"systematic_error_mha": rp["delta_mha"], "systematic_error_stderr_mha": rp["stderr_mha"],
                      "error_source": f"profiles/precision_study_{name}.json (restart-paired blocks)"}
    out["fp64"] = {"systematic_error_mha": 0.0}
    return out


def time_config(system, name, precision, steps, warmup, dominant_roofline=False, breakdown=False, nsteps=None):
    """Times `steps` device blocks (afq_sr_block: n steps + QR + energy + comb
    resampling + overlaps) of workload `name` in `precision`, after `warmup`
    blocks.  With dominant_roofline the warm-up times every kernel class and
    the timed region brackets the dominant class with HIP events on the launch
    stream."""
    import numpy as np
    import torch

    from ad_afqmc_amd import engine as E

    cfg = syn_config(name)
    N, G, W = cfg["norb"], cfg["nchol"], cfg["n_walkers"]
    uhf = cfg["uhf"]
    Oa = cfg["nocc_up"]
    Ob = cfg.get("nocc_dn", Oa)
    nsteps = nsteps or cfg["n_prop_steps"]
    job, st, keep = _new_job(system, name, precision, nsteps)

    classes = E.nat.KERNEL_CLASSES if dominant_roofline or breakdown else []
    for c in classes:
        job.profile(c, True)
    for i in range(warmup):
        _sr_block(job, st, nsteps, 1234)
    torch.cuda.synchronize()
    bd = {c: job.profile_read(c) for c in classes}
    for c in classes:
        job.profile(c, False)
    res = {"breakdown": {c: v[0] / max(warmup, 1) for c, v in bd.items()} or None, "roofline": None}
    if breakdown:
        tot = sum(v[0] for v in bd.values())
        for c, (ms, n) in sorted(bd.items(), key=lambda kv: -kv[1][0]):
            print(f"[bench] {name} {precision} {c:14s} {ms / max(warmup, 1):9.3f} ms/block  "
                  f"{n // max(warmup, 1):5d} launches/block  {100 * ms / max(tot, 1e-9):5.1f}%", file=sys.stderr)
    dominant = max(bd, key=lambda c: bd[c][0]) if dominant_roofline else None
    if dominant:
        job.profile(dominant, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = [_sr_block(job, st, nsteps, 1234) for _ in range(steps)]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if dominant:
        dom_ms, dom_n = job.profile_read(dominant)
        job.profile(dominant, False)
        per_launch = launch_flops(dominant, N, Oa, Ob, G, W, unrestricted=uhf,
                                  vhs_cols=vhs_columns(system.chol, N) if dominant == "vhs" else None)
        avg_s = dom_ms / max(dom_n, 1) / 1e3
        # bf16x3 does three bf16 MFMA products per FP32-equivalent product: its
        # ceiling is a third of the dense bf16 peak
        peak = {"fp64": PEAKS["fp64"], "fp32": PEAKS["fp32"], "bf16x3": PEAKS["bf16"] / 3.0}[precision]
        if per_launch is not None and avg_s > 0:
            ach = per_launch / avg_s
            roof = {"kernel": dominant, "bound": "mfma", "achieved": ach / 1e12, "peak": peak / 1e12,
                    "unit": "TFLOP/s", "frac": ach / peak, "traffic": None, "avg_launch_ms": avg_s * 1e3,
                    "launches": dom_n, "flops_per_launch": per_launch}
            # HBM bytes per launch and the MFMA pipe's busy fraction, both from
            # rocprofv3 counter runs of this kernel on this workload
            for fname, field in (("traffic.json", "traffic"), ("mfma_busy.json", "mfma_busy")):
                tfile = os.path.join(ROOT, "profiles", fname)
                if not os.path.exists(tfile):
                    continue
                try:
                    t = json.load(open(tfile)).get(f"{name}:{precision}:{dominant}")
                    if t is not None:
                        roof[field] = t
                except Exception:
                    pass
            res["roofline"] = roof
    blocks = torch.stack(outs).cpu().numpy()
    res.update(value=W * nsteps * steps / elapsed, elapsed=elapsed, blocks=blocks,
               ms_per_block=elapsed / steps * 1e3)
    job.close()
    del job, st, keep
    torch.cuda.empty_cache()
    return res


def h10_timing(steps=20, warmup=3):
    """The H10 chain (STO-6G, 1.6 bohr, own integrals, 1e-5 Cholesky
    threshold), RHF trial, 50 walkers, dt=0.005, blocks of 10 steps, FP64."""
    import numpy as np
    import torch

    from ad_afqmc_amd import hamiltonian as hm
    from ad_afqmc_amd import integrals as I
    from ad_afqmc_amd import propagation as pr
    from ad_afqmc_amd import wavefunctions as wf
    from ad_afqmc_amd.binding import engine_for

    h = I.hydrogen_chain(10, 1.6, "sto-6g", 1e-5)
    W, dt, nsteps = 50, 0.005, 10
    trial = wf.rhf(10, (5, 5))
    prop = pr.propagator_afqmc(dt=dt, n_walkers=W, walker_type="restricted")
    wave, ham = h.wave_data(), h.ham_data()
    H = hm.hamiltonian(10)
    ham = H.build_measurement_intermediates(ham, trial, wave)
    ham = H.build_propagation_intermediates(ham, prop, trial, wave)
    pd = prop.init_prop_data(trial, wave, ham, seed=0)
    job = engine_for(trial, wave, ham, "restricted", W, prop)
    job.reserve_fields(nsteps)
    up = pd["walkers"].data.clone()
    st = dict(up=up, dn=None, weights=pd["weights"].clone(), overlaps=job.overlap(up, None),
              scalars=torch.stack([pd["e_estimate"], pd["e_estimate"].clone()]).contiguous(),
              counter=torch.zeros(1, dtype=torch.int64, device=up.device),
              out=torch.empty(2, dtype=torch.float64, device=up.device))
    for _ in range(warmup):
        _sr_block(job, st, nsteps, 5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = [_sr_block(job, st, nsteps, 5) for _ in range(steps)]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    b = torch.stack(outs).cpu().numpy()
    # the same blocks replayed from one captured HIP graph per block (as the
    # serving executor runs them): a system this small is bound by launches
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        job.sr_block(st["up"], st["dn"], st["weights"], st["overlaps"], st["scalars"], nsteps, 5, st["counter"],
                     st["out"])
        st["scalars"][:1].mul_(0.9).add_(st["out"][:1], alpha=0.1)
    outs_g = torch.empty((steps, 2), dtype=torch.float64, device=up.device)
    graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        graph.replay()
        outs_g[i].copy_(st["out"])
    torch.cuda.synchronize()
    el_g = time.perf_counter() - t0
    bg = outs_g.cpu().numpy()
    del graph
    job.close()
    return {"value": W * nsteps * steps / el, "unit": "walker-steps/s", "ms_per_block": el / steps * 1e3,
            "blocks": steps, "nchol": int(h.chol.shape[0]), "e_rhf": float(h.e_rhf),
            "mean_block_energy": float(np.sum(b[:, 0] * b[:, 1]) / np.sum(b[:, 1])),
            "graph_replay": {"value": W * nsteps * steps / el_g, "ms_per_block": el_g / steps * 1e3,
                             "mean_block_energy": float(np.sum(bg[:, 0] * bg[:, 1]) / np.sum(bg[:, 1]))},
            "workload": "H10 chain, STO-6G, 1.6 bohr, RHF trial, 50 walkers, dt 0.005, 10 steps per block, FP64"}


def adjoint_timing(name="rhf_n200_o40_g1200_w512", n_blocks=2, n_steps=10, reps=2, l_coef=0.3):
    """Reverse pass of a phaseless segment (afq_segment_energy_vjp) on
    workload `name` (RHF trial with restricted walkers, or UHF trial with
    unrestricted walkers; FP64): n_blocks blocks of n_steps steps with given
    fields, the value alone and the value with dF/dh1 and dF/dchol."""
    import numpy as np
    import torch

    from ad_afqmc_amd import synthetic as syn
    from ad_afqmc_amd.engine import EngineJob

    cfg = syn_config(name)
    N, G, W = cfg["norb"], cfg["nchol"], cfg["n_walkers"]
    uhf = cfg["uhf"]
    s = make_system(name, l_coef)
    mo_dn = s.mo_dn if uhf else s.mo_up
    up = torch.as_tensor(syn.perturbed_walkers(s.mo_up, W, 0.3 / math.sqrt(N), seed=1)).cuda()
    dn = torch.as_tensor(syn.perturbed_walkers(mo_dn, W, 0.3 / math.sqrt(N), seed=2)).cuda() if uhf else None
    weights = torch.ones(W, dtype=torch.float64, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(5)
    fields = torch.randn((n_blocks, n_steps, W, G), dtype=torch.float64, device="cuda", generator=gen)
    wtype, ttype = ("unrestricted", "uhf") if uhf else ("restricted", "rhf")
    job = EngineJob(N, (s.mo_up.shape[1], mo_dn.shape[1]), G, W, wtype, ttype, 6, cfg["dt"], "fp64")
    try:
        job.set_hamiltonian(float(s.h0), np.array(s.h1), np.array(s.chol).reshape(G, N * N))
        job.set_trial(s.mo_up, mo_dn if uhf else None, s.rdm1())
        job.build_measurement()
        job.build_propagation()
        e_est = float(torch.mean(job.energy(up, dn).real))
        out = {}
        for want in (False, True, False, True):  # the first pair warms up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                F, _, _ = job.segment_energy_vjp(up, weights, e_est, fields, want_grad=want, walkers_dn=dn)
            torch.cuda.synchronize()
            out["with_gradient" if want else "value_only"] = (time.perf_counter() - t0) / reps * 1e3
    finally:
        job.close()
    ws = W * n_blocks * n_steps
    return {"workload": f"{name}, {n_blocks} blocks x {n_steps} steps, FP64",
            "ms_value_only": out["value_only"], "ms_with_gradient": out["with_gradient"],
            "gradient_over_value": out["with_gradient"] / out["value_only"],
            "walker_steps_per_s_with_gradient": ws / out["with_gradient"] * 1e3, "segment_energy": float(F)}


def h10_cpu_baseline(n_steps=20):
    """The FP64 oracle on the H10 chain, 50 walkers, one core: n_steps steps
    and one measurement, scaled to blocks of 10 steps."""
    import numpy as np

    from ad_afqmc_amd import integrals as I
    from oracle import afqmc_oracle as O

    limits = _one_thread()
    try:
        h = I.hydrogen_chain(10, 1.6, "sto-6g", 1e-5)
        W, dt, G = 50, 0.005, h.chol.shape[0]
        mo = h.trial_mo()
        trial = O.RHFTrial(10, (5, 5), mo)
        h1s, rh, rc = O.rhf_measurement_intermediates(h.ham_data()["h1"], h.chol, mo)
        rdm1 = O.rdm1_from_mo(mo)
        pi = O.propagation_intermediates(h.h0, h1s, h.chol, rdm1, dt)
        ham = dict(h0=h.h0, chol=h.chol, rot_h1=rh, rot_chol=rc, mf_shifts=pi["mf_shifts"], h0_prop=pi["h0_prop"],
                   exp_h1=pi["exp_h1"])
        w0 = O.init_walkers(rdm1, (5, 5), W, "restricted")
        e0 = float(np.mean(trial.calc_energy(w0, h.h0, rh, rc).real))
        st = dict(walkers=w0, weights=np.ones(W), overlaps=trial.calc_overlap(w0), e_estimate=e0,
                  pop_control_ene_shift=e0)
        rng = np.random.default_rng(0)
        t0 = time.perf_counter()
        for _ in range(n_steps):
            st = O.propagate_constrained(trial, st, rng.standard_normal((W, G)), ham, dt)
        t_steps = time.perf_counter() - t0
        t0 = time.perf_counter()
        st["walkers"] = O.orthonormalize(st["walkers"])
        st["overlaps"] = trial.calc_overlap(st["walkers"])
        trial.calc_energy(st["walkers"], h.h0, rh, rc)
        t_meas = time.perf_counter() - t0
    finally:
        if limits is not None:
            limits.restore_original_limits()
    t_step = t_steps / n_steps
    return {"value": W * 10 / (10 * t_step + t_meas), "unit": "walker-steps/s", "cores": 1, "kind": "port",
            "sample": f"50 walkers x {n_steps} steps + one measurement, FP64 NumPy oracle, 1 BLAS thread"}


def _mode_entry(name, r, base_value, errs, m):
    cfg = syn_config(name)
    Oa = cfg["nocc_up"]
    Ob = cfg.get("nocc_dn", Oa)
    e = {"value": r["value"], "ms_per_step": r["ms_per_block"], "blocks_timed": len(r["blocks"]),
         "speedup_vs_fp64": r["value"] / base_value}
    e.update(errs)
    e["whole_step_roofline"] = whole_step_roofline(cfg["norb"], Oa, Ob, cfg["nchol"], m, not cfg["uhf"], r["precision"],
                                                   r["value"])
    return e


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed blocks")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="uhf_n300_o60_58_g2000_w256")
    ap.add_argument("--precision", default="bf16x3")
    ap.add_argument("--lcoef", type=float, default=0.3,
                    help="Cholesky scale coefficient c: s_g multiplied by c/sqrt(N) (0: the stated scale)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-walkers", type=int, default=4)
    ap.add_argument("--cpu-sample-steps", type=int, default=6)
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="walker-sharded oracle processes for the all-cores CPU baseline (1: skip it)")
    ap.add_argument("--breakdown", action="store_true", help="print per-class kernel times to stderr")
    ap.add_argument("--only", action="store_true", help="time the headline workload only (for profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (to scale HBM counters) and exit")
    ap.add_argument("--serving-jobs", type=int, default=64,
                    help="jobs in the serving stream (Poisson arrivals at 4/s, N~U{30..250}, 300 steps); 0 skips it")
    ap.add_argument("--serving-precision", default="bf16x3")
    ap.add_argument("--precision-study", default="", metavar="WORKLOAD",
                    help="only run the FP64 / FP32 / bf16x3 energy comparison on WORKLOAD and exit")
    ap.add_argument("--study-blocks", type=int, default=50)
    ap.add_argument("--study-eql", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)
    l_coef = args.lcoef if args.lcoef > 0 else None

    import numpy as np
    import torch

    if args.precision_study:
        print(json.dumps({"precision_study": precision_study(args.precision_study, args.study_blocks,
                                                             args.study_eql, args.lcoef)}), flush=True)
        return

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8}))
        return

    name = args.config
    cfg = syn_config(name)
    N, G, W = cfg["norb"], cfg["nchol"], cfg["n_walkers"]
    uhf = cfg["uhf"]
    Oa = cfg["nocc_up"]
    Ob = cfg.get("nocc_dn", Oa)
    nsteps, dt = cfg["n_prop_steps"], cfg["dt"]
    cpu = cpu_all = h10_cpu = None
    if not args.no_cpu_baseline and not args.only:
        # before anything touches the GPU: the parallel baseline spawns its workers
        print(f"[bench] cpu baseline (oracle sample of {name})", file=sys.stderr, flush=True)
        cpu = cpu_baseline(name, args.cpu_sample_walkers, args.cpu_sample_steps, l_coef or 0.3)
        if args.cpu_procs > 1:
            print("[bench] cpu baseline, all cores", file=sys.stderr, flush=True)
            cpu_all = cpu_baseline_parallel(name, args.cpu_sample_walkers, args.cpu_sample_steps,
                                            min(args.cpu_procs, os.cpu_count() or 1), l_coef or 0.3)
        print("[bench] cpu baseline (H10)", file=sys.stderr, flush=True)
        h10_cpu = h10_cpu_baseline()
    print(f"[bench] building {name}", file=sys.stderr, flush=True)
    system = make_system(name, l_coef)
    r = time_config(system, name, args.precision, args.steps, args.warmup, dominant_roofline=True,
                    breakdown=args.breakdown)
    r["precision"] = args.precision
    value, elapsed, blocks = r["value"], r["elapsed"], r["blocks"]
    finite = bool(np.all(np.isfinite(blocks)))
    f_total, _ = flops_per_walker_step(N, Oa, Ob, G, nsteps, not uhf)
    line = {
        "metric": "walker-steps/s at <=0.1 mHa vs FP64 oracle; p50/p99 job latency under load",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else args.precision,
        "data": "synthetic",
        "config": {
            "workload": name,
            "norb": N,
            "nocc": [Oa, Ob] if uhf else [Oa],
            "nchol": G,
            "walkers": W,
            "dt": dt,
            "steps_per_block": nsteps,
            "trial": "uhf" if uhf else "rhf",
            "walker_type": "unrestricted" if uhf else "restricted",
            "precision": args.precision,
            "cholesky_scale": "stated" if l_coef is None else f"stated x {l_coef}/sqrt(N)",
            "block": "afq_sr_block: 10 steps + QR + overlap/energy + comb resampling + overlaps",
            "parallelism": "single-gpu",
        },
        "model_flops_per_walker_step": f_total,
        "achieved_tflops_whole_step": f_total * value / 1e12,
        "roofline": r["roofline"],
        "whole_step_roofline": whole_step_roofline(N, Oa, Ob, G, nsteps, not uhf, args.precision, value),
        "kernel_breakdown_ms_per_block": r["breakdown"],
        "population_finite": finite,
        "mean_block_energy": float(np.sum(blocks[:, 0] * blocks[:, 1]) / np.sum(blocks[:, 1])) if finite else None,
    }
    errs = _precision_errors(name)
    line.update(errs.get(args.precision, {}))
    if not args.only:
        # the other modes of the same workload over the same number of blocks
        modes = {args.precision: r}
        for p in ("fp64", "fp32", "bf16x3"):
            if p not in modes:
                print(f"[bench] {name} {p}", file=sys.stderr, flush=True)
                modes[p] = time_config(system, name, p, args.steps, args.warmup, breakdown=args.breakdown)
                modes[p]["precision"] = p
        base = modes["fp64"]["value"]
        line["speedup_vs_fp64"] = value / base
        line["precision_modes"] = {p: _mode_entry(name, m, base, errs.get(p, {}), nsteps) for p, m in modes.items()}
        # the reference's default measurement interval (50 steps, launch_script.py:120)
        m50 = {}
        for p in ("fp64", "fp32", "bf16x3"):
            print(f"[bench] {name} {p} at 50 steps per block", file=sys.stderr, flush=True)
            m = time_config(system, name, p, 3, 1, nsteps=50)
            m["precision"] = p
            m50[p] = m
        line["measurement_interval_50"] = {p: _mode_entry(name, m, m50["fp64"]["value"], {}, 50)
                                           for p, m in m50.items()}
        del system
        torch.cuda.empty_cache()
        # the stated Cholesky scale: same dense work per block, dying population
        print(f"[bench] {name} stated scale", file=sys.stderr, flush=True)
        system = make_system(name, None)
        st = time_config(system, name, args.precision, 3, 1)
        line["stated_generator"] = {"value": st["value"], "ms_per_step": st["ms_per_block"], "blocks_timed": 3,
                                    "population_finite": bool(np.all(np.isfinite(st["blocks"])))}
        del system
        torch.cuda.empty_cache()
        # the RHF workload (N=200, O=40, G=1200, 512 walkers) in every mode
        rhf = "rhf_n200_o40_g1200_w512"
        print(f"[bench] building {rhf}", file=sys.stderr, flush=True)
        system = make_system(rhf, l_coef)
        rerrs = _precision_errors(rhf)
        rm = {}
        for p in ("fp64", "fp32", "bf16x3"):
            rm[p] = time_config(system, rhf, p, args.steps, args.warmup, dominant_roofline=(p == "fp64"),
                                breakdown=args.breakdown)
            rm[p]["precision"] = p
        line[rhf] = {p: _mode_entry(rhf, m, rm["fp64"]["value"], rerrs.get(p, {}), 10) for p, m in rm.items()}
        line[rhf]["fp64"]["roofline"] = rm["fp64"]["roofline"]
        line[rhf]["fp64"]["kernel_breakdown_ms_per_block"] = rm["fp64"]["breakdown"]
        del system
        torch.cuda.empty_cache()
        print("[bench] H10 chain", file=sys.stderr, flush=True)
        line["h10_chain"] = h10_timing()
        line["h10_chain"]["cpu_baseline"] = h10_cpu
        print("[bench] segment reverse pass", file=sys.stderr, flush=True)
        line["segment_adjoint"] = adjoint_timing(l_coef=l_coef or 0.3)
        torch.cuda.empty_cache()
        line["segment_adjoint_uhf"] = adjoint_timing("uhf_n300_o60_58_g2000_w256", l_coef=l_coef or 0.3)
        torch.cuda.empty_cache()
    if args.serving_jobs > 0 and not args.only:
        from ad_afqmc_amd import serving

        def serve(make_executor, realtime, nmax=250, rate=4.0):
            specs = serving.poisson_stream(args.serving_jobs, rate=rate, seed=0, nmax=nmax,
                                           precision=args.serving_precision)
            ex = make_executor()
            sch = serving.Scheduler(ex, hbm_cap_bytes=200e9, max_active=64)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = sch.run(specs, realtime=realtime)
            torch.cuda.synchronize()
            summ = serving.summarize(res, time.perf_counter() - t0)
            summ.update(max_concurrent_jobs=sch.max_concurrent, peak_admitted_gb=sch.peak_used / 1e9)
            if hasattr(ex, "rounds"):
                summ["grouped_rounds"] = ex.rounds
            return summ

        executors = {
            # one stream and one captured graph per job
            "streams": lambda: serving.GpuExecutor(n_prop_steps=10),
            # every job in grouped blocks (afq_group_sr_block), size classes N <= 100, <= 175, > 175
            "grouped": lambda: serving.GroupedExecutor(n_prop_steps=10),
            # jobs with N <= 100 grouped, larger ones on their own streams
            "grouped_small_solo_large": lambda: serving.GroupedExecutor(n_prop_steps=10, lanes=(), solo_above=100),
        }
        torch.cuda.empty_cache()
        out = {}
        for mode, make in executors.items():
            print(f"[bench] serving stream ({mode}): {args.serving_jobs} jobs", file=sys.stderr, flush=True)
            out[mode] = serve(make, True)
            print(f"[bench] serving capacity ({mode}): all jobs at t=0", file=sys.stderr, flush=True)
            out[mode]["capacity_all_at_t0"] = serve(make, False)
        # a stream of small jobs only (N <= 80, 16 jobs/s), where launches weigh most
        small = {}
        for mode in ("streams", "grouped"):
            print(f"[bench] small-job stream ({mode})", file=sys.stderr, flush=True)
            small[mode] = serve(executors[mode], True, nmax=80, rate=16.0)
            small[mode]["capacity_all_at_t0"] = serve(executors[mode], False, nmax=80, rate=16.0)
        summ = dict(out["streams"])
        summ.update(arrival_rate_per_s=4.0, precision=args.serving_precision, hbm_cap_gb=200, hip_graphs=True,
                    executor="streams: one HIP stream and one captured afq_sr_block graph per job",
                    grouped=out["grouped"], grouped_small_solo_large=out["grouped_small_solo_large"],
                    small_jobs=dict(small, workload="N~U{30..80}, O=N/5, G=5N, W~U{64..512}, 300 steps, Poisson 16 jobs/s"),
                    workload="N~U{30..250}, O=N/5, G=5N, W~U{64..512}, 300 steps, Poisson 4 jobs/s")
        line["serving"] = summ
    line["cpu_baseline"] = cpu
    line["cpu_baseline_all_cores"] = cpu_all
    line["host"] = host_description()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

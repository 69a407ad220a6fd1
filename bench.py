#!/usr/bin/env python
"""Walker-step throughput of the phaseless-AFQMC engine on one MI355X.

One bench step = one device sampler block on the synthetic RHF workload
(N=200, O=40, G=1200, 512 walkers, dt=0.005): 10 phaseless steps with
on-device fields, QR re-orthonormalisation, overlap and local-energy
measurement, population-control update, then comb resampling.
value = walkers * 10 * steps / wall time of the timed region (walker-steps/s).

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


def _oracle_block(cfg, W, seed):
    """Set up the FP64 oracle on the workload and return a closure that runs one
    block (n_prop_steps steps + QR + energy) of W walkers, and the step count."""
    import numpy as np

    from ad_afqmc_amd import synthetic as syn
    from oracle import afqmc_oracle as O

    s = syn.make_system(cfg["norb"], cfg["nocc_up"], cfg["nchol"], seed=0)
    dt, nsteps = cfg["dt"], cfg["n_prop_steps"]
    h1s, rh, rc = O.rhf_measurement_intermediates(s.h1, s.chol, s.mo_up)
    pi = O.propagation_intermediates(s.h0, h1s, s.chol, s.rdm1(), dt)
    trial = O.RHFTrial(s.norb, s.nelec, s.mo_up)
    ham = dict(h0=s.h0, chol=s.chol, rot_h1=rh, rot_chol=rc, mf_shifts=pi["mf_shifts"], h0_prop=pi["h0_prop"],
               exp_h1=pi["exp_h1"])
    w0 = O.init_walkers(s.rdm1(), s.nelec, W, "restricted")
    e0 = float(np.mean(trial.calc_energy(w0, s.h0, rh, rc).real))
    state = dict(walkers=w0, weights=np.ones(W), overlaps=trial.calc_overlap(w0), e_estimate=e0,
                 pop_control_ene_shift=e0)
    fields = np.random.default_rng(seed).standard_normal((nsteps, W, s.nchol))
    return (lambda: O.block_scan(trial, state, fields, ham, dt)), nsteps


def _one_thread():
    """Limits BLAS to one thread; numpy is imported first so that its BLAS is
    loaded when the limit is applied (a fresh worker process has not loaded it)."""
    try:
        import numpy  # noqa: F401
        from threadpoolctl import threadpool_limits

        return threadpool_limits(limits=1)
    except Exception:  # pragma: no cover
        return None


def cpu_baseline(cfg, sample_walkers=8, seed=0):
    """The FP64 oracle timed on one host core on a bounded sample: a few walkers
    of the same workload through one block (10 steps + QR + energy)."""
    limits = _one_thread()
    try:
        run, nsteps = _oracle_block(cfg, sample_walkers, seed)
        t0 = time.perf_counter()
        run()
        t = time.perf_counter() - t0
    finally:
        if limits is not None:
            limits.restore_original_limits()
    W = sample_walkers
    return {
        "value": W * nsteps / t,
        "unit": "walker-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{W} walkers x 1 block ({nsteps} steps + QR + energy) of the same workload, FP64 NumPy oracle, "
        f"1 BLAS thread, {t:.1f} s",
    }


def _cpu_worker(cfg, W, seed, barrier, queue):
    _limits = _one_thread()
    run, nsteps = _oracle_block(cfg, W, seed)
    barrier.wait(timeout=900)
    t0 = time.perf_counter()
    run()
    queue.put((W * nsteps, t0, time.perf_counter()))


def cpu_baseline_parallel(cfg, sample_walkers=8, procs=16):
    """The same oracle sample in `procs` walker-sharded processes of one BLAS
    thread each (the reference's MPI walker parallelism with seed + rank):
    aggregate walker-steps over the span from the first start to the last end.
    Worker processes are spawned fresh, before the caller touches the GPU."""
    import multiprocessing as mp

    ctx = mp.get_context("spawn")
    barrier, queue = ctx.Barrier(procs), ctx.Queue()
    ps = [ctx.Process(target=_cpu_worker, args=(cfg, sample_walkers, 1 + r, barrier, queue)) for r in range(procs)]
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
    work = sum(r[0] for r in res)
    span = max(r[2] for r in res) - min(r[1] for r in res)
    return {
        "value": work / span,
        "unit": "walker-steps/s",
        "cores": procs,
        "kind": "port",
        "sample": f"{procs} processes x {sample_walkers} walkers x 1 block of the same workload, FP64 NumPy oracle, "
        f"1 BLAS thread each, {span:.1f} s span",
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


def precision_study(cfg_name: str, n_blocks: int, n_eql: int, l_coef: float = 0.3):
    """Energy of each precision mode against FP64 on identical Philox fields and
    SR offsets: the trajectories stay correlated, so the paired block-energy
    differences measure the systematic error with a small error bar."""
    import numpy as np
    import torch

    from ad_afqmc_amd import hamiltonian as hm
    from ad_afqmc_amd import propagation as pr
    from ad_afqmc_amd import stat_utils
    from ad_afqmc_amd import synthetic as syn
    from ad_afqmc_amd import wavefunctions as wf
    from ad_afqmc_amd.binding import engine_for

    cfg = syn.CONFIGS[cfg_name]
    N, G, W = cfg["norb"], cfg["nchol"], cfg["n_walkers"]
    Oa = cfg["nocc_up"]
    Ob = cfg.get("nocc_dn", Oa)
    uhf = cfg["uhf"]
    # Cholesky scale l_coef/sqrt(N): a correlation energy small enough for a stable
    # walker population at dt = 0.005 (the stated generator, l_coef = sqrt(N), kills it)
    lsc = l_coef / math.sqrt(N)
    s = syn.make_system(N, Oa, G, nocc_dn=Ob, uhf=uhf, seed=0, l_scale=lsc)
    nsteps, dt = cfg["n_prop_steps"], cfg["dt"]
    trial = wf.uhf(N, (Oa, Ob)) if uhf else wf.rhf(N, (Oa, Oa))
    wtype = "unrestricted" if uhf else "restricted"
    zetas = torch.rand(n_eql + n_blocks, dtype=torch.float64, device="cuda",
                       generator=torch.Generator(device="cuda").manual_seed(7))
    out = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        wave, ham = s.wave_data(), s.ham_data()
        prop = pr.propagator_afqmc(dt=dt, n_walkers=W, walker_type=wtype, precision=prec)
        H = hm.hamiltonian(N)
        ham = H.build_measurement_intermediates(ham, trial, wave)
        ham = H.build_propagation_intermediates(ham, prop, trial, wave)
        pd = prop.init_prop_data(trial, wave, ham, seed=0)
        job = engine_for(trial, wave, ham, wtype, W, prop)
        up, dn = (t.clone() if t is not None else None for t in pd["walkers"].tensors())
        weights = pd["weights"].clone()
        overlaps = job.overlap(up, dn)
        scalars = torch.stack([pd["e_estimate"], pd["e_estimate"].clone()]).contiguous()
        counter, blocks = 0, []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(n_eql + n_blocks):
            o, counter = job.block(up, dn, weights, overlaps, scalars, nsteps, 99, counter)
            job.stochastic_reconfiguration(up, dn, weights, zetas[i : i + 1])
            blocks.append(o)
            # e_estimate follows the block energy as in the reference driver (driver.py:168)
            scalars[0] = 0.9 * scalars[0] + 0.1 * o[0]
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        b = torch.stack(blocks).cpu().numpy()[n_eql:]
        mean, err = stat_utils.blocking_analysis(b[:, 1], b[:, 0])
        out[prec] = dict(blocks=b, energy=float(mean), error=float(err) if err is not None else None,
                         walker_steps_per_s=W * nsteps * (n_eql + n_blocks) / t, finite=bool(np.all(np.isfinite(b))))
        del job, wave, ham, pd, up, dn, weights, overlaps, scalars
        torch.cuda.empty_cache()
    ref = out["fp64"]["blocks"]
    summary = {"workload": cfg_name, "l_scale": f"{l_coef}/sqrt(N)",
               "blocks": n_blocks, "equilibration_blocks": n_eql, "steps_per_block": nsteps}
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
    summary.update(_restart_paired(s, cfg, trial, wtype, n_blocks, n_eql, zetas))
    return summary


def _restart_paired(s, cfg, trial, wtype, n_blocks, n_eql, zetas):
    """Per-block error without trajectory divergence: along the FP64 trajectory,
    each block is re-run from the same FP64 state (walkers, weights, overlaps,
    shift, field counter) in each reduced-precision mode, and the local energies
    of the same FP64 walkers are measured in each mode.  The means of these
    differences are the systematic errors per block; their spread is small
    because both sides start from identical states."""
    import numpy as np
    import torch

    from ad_afqmc_amd import hamiltonian as hm
    from ad_afqmc_amd import propagation as pr
    from ad_afqmc_amd.binding import engine_for

    N, W = cfg["norb"], cfg["n_walkers"]
    nsteps, dt = cfg["n_prop_steps"], cfg["dt"]
    jobs, waves = {}, {}
    for prec in ("fp64", "fp32", "bf16x3"):
        wave, ham = s.wave_data(), s.ham_data()
        prop = pr.propagator_afqmc(dt=dt, n_walkers=W, walker_type=wtype, precision=prec)
        H = hm.hamiltonian(N)
        ham = H.build_measurement_intermediates(ham, trial, wave)
        ham = H.build_propagation_intermediates(ham, prop, trial, wave)
        jobs[prec] = engine_for(trial, wave, ham, wtype, W, prop)
        waves[prec] = (wave, ham, prop)
    wave, ham, prop = waves["fp64"]
    pd = prop.init_prop_data(trial, wave, ham, seed=0)
    up, dn = (t.clone() if t is not None else None for t in pd["walkers"].tensors())
    weights = pd["weights"].clone()
    J = jobs["fp64"]
    overlaps = J.overlap(up, dn)
    scalars = torch.stack([pd["e_estimate"], pd["e_estimate"].clone()]).contiguous()
    counter = 0
    blk = {p: [] for p in jobs}
    ene = {p: [] for p in jobs}
    cl = lambda t: None if t is None else t.clone()
    for i in range(n_eql + n_blocks):
        state = (cl(up), cl(dn), weights.clone(), overlaps.clone(), scalars.clone(), counter)
        if i >= n_eql:
            for p, job in jobs.items():
                e = job.energy(up, dn).real
                ene[p].append(float((weights * e).sum() / weights.sum()))
        o, counter = J.block(up, dn, weights, overlaps, scalars, nsteps, 99, counter)
        if i >= n_eql:
            blk["fp64"].append(float(o[0]))
            for p in ("fp32", "bf16x3"):
                u2, d2, w2, ov2, sc2, c0 = (cl(x) if isinstance(x, torch.Tensor) else x for x in state)
                o2, _ = jobs[p].block(u2, d2, w2, ov2, sc2, nsteps, 99, c0)
                blk[p].append(float(o2[0]))
        J.stochastic_reconfiguration(up, dn, weights, zetas[i : i + 1])
        scalars[0] = 0.9 * scalars[0] + 0.1 * o[0]
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


def make_system(name):
    """The stated synthetic generator for workload `name` (seed 0)."""
    from ad_afqmc_amd import synthetic as syn

    cfg = syn.CONFIGS[name]
    Oa = cfg["nocc_up"]
    return syn.make_system(cfg["norb"], Oa, cfg["nchol"], nocc_dn=cfg.get("nocc_dn", Oa), uhf=cfg["uhf"], seed=0)


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
    """Times `steps` device blocks (10 phaseless steps + QR + energy + comb
    resampling each) of workload `name` in `precision`, after `warmup` blocks.
    With dominant_roofline the warm-up times every kernel class and the timed
    region brackets the dominant class with HIP events on the launch stream."""
    import numpy as np
    import torch

    from ad_afqmc_amd import engine as E
    from ad_afqmc_amd import hamiltonian as hm
    from ad_afqmc_amd import propagation as pr
    from ad_afqmc_amd import wavefunctions as wf
    from ad_afqmc_amd.binding import engine_for

    cfg = syn_config(name)
    N, G, W = cfg["norb"], cfg["nchol"], cfg["n_walkers"]
    uhf = cfg["uhf"]
    Oa = cfg["nocc_up"]
    Ob = cfg.get("nocc_dn", Oa)
    nsteps, dt = nsteps or cfg["n_prop_steps"], cfg["dt"]
    wave, ham = system.wave_data(), system.ham_data()
    trial = wf.uhf(N, (Oa, Ob)) if uhf else wf.rhf(N, (Oa, Oa))
    wtype = "unrestricted" if uhf else "restricted"
    prop = pr.propagator_afqmc(dt=dt, n_walkers=W, walker_type=wtype, precision=precision)
    H = hm.hamiltonian(N)
    ham = H.build_measurement_intermediates(ham, trial, wave)
    ham = H.build_propagation_intermediates(ham, prop, trial, wave)
    pd = prop.init_prop_data(trial, wave, ham, seed=0)
    job = engine_for(trial, wave, ham, wtype, W, prop)
    up, dn = (t.clone() if t is not None else None for t in pd["walkers"].tensors())
    weights = pd["weights"].clone()
    overlaps = job.overlap(up, dn)
    scalars = torch.stack([pd["e_estimate"], pd["e_estimate"].clone()]).contiguous()
    zetas = torch.rand(warmup + steps + 8, dtype=torch.float64, device=up.device)
    counter = 0

    def one_block(i):
        nonlocal counter
        out, counter = job.block(up, dn, weights, overlaps, scalars, nsteps, 1234, counter)
        job.stochastic_reconfiguration(up, dn, weights, zetas[i : i + 1])
        return out

    classes = E.nat.KERNEL_CLASSES if dominant_roofline or breakdown else []
    for c in classes:
        job.profile(c, True)
    for i in range(warmup):
        one_block(i)
    torch.cuda.synchronize()
    bd = {c: job.profile_read(c) for c in classes}
    for c in classes:
        job.profile(c, False)
    res = {"breakdown": {c: v[0] / max(warmup, 1) for c, v in bd.items()} or None, "roofline": None}
    if breakdown:
        tot = sum(v[0] for v in bd.values())
        for c, (ms, n) in sorted(bd.items(), key=lambda kv: -kv[1][0]):
            print(f"[bench] {c:14s} {ms / max(warmup, 1):9.3f} ms/block  {n // max(warmup, 1):5d} launches/block"
                  f"  {100 * ms / max(tot, 1e-9):5.1f}%", file=sys.stderr)
    dominant = max(bd, key=lambda c: bd[c][0]) if dominant_roofline else None
    if dominant:
        job.profile(dominant, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = [one_block(warmup + i) for i in range(steps)]
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
                    "flops_per_launch": per_launch}
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
    res.update(value=W * nsteps * steps / elapsed, elapsed=elapsed, blocks=torch.stack(outs).cpu().numpy())
    job.close()
    del job, up, dn, weights, overlaps, scalars, pd, ham, wave
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed blocks")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="rhf_n200_o40_g1200_w512")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-walkers", type=int, default=8)
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="walker-sharded oracle processes for the all-cores CPU baseline (1: skip it)")
    ap.add_argument("--breakdown", action="store_true", help="print per-class kernel times to stderr")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (to scale HBM counters) and exit")
    ap.add_argument("--serving-jobs", type=int, default=64,
                    help="jobs in the serving stream (Poisson arrivals at 4/s, N~U{30..250}, 300 steps); 0 skips it")
    ap.add_argument("--serving-precision", default="bf16x3")
    ap.add_argument("--precision-study", default="", metavar="WORKLOAD",
                    help="only run the FP64 / FP32 / bf16x3 energy comparison on WORKLOAD and exit")
    ap.add_argument("--study-blocks", type=int, default=50)
    ap.add_argument("--study-eql", type=int, default=10)
    ap.add_argument("--study-lcoef", type=float, default=0.3, help="Cholesky scale coefficient c in c/sqrt(N)")
    ap.add_argument("--no-modes", action="store_true",
                    help="skip the FP32 / bf16x3 runs of the workload and the N=300, G=2000 mode comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.precision_study:
        print(json.dumps({"precision_study": precision_study(args.precision_study, args.study_blocks,
                                                             args.study_eql, args.study_lcoef)}), flush=True)
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

    cfg = syn_config(args.config)
    N, G, W = cfg["norb"], cfg["nchol"], cfg["n_walkers"]
    uhf = cfg["uhf"]
    Oa = cfg["nocc_up"]
    Ob = cfg.get("nocc_dn", Oa)
    nsteps, dt = cfg["n_prop_steps"], cfg["dt"]
    cpu = cpu_all = None
    if not args.no_cpu_baseline and not uhf:
        # before anything touches the GPU: the parallel baseline spawns its workers
        print("[bench] cpu baseline (oracle sample)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(cfg, args.cpu_sample_walkers)
        if args.cpu_procs > 1:
            cpu_all = cpu_baseline_parallel(cfg, args.cpu_sample_walkers, min(args.cpu_procs, os.cpu_count() or 1))
    print(f"[bench] building {args.config}", file=sys.stderr, flush=True)
    system = make_system(args.config)
    r = time_config(system, args.config, args.precision, args.steps, args.warmup, dominant_roofline=True,
                    breakdown=args.breakdown)
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
            "workload": args.config,
            "norb": N,
            "nocc": [Oa, Ob] if uhf else [Oa],
            "nchol": G,
            "walkers": W,
            "dt": dt,
            "steps_per_block": nsteps,
            "trial": "uhf" if uhf else "rhf",
            "precision": args.precision,
            "parallelism": "single-gpu",
        },
        "model_flops_per_walker_step": f_total,
        "achieved_tflops_whole_step": f_total * value / 1e12,
        "roofline": r["roofline"],
        "whole_step_roofline": whole_step_roofline(N, Oa, Ob, G, nsteps, not uhf, args.precision, value),
        "kernel_breakdown_ms_per_block": r["breakdown"],
        # The stated generator kills the walker population (see synthetic.make_system);
        # the work per block is the same dense arithmetic either way.
        "population_finite": finite,
        "mean_block_energy": float(np.mean(blocks[:, 0])) if finite else None,
        "population_note": ("the generator as stated has a two-body scale growing like sqrt(N) and clips every "
                            "walker within a few steps; every timed kernel is dense and does the same work either "
                            "way. Energies and precision errors use Cholesky scale 0.3/sqrt(N) (profiles/precision_study_*.json)."),
    }
    del system
    torch.cuda.empty_cache()
    if not args.no_modes:
        # the reduced-precision modes on the same workload, with their measured
        # systematic energy error (restart-paired blocks, profiles/precision_study_*.json)
        modes = {}
        errs = _precision_errors(args.config)
        system = make_system(args.config)
        for p in ("fp64", "fp32", "bf16x3"):
            if p == args.precision:
                modes[p] = {"value": value, "ms_per_step": elapsed / args.steps * 1e3}
            else:
                m = time_config(system, args.config, p, max(3, args.steps // 2), 1)
                modes[p] = {"value": m["value"], "ms_per_step": m["elapsed"] / max(3, args.steps // 2) * 1e3}
            modes[p].update(errs.get(p, {}))
            modes[p]["whole_step_roofline"] = whole_step_roofline(N, Oa, Ob, G, nsteps, not uhf, p,
                                                                  modes[p]["value"])
        line["precision_modes"] = modes
        del system
        torch.cuda.empty_cache()
        # the mixed-precision target is stated at N=300, G=2000 (UHF, 256 walkers)
        big = "uhf_n300_o60_58_g2000_w256"
        print(f"[bench] building {big}", file=sys.stderr, flush=True)
        system = make_system(big)
        bc = syn_config(big)
        errs = _precision_errors(big)
        # measurement every 10 steps (as the RHF workload) and every 50 (the
        # reference's default interval)
        for m_int, key in ((10, "uhf_n300_g2000_modes"), (50, "uhf_n300_g2000_modes_m50")):
            ns = {}
            for p in ("fp64", "fp32", "bf16x3"):
                m = time_config(system, big, p, 2, 1, nsteps=m_int)
                ns[p] = {"value": m["value"], "ms_per_step": m["elapsed"] / 2 * 1e3, "steps_per_block": m_int}
            for p in ns:
                if m_int == bc["n_prop_steps"]:
                    ns[p].update(errs.get(p, {}))
                ns[p]["speedup_vs_fp64"] = ns[p]["value"] / ns["fp64"]["value"]
                ns[p]["whole_step_roofline"] = whole_step_roofline(bc["norb"], bc["nocc_up"], bc["nocc_dn"],
                                                                   bc["nchol"], m_int, False, p, ns[p]["value"])
            line[key] = ns
        del system
        torch.cuda.empty_cache()
    if args.serving_jobs > 0:
        from ad_afqmc_amd import serving

        torch.cuda.empty_cache()
        print(f"[bench] serving stream: {args.serving_jobs} jobs", file=sys.stderr, flush=True)
        specs = serving.poisson_stream(args.serving_jobs, rate=4.0, seed=0, precision=args.serving_precision)
        sch = serving.Scheduler(serving.GpuExecutor(n_prop_steps=10), hbm_cap_bytes=200e9, max_active=64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = sch.run(specs, realtime=True)
        torch.cuda.synchronize()
        summ = serving.summarize(res, time.perf_counter() - t0)
        summ.update(arrival_rate_per_s=4.0, precision=args.serving_precision, hbm_cap_gb=200,
                    max_concurrent_jobs=sch.max_concurrent, peak_admitted_gb=sch.peak_used / 1e9, hip_graphs=True,
                    workload="N~U{30..250}, O=N/5, G=5N, W~U{64..512}, 300 steps, Poisson 4 jobs/s")
        line["serving"] = summ
    line["cpu_baseline"] = cpu
    line["cpu_baseline_all_cores"] = cpu_all
    line["host"] = host_description()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

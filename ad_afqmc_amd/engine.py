"""Job handle over libafq.so with PyTorch-ROCm tensors as device buffers.

Every call is enqueued on torch's current HIP stream; nothing synchronises the
host except `build_propagation` (the expm scaling exponent is chosen on the
host) and explicit `.cpu()` conversions by the caller.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np
import torch

from . import _native as nat


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("ad_afqmc_amd runs on a HIP device; torch.cuda.is_available() is False")
    return torch.device("cuda")


def stream_ptr() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: Optional[torch.Tensor]) -> C.c_void_p:
    return C.c_void_p(0 if t is None else t.data_ptr())


def check_tensor(name: str, t, shape, dtype) -> None:
    """Raises ValueError unless t is a contiguous device tensor of exactly this
    shape and dtype: the kernels index raw pointers by the job's sizes, so a
    mismatched buffer must be refused before it reaches them."""
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch.Tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name} has dtype {t.dtype}, expected {dtype}")
    if tuple(t.shape) != tuple(int(x) for x in shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(int(x) for x in shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor (it is on {t.device})")


def walker_shapes(norb: int, nelec: Tuple[int, int], walker_type: str, n: int):
    """Buffer shapes of n walkers: restricted (n, N, nelec_up) and no second
    buffer; unrestricted (n, N, nelec_up) and (n, N, nelec_dn)."""
    if walker_type == "restricted":
        return (n, norb, nelec[0]), None
    return (n, norb, nelec[0]), (n, norb, nelec[1])


def as_f64(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float64))).to(device())


def as_c128(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=torch.complex128).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.complex128))).to(device())


class EngineJob:
    """One resident AFQMC job (Hamiltonian, trial, intermediates, scratch)."""

    def __init__(
        self,
        norb: int,
        nelec: Tuple[int, int],
        nchol: int,
        n_walkers: int,
        walker_type: str = "restricted",
        trial_type: str = "rhf",
        n_exp_terms: int = 6,
        dt: float = 0.01,
        precision: str = "fp64",
    ):
        device()
        if walker_type not in ("restricted", "unrestricted"):
            raise NotImplementedError(f"walker_type {walker_type!r} is not supported by the engine")
        if trial_type not in ("rhf", "uhf"):
            raise NotImplementedError(f"trial {trial_type!r} is not supported by the engine")
        self.norb, self.nelec, self.nchol = int(norb), (int(nelec[0]), int(nelec[1])), int(nchol)
        self.capacity = self.batch = int(n_walkers)
        self.walker_type, self.trial_type = walker_type, trial_type
        self.n_exp_terms, self.dt, self.precision = int(n_exp_terms), float(dt), precision
        desc = nat.JobDesc(
            norb=self.norb,
            nelec_up=self.nelec[0],
            nelec_dn=self.nelec[1],
            nchol=self.nchol,
            n_walkers=self.capacity,
            walker_type=nat.AFQ_WALKERS_RESTRICTED if walker_type == "restricted" else nat.AFQ_WALKERS_UNRESTRICTED,
            trial_type=nat.AFQ_TRIAL_RHF if trial_type == "rhf" else nat.AFQ_TRIAL_UHF,
            n_exp_terms=self.n_exp_terms,
            precision=nat.PRECISIONS[precision],
            reserved=0,
            dt=self.dt,
        )
        h = C.c_void_p()
        nat.call("afq_job_create", C.byref(h), C.byref(desc))
        self._h = h
        self._lib = nat.lib()

    # ------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.afq_job_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def restricted(self) -> bool:
        return self.walker_type == "restricted"

    def set_batch(self, n: int):
        if n != self.batch:
            nat.call("afq_job_set_batch", self._h, int(n))
            self.batch = int(n)

    # ------------------------------------------------------------ setup
    def set_hamiltonian(self, h0: float, h1=None, chol=None):
        if h1 is None:
            nat.call("afq_job_set_hamiltonian", self._h, float(h0), None, None, stream_ptr())
            return
        h1 = as_f64(h1).reshape(-1, self.norb, self.norb)
        if h1.shape[0] == 1:
            h1 = torch.cat([h1, h1])
        chol = as_f64(chol).reshape(self.nchol, self.norb * self.norb)
        nat.call("afq_job_set_hamiltonian", self._h, float(h0), ptr(h1), ptr(chol), stream_ptr())

    def set_trial(self, mo_up, mo_dn, rdm1):
        N = self.norb
        up = as_f64(mo_up)
        check_tensor("mo_up", up, (N, self.nelec[0]), torch.float64)
        dn = None if mo_dn is None else as_f64(mo_dn)
        if self.trial_type == "uhf":
            if dn is None:
                raise ValueError("a UHF trial needs mo_dn")
            check_tensor("mo_dn", dn, (N, self.nelec[1]), torch.float64)
        rdm = as_f64(rdm1)
        check_tensor("rdm1", rdm, (2, N, N), torch.float64)
        nat.call("afq_job_set_trial", self._h, ptr(up), ptr(dn), ptr(rdm), stream_ptr())

    def build_measurement(self):
        nat.call("afq_build_measurement_intermediates", self._h, stream_ptr())

    def build_propagation(self):
        nat.call("afq_build_propagation_intermediates", self._h, stream_ptr())

    def _int_shape(self, which: int, spin: int):
        N, G, O = self.norb, self.nchol, self.nelec[spin]
        return {
            nat.AFQ_INT_H1: ((2, N, N), torch.float64),
            nat.AFQ_INT_ROT_H1: ((O, N), torch.float64),
            nat.AFQ_INT_ROT_CHOL: ((G, O, N), torch.float64),
            nat.AFQ_INT_MF_SHIFTS: ((G,), torch.complex128),
            nat.AFQ_INT_H0_PROP: ((), torch.complex128),
            nat.AFQ_INT_EXP_H1: ((N, N), torch.float64),
        }[which]

    def get(self, which: int, spin: int = 0) -> torch.Tensor:
        shape, dtype = self._int_shape(which, spin)
        out = torch.empty(shape, dtype=dtype, device=device())
        nat.call("afq_get_intermediate", self._h, which, spin, ptr(out), stream_ptr())
        return out

    def set(self, which: int, spin: int, value):
        shape, dtype = self._int_shape(which, spin)
        t = as_c128(value) if dtype == torch.complex128 else as_f64(value)
        if tuple(t.shape) != tuple(shape):
            t = t.reshape(shape)
        nat.call("afq_set_intermediate", self._h, which, spin, ptr(t), stream_ptr())

    # ------------------------------------------------------------ measurements
    def _w(self, up: torch.Tensor, dn: Optional[torch.Tensor]):
        if not isinstance(up, torch.Tensor) or up.dim() != 3:
            raise ValueError("walkers must be a (n_walkers, norb, nocc) tensor")
        n = int(up.shape[0])
        if n < 1 or n > self.capacity:
            raise ValueError(f"{n} walkers outside [1, job capacity {self.capacity}]")
        su, sd = walker_shapes(self.norb, self.nelec, self.walker_type, n)
        check_tensor("walkers_up" if sd is not None else "walkers", up, su, torch.complex128)
        if sd is None:
            if dn is not None:
                raise ValueError("restricted walkers take one buffer (walkers_dn must be None)")
        else:
            if dn is None:
                raise ValueError("unrestricted walkers need walkers_dn")
            check_tensor("walkers_dn", dn, sd, torch.complex128)
        self.set_batch(n)
        return ptr(up), ptr(dn)

    def _state(self, n, weights=None, overlaps=None, scalars=None):
        if weights is not None:
            check_tensor("weights", weights, (n,), torch.float64)
        if overlaps is not None:
            check_tensor("overlaps", overlaps, (n,), torch.complex128)
        if scalars is not None:
            check_tensor("scalars", scalars, (2,), torch.float64)

    def overlap(self, up, dn=None) -> torch.Tensor:
        pu, pd = self._w(up, dn)
        out = torch.empty(up.shape[0], dtype=torch.complex128, device=up.device)
        nat.call("afq_calc_overlap", self._h, pu, pd, ptr(out), stream_ptr())
        return out

    def green(self, up, dn=None):
        pu, pd = self._w(up, dn)
        W, N = up.shape[0], self.norb
        gu = torch.empty((W, self.nelec[0], N), dtype=torch.complex128, device=up.device)
        gd = torch.empty((W, self.nelec[1], N), dtype=torch.complex128, device=up.device)
        nat.call("afq_calc_green", self._h, pu, pd, ptr(gu), ptr(gd), stream_ptr())
        return gu, gd

    def force_bias(self, up, dn=None) -> torch.Tensor:
        pu, pd = self._w(up, dn)
        out = torch.empty((up.shape[0], self.nchol), dtype=torch.complex128, device=up.device)
        nat.call("afq_calc_force_bias", self._h, pu, pd, ptr(out), stream_ptr())
        return out

    def energy(self, up, dn=None) -> torch.Tensor:
        pu, pd = self._w(up, dn)
        out = torch.empty(up.shape[0], dtype=torch.complex128, device=up.device)
        nat.call("afq_calc_energy", self._h, pu, pd, ptr(out), stream_ptr())
        return out

    # ------------------------------------------------------------ propagation (in place)
    def trotprop(self, up, dn, shifted_fields: torch.Tensor):
        pu, pd = self._w(up, dn)
        check_tensor("shifted_fields", shifted_fields, (up.shape[0], self.nchol), torch.complex128)
        nat.call("afq_apply_trotprop", self._h, pu, pd, ptr(shifted_fields), stream_ptr())

    def propagate(self, up, dn, fields, weights, overlaps, scalars, fb_out=None):
        pu, pd = self._w(up, dn)
        n = up.shape[0]
        check_tensor("fields", fields, (n, self.nchol), torch.float64)
        self._state(n, weights, overlaps, scalars)
        if fb_out is not None:
            check_tensor("force_bias_out", fb_out, (n, self.nchol), torch.complex128)
        nat.call(
            "afq_propagate_constrained",
            self._h,
            pu,
            pd,
            ptr(fields),
            ptr(weights),
            ptr(overlaps),
            ptr(scalars),
            ptr(fb_out),
            stream_ptr(),
        )

    def orthonormalize(self, up, dn=None):
        pu, pd = self._w(up, dn)
        nat.call("afq_orthonormalize", self._h, pu, pd, stream_ptr())

    def stochastic_reconfiguration(self, up, dn, weights, zeta: torch.Tensor, idx_out=None):
        pu, pd = self._w(up, dn)
        self._state(up.shape[0], weights)
        check_tensor("zeta", zeta, (1,), torch.float64)
        if idx_out is not None:
            check_tensor("indices_out", idx_out, (up.shape[0],), torch.int32)
        nat.call("afq_stochastic_reconfiguration", self._h, pu, pd, ptr(weights), ptr(zeta), ptr(idx_out), stream_ptr())

    # ------------------------------------------------------------ live kernel timing
    def profile(self, kernel_class: str, enable: bool = True):
        nat.call("afq_job_profile", self._h, nat.KERNEL_CLASSES.index(kernel_class), int(enable))

    def profile_read(self, kernel_class: str):
        """(summed ms, launches) of the class since the last read (synchronises)."""
        ms, n = C.c_double(), C.c_int64()
        nat.call("afq_job_profile_read", self._h, nat.KERNEL_CLASSES.index(kernel_class), C.byref(ms), C.byref(n))
        return float(ms.value), int(n.value)

    def block(self, up, dn, weights, overlaps, scalars, n_prop_steps: int, seed: int, counter: int):
        """One sampler block on the device; returns ({block_energy, block_weight} tensor, new counter)."""
        pu, pd = self._w(up, dn)
        self._state(up.shape[0], weights, overlaps, scalars)
        if int(n_prop_steps) < 1:
            raise ValueError("n_prop_steps must be >= 1")
        out = torch.empty(2, dtype=torch.float64, device=up.device)
        ctr = C.c_uint64(counter)
        nat.call(
            "afq_block",
            self._h,
            pu,
            pd,
            ptr(weights),
            ptr(overlaps),
            ptr(scalars),
            int(n_prop_steps),
            C.c_uint64(seed),
            C.byref(ctr),
            ptr(out),
            stream_ptr(),
        )
        return out, int(ctr.value)

    def reserve_fields(self, n_prop_steps: int):
        """Allocates the field buffer of n_prop_steps steps (before graph capture)."""
        nat.call("afq_job_reserve_fields", self._h, int(n_prop_steps), stream_ptr())

    def sr_block(self, up, dn, weights, overlaps, scalars, n_prop_steps: int, seed: int, counter_dev: torch.Tensor,
                 out: torch.Tensor):
        """One block + comb resampling with the stream counter in device memory
        (counter_dev: one int64 element, advanced on the device); capturable in a
        HIP graph.  out (2,) receives {block_energy, block_weight}."""
        check_tensor("counter_dev", counter_dev, (1,), torch.int64)
        check_tensor("block_out", out, (2,), torch.float64)
        pu, pd = self._w(up, dn)
        self._state(up.shape[0], weights, overlaps, scalars)
        nat.call(
            "afq_sr_block",
            self._h,
            pu,
            pd,
            ptr(weights),
            ptr(overlaps),
            ptr(scalars),
            int(n_prop_steps),
            C.c_uint64(seed),
            ptr(counter_dev),
            ptr(out),
            stream_ptr(),
        )

    def segment_energy_vjp(self, walkers, weights, e_estimate: float, fields, want_grad: bool = True,
                           walkers_dn=None):
        """Segment energy F over fields (n_blocks, n_steps, W, G) and, with
        want_grad, dF/dh1 (2,N,N) and dF/dchol (G,N*N) (afq_segment_energy_vjp).
        RHF trial with restricted walkers (W,N,nocc), or UHF trial with
        unrestricted walkers (walkers (W,N,nocc_up) and walkers_dn (W,N,nocc_dn)).
        Synchronises the stream."""
        if self.restricted != (self.trial_type == "rhf"):
            raise NotImplementedError("the segment derivative takes an RHF trial with restricted walkers "
                                      "or a UHF trial with unrestricted walkers")
        f = as_f64(fields)
        if f.dim() != 4 or f.shape[2] != self.batch or f.shape[3] != self.nchol:
            raise ValueError(f"fields must be (n_blocks, n_steps, {self.batch}, {self.nchol}), got {tuple(f.shape)}")
        su, sd = walker_shapes(self.norb, self.nelec, self.walker_type, self.batch)
        w = as_c128(walkers)
        check_tensor("walkers" if sd is None else "walkers_up", w, su, torch.complex128)
        if sd is None:
            if walkers_dn is not None:
                raise ValueError("restricted walkers take one buffer (walkers_dn must be None)")
        else:
            if walkers_dn is None:
                raise ValueError("unrestricted walkers need walkers_dn")
            wd = as_c128(walkers_dn)
            check_tensor("walkers_dn", wd, sd, torch.complex128)
            w = torch.cat([w.reshape(-1), wd.reshape(-1)])  # spin 1's panel after spin 0's
        wt = as_f64(weights)
        check_tensor("weights", wt, (self.batch,), torch.float64)
        out = torch.empty(1, dtype=torch.float64, device=w.device)
        gh1 = gL = None
        if want_grad:
            gh1 = torch.empty((2, self.norb, self.norb), dtype=torch.float64, device=w.device)
            gL = torch.empty((self.nchol, self.norb * self.norb), dtype=torch.float64, device=w.device)
        nat.call(
            "afq_segment_energy_vjp",
            self._h,
            ptr(w),
            ptr(wt),
            float(e_estimate),
            ptr(f),
            int(f.shape[0]),
            int(f.shape[1]),
            ptr(out),
            ptr(gh1),
            ptr(gL),
            stream_ptr(),
        )
        return out[0], gh1, gL


class JobGroup:
    """Several engine jobs stepped together: each `sr_block()` runs one
    afq_sr_block for every member, every kernel of the block launched once
    over all members (afq_group_sr_block).  Members: restricted walkers of an
    RHF trial in bf16x3 (N <= 512) with fields reserved for n_prop_steps.  The
    group keeps each member's walkers, weights, overlaps, scalars, counter and
    block output tensors (it holds references to them); a block also advances
    each e_estimate to 0.9 e_estimate + 0.1 E_block.  Capturable in a HIP
    graph."""

    def __init__(self, members, n_prop_steps: int):
        """members: sequence of (job, walkers, weights, overlaps, scalars,
        counter_dev, block_out, seed)."""
        if not members:
            raise ValueError("a group needs at least one job")
        self._h = C.c_void_p()
        self._keep = []
        n = len(members)
        P = C.c_void_p
        jobs, walkers, weights, overlaps, scalars, counters, outs = ((P * n)() for _ in range(7))
        seeds = (C.c_uint64 * n)()
        for i, (job, up, w, ov, sc, cnt, out, seed) in enumerate(members):
            pu, _ = job._w(up, None)
            job._state(up.shape[0], w, ov, sc)
            check_tensor("counter_dev", cnt, (1,), torch.int64)
            check_tensor("block_out", out, (2,), torch.float64)
            jobs[i] = job._h.value
            walkers[i] = pu.value
            weights[i] = w.data_ptr()
            overlaps[i] = ov.data_ptr()
            scalars[i] = sc.data_ptr()
            counters[i] = cnt.data_ptr()
            outs[i] = out.data_ptr()
            seeds[i] = int(seed)
            self._keep.append((job, up, w, ov, sc, cnt, out))
        nat.call("afq_group_create", C.byref(self._h), jobs, n, walkers, weights, overlaps, scalars, counters, outs,
                 seeds, int(n_prop_steps))

    def sr_block(self):
        nat.call("afq_group_sr_block", self._h, stream_ptr())

    def close(self):
        if self._h:
            nat.lib().afq_group_destroy(self._h)
            self._h = C.c_void_p()
        self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def normal_fill(n: int, seed: int, counter: int) -> torch.Tensor:
    """N(0,1) draws of the engine's Philox4x32-10 field generator."""
    out = torch.empty(int(n), dtype=torch.float64, device=device())
    nat.call("afq_normal_fill", ptr(out), int(n), C.c_uint64(seed), C.c_uint64(counter), stream_ptr())
    return out


def _walker_buffer(name: str, w) -> Tuple[int, int, int]:
    if not isinstance(w, torch.Tensor) or w.dim() != 3:
        raise ValueError(f"{name} must be a (n_walkers, norb, ncols) tensor")
    check_tensor(name, w, tuple(w.shape), torch.complex128)
    return tuple(int(x) for x in w.shape)


def sr_zeta(seed: int, counter: int) -> torch.Tensor:
    """The comb offset afq_sr_block draws for the block keyed by (seed, counter)."""
    out = torch.empty(1, dtype=torch.float64, device=device())
    nat.call("afq_sr_zeta", ptr(out), C.c_uint64(seed), C.c_uint64(counter), stream_ptr())
    return out


def walkers_orthonormalize(w: torch.Tensor):
    W, N, c = _walker_buffer("walkers", w)
    nat.call("afq_walkers_orthonormalize", ptr(w), int(W), int(N), int(c), stream_ptr())


def walkers_sr(up: torch.Tensor, dn: Optional[torch.Tensor], weights: torch.Tensor, zeta: torch.Tensor):
    W = _walker_buffer("walkers_up", up)[0]
    if dn is not None and _walker_buffer("walkers_dn", dn)[0] != W:
        raise ValueError("walkers_up and walkers_dn hold different walker counts")
    check_tensor("weights", weights, (W,), torch.float64)
    check_tensor("zeta", zeta, (1,), torch.float64)
    eu = up[0].numel()
    ed = 0 if dn is None else dn[0].numel()
    ws_bytes = nat.lib().afq_walkers_sr_workspace(int(W), int(max(eu, ed)))
    ws = torch.empty(int(ws_bytes), dtype=torch.uint8, device=up.device)
    idx = torch.empty(W, dtype=torch.int32, device=up.device)
    nat.call(
        "afq_walkers_sr", ptr(up), int(eu), ptr(dn), int(ed), int(W), ptr(weights), ptr(zeta), ptr(idx), ptr(ws), stream_ptr()
    )
    return idx

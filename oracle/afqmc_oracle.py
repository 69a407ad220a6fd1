"""FP64 CPU oracle for the phaseless-AFQMC walker step (RHF/UHF trials).

TEST INFRASTRUCTURE ONLY.  This module is a plain-NumPy restatement of the
reference's hot path, used as the checker for the HIP engine.  Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import
it; the product package (`ad_afqmc_amd`) never does, and it has no fallback to
it.

Every function cites the reference file:line it restates (paths are relative to
the reference checkout).  Pinning (what checks each function independently of
this restatement) is recorded per function in the `PINS` table at the bottom;
functions listed there as "parity unpinned" have no independent check beyond
the reference's own semantics as read.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple, Union

import numpy as np
import scipy.linalg

Walkers = Union[np.ndarray, List[np.ndarray]]


# ---------------------------------------------------------------------------
# Setup intermediates
# ---------------------------------------------------------------------------
def symmetrize_h1(h1: np.ndarray) -> np.ndarray:
    """(h1[s] + h1[s].T)/2 per spin.  wavefunctions.py:1218-1223, 1473-1478."""
    h1 = np.asarray(h1, dtype=np.float64)
    return np.stack([(h1[0] + h1[0].T) / 2.0, (h1[1] + h1[1].T) / 2.0])


def rhf_measurement_intermediates(h1, chol, mo):
    """rhf._build_measurement_intermediates, wavefunctions.py:1215-1232.

    Returns (h1_symmetrised (2,N,N), rot_h1 (O,N), rot_chol (G,O,N))."""
    h1 = symmetrize_h1(h1)
    n = h1.shape[-1]
    L = np.asarray(chol, dtype=np.float64).reshape(-1, n, n)
    rot_h1 = mo.T.conj() @ ((h1[0] + h1[1]) / 2.0)
    rot_chol = np.einsum("pi,gij->gpj", mo.T.conj(), L, optimize=True)
    return h1, rot_h1, rot_chol


def uhf_measurement_intermediates(h1, chol, mo_up, mo_dn):
    """uhf._build_measurement_intermediates, wavefunctions.py:1472-1496."""
    h1 = symmetrize_h1(h1)
    n = h1.shape[-1]
    L = np.asarray(chol, dtype=np.float64).reshape(-1, n, n)
    rot_h1 = [mo_up.T.conj() @ h1[0], mo_dn.T.conj() @ h1[1]]
    rot_chol = [
        np.einsum("pi,gij->gpj", mo_up.T.conj(), L, optimize=True),
        np.einsum("pi,gij->gpj", mo_dn.T.conj(), L, optimize=True),
    ]
    return h1, rot_h1, rot_chol


def propagation_intermediates(h0, h1, chol, rdm1, dt, walker_type="restricted"):
    """propagator_afqmc._build_propagation_intermediates, propagation.py:378-438
    (restricted and unrestricted walkers; the free-projection fields :393-410 are
    not part of the phaseless path).

    Returns dict(mf_shifts (G,) complex, h0_prop complex, exp_h1 (N,N) or (2,N,N),
    v0, v1, h1_mod)."""
    h1 = np.asarray(h1, dtype=np.float64)
    n = h1.shape[-1]
    L = np.asarray(chol, dtype=np.float64).reshape(-1, n, n)
    rho = rdm1[0] + rdm1[1]  # :386
    mf_shifts = 1.0j * np.einsum("gpq,pq->g", L, rho)  # :387-389
    h0_prop = -h0 - np.sum(mf_shifts**2) / 2.0  # :390-392
    v0 = 0.5 * np.einsum("gik,gkj->ij", L, L, optimize=True)  # :412-417
    mf_r = (1.0j * mf_shifts).real  # :418
    v1 = np.einsum("g,gik->ik", mf_r, L)  # :419-423
    if walker_type == "unrestricted":  # :427-434
        h1_mod = h1 - np.array([v0 + v1, v0 + v1])
        exp_h1 = np.array(
            [scipy.linalg.expm(-dt * h1_mod[0] / 2.0), scipy.linalg.expm(-dt * h1_mod[1] / 2.0)]
        )
    else:  # :435-437
        h1_mod = (h1[0] + h1[1]) / 2.0 - v0 - v1
        exp_h1 = scipy.linalg.expm(-dt * h1_mod / 2.0)
    return dict(mf_shifts=mf_shifts, h0_prop=h0_prop, exp_h1=exp_h1, v0=v0, v1=v1, h1_mod=h1_mod)


def rdm1_from_mo(mo_up, mo_dn=None):
    """rhf._calc_rdm1 / uhf._calc_rdm1, wavefunctions.py:1176-1178, 1400-1403."""
    if mo_dn is None:
        d = mo_up @ mo_up.T
        return np.array([d, d])
    return np.array([mo_up @ mo_up.T.conj(), mo_dn @ mo_dn.T.conj()])


def init_walkers(rdm1, nelec, n_walkers, walker_type):
    """wave_function.get_init_walkers, wavefunctions.py:427-466 (natural orbitals,
    descending occupation, replicated, complex)."""
    if walker_type == "restricted":
        nat = np.linalg.eigh(rdm1[0] + rdm1[1])[1][:, ::-1]
        return np.array([nat[:, : nelec[0]] + 0.0j] * n_walkers)
    up = np.linalg.eigh(rdm1[0])[1][:, ::-1][:, : nelec[0]]
    dn = np.linalg.eigh(rdm1[1])[1][:, ::-1][:, : nelec[1]]
    return [np.array([up + 0.0j] * n_walkers), np.array([dn + 0.0j] * n_walkers)]


# ---------------------------------------------------------------------------
# Trial measurements (batched over walkers)
# ---------------------------------------------------------------------------
def green(walkers: np.ndarray, mo: np.ndarray) -> np.ndarray:
    """Half-rotated Green's function G = (W (T^H W)^-1)^T per walker,
    rhf._calc_green wavefunctions.py:1089-1100 and uhf._calc_green :1289-1312.
    walkers (W,N,O) -> (W,O,N)."""
    S = np.einsum("pi,wpj->wij", mo.conj(), walkers)
    Sinv = np.linalg.inv(S)
    return np.einsum("wpk,wki->wip", walkers, Sinv)


def det_overlap(walkers: np.ndarray, mo: np.ndarray) -> np.ndarray:
    """det(T^H W) per walker (jnp.linalg.det; wavefunctions.py:1066, 1266-1268)."""
    S = np.einsum("pi,wpj->wij", mo.conj(), walkers)
    return np.linalg.det(S)


def _real_times_complex(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """A @ B for real A and complex B as two real products (the same values as
    the complex product, without upcasting A)."""
    return (A @ np.ascontiguousarray(B.real)) + 1.0j * (A @ np.ascontiguousarray(B.imag))


def _exchange_and_coulomb(rot_chol: np.ndarray, G: np.ndarray):
    """c_g = tr f_g and K = sum_g sum_ij f_gij f_gji with f_g = rot_chol_g . G^T,
    one walker at a time (wavefunctions.py:1137-1139, 1347-1356)."""
    Wn = G.shape[0]
    ng, no, n = rot_chol.shape
    rc2 = np.asarray(rot_chol, dtype=np.float64).reshape(ng * no, n)
    c = np.empty((Wn, ng), dtype=np.complex128)
    K = np.empty(Wn, dtype=np.complex128)
    for w in range(Wn):
        f = _real_times_complex(rc2, G[w].T).reshape(ng, no, G.shape[1])
        c[w] = np.einsum("gii->g", f)
        K[w] = np.sum(f * f.transpose(0, 2, 1))
    return c, K


def _fb_contract(rot_chol: np.ndarray, G: np.ndarray) -> np.ndarray:
    """sum_ij rot_chol_gij G_wij -> (W, G) (wavefunctions.py:1107, 1323-1328)."""
    rc2 = np.asarray(rot_chol, dtype=np.float64).reshape(rot_chol.shape[0], -1)
    g2 = G.reshape(G.shape[0], -1)
    return _real_times_complex(rc2, g2.T).T


class RHFTrial:
    """rhf trial (wavefunctions.py:1038-1236) with RHF (restricted) and UHF
    (unrestricted) walker batches."""

    def __init__(self, norb: int, nelec: Tuple[int, int], mo: np.ndarray):
        assert nelec[0] == nelec[1]  # :1060-1063
        self.norb, self.nelec, self.mo = norb, tuple(nelec), np.asarray(mo)

    def split(self, walkers: Walkers):
        if isinstance(walkers, (list, tuple)):
            return walkers[0], walkers[1], False
        return walkers, walkers, True

    def calc_overlap(self, walkers: Walkers) -> np.ndarray:
        """_calc_overlap_restricted / _unrestricted, wavefunctions.py:1064-1074."""
        up, dn, restricted = self.split(walkers)
        if restricted:
            return det_overlap(up, self.mo) ** 2
        return det_overlap(up, self.mo) * det_overlap(dn, self.mo)

    def calc_green(self, walkers: Walkers):
        up, dn, restricted = self.split(walkers)
        gu = green(up, self.mo)
        return (gu, gu) if restricted else (gu, green(dn, self.mo))

    def calc_force_bias(self, walkers: Walkers, rot_chol: np.ndarray) -> np.ndarray:
        """_calc_force_bias_restricted/_unrestricted, wavefunctions.py:1102-1126."""
        up, dn, restricted = self.split(walkers)
        if restricted:
            return 2.0 * _fb_contract(rot_chol, green(up, self.mo))
        g = green(up, self.mo) + green(dn, self.mo)
        return _fb_contract(rot_chol, g)

    def calc_energy(self, walkers: Walkers, h0, rot_h1, rot_chol) -> np.ndarray:
        """_calc_energy_restricted/_unrestricted, wavefunctions.py:1128-1174."""
        up, dn, restricted = self.split(walkers)
        if restricted:
            G = green(up, self.mo)
            ene1 = 2.0 * np.einsum("wij,ij->w", G, rot_h1)
            c, exc = _exchange_and_coulomb(rot_chol, G)
            ene2 = 2.0 * np.sum(c * c, axis=1) - exc
            return ene2 + ene1 + h0
        Gu, Gd = green(up, self.mo), green(dn, self.mo)
        ene1 = np.einsum("wij,ij->w", Gu + Gd, rot_h1)
        cu, ku = _exchange_and_coulomb(rot_chol, Gu)
        cd, kd = _exchange_and_coulomb(rot_chol, Gd)
        ene2 = (np.sum(cu * cu, 1) + np.sum(cd * cd, 1) + 2.0 * np.sum(cu * cd, 1) - ku - kd) / 2.0
        return ene2 + ene1 + h0


class UHFTrial:
    """uhf trial (wavefunctions.py:1238-1500).  Restricted walker batches go
    through the base-class mapping walker[:, :nelec[s]] (wavefunctions.py:181-186,
    237-244, 366-373)."""

    def __init__(self, norb: int, nelec: Tuple[int, int], mo_up: np.ndarray, mo_dn: np.ndarray):
        self.norb, self.nelec = norb, tuple(nelec)
        self.mo = [np.asarray(mo_up), np.asarray(mo_dn)]

    def split(self, walkers: Walkers):
        if isinstance(walkers, (list, tuple)):
            return walkers[0], walkers[1]
        return walkers[:, :, : self.nelec[0]], walkers[:, :, : self.nelec[1]]

    def calc_overlap(self, walkers: Walkers) -> np.ndarray:
        """_calc_overlap_unrestricted, wavefunctions.py:1259-1268."""
        up, dn = self.split(walkers)
        return det_overlap(up, self.mo[0]) * det_overlap(dn, self.mo[1])

    def calc_green(self, walkers: Walkers):
        up, dn = self.split(walkers)
        return green(up, self.mo[0]), green(dn, self.mo[1])

    def calc_force_bias(self, walkers: Walkers, rot_chol: Sequence[np.ndarray]) -> np.ndarray:
        """_calc_force_bias_unrestricted, wavefunctions.py:1314-1329."""
        gu, gd = self.calc_green(walkers)
        return _fb_contract(rot_chol[0], gu) + _fb_contract(rot_chol[1], gd)

    def calc_energy(self, walkers: Walkers, h0, rot_h1, rot_chol) -> np.ndarray:
        """_calc_energy_unrestricted, wavefunctions.py:1331-1363."""
        gu, gd = self.calc_green(walkers)
        ene1 = np.einsum("wij,ij->w", gu, rot_h1[0]) + np.einsum("wij,ij->w", gd, rot_h1[1])
        cu, ku = _exchange_and_coulomb(rot_chol[0], gu)
        cd, kd = _exchange_and_coulomb(rot_chol[1], gd)
        ene2 = (np.sum(cu * cu, 1) + np.sum(cd * cd, 1) + 2.0 * np.sum(cu * cd, 1) - ku - kd) / 2.0
        return ene2 + ene1 + h0


# ---------------------------------------------------------------------------
# Propagation
# ---------------------------------------------------------------------------
def vhs_from_fields(shifted_fields: np.ndarray, chol: np.ndarray, dt: float) -> np.ndarray:
    """VHS = i sqrt(dt) (x~ @ L).reshape(N,N), propagation.py:196-203, 215-222."""
    n2 = chol.reshape(chol.shape[0], -1).shape[1]
    n = int(round(math.sqrt(n2)))
    L = np.asarray(chol, dtype=np.float64).reshape(-1, n2)
    x = np.asarray(shifted_fields, dtype=np.complex128)
    # complex @ real as one real product of the stacked real and imaginary rows
    W = x.shape[0]
    v = np.concatenate([x.real, x.imag], axis=0) @ L
    v = v[:W] + 1.0j * v[W:]
    return 1.0j * math.sqrt(dt) * v.reshape(-1, n, n)


def apply_exp_vhs(walker: np.ndarray, vhs: np.ndarray, exp_h1: np.ndarray, n_exp_terms: int):
    """_apply_trotprop_single_det_vhs, propagation.py:154-176 (one walker)."""
    walker = exp_h1 @ walker
    carry = walker
    terms = []
    for _ in range(1, n_exp_terms):
        carry = vhs @ carry
        terms.append(carry)
    fact = np.array([1.0 / math.factorial(n + 1) for n in range(n_exp_terms - 1)]).reshape(-1, 1, 1)
    if terms:
        walker = walker + np.sum(np.array(terms) * fact, axis=0)
    return exp_h1 @ walker


def apply_trotprop(walkers: Walkers, shifted_fields, chol, exp_h1, dt, n_exp_terms=6) -> Walkers:
    """_apply_trotprop for RHF/UHF walker batches, propagation.py:178-272."""
    vhs = vhs_from_fields(shifted_fields, chol, dt)
    if isinstance(walkers, (list, tuple)):
        up = np.array([apply_exp_vhs(w, v, exp_h1[0], n_exp_terms) for w, v in zip(walkers[0], vhs)])
        dn = np.array([apply_exp_vhs(w, v, exp_h1[1], n_exp_terms) for w, v in zip(walkers[1], vhs)])
        return [up, dn]
    return np.array([apply_exp_vhs(w, v, exp_h1, n_exp_terms) for w, v in zip(walkers, vhs)])


def step_terms(fields: np.ndarray, fb: np.ndarray, mf_shifts: np.ndarray, dt: float):
    """Field shifts, shifted fields, shift_term and fb_term of one phaseless
    step, propagation.py:296-301.  Returns (field_shifts, shifted, shift_term,
    fb_term)."""
    sqdt = math.sqrt(dt)
    field_shifts = -sqdt * (1.0j * fb - mf_shifts)  # :296
    shifted = fields - field_shifts  # :297
    shift_term = np.sum(shifted * mf_shifts, axis=1)  # :298
    fb_term = np.sum(fields * field_shifts - field_shifts * field_shifts / 2.0, axis=1)  # :299-301
    return field_shifts, shifted, shift_term, fb_term


def importance_function(shift_term, fb_term, ov_new, ov_old, e_shift, h0_prop, dt: float):
    """Unclipped complex importance function and the phase angle of the
    phaseless constraint, propagation.py:306-319.  Returns (imp, theta)."""
    sqdt = math.sqrt(dt)
    ratio = ov_new / ov_old
    imp = np.exp(-sqdt * shift_term + fb_term + dt * (e_shift + h0_prop)) * ratio  # :306-314
    theta = np.angle(np.exp(-sqdt * shift_term) * ratio)  # :315-319
    return imp, theta


def phaseless_factor(imp, theta):
    """|I| cos(theta) with the reference's clipping, propagation.py:320-327."""
    ph = np.abs(imp) * np.cos(theta)  # :320
    ph = np.where(np.isnan(ph), 0.0, ph)
    ph = np.where(ph < 1.0e-3, 0.0, ph)
    return np.where(ph > 100.0, 0.0, ph)  # :321-327


def propagate_constrained(trial, state: dict, fields: np.ndarray, ham: dict, dt: float, n_exp_terms=6) -> dict:
    """propagator_afqmc.propagate_constrained, propagation.py:274-336.

    state: walkers, weights (W,), overlaps (W,) complex, e_estimate, pop_control_ene_shift.
    ham: chol (G,N*N), rot_chol, mf_shifts, h0_prop, exp_h1.
    Returns a new state dict (also carrying the force bias, the shifted fields,
    the unclipped complex importance function and the phaseless factor)."""
    W = fields.shape[0]
    fb = trial.calc_force_bias(state["walkers"], ham["rot_chol"])  # :295
    _, shifted, shift_term, fb_term = step_terms(fields, fb, ham["mf_shifts"], dt)  # :296-301
    walkers = apply_trotprop(state["walkers"], shifted, ham["chol"], ham["exp_h1"], dt, n_exp_terms)  # :302-304
    ov_new = trial.calc_overlap(walkers)  # :305
    imp, theta = importance_function(shift_term, fb_term, ov_new, state["overlaps"], state["pop_control_ene_shift"],
                                     ham["h0_prop"], dt)
    ph = phaseless_factor(imp, theta)
    weights = ph * state["weights"]
    weights = np.where(weights > 100.0, 0.0, weights)  # :328-331
    e_shift = state["e_estimate"] - 0.1 * np.log(np.sum(weights) / W) / dt  # :332-334
    out = dict(state)
    out.update(
        walkers=walkers,
        weights=weights,
        overlaps=ov_new,
        pop_control_ene_shift=e_shift,
        force_bias=fb,
        importance=imp,
        phaseless_factor=ph,
        shifted_fields=shifted,
    )
    return out


# ---------------------------------------------------------------------------
# Re-orthonormalisation and population control
# ---------------------------------------------------------------------------
def qr_restricted(walkers: np.ndarray):
    """qr_vmap_restricted, linalg_utils.py:51-55 (Householder QR; R discarded
    by orthonormalize, walkers.py:155-157)."""
    q, r = np.linalg.qr(walkers)
    return q, np.prod(np.diagonal(r, axis1=1, axis2=2), axis=1)


def qr_unrestricted(walkers: Sequence[np.ndarray]):
    """qr_vmap_unrestricted, linalg_utils.py:58-72: first column multiplied by
    sign(prod diag R) (sign(z) = z/|z|)."""
    out, norms = [], []
    for w in walkers:
        q, r = np.linalg.qr(w)
        nf = np.prod(np.diagonal(r, axis1=1, axis2=2), axis=1)
        sgn = nf / np.abs(nf)
        q = q.copy()
        q[:, :, 0] *= sgn.reshape(-1, 1)
        out.append(q)
        norms.append(np.abs(nf))
    return out, np.array(norms)


def orthonormalize(walkers: Walkers) -> Walkers:
    if isinstance(walkers, (list, tuple)):
        return qr_unrestricted(walkers)[0]
    return qr_restricted(walkers)[0]


def sr_indices(weights: np.ndarray, zeta: float):
    """stochastic_reconfiguration_restricted/_unrestricted, sr.py:24-48.
    Returns (indices, new_weights).  Out-of-range indices clamp to W-1 as
    jax gathers do."""
    W = weights.shape[0]
    cum = np.cumsum(np.abs(weights))
    total = cum[-1]
    z = total * (np.arange(W) + zeta) / W
    idx = np.searchsorted(cum, z, side="left")
    idx = np.minimum(idx, W - 1)
    return idx, np.ones(W) * (total / W)


def stochastic_reconfiguration(walkers: Walkers, weights: np.ndarray, zeta: float):
    idx, w = sr_indices(weights, zeta)
    if isinstance(walkers, (list, tuple)):
        return [walkers[0][idx], walkers[1][idx]], w
    return walkers[idx], w


# ---------------------------------------------------------------------------
# Sampler block loop (injected fields / zetas)
# ---------------------------------------------------------------------------
def block_scan(trial, state, fields_block, ham, dt, n_exp_terms=6):
    """sampler._block_scan, sampling.py:52-94, with fields supplied
    (n_prop_steps, W, G) instead of drawn from jax.random."""
    for f in fields_block:
        state = propagate_constrained(trial, state, f, ham, dt, n_exp_terms)
    state["n_killed_walkers"] = state.get("n_killed_walkers", 0) + int(np.sum(state["weights"] == 0))
    state["walkers"] = orthonormalize(state["walkers"])  # :79
    state["overlaps"] = trial.calc_overlap(state["walkers"])  # :80
    e = np.real(trial.calc_energy(state["walkers"], ham["h0"], ham["rot_h1"], ham["rot_chol"]))
    e = np.where(np.abs(e - state["e_estimate"]) > np.sqrt(2.0 / dt), state["e_estimate"], e)  # :84-88
    bw = np.sum(state["weights"])
    be = np.sum(e * state["weights"]) / bw
    state["pop_control_ene_shift"] = 0.9 * state["pop_control_ene_shift"] + 0.1 * be  # :90-92
    return state, be, bw


def propagate_phaseless(trial, state, ham, dt, n_prop_steps, n_ene_blocks, n_sr_blocks, fields_fn, zeta_fn, n_exp_terms=6):
    """sampler.propagate_phaseless + _sr_block_scan, sampling.py:146-168, 343-365.
    fields_fn() -> (n_prop_steps, W, G); zeta_fn() -> float in [0,1)."""
    state["overlaps"] = trial.calc_overlap(state["walkers"])
    state["n_killed_walkers"] = 0
    state["pop_control_ene_shift"] = state["e_estimate"]
    es, ws = [], []
    for _ in range(n_sr_blocks):
        for _ in range(n_ene_blocks):
            state, be, bw = block_scan(trial, state, fields_fn(), ham, dt, n_exp_terms)
            es.append(be)
            ws.append(bw)
        state["walkers"], state["weights"] = stochastic_reconfiguration(state["walkers"], state["weights"], zeta_fn())
        state["overlaps"] = trial.calc_overlap(state["walkers"])
    es, ws = np.array(es), np.array(ws)
    return np.sum(es * ws) / np.sum(ws), state


# ---------------------------------------------------------------------------
# Statistics (driver post-processing)
# ---------------------------------------------------------------------------
def blocking_analysis(weights, energies, neql=0):
    """stat_utils.blocking_analysis, stat_utils.py:8-64 (printing omitted)."""
    n = weights.shape[0] - neql
    weights, energies = weights[neql:], energies[neql:]
    we = weights * energies
    mean = we.sum() / weights.sum()
    sizes = np.array([1, 2, 5, 10, 20, 50, 100, 200, 300, 400, 500, 1000, 10000])
    prev, plateau, errors = 0.0, None, []
    for i in sizes[sizes < n / 2.0]:
        nb = int(n // i)
        bw = np.array([weights[j * i : (j + 1) * i].sum() for j in range(nb)])
        be = np.array([we[j * i : (j + 1) * i].sum() for j in range(nb)]) / bw
        v1, v2 = bw.sum(), (bw**2).sum()
        m = (bw * be).sum() / v1
        err = ((bw * (be - m) ** 2).sum() / (v1 - v2 / v1) / (nb - 1)) ** 0.5
        errors.append(err)
        if err < 1.05 * prev and plateau is None:
            plateau = max(err, prev)
        prev = err
    if plateau is None and errors:
        plateau = 2 * errors[0]
    return mean, plateau


def reject_outliers(data, obs, m=10.0, min_threshold=1e-5):
    """stat_utils.reject_outliers, stat_utils.py:67-78."""
    t = data[:, obs]
    med = np.median(t)
    d = np.abs(t - med)
    mdev = np.median(d)
    q1, q3 = np.percentile(t, [25, 75])
    disp = max(mdev, (q3 - q1) / 1.349, min_threshold)
    mask = d / disp < m
    return data[mask], mask


# ---------------------------------------------------------------------------
# Pinning table (what checks each function independently of this file)
# ---------------------------------------------------------------------------
PINS = {
    "green/det_overlap/RHFTrial/UHFTrial": "reference golden scalars tests/test_wavefunctions.py:128-184 "
    "(replayed inputs, tests/golden/golden_wavefunctions.npz; tests/test_oracle.py::test_golden_*); brute-force "
    "second quantisation on 3 orbitals (test_energy_bruteforce_*, test_force_bias_bruteforce); W->W.R invariance",
    "rhf/uhf_measurement_intermediates": "W=T closed-form RHF energy (test_energy_at_trial_closed_form)",
    "propagation_intermediates": "mf_shifts: fb(W=T) = -i mf_shifts (test_fb_equals_minus_i_mf_at_trial); "
    "h0_prop, v0, v1: the Fock-space operator identity H = -h0_prop + sum h1_mod E + 1/2 sum (V_g - m_g)^2 "
    "(tests/test_oracle_importance.py::test_mean_field_decomposition_is_the_hamiltonian); exp_h1 vs the "
    "eigendecomposition of the symmetric h1_mod (test_propagation_intermediates_closed_forms)",
    "apply_exp_vhs/apply_trotprop": "error equals the first omitted Taylor term VHS^6 W/6! "
    "(test_taylor_within_remainder_bound)",
    "step_terms/importance_function/propagate_constrained": "Gaussian mean of the unclipped importance function "
    "equals 1 - dt (E_L - E_shift) + O(dt^2) (Gauss-Hermite quadrature, Richardson in dt; "
    "test_girsanov_mean_of_importance_function); W=T, x=0 closed form exp(dt(E_shift+h0_prop)) "
    "det(T^T exp(-dt h1_mod) T) (test_trial_walker_zero_field_closed_form); a sign error planted in fb_term, "
    "shift_term, h0_prop or E_shift fails the Girsanov pin (test_mutations_fail_the_pins); dt->0 leaves weights "
    "at 1 (test_propagate_dt_to_zero_keeps_weights); clipping known answers (GPU test)",
    "phaseless_factor": "clipping known answers (tests/test_gpu_parity.py::test_phaseless_clipping_known_answers) "
    "and |I| at W=T (test_trial_walker_zero_field_closed_form)",
    "qr_restricted/qr_unrestricted": "Q^H Q = I, span preserved, overlap phase preserved (UHF sign fix) "
    "(test_qr_invariants)",
    "sr_indices": "equal weights -> identity; brute-force comb on small W (test_sr_known_answers)",
    "blocking_analysis/reject_outliers": "i.i.d. samples: error ~ sigma/sqrt(n); planted outlier removed "
    "(test_blocking_and_outliers)",
    "propagate_phaseless/block_scan": "composition of the pinned pieces above in the order of sampling.py:52-94, "
    "146-168; H6 fixture mixed energy within the error bar of -3.2382 (tests/test_driver.py:74-86; statistical "
    "only: the threefry field stream is not reproducible, so bit-level driver energies are parity unpinned)",
}

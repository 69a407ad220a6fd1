"""FP64 CPU restatement of the reverse-mode derivative of a phaseless segment.

TEST INFRASTRUCTURE ONLY (same rules as afqmc_oracle: imported by `tests/`
only; the product package never calls it).

The function differentiated is the reference's
`sampler.propagate_phaseless_ad_nosr_norot` (sampling.py:300-341) for an RHF
trial with restricted walkers:

    h1 -> symmetrised h1, rot_h1, rot_chol      (wavefunctions.py:1215-1232)
    -> mf_shifts, h0_prop, v0, v1, exp_h1       (propagation.py:378-438)
    -> overlaps recomputed, E_shift = e_estimate (sampling.py:326-328)
    -> n_blocks x _block_scan                    (sampling.py:52-94)
    F = sum_b E_b W_b / sum_b W_b                (sampling.py:341)

with the fields of every step given (the reference draws them from the
threefry key inside prop_data, which carries no tangent) and e_estimate held
fixed.  The reference obtains dF/dh1 and dF/dchol with jax.vjp
(driver.py:865-876); here the reverse pass is written out by hand:

* complex intermediates carry the adjoint g_z = dF/dRe z - i dF/dIm z, so a
  holomorphic map y = A x pulls back as g_x = A^T g_y (plain transposes);
* a real parameter p entering z gets Re(g_z dz/dp);
* the re-orthonormalisation is a change of gauge W -> W R^-1 under which
  everything downstream is invariant (G, energy, overlap ratios), so its
  adjoint is g_W = g_Q R^-T with R = Q^H W;
* the clipping masks (phaseless factor, weight cap, energy window) are held
  fixed, as jnp.where differentiates them;
* exp_h1's adjoint is the Frechet derivative of expm at A^T
  (scipy.linalg.expm_frechet).

Pinned by central finite differences of `segment_energy` (tests/test_oracle_adjoint.py).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg

from . import afqmc_oracle as orc


def _setup(h0, h1, chol, mo, dt):
    n = mo.shape[0]
    L = np.asarray(chol, dtype=np.float64).reshape(-1, n, n)
    h1s_full, rot_h1, rot_chol = orc.rhf_measurement_intermediates(h1, L, mo)
    rdm1 = orc.rdm1_from_mo(mo)
    pi = orc.propagation_intermediates(h0, h1s_full, L, rdm1, dt)
    return dict(
        h0=h0,
        L=L,
        chol=L.reshape(L.shape[0], -1),
        rot_h1=rot_h1,
        rot_chol=rot_chol,
        mf_shifts=pi["mf_shifts"],
        m=pi["mf_shifts"].imag.copy(),
        h0_prop=pi["h0_prop"],
        exp_h1=pi["exp_h1"],
        h1_mod=pi["h1_mod"],
        rho=rdm1[0] + rdm1[1],
        mo=mo,
    )


def segment_energy(h0, h1, chol, mo, walkers, weights, e_estimate, fields, dt, n_exp_terms=6):
    """Forward value F (the reference's propagate_phaseless_ad_nosr_norot);
    fields: (n_blocks, n_steps, W, G)."""
    P = _setup(h0, h1, chol, mo, dt)
    trial = orc.RHFTrial(mo.shape[0], (mo.shape[1], mo.shape[1]), mo)
    state = dict(walkers=np.array(walkers), weights=np.array(weights, dtype=np.float64), e_estimate=e_estimate)
    state["overlaps"] = trial.calc_overlap(state["walkers"])
    state["pop_control_ene_shift"] = e_estimate
    es, ws = [], []
    for fb in fields:
        state, be, bw = orc.block_scan(trial, state, fb, P, dt, n_exp_terms)
        es.append(be)
        ws.append(bw)
    es, ws = np.array(es), np.array(ws)
    return float(np.sum(es * ws) / np.sum(ws))


# ---------------------------------------------------------------------------
# forward pieces with their tapes
# ---------------------------------------------------------------------------
def _green_parts(W, T):
    S = np.einsum("pi,wpj->wij", T, W)
    Sinv = np.linalg.inv(S)
    Theta = W @ Sinv  # (W,N,O) = G^T
    return S, Sinv, Theta.transpose(0, 2, 1)


def _step_forward(W, w, ov, esh, x, P, dt, K, e_est):
    sqdt = math.sqrt(dt)
    T, rc, m, L, E = P["mo"], P["rot_chol"], P["m"], P["L"], P["exp_h1"]
    _, _, G = _green_parts(W, T)
    fb = 2.0 * np.einsum("gij,wij->wg", rc, G)
    s = -1.0j * sqdt * (fb - m)  # -sqrt(dt)(i fb - mf), mf = i m
    xt = x - s
    sh = xt @ (1.0j * m)  # shift_term
    ft = np.sum(x * s - s * s / 2.0, axis=1)
    vhs = 1.0j * sqdt * np.einsum("wg,gpq->wpq", xt, L)
    W1 = E @ W
    terms = [W1]
    for n in range(1, K + 1):
        terms.append(vhs @ terms[-1] / n)
    W2 = np.sum(terms, axis=0)
    Wn = E @ W2
    ovn = np.linalg.det(np.einsum("pi,wpj->wij", T, Wn)) ** 2
    ratio = ovn / ov
    imp = np.exp(-sqdt * sh + ft + dt * (esh + P["h0_prop"])) * ratio
    theta = np.angle(np.exp(-sqdt * sh) * ratio)
    ph_raw = np.abs(imp) * np.cos(theta)
    mask_ph = ~(np.isnan(ph_raw) | (ph_raw < 1.0e-3) | (ph_raw > 100.0))
    ph = np.where(mask_ph, ph_raw, 0.0)
    wn = ph * w
    mask_w = wn <= 100.0
    wn = np.where(mask_w, wn, 0.0)
    eshn = e_est - 0.1 * np.log(np.sum(wn) / W.shape[0]) / dt
    tape = dict(W=W, w=w, x=x, s=s, xt=xt, vhs=vhs, terms=terms, W2=W2, Wn=Wn, imp=imp, theta=theta,
                ph_raw=ph_raw, mask_ph=mask_ph, ph=ph, wn=wn, mask_w=mask_w, G=G)
    return Wn, wn, ovn, eshn, tape


def _energy_parts(G, P):
    rc, rot_h1 = P["rot_chol"], P["rot_h1"]
    f = np.einsum("gin,wjn->wgij", rc, G)  # f_wg = rc_g G_w^T
    c = np.einsum("wgii->wg", f)
    exc = np.einsum("wgij,wgji->w", f, f)
    e = P["h0"] + 2.0 * np.einsum("wij,ij->w", G, rot_h1) + 2.0 * np.sum(c * c, axis=1) - exc
    return e, f, c


# ---------------------------------------------------------------------------
# reverse pass
# ---------------------------------------------------------------------------
def _step_backward(tape, gWn, gwn, geshn, P, dt, K, acc):
    """Adjoint of one propagate_constrained step (propagation.py:274-336)."""
    sqdt = math.sqrt(dt)
    T, rc, m, L, E = P["mo"], P["rot_chol"], P["m"], P["L"], P["exp_h1"]
    W, w, Wn = tape["W"], tape["w"], tape["Wn"]
    nw = W.shape[0]
    # E_shift' = e_est - 0.1 log(sum w'/W)/dt
    gwn = gwn + geshn * (-0.1 / dt) / np.sum(tape["wn"])
    # w' = mask_w ph w
    gph = np.where(tape["mask_w"], gwn * w, 0.0)
    gw = np.where(tape["mask_w"], gwn * tape["ph"], 0.0)
    # ph = mask_ph |imp| cos(theta)
    alpha = np.where(tape["mask_ph"], gph * tape["ph_raw"], 0.0)
    beta = np.where(tape["mask_ph"], gph * (-np.abs(tape["imp"]) * np.sin(tape["theta"])), 0.0)
    gd = alpha - 1.0j * beta  # adjoint of log(ov'/ov) and of -sqrt(dt) shift_term
    gesh = dt * np.sum(alpha)
    acc["h0_prop"] += dt * np.sum(alpha)
    # log ov' = 2 log det(T^T W')
    _, Sinv_n, _ = _green_parts(Wn, T)
    gWn = gWn + np.einsum("pi,wji->wpj", T, 2.0 * gd[:, None, None] * Sinv_n)
    # shift_term = sum xt * (i m), u = -sqrt(dt) shift_term
    gsh = -sqdt * gd
    gxt = gsh[:, None] * (1.0j * m)[None, :]
    acc["m"] += np.real(np.einsum("w,wg->g", gsh, 1.0j * tape["xt"]))
    # fb_term = sum (x s - s^2/2)
    gs = alpha[:, None] * (tape["x"] - tape["s"])
    # W' = E (sum_n terms_n) with terms_0 = E W, terms_n = vhs terms_{n-1} / n
    gW2 = np.einsum("qp,wqo->wpo", E, gWn)
    acc["E"] += np.real(np.einsum("wpo,wqo->pq", gWn, tape["W2"]))
    vhs, terms = tape["vhs"], tape["terms"]
    gT = gW2
    gvhs = np.zeros_like(vhs)
    for n in range(K, 0, -1):
        gvhs += np.einsum("wpo,wqo->wpq", gT, terms[n - 1]) / n
        gT = gW2 + np.einsum("wqp,wqo->wpo", vhs, gT) / n
    gW1 = gT
    acc["E"] += np.real(np.einsum("wpo,wqo->pq", gW1, W))
    gW = np.einsum("qp,wqo->wpo", E, gW1)
    # vhs = i sqrt(dt) sum_g xt_g L_g
    gxt = gxt + 1.0j * sqdt * np.einsum("gpq,wpq->wg", L, gvhs)
    acc["L"] += np.real(1.0j * sqdt * np.einsum("wg,wpq->gpq", tape["xt"], gvhs))
    # xt = x - s ; s = -i sqrt(dt) (fb - m)
    gs = gs - gxt
    gfb = -1.0j * sqdt * gs
    acc["m"] += np.real(np.sum(1.0j * sqdt * gs, axis=0))
    # fb = 2 sum rc * G
    G = tape["G"]
    gG = 2.0 * np.einsum("wg,gij->wij", gfb, rc)
    acc["rc"] += np.real(2.0 * np.einsum("wg,wij->gij", gfb, G))
    # G^T = W S^-1, S = T^T W ; log ov = 2 log det S enters with -gd
    _, Sinv, _ = _green_parts(W, T)
    gTheta = gG.transpose(0, 2, 1)
    gW = gW + gTheta @ Sinv.transpose(0, 2, 1)
    gS = -(G @ gTheta) @ Sinv.transpose(0, 2, 1) - 2.0 * gd[:, None, None] * Sinv.transpose(0, 2, 1)
    gW = gW + np.einsum("pi,wij->wpj", T, gS)
    return gW, gw, gesh


def _energy_backward(Q, kappa, P, acc):
    """Adjoint of kappa_w * Re E_L(Q_w) (wavefunctions.py:1128-1150)."""
    T, rc, rot_h1 = P["mo"], P["rot_chol"], P["rot_h1"]
    _, Sinv, G = _green_parts(Q, T)
    _, f, c = _energy_parts(G, P)
    k = kappa.astype(np.complex128)
    # dE/dG = 2 rot_h1 + sum_g (4 c_g rc_g - 2 f_g rc_g)
    gG = k[:, None, None] * (
        2.0 * rot_h1[None] + 4.0 * np.einsum("wg,gin->win", c, rc) - 2.0 * np.einsum("wgij,gjn->win", f, rc)
    )
    # dE/drc_g = 4 c_g G - 2 f_g^T G
    acc["rc"] += np.real(
        4.0 * np.einsum("w,wg,win->gin", k, c, G) - 2.0 * np.einsum("w,wgji,wjn->gin", k, f, G)
    )
    acc["rot_h1"] += np.real(2.0 * np.einsum("w,win->in", k, G))
    gTheta = gG.transpose(0, 2, 1)
    gQ = gTheta @ Sinv.transpose(0, 2, 1)
    gS = -(G @ gTheta) @ Sinv.transpose(0, 2, 1)
    return gQ + np.einsum("pi,wij->wpj", T, gS)


def segment_energy_vjp(h0, h1, chol, mo, walkers, weights, e_estimate, fields, dt, n_exp_terms=6):
    """Returns (F, dF/dh1 (2,N,N), dF/dchol (G,N*N))."""
    K = n_exp_terms - 1
    P = _setup(h0, h1, chol, mo, dt)
    T = mo
    n = T.shape[0]
    Lm = P["L"]
    clip = math.sqrt(2.0 / dt)
    nb = len(fields)
    # forward with per-block checkpoints
    W = np.array(walkers, dtype=np.complex128)
    w = np.array(weights, dtype=np.float64)
    ov = np.linalg.det(np.einsum("pi,wpj->wij", T, W)) ** 2
    esh = float(e_estimate)
    ck, blocks = [], []
    for b in range(nb):
        ck.append((W.copy(), w.copy(), ov.copy(), esh))
        tapes = []
        for x in fields[b]:
            W, w, ov, esh, tp = _step_forward(W, w, ov, esh, x, P, dt, K, e_estimate)
            tapes.append(tp)
        Wend = W
        Q = orc.qr_restricted(Wend)[0]
        ov = np.linalg.det(np.einsum("pi,wpj->wij", T, Q)) ** 2
        e_raw = np.real(_energy_parts(_green_parts(Q, T)[2], P)[0])
        cm = np.abs(e_raw - e_estimate) <= clip
        ec = np.where(cm, e_raw, e_estimate)
        bw = np.sum(w)
        be = np.sum(ec * w) / bw
        blocks.append(dict(Wend=Wend, Q=Q, cm=cm, ec=ec, bw=bw, be=be, w=w.copy(), esh_before=esh))
        esh = 0.9 * esh + 0.1 * be
        W = Q
    D = sum(bl["bw"] for bl in blocks)
    F = sum(bl["be"] * bl["bw"] for bl in blocks) / D
    # reverse
    acc = dict(
        E=np.zeros((n, n)), L=np.zeros_like(Lm), rc=np.zeros_like(P["rot_chol"]), rot_h1=np.zeros_like(P["rot_h1"]),
        m=np.zeros(Lm.shape[0]), h0_prop=0.0,
    )
    gW = np.zeros_like(W)
    gw = np.zeros(W.shape[0])
    gesh = 0.0
    for b in range(nb - 1, -1, -1):
        bl = blocks[b]
        # recompute the block's tapes from its checkpoint
        W0, w0, ov0, esh0 = ck[b]
        Wk, wk, ovk, eshk = W0, w0, ov0, esh0
        tapes = []
        for x in fields[b]:
            Wk, wk, ovk, eshk, tp = _step_forward(Wk, wk, ovk, eshk, x, P, dt, K, e_estimate)
            tapes.append(tp)
        # block end: E_shift <- 0.9 E_shift + 0.1 E_b ; F = sum E_b W_b / sum W_b
        gbe = 0.1 * gesh + bl["bw"] / D
        gbw = (bl["be"] - F) / D
        gesh = 0.9 * gesh
        kappa = np.where(bl["cm"], gbe * bl["w"] / bl["bw"], 0.0)
        gw = gw + gbe * (bl["ec"] - bl["be"]) / bl["bw"] + gbw
        gQ = gW + _energy_backward(bl["Q"], kappa, P, acc)
        R = np.einsum("wpi,wpj->wij", bl["Q"].conj(), bl["Wend"])
        gW = gQ @ np.linalg.inv(R).transpose(0, 2, 1)
        for tp in reversed(tapes):
            gW, gw, gesh = _step_backward(tp, gW, gw, gesh, P, dt, K, acc)
    # setup adjoints
    dt2 = -0.5 * dt
    A = dt2 * P["h1_mod"]
    gA = scipy.linalg.expm_frechet(A.T, acc["E"], compute_expm=False)
    ghmod = dt2 * gA
    m = P["m"]
    gm = acc["m"] + acc["h0_prop"] * m  # h0_prop = -h0 + sum m^2 / 2
    gv0 = -ghmod
    gv1 = -ghmod  # v1 = -sum m_g L_g
    gm = gm - np.einsum("gpq,pq->g", Lm, gv1)
    gL = acc["L"].copy()
    gL += 0.5 * (np.einsum("pq,grq->gpr", gv0, Lm) + np.einsum("gqp,qr->gpr", Lm, gv0))
    gL += -m[:, None, None] * gv1[None]
    gL += np.einsum("pi,gin->gpn", T, acc["rc"])
    gL += gm[:, None, None] * P["rho"][None]
    gh1s = ghmod + T @ acc["rot_h1"]
    sym = 0.5 * (gh1s + gh1s.T)
    gh1 = np.array([0.5 * sym, 0.5 * sym])
    return F, gh1, gL.reshape(gL.shape[0], -1)

"""The reference's CPU-runnable case on the device: linear H10, 1.6 bohr,
STO-6G (own integrals), Cholesky threshold 1e-5, RHF trial, 50 walkers,
dt = 0.005, FP64.  The device driver's mixed energy must agree with the CPU
oracle's (its own field stream) within the combined blocking error bars, and
lie below the RHF energy."""
import math

import numpy as np
import pytest

from ad_afqmc_amd import driver, sampling
from ad_afqmc_amd import hamiltonian as hm
from ad_afqmc_amd import integrals as I
from ad_afqmc_amd import propagation as pr
from ad_afqmc_amd import wavefunctions as wf
from oracle import afqmc_oracle as O

pytestmark = pytest.mark.gpu

N_BLOCKS, N_STEPS, W, DT = 60, 10, 50, 0.005


def _oracle_energy(h, seed=0):
    mo = h.trial_mo()
    trial = O.RHFTrial(10, (5, 5), mo)
    h1s, rot_h1, rot_chol = O.rhf_measurement_intermediates(h.ham_data()["h1"], h.chol, mo)
    rdm1 = O.rdm1_from_mo(mo)
    pi = O.propagation_intermediates(h.h0, h1s, h.chol, rdm1, DT)
    ham = dict(h0=h.h0, chol=h.chol, rot_h1=rot_h1, rot_chol=rot_chol, mf_shifts=pi["mf_shifts"],
               h0_prop=pi["h0_prop"], exp_h1=pi["exp_h1"])
    rng = np.random.default_rng(seed)
    walkers = O.init_walkers(rdm1, (5, 5), W, "restricted")
    e0 = float(np.mean(np.real(trial.calc_energy(walkers, h.h0, rot_h1, rot_chol))))
    state = dict(walkers=walkers, weights=np.ones(W), e_estimate=e0)
    G = h.chol.shape[0]
    es, ws = [], []
    for b in range(5 + N_BLOCKS):  # 50 equilibration steps, then the sampling blocks
        e, state = O.propagate_phaseless(trial, state, ham, DT, N_STEPS, 1, 1,
                                         lambda: rng.standard_normal((N_STEPS, W, G)), lambda: rng.random())
        state["e_estimate"] = 0.9 * state["e_estimate"] + 0.1 * float(np.float32(e))
        if b >= 5:
            es.append(e)
            ws.append(np.sum(state["weights"]))
    return O.blocking_analysis(np.array(ws), np.array(es))


def test_h10_sto6g_device_driver_against_oracle():
    print("[h10] integrals + SCF + Cholesky", flush=True)
    h = I.hydrogen_chain(10, 1.6, "sto-6g", 1e-5)
    trial = wf.rhf(10, (5, 5))
    prop = pr.propagator_afqmc(dt=DT, n_walkers=W, walker_type="restricted")
    smp = sampling.sampler(n_prop_steps=N_STEPS, n_ene_blocks=1, n_sr_blocks=1, n_blocks=N_BLOCKS)
    print("[h10] device driver", flush=True)
    e, err = driver.afqmc_energy(h.ham_data(), hm.hamiltonian(10), prop, trial, h.wave_data(), smp,
                                 dict(n_eql=1, n_ene_blocks_eql=1, n_sr_blocks_eql=1, seed=0), verbose=False)
    print("[h10] oracle", flush=True)
    eo, erro = _oracle_energy(h)
    print(f"[h10] device {e:.6f} +/- {err:.6f}  oracle {eo:.6f} +/- {erro:.6f}  RHF {h.e_rhf:.6f}", flush=True)
    assert math.isfinite(e) and err > 0
    assert abs(e - eo) < 4 * math.hypot(err, erro) + 2e-3
    assert e < h.e_rhf

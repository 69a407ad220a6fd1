"""Maps the reference's dict data contract onto a resident engine job.

The reference passes plain dicts (`ham_data`, `wave_data`, `prop_data`) to its
trial and propagator methods (ad_afqmc/propagation.py:274-282,
ad_afqmc/wavefunctions.py:93-395) and reads whatever arrays they hold.  Here a
job is attached to `wave_data` (every trial method receives it) and kept in
sync with the arrays the dicts currently hold: an array is (re-)uploaded when
the object stored under its key is not the one uploaded last.  Intermediates
the job builds itself are stored back into `ham_data` as device tensors, so a
later call finds them identical and uploads nothing.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _native as nat
from .engine import EngineJob

_KEY = "_afq_binding"

# ham_data key -> (intermediate id, per-spin?)
_INTERMEDIATES = {
    "rot_h1": nat.AFQ_INT_ROT_H1,
    "rot_chol": nat.AFQ_INT_ROT_CHOL,
    "mf_shifts": nat.AFQ_INT_MF_SHIFTS,
    "h0_prop": nat.AFQ_INT_H0_PROP,
    "exp_h1": nat.AFQ_INT_EXP_H1,
}


class Binding:
    def __init__(self):
        self.job: Optional[EngineJob] = None
        self.key = None
        self.sources = {}

    def same(self, name, obj) -> bool:
        old = self.sources.get(name, None)
        if old is None:
            return False
        if isinstance(obj, (float, int, np.floating, np.integer)):
            return isinstance(old, (float, int, np.floating, np.integer)) and float(old) == float(obj)
        if isinstance(obj, (list, tuple)) and isinstance(old, (list, tuple)):
            return len(obj) == len(old) and all(a is b for a, b in zip(obj, old))
        return old is obj


def _nchol(ham_data: dict) -> Optional[int]:
    if ham_data is None:
        return None
    if "chol" in ham_data:
        return int(ham_data["chol"].shape[0])
    rc = ham_data.get("rot_chol", None)
    if rc is None:
        return None
    if isinstance(rc, (list, tuple)):
        rc = rc[0]
    return int(rc.shape[0])


def _trial_rdm1(trial, wave_data):
    if "rdm1" in wave_data:
        return wave_data["rdm1"]
    return trial._calc_rdm1(wave_data)


def _spin_list(x, rhf: bool):
    if rhf:
        return [x]
    if isinstance(x, (list, tuple)):
        return list(x)
    x = torch.as_tensor(x) if not isinstance(x, torch.Tensor) else x
    return [x[0], x[1]]


def engine_for(trial, wave_data: dict, ham_data: Optional[dict], walker_type: str, n_walkers: int, prop=None) -> EngineJob:
    """Returns the job for (trial, walker type, sizes, propagator settings),
    creating or refreshing it so that it reflects the current dict contents."""
    b: Binding = wave_data.get(_KEY, None)
    if b is None:
        b = Binding()
        wave_data[_KEY] = b
    nchol = _nchol(ham_data)
    old = b.key
    if nchol is None:
        nchol = old[2] if old is not None else 1
    if prop is not None:
        dt, nexp, prec = float(prop.dt), int(prop.n_exp_terms), getattr(prop, "precision", "fp64")
    elif old is not None:
        dt, nexp, prec = old[4], old[5], old[6]
    else:
        dt, nexp, prec = 0.01, 6, "fp64"
    trial_type = "rhf" if trial.kind == "rhf" else "uhf"
    key = (trial.norb, tuple(trial.nelec), nchol, walker_type, dt, nexp, prec, trial_type)
    cap = n_walkers
    if b.job is not None and b.key == key and b.job.capacity >= n_walkers:
        job = b.job
    else:
        if b.job is not None and b.key == key:
            cap = max(cap, b.job.capacity)
        if b.job is not None:
            b.job.close()
        job = EngineJob(trial.norb, trial.nelec, nchol, cap, walker_type, trial_type, nexp, dt, prec)
        b.job, b.key, b.sources = job, key, {}
    rhf = trial_type == "rhf"

    # trial
    mo = wave_data["mo_coeff"]
    rdm1 = _trial_rdm1(trial, wave_data)
    if not (b.same("mo_coeff", mo) and b.same("rdm1", rdm1)):
        mos = _spin_list(mo, rhf)
        job.set_trial(mos[0], None if rhf else mos[1], rdm1)
        b.sources["mo_coeff"] = mo if not isinstance(mo, list) else tuple(mo)
        b.sources["rdm1"] = rdm1

    if ham_data is not None:
        h0 = ham_data.get("h0", 0.0)
        h0v = float(np.asarray(h0.cpu() if isinstance(h0, torch.Tensor) else h0).reshape(-1)[0])
        if "h1" in ham_data and "chol" in ham_data:
            if not (b.same("h1", ham_data["h1"]) and b.same("chol", ham_data["chol"]) and b.same("h0", h0v)):
                job.set_hamiltonian(h0v, ham_data["h1"], ham_data["chol"])
                b.sources.update(h1=ham_data["h1"], chol=ham_data["chol"], h0=h0v)
        elif not b.same("h0", h0v):
            job.set_hamiltonian(h0v)
            b.sources["h0"] = h0v
        for name, which in _INTERMEDIATES.items():
            if name not in ham_data:
                continue
            val = ham_data[name]
            if b.same(name, val):
                continue
            if name in ("rot_h1", "rot_chol"):
                for s, v in enumerate(_spin_list(val, rhf)):
                    job.set(which, s, v)
            elif name == "exp_h1":
                v = val if isinstance(val, torch.Tensor) else torch.as_tensor(np.asarray(val))
                if v.dim() == 3:
                    for s in range(v.shape[0] if walker_type == "unrestricted" else 1):
                        job.set(which, s, v[s])
                else:
                    job.set(which, 0, v)
            else:
                job.set(which, 0, val)
            b.sources[name] = val if not isinstance(val, list) else tuple(val)
    return job


def record(wave_data: dict, name: str, value):
    """Notes that `value` (now stored in a dict by the caller) is what the job holds."""
    b: Binding = wave_data.get(_KEY, None)
    if b is not None:
        b.sources[name] = value if not isinstance(value, list) else tuple(value)

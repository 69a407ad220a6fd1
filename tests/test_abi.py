"""libafq.so loads and exports exactly what include/afq.h declares; argument
validation that happens before any device work."""
import ctypes as C
import os
import re

import pytest

from ad_afqmc_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "afq.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const char\*|int64_t|int)\s+(afq_\w+)\s*\(", src, re.M)))


def test_header_symbols_exported():
    declared = _declared()
    assert len(declared) >= 20
    L = C.CDLL(nat.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), f"{name} declared in afq.h but not exported"
    assert sorted(nat.exported_symbols()) == declared


def test_version_and_validation():
    L = nat.lib()
    assert L.afq_version().startswith(b"afq")
    h = C.c_void_p()
    bad = nat.JobDesc(norb=10, nelec_up=5, nelec_dn=4, nchol=3, n_walkers=4, walker_type=0, trial_type=nat.AFQ_TRIAL_RHF,
                      n_exp_terms=6, precision=0, reserved=0, dt=0.01)
    rc = L.afq_job_create(C.byref(h), C.byref(bad))
    assert rc == -1 and b"RHF requires" in L.afq_last_error()
    bad.trial_type = nat.AFQ_TRIAL_UHF
    bad.nelec_dn = 6  # restricted walkers need nelec_dn <= nelec_up
    assert L.afq_job_create(C.byref(h), C.byref(bad)) == -1
    bad.nelec_dn = 4
    bad.dt = -1.0
    assert L.afq_job_create(C.byref(h), C.byref(bad)) == -1
    bad.dt = 0.01
    bad.nelec_up = 300
    bad.norb = 320
    assert L.afq_job_create(C.byref(h), C.byref(bad)) == -5  # nocc > 256 unsupported
    assert L.afq_walkers_sr_workspace(0, 10) == 0
    assert L.afq_walkers_sr_workspace(100, 10) >= 100 * 10 * 16


def test_null_arguments_rejected():
    L = nat.lib()
    assert L.afq_calc_overlap(None, None, None, None, None) == -1
    assert L.afq_job_destroy(None) == 0
    assert L.afq_walkers_orthonormalize(None, 1, 1, 1, None) == -1

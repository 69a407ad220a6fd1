"""afq_walkers_sr_noalloc (workspace in the library's static device array, no
allocation) against afq_walkers_sr (caller workspace), through the C ABI: the
resampled walkers, weights and indices must be bitwise equal, on two calls in
a row, the second with a smaller input."""
import ctypes as C

import numpy as np
import pytest
import torch

from ad_afqmc_amd import _native as nat

pytestmark = pytest.mark.gpu


def _ptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _case(rng, W, eu, ed):
    cplx = lambda n: torch.as_tensor(rng.standard_normal((W, n)) + 1j * rng.standard_normal((W, n)), device="cuda")
    up = cplx(eu)
    dn = cplx(ed) if ed else None
    w = rng.random(W) * 2.0
    w[::7] = 0.0
    return up, dn, torch.as_tensor(w, device="cuda"), torch.tensor([float(rng.random())], dtype=torch.float64,
                                                                    device="cuda")


def _with_workspace(L, up, dn, w, zeta):
    W, eu = up.shape
    ed = 0 if dn is None else dn.shape[1]
    ws = torch.empty(int(L.afq_walkers_sr_workspace(W, max(eu, ed))), dtype=torch.uint8, device="cuda")
    idx = torch.empty(W, dtype=torch.int32, device="cuda")
    rc = L.afq_walkers_sr(_ptr(up), eu, _ptr(dn), ed, W, _ptr(w), _ptr(zeta), _ptr(idx), _ptr(ws), _stream())
    nat.check("afq_walkers_sr", rc)
    return idx


def _no_alloc(L, up, dn, w, zeta):
    W, eu = up.shape
    ed = 0 if dn is None else dn.shape[1]
    idx = torch.empty(W, dtype=torch.int32, device="cuda")
    rc = L.afq_walkers_sr_noalloc(_ptr(up), eu, _ptr(dn), ed, W, _ptr(w), _ptr(zeta), _ptr(idx), _stream())
    nat.check("afq_walkers_sr_noalloc", rc)
    return idx


def test_noalloc_sr_matches_workspace_variant():
    L = nat.lib()
    rng = np.random.default_rng(2024)
    # first call: 300 walkers of 8000 + 7600 elements, several gather chunks of the
    # 16 MiB array; second call: fewer and smaller walkers, one chunk
    for W, eu, ed in ((300, 200 * 40, 200 * 38), (37, 24, 0)):
        up, dn, w, zeta = _case(rng, W, eu, ed)
        ref = [t.clone() if t is not None else None for t in (up, dn, w)]
        idx_ref = _with_workspace(L, ref[0], ref[1], ref[2], zeta)
        idx = _no_alloc(L, up, dn, w, zeta)
        torch.cuda.synchronize()
        assert torch.equal(idx, idx_ref), (W, eu, ed)
        assert torch.equal(up, ref[0]) and (dn is None or torch.equal(dn, ref[1]))
        assert torch.equal(w, ref[2])
        # the comb actually moved walkers
        assert not torch.equal(idx.cpu(), torch.arange(W, dtype=torch.int32))


def test_noalloc_sr_refuses_more_walkers_than_its_array_holds():
    L = nat.lib()
    cap = int(L.afq_walkers_sr_noalloc_max_walkers())
    assert cap == 65535
    one = torch.zeros(4, dtype=torch.complex128, device="cuda")
    w = torch.ones(4, dtype=torch.float64, device="cuda")
    zeta = torch.zeros(1, dtype=torch.float64, device="cuda")
    # rejected on the host before any device work
    rc = L.afq_walkers_sr_noalloc(_ptr(one), 1, None, 0, cap + 1, _ptr(w), _ptr(zeta), None, _stream())
    assert rc == -1 and b"static scratch limit" in L.afq_last_error()

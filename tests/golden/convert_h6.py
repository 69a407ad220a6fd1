"""Converts the reference's H6 driver-test integrals to tests/golden/h6_fcidump_chol.npz.

Input (data files of the reference test suite, read-only):
  tests/driver_test_files/FCIDUMP_chol  (HDF5: header [nelec, norb, ms, nchol],
                                          energy_core, hcore (N*N), chol (G*N*N))
  tests/driver_test_files/mo_coeff.npz  (mo_coeff (2, N, N))
Layout as read by the reference (ad_afqmc/launch_script.py:60-94).

Needs h5py (run with an interpreter that has it), e.g.
  /opt/conda/bin/python3.9 tests/golden/convert_h6.py /root/reference/tests/driver_test_files
"""
import os
import sys

import h5py
import numpy as np

src = sys.argv[1]
with h5py.File(os.path.join(src, "FCIDUMP_chol"), "r") as f:
    header = np.array(f["header"])
    energy_core = float(np.array(f.get("energy_core")))
    hcore = np.array(f.get("hcore"), dtype=np.float64)
    chol = np.array(f.get("chol"), dtype=np.float64)
mo = np.load(os.path.join(src, "mo_coeff.npz"))["mo_coeff"]
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "h6_fcidump_chol.npz")
np.savez(out, header=header.astype(np.int64), energy_core=energy_core, hcore=hcore, chol=chol, mo_coeff=mo)
print("wrote", out, header, chol.shape, mo.shape)

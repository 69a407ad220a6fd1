"""Host-side statistics of the block energies (ad_afqmc/stat_utils.py:8-78)."""
from __future__ import annotations

import numpy as np


def blocking_analysis(weights, energies, neql: int = 0, printQ: bool = False):
    """Weighted blocking analysis with plateau detection (stat_utils.py:8-64).
    Returns (mean, error); error is 2x the unblocked error without a plateau."""
    weights = np.asarray(weights, dtype=np.float64)[neql:]
    energies = np.asarray(energies, dtype=np.float64)[neql:]
    n = weights.shape[0]
    we = weights * energies
    mean = we.sum() / weights.sum()
    sizes = np.array([1, 2, 5, 10, 20, 50, 100, 200, 300, 400, 500, 1000, 10000])
    prev, plateau, errors = 0.0, None, []
    for i in sizes[sizes < n / 2.0]:
        nb = int(n // i)
        bw = weights[: nb * i].reshape(nb, i).sum(axis=1)
        be = we[: nb * i].reshape(nb, i).sum(axis=1) / bw
        v1, v2 = bw.sum(), (bw**2).sum()
        m = (bw * be).sum() / v1
        err = ((bw * (be - m) ** 2).sum() / (v1 - v2 / v1) / (nb - 1)) ** 0.5
        errors.append(err)
        if printQ:
            print(f"  {i:5d}           {nb:6d}       {m:.8e}       {err:.6e}")
        if err < 1.05 * prev and plateau is None:
            plateau = max(err, prev)
        prev = err
    if plateau is None and errors:
        plateau = 2 * errors[0]
    return mean, plateau


def reject_outliers(data, obs: int, m: float = 10.0, min_threshold: float = 1e-5):
    """Median / IQR outlier rejection (stat_utils.py:67-78)."""
    data = np.asarray(data)
    t = data[:, obs]
    med = np.median(t)
    d = np.abs(t - med)
    mdev = np.median(d)
    q1, q3 = np.percentile(t, [25, 75])
    disp = max(mdev, (q3 - q1) / 1.349, min_threshold)
    mask = d / disp < m
    return data[mask], mask

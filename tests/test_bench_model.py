"""Host-side pieces of bench.py that feed the roofline object: the packed
column count of the symmetric VHS product, the symmetry test that picks it,
and the per-launch flop model (CPU only)."""
import numpy as np

import bench


def _lower_block_pairs(N):
    """Brute force: entries (i, j) of the zero-padded (16 nb)^2 matrix whose
    16 x 16 block lies on or below the block diagonal."""
    nb = (N + 15) // 16
    idx = np.arange(16 * nb) // 16
    return int(np.count_nonzero(idx[:, None] >= idx[None, :]))


def test_packed_columns_count_the_lower_block_triangle():
    for N in (1, 16, 17, 70, 96, 200, 300):
        assert bench.sym_packed_cols(N) == _lower_block_pairs(N), N
    # the headline and N=300 workloads: 91 of 169 and 190 of 361 blocks
    assert bench.sym_packed_cols(200) == 91 * 256
    assert bench.sym_packed_cols(300) == 190 * 256


def test_vhs_columns_follow_the_symmetry_of_L():
    rng = np.random.default_rng(0)
    N, G = 64, 5
    a = rng.standard_normal((G, N, N))
    sym = (a + a.transpose(0, 2, 1)).reshape(G, N * N)
    assert bench.vhs_columns(sym, N) == bench.sym_packed_cols(N)
    # an antisymmetric part above 1e-12 of max |L| selects the full product
    asym = sym + 1e-9 * (a - a.transpose(0, 2, 1)).reshape(G, N * N)
    assert bench.vhs_columns(asym, N) == N * N
    # at small N the padded blocks would cost more than the N^2 columns
    small = sym.reshape(G, N, N)[:, :24, :24].reshape(G, 24 * 24)
    assert bench.vhs_columns(small, 24) == 24 * 24


def test_launch_flops_follow_panels_and_spins():
    N, G, W = 300, 2000, 256
    # restricted walkers: one spin buffer of Oa columns
    assert bench.launch_flops("taylor", N, 40, 40, G, W) == 8.0 * N * N * 40 * W
    assert bench.launch_flops("exchange", N, 40, 40, G, W) == 4.0 * G * 40 * 40 * N * W
    # unrestricted walkers: one Taylor launch covers both spins' columns, the
    # exchange launches alternate between the spins
    assert bench.launch_flops("taylor", N, 60, 58, G, W, unrestricted=True) == 8.0 * N * N * 118 * W
    ex = bench.launch_flops("exchange", N, 60, 58, G, W, unrestricted=True)
    assert ex == 0.5 * (4.0 * G * 60 * 60 * N * W + 4.0 * G * 58 * 58 * N * W)
    assert bench.launch_flops("exp_h1", N, 60, 58, G, W, unrestricted=True) == 4.0 * N * N * 59 * W
    # the VHS product over packed columns
    npk = bench.sym_packed_cols(N)
    assert bench.launch_flops("vhs", N, 60, 58, G, W, vhs_cols=npk) == 4.0 * G * npk * W
    assert bench.launch_flops("vhs", N, 60, 58, G, W) == 4.0 * G * N * N * W
    assert bench.launch_flops("sr", N, 60, 58, G, W) is None

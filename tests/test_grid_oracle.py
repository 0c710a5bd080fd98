"""Pins of the gridder oracle (oracle/orc_grid.c: vcg, 3_vcg/tgefuncs.c
:290-376, 377-626, 816-839) by facts that do not retype its loop:
  * the shipped inputparameters (Umax 602, FWHM 1480.73) give the grid the
    reference's own files record: du_phy = 1.0254384902642049
    (mabs_funcs.py:30) and the archived side 1176 (README.md:137,
    mfunc_winds.py:286);
  * hand-computed cells, the v < 0 flip equal to the conjugate group,
    constant visibilities summing to count x value, flags, the Umax
    boundary, channel offsets and collapse sums.
"""
import numpy as np
import pytest

import oracle as O


def test_setup_matches_published_grid():
    dU, NN = O.grid_setup(602.0, 1480.73)
    assert dU == 1.0254384902642049
    assert NN == 1176


def one(u, v, vis, nc=3, nstokes=2, nchan=None):
    nchan = nchan or nc
    d = np.zeros((1, nchan, nstokes, 3), np.float32)
    d[0, :, 0, :] = vis
    return np.array([[u, v, 0.0]], np.float32), d


def cell(NN, ii, jj):
    return ii * (NN // 2 + 1) + jj


def test_hand_cells_and_flip():
    NN, dU = 16, 2.0
    vis = np.array([[1.5, 0.25, 1.0]] * 3, np.float32)
    rp, d = one(3.1, 4.9, vis)                   # round(1.55) = 2, round(2.45) = 2
    gv, s, n = O.grid(NN, 3, rp, d, 1.0, dU, 100.0)
    assert n == 1 and s[cell(NN, 2, 2)].tolist() == [1, 1, 1]
    assert gv[cell(NN, 2, 2), 0].tolist() == [1.5, 0.25] and np.count_nonzero(s) == 3
    rp, d = one(-3.1, 4.9, vis)                  # round(-1.55) = -2 -> row NN - 2
    gv, s, n = O.grid(NN, 3, rp, d, 1.0, dU, 100.0)
    assert s[cell(NN, 14, 2)].tolist() == [1, 1, 1]
    # v < 0: (u, v, V) lands where (-u, -v, conj V) does, with the same values
    rp, d = one(3.1, -4.9, vis)
    gv1, s1, _ = O.grid(NN, 3, rp, d, 1.0, dU, 100.0)
    vc = vis.copy(); vc[:, 1] *= -1
    rp, d = one(-3.1, 4.9, vc)
    gv2, s2, _ = O.grid(NN, 3, rp, d, 1.0, dU, 100.0)
    assert np.array_equal(gv1, gv2) and np.array_equal(s1, s2)
    assert gv1[cell(NN, 14, 2), 0].tolist() == [1.5, -0.25]


def test_constant_sums_flags_and_boundary():
    NN, dU, K = 16, 1.0, 7
    rp = np.array([[3.0, 4.0, 0.0]] * K, np.float32)  # |U| = 5 exactly
    d = np.zeros((K, 4, 1, 3), np.float32)
    d[..., 0] = 0.5
    d[..., 1] = -2.0
    d[..., 2] = 1.0
    d[2, 1, 0, 2] = 0.0                               # flagged visibility of channel 1
    d[5, 3, 0, 2] = -1.0
    gv, s, n = O.grid(NN, 4, rp, d, 1.0, dU, 5.0, flag_key=1)
    c = cell(NN, 3, 4)
    assert n == K and s[c].tolist() == [K, K - 1, K, K - 1]
    assert gv[c, :, 0].tolist() == [0.5 * K, 0.5 * (K - 1), 0.5 * K, 0.5 * (K - 1)]
    assert gv[c, :, 1].tolist() == [-2.0 * K, -2.0 * (K - 1), -2.0 * K, -2.0 * (K - 1)]
    gv, s, n = O.grid(NN, 4, rp, d, 1.0, dU, 5.0, flag_key=0)   # weights ignored
    assert s[c].tolist() == [K] * 4 and gv[c, :, 0].tolist() == [0.5 * K] * 4
    gv, s, n = O.grid(NN, 4, rp, d, 1.0, dU, 4.999)              # |U| > Umax: skipped
    assert n == 0 and not s.any() and not gv.any()


def test_channel_offset_and_collapse():
    NN, dU = 8, 1.0
    rng = np.random.default_rng(4)
    d = rng.standard_normal((3, 12, 2, 3)).astype(np.float32)
    d[..., 2] = 1.0
    rp = np.array([[1.0, 1.0, 0], [1.2, 0.9, 0], [-2.0, -1.0, 0]], np.float32)
    gv, s, _ = O.grid(NN, 10, rp, d, 1.0, dU, 10.0, chan1=3)
    gv1, s1, _ = O.grid(NN, 10, rp, np.ascontiguousarray(d[:, 2:12]), 1.0, dU, 10.0)
    assert np.array_equal(gv, gv1) and np.array_equal(s, s1)
    gc, sc, _ = O.grid(NN, 10, rp, d, 1.0, dU, 10.0, n_collapse=4, chan1=3)
    assert gc.shape[1] == 2                            # 10 // 4 channels, the remainder dropped
    for col in range(2):
        want = ((0.0 + gv[:, 4 * col]) + gv[:, 4 * col + 1] + gv[:, 4 * col + 2]) + gv[:, 4 * col + 3]
        assert np.array_equal(gc[:, col], want)
        assert np.array_equal(sc[:, col], s[:, 4 * col:4 * col + 4].sum(1))

"""Pins of the MAPS oracle (oracle/orc_maps.c: maps.c:100-229 with cl_dnu,
amber_funcs.c:1418-1488) by facts that do not retype its loops:
  * bins from the bin edges 10^(k binsiz) umin by comparison (no logarithm),
    the row fold ia = N - i and the half-open [umin, umax) range;
  * a cube V(x, n) = S(x, n) c(x) e^{i phi n} normalises to c(x) e^{i phi n}
    where S > 0, so with every cell sampled cldnu[b][d] is
    (nc - d) cos(phi d) <|c|^2>_b; with channel masks it is
    cos(phi d) sum_n1 sum_{x in b} [S1>0][S2>0] |c|^2 / no_b(n1), where the
    division uses the lower channel's count;
  * numdnu[d] = nc - d, and binu/num of the last pair by enumeration.
"""
import math

import numpy as np
import pytest

import oracle

NB, UMIN, UMAX = 15, 1.0, 500.0


def bin_by_edges(uu, NB=NB, umin=UMIN, umax=UMAX):
    if not (umin <= uu < umax):
        return -1
    edges = [umin * (umax / umin) ** (k / NB) for k in range(NB + 1)]
    for k in range(NB):
        if edges[k] <= uu < edges[k + 1]:
            return k
    return NB - 1


def test_bin_hand_cases():
    # 3-4-5 cell: |U| = 5 lies in [500^(3/15), 500^(4/15)) = [3.47, 5.25)
    assert oracle.maps_bin(64, 1.0, 3, 4) == (3, 5.0)
    assert oracle.maps_bin(64, 1.0, 61, 4) == (3, 5.0)       # row 61 folds to ia = 3
    assert oracle.maps_bin(64, 1.0, 0, 0)[0] == -1           # DC: |U| = 0 < umin
    assert oracle.maps_bin(64, 1.0, 0, 1) == (0, 1.0)        # |U| = umin is inside
    assert oracle.maps_bin(64, 100.0, 0, 5)[0] == -1         # |U| = umax is outside
    assert oracle.maps_bin(64, 99.9, 3, 4)[0] == NB - 1
    # row N/2 keeps ia = N/2; |U| = 32 lies in [500^(8/15), 500^(9/15)) = [27.5, 41.6)
    assert oracle.maps_bin(64, 1.0, 32, 0) == (8, 32.0)


@pytest.mark.parametrize("N,Du", [(32, 1.0), (48, 3.7), (64, 9.5)])
def test_bins_match_edges(N, Du):
    for i in range(N):
        for j in range(N // 2 + 1):
            b, uu = oracle.maps_bin(N, Du, i, j)
            ia = N - i if i > N // 2 else i
            assert uu == math.sqrt(ia * ia + j * j) * Du
            near = any(abs(uu - UMIN * (UMAX / UMIN) ** (k / NB)) < 1e-9 * uu for k in range(NB + 1))
            if not near:
                assert b == bin_by_edges(uu), (i, j, uu)


def phase_cube(N, nc, rng, phi, samp):
    Nh = N // 2 + 1
    c = rng.normal(size=(N * Nh,)) + 1j * rng.normal(size=(N * Nh,))
    v = (samp * c[:, None]) * np.exp(1j * phi * np.arange(nc))[None, :]
    gv = np.stack([v.real, v.imag], axis=-1)
    return gv, c


def cells_by_bin(N, Du):
    Nh = N // 2 + 1
    out = {}
    for i in range(N):
        for j in range(Nh):
            b, uu = oracle.maps_bin(N, Du, i, j)
            out[i * Nh + j] = (b, uu)
    return out


def test_cldnu_closed_form_all_sampled():
    N, nc, Du, phi = 32, 12, 9.0, 0.37
    rng = np.random.default_rng(1)
    Nh = N // 2 + 1
    samp = rng.integers(1, 4, size=(N * Nh, nc)).astype(np.float64)   # counts >= 1: all sampled
    gv, c = phase_cube(N, nc, rng, phi, samp)
    cl, nd, bu, nm = oracle.cldnu(N, nc, Du, gv, samp)
    cells = cells_by_bin(N, Du)
    a2 = np.abs(c) ** 2
    for b in range(NB):
        xs = [x for x, (bb, _) in cells.items() if bb == b]
        if not xs:
            assert np.all(cl[b] == 0) and nm[b] == 0
            continue
        mean = a2[xs].sum() / len(xs)
        for d in range(nc):
            want = (nc - d) * math.cos(phi * d) * mean
            assert abs(cl[b, d] - want) <= 1e-12 * (nc - d) * mean, (b, d)
        assert nm[b] == len(xs)
        assert abs(bu[b] - sum(cells[x][1] for x in xs) / len(xs)) <= 1e-13 * bu[b]
    assert np.array_equal(nd, nc - np.arange(nc, dtype=np.float64))


def test_cldnu_closed_form_masked():
    N, nc, Du, phi = 32, 10, 9.0, -0.81
    rng = np.random.default_rng(2)
    Nh = N // 2 + 1
    samp = rng.integers(0, 3, size=(N * Nh, nc)).astype(np.float64)   # 0 = unsampled
    gv, c = phase_cube(N, nc, rng, phi, samp)
    cl, nd, bu, nm = oracle.cldnu(N, nc, Du, gv, samp)
    cells = cells_by_bin(N, Du)
    a2 = np.abs(c) ** 2
    on = samp > 0
    for b in range(NB):
        xs = np.array([x for x, (bb, _) in cells.items() if bb == b], int)
        for d in range(nc):
            want, scale = 0.0, 0.0
            for n1 in range(nc - d):
                no = int(on[xs, n1].sum()) if len(xs) else 0
                s = float((on[xs, n1] & on[xs, n1 + d]) @ a2[xs]) if len(xs) else 0.0
                if no:
                    want += s / no
                    scale += s / no
            want *= math.cos(phi * d)
            assert abs(cl[b, d] - want) <= 1e-12 * max(scale, 1e-300), (b, d)
        last = xs[on[xs, nc - 1]] if len(xs) else xs
        assert nm[b] == len(last)
        if len(last):
            assert abs(bu[b] - sum(cells[x][1] for x in last) / len(last)) <= 1e-13 * bu[b]

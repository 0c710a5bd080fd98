"""Pins amb_compute_grid, the grid the lag-Fourier form sums a triangle on.

A ring image is an unnormalised c2r of the ring's cells, so the same
frequencies sampled on any grid n that holds them give the image of that grid.
The triple sum over an n x n grid is n^2 times the sum of c1 c2 c3 over
frequency triples with k1 + k2 + k3 = 0 (mod n); when no such sum of ring
frequencies reaches a nonzero multiple of n, that is the closed-triangle sum,
the same for every such grid.  These tests check, with the oracle's images on
the reference grid Ng (amber_funcs.c:573-595) moved to n by their spectra,
that the reference's sum over Ng^2 pixels equals (Ng/n)^2 times the sum over
n^2 pixels for every triangle of a ring set, that a grid one step below the
rule breaks that for some triangle, and the rule's own properties.  CPU only."""
import numpy as np
import pytest

import oracle as O
import amber_amd as A


def smooth7(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def reach(uh):
    """largest |component| of an integer cell with sqrt(ia^2 + j^2) < uh"""
    m = int(np.ceil(uh)) + 1
    while m > 0 and np.sqrt(m * m) >= uh:
        m -= 1
    return m


def resample(img, n, m):
    """the image of the same frequencies on an n x n grid (|k| <= m)"""
    Ng = img.shape[0]
    H = np.fft.rfft2(img) / (Ng * Ng)
    Hn = np.zeros((n, n // 2 + 1), complex)
    for ia in range(-m, m + 1):
        Hn[ia % n, :m + 1] = H[ia % Ng, :m + 1]
    return np.fft.irfft2(Hn, s=(n, n)) * (n * n)


def triple(a, b, c):
    t = a * b * c
    return t.sum(), np.abs(t).sum()


@pytest.mark.parametrize("uh", [0.5, 1.0, 1.5, 2.0, 7.3, 10.0, 49.99, 50.0, 50.01, 200.0, 270.0])
def test_ring_reach(uh):
    assert A.ring_reach(uh) == reach(uh)


def test_rule_properties():
    for lo, hi, du in [(*A.log_rings(1, 200, 24), 1.0), (*A.log_rings(2, 60, 8), 1.0),
                       (*O.annular_rings(22), 1.0254384902642049)]:
        for r1, r2, r3 in A.triangles(lo, hi):
            ng = A.ng(hi[r1], du)
            n = A.compute_grid(hi[r1], hi[r2], hi[r3], du)
            m = [reach(hi[r]) for r in (r1, r2, r3)]
            assert 0 < n <= ng and n % 2 == 0
            if n < ng:
                assert smooth7(n) and n > sum(m) and n >= 2 * max(m) + 2
                # the reference grid itself is then alias-free too
                assert sum(m) < ng
            else:
                # nothing smaller qualifies below Ng
                need = max(sum(m) + 1, 2 * max(m) + 2)
                assert all(not (smooth7(k) and k >= need) for k in range(2, ng, 2))


def test_metric_rings_work_fraction():
    # the metric workload's 506 triangles: pixels summed on compute grids
    # against the reference grids (each triangle weighted alike)
    lo, hi = A.log_rings(1, 200, 24)
    t = A.triangles(lo, hi)
    ref = sum(A.ng(hi[a]) ** 2 for a, b, c in t)
    red = sum(A.compute_grid(hi[a], hi[b], hi[c]) ** 2 for a, b, c in t)
    assert len(t) == 506 and 0.5 < red / ref < 0.6


@pytest.mark.parametrize("du", [1.0, 1.0254384902642049])
def test_sum_on_compute_grid_equals_reference_grid(du):
    N, nc = 128, 3
    lo, hi = A.log_rings(2, 40, 8)
    gv, samp, uf = A.generate(N, nc, nflag=0, seed=77)
    tris = A.triangles(lo, hi)
    reduced = 0
    imgs = {}

    def img(r, Ng, p):
        key = (r, Ng, p)
        if key not in imgs:
            imgs[key] = O.ring_c2r(N, nc, Ng, lo[r], hi[r], p, gv=gv)
        return imgs[key]

    for r1, r2, r3 in tris:
        Ng = A.ng(hi[r1], du)
        n = A.compute_grid(hi[r1], hi[r2], hi[r3], du)
        m = [reach(hi[r]) for r in (r1, r2, r3)]
        for p in range(nc):
            a, b, c = img(r1, Ng, p), img(r2, Ng, (p + 1) % nc), img(r3, Ng, (p + 2) % nc)
            s_ref, mag = triple(a, b, c)
            s_n, _ = triple(resample(a, n, m[0]), resample(b, n, m[0]), resample(c, n, m[0]))
            assert abs(s_ref - (Ng / n) ** 2 * s_n) <= 1e-11 * mag, (r1, r2, r3, p, Ng, n)
        reduced += n < Ng
    assert reduced >= len(tris) // 2


def test_grid_below_the_rule_aliases():
    # one even grid below m1 + m2 + m3 + 1 (where that bound is the binding
    # one) wraps sums of ring frequencies onto zero: the proportionality
    # fails for some triangle, so the bound is not loose by a term
    N, nc = 128, 1
    lo, hi = A.log_rings(2, 40, 8)
    gv, samp, uf = A.generate(N, nc, nflag=0, seed=78)
    broke = 0
    tried = 0
    for r1, r2, r3 in A.triangles(lo, hi):
        m = [reach(hi[r]) for r in (r1, r2, r3)]
        nb = sum(m) - (sum(m) % 2)  # even, <= m1 + m2 + m3
        if nb < 2 * max(m) + 2:
            continue
        tried += 1
        Ng = A.ng(hi[r1])
        a, b, c = (O.ring_c2r(N, nc, Ng, lo[r], hi[r], 0, gv=gv) for r in (r1, r2, r3))
        s_ref, mag = triple(a, b, c)
        s_n, _ = triple(resample(a, nb, m[0]), resample(b, nb, m[0]), resample(c, nb, m[0]))
        broke += abs(s_ref - (Ng / nb) ** 2 * s_n) > 1e-6 * mag
    assert tried >= 3 and broke >= 1

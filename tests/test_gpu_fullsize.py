"""GPU at the metric workload's own size (768 channels x 512^2 grid, all
767 lags, rings up to 200 cells, Ng = 612) through properties that hold at
any size, since the FP64 oracle cannot run whole triangles here:
  - counts equal the closed form nc - (max(0,d2,d3) - min(0,d2,d3));
  - doubling the cube scales B by exactly 8: a power-of-two scale passes
    unchanged through the FFT, the hi/lo split, the MFMA products, the
    segmented sums and the compensated reduction, so the identity is bit
    for bit.
A few lag cells of the largest triangle are also checked against the FP64
oracle (per cell relative to the sum of |terms|)."""
import numpy as np
import pytest

import amber_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

N, NC, NC10 = 512, 768, 767


@pytest.fixture(scope="module")
def big():
    gv, samp, uf = A.generate(N, NC, nflag=0, seed=3001)
    return gv, samp, {}


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3", "fp32", "fp8", "lagft"])
def test_fullsize_counts_and_scaling(ctx, big, prec):
    gv, samp, cache = big
    lo, hi = A.log_rings(1, 200, 24)
    tris = A.triangles(lo, hi)
    pick = [t for t in tris if t[0] == 23][:1] + [tris[len(tris) // 2]]
    ctx.set_precision(prec)
    ctx.set_rings(lo, hi)
    c1 = ctx.load_cube(N, NC, 1.0, 0, gv, samp)
    ctx.set_channels(c1, None, NC10)
    c2 = ctx.load_cube(N, NC, 1.0, 0, 2.0 * gv, samp)
    ctx.set_channels(c2, None, NC10)
    print("case fullsize prec=%s triangles=%s" % (prec, [tuple(t) for t in pick]), flush=True)
    d = np.arange(-NC10, NC10 + 1)
    d2, d3 = np.meshgrid(d, d, indexing="ij")
    spread = np.maximum(0, np.maximum(d2, d3)) - np.minimum(0, np.minimum(d2, d3))
    closed = np.maximum(NC - spread, 0)
    for t in pick:
        a = ctx.bispec_ex(c1, *map(int, t))
        b = ctx.bispec_ex(c2, *map(int, t))
        assert np.array_equal(a["cnt"], closed)
        assert np.array_equal(b["bl64"], 8.0 * a["bl64"])
        assert np.array_equal(b["bl"], 8.0 * a["bl"])
    # FP64 oracle on a few cells of the largest triangle
    t = pick[0]
    res = ctx.bispec_ex(c1, *map(int, t))
    rng = np.random.default_rng(9)
    i2 = rng.integers(0, 2 * NC10 + 1, 6)
    i3 = rng.integers(0, 2 * NC10 + 1, 6)
    if "cells" not in cache:  # FP64 planes of 768 channels at Ng = 612: built once for both forms
        mask = np.ones(NC, np.uint8)
        cache["cells"] = O.bispec_cells(N, NC, 1.0, gv, samp, lo, hi, *map(int, t), mask, NC10, i2 - NC10,
                                        i3 - NC10, want_abs=True)
    b64, c64, _, ab = cache["cells"]
    got = res["bl64"][i2, i3]
    cell = np.max(np.abs(got - b64) / np.maximum(ab, 1e-300))
    print("fullsize Ng=%d prec=%s per-cell/|terms| %.3g" % (res["ng"], prec, cell), flush=True)
    # FP8 (3 mantissa bits, a screening form): its own bar
    assert cell <= (1e-4 if prec == "fp8" else 2e-8)
    ctx.release_cube(c1)
    ctx.release_cube(c2)
    ctx.set_precision("bf16x3")


# normwise bar at the metric workload, every one of its 768 channels
FULL_TOL = {"fp16x3": 1e-5, "fp32": 1e-5, "bf16x3": 1e-5, "lagft": 1e-5}


@pytest.fixture(scope="module")
def metric_cells(big):
    """FP64 oracle on 48 random lag cells (plus the zero lag) of the
    metric workload's Ng = 320 and Ng = 612 triangles, all 768 channels."""
    gv, samp, _ = big
    lo, hi = A.log_rings(1, 200, 24)
    tris = A.triangles(lo, hi)
    ngs = np.array([A.ng(hi[t[0]]) for t in tris])
    picks = [int(np.flatnonzero(ngs == 320)[0]), int(np.flatnonzero(ngs == 612)[0])]
    rng = np.random.default_rng(11)
    out = []
    for k in picks:
        t = tuple(int(x) for x in tris[k])
        d2 = np.append(rng.integers(-NC10, NC10 + 1, 48), 0)
        d3 = np.append(rng.integers(-NC10, NC10 + 1, 48), 0)
        print("oracle cells of triangle %s Ng=%d" % (t, ngs[k]), flush=True)
        b64, c64, _, ab = O.bispec_cells(N, NC, 1.0, gv, samp, lo, hi, *t, np.ones(NC, np.uint8), NC10, d2, d3,
                                         want_abs=True)
        out.append((t, int(ngs[k]), d2 + NC10, d3 + NC10, b64, c64))
    return lo, hi, out


@pytest.mark.parametrize("prec", ["fp16x3", "fp32", "bf16x3", "lagft"])
def test_metric_workload_error_vs_fp64(ctx, big, metric_cells, prec):
    """The contraction's error against the FP64 oracle at the metric
    workload (768 ch x 512^2, all lags), on the Ng = 320 and Ng = 612 bins
    where cancellation is deepest: normwise over 49 cells per triangle."""
    gv, samp, _ = big
    lo, hi, cells = metric_cells
    ctx.set_precision(prec)
    ctx.set_rings(lo, hi)
    c1 = ctx.load_cube(N, NC, 1.0, 0, gv, samp)
    ctx.set_channels(c1, None, NC10)
    try:
        for t, ng, i2, i3, b64, c64 in cells:
            res = ctx.bispec_ex(c1, *t)
            got = res["bl64"][i2, i3]
            assert np.array_equal(res["cnt"][i2, i3], c64)
            e = float(np.linalg.norm(got - b64) / np.linalg.norm(b64))
            print("metric workload %s triangle %s Ng=%d normwise %.3g" % (prec, t, ng, e), flush=True)
            assert e <= FULL_TOL[prec], (prec, t, e)
    finally:
        ctx.release_cube(c1)
        ctx.set_precision("bf16x3")


@pytest.mark.parametrize("prec,t", [("bf16x3", (23, 20, 20)), ("lagft", (23, 20, 20)), ("lagft", (23, 23, 20)),
                                    ("lagft", (23, 23, 23))])
def test_metric_workload_normalisation_is_exact(ctx, big, prec, t):
    """The amber_v2 normalisation at the metric workload (768 ch x 512^2,
    Ng = 612): equal bit for bit to the FP64 oracle's sums snapped to their
    1/8 grid of Ng^2, on 49 lag cells, whether the fast path certified the
    triangle or it was recomputed in FP64 (split form), or computed by the
    FP64 lag-Fourier stages (lagft, certificate margin far inside 1/32; the
    r1 == r2 triangle on the FP64 sheared tiles)."""
    gv, samp, _ = big
    lo, hi = A.log_rings(1, 200, 24)
    ctx.set_precision(prec)
    ctx.set_rings(lo, hi)
    c1 = ctx.load_cube(N, NC, 1.0, 0, gv, samp)
    ctx.set_channels(c1, None, NC10)
    rng = np.random.default_rng(13)
    d2 = np.append(rng.integers(-NC10, NC10 + 1, 48), 0)
    d3 = np.append(rng.integers(-NC10, NC10 + 1, 48), 0)
    try:
        ctx.reset_stats()
        res = ctx.bispec_ex(c1, *t, want_norm=True)
        st = ctx.stats()
        print("metric workload normalisation Ng=%d fast margin %.3g recomputed %d" %
              (res["ng"], st["cert_margin_max"], st["n_norm_exact"]), flush=True)
        _, c64, n64 = O.bispec_cells(N, NC, 1.0, gv, samp, lo, hi, *t, np.ones(NC, np.uint8), NC10, d2, d3,
                                     want_bl=False, want_norm=True)
        P = res["ng"] ** 2
        assert np.max(np.abs(n64 / P * 8 - np.rint(n64 / P * 8))) < 1e-6
        assert np.array_equal(res["cnt"][d2 + NC10, d3 + NC10], c64)
        assert np.array_equal(res["norm64"][d2 + NC10, d3 + NC10], P * np.rint(n64 / P * 8) / 8)
        if prec == "lagft":
            assert res["cert_margin"] < 1e-3
    finally:
        ctx.release_cube(c1)
        ctx.set_precision("bf16x3")


def test_metric_workload_symmetric_tilings(ctx, big):
    """The lag-Fourier form's symmetric tilings at the metric workload's
    transform length (768 channels, all lags: L = 1536), on its compute
    grids: an r1 == r2 triangle (sheared tiles with their k3 = L/2 row and
    fixed-point column), an r2 == r3 one (mode 4: five anti-diagonal corner
    cells at this L) and an equilateral one.  49 cells each against the FP64
    oracle, and every cell against the full (k2, k3) tiling of a context
    with the symmetries off (AMB_NO_SYM), counts equal."""
    import os
    gv, samp, _ = big
    lo, hi = A.log_rings(1, 200, 24)
    tris = [tuple(int(x) for x in t) for t in A.triangles(lo, hi)]
    small = [t for t in tris if A.ng(hi[t[0]]) <= 130]
    pick = [[t for t in small if t[0] == t[1] and t[1] != t[2]][-1],
            [t for t in small if t[1] == t[2] and t[0] != t[1]][-1],
            [t for t in small if t[0] == t[1] == t[2]][-1]]
    os.environ["AMB_NO_SYM"] = "1"
    try:
        plain = A.Context(device=0, precision="lagft")
    finally:
        del os.environ["AMB_NO_SYM"]
    ctx.set_precision("lagft")
    rng = np.random.default_rng(17)
    try:
        for c in (ctx, plain):
            c.set_rings(lo, hi)
        c1 = ctx.load_cube(N, NC, 1.0, 0, gv, samp)
        ctx.set_channels(c1, None, NC10)
        c2 = plain.load_cube(N, NC, 1.0, 0, gv, samp)
        plain.set_channels(c2, None, NC10)
        for t in pick:
            a = ctx.bispec_ex(c1, *t)
            b = plain.bispec_ex(c2, *t)
            assert np.array_equal(a["cnt"], b["cnt"])
            e = np.linalg.norm(a["bl64"] - b["bl64"]) / np.linalg.norm(b["bl64"])
            d2 = np.append(rng.integers(-NC10, NC10 + 1, 48), 0)
            d3 = np.append(rng.integers(-NC10, NC10 + 1, 48), 0)
            b64, c64, _ = O.bispec_cells(N, NC, 1.0, gv, samp, lo, hi, *t, np.ones(NC, np.uint8), NC10, d2, d3)
            got = a["bl64"][d2 + NC10, d3 + NC10]
            eo = np.linalg.norm(got - b64) / np.linalg.norm(b64)
            print("metric workload lagft tiling of %s (grid %d of Ng %d): vs full tiling %.3g, vs oracle %.3g"
                  % (t, A.compute_grid(hi[t[0]], hi[t[1]], hi[t[2]]), A.ng(hi[t[0]]), e, eo), flush=True)
            assert np.array_equal(a["cnt"][d2 + NC10, d3 + NC10], c64)
            assert e <= 2e-6 and eo <= FULL_TOL["lagft"], (t, e, eo)
        ctx.release_cube(c1)
    finally:
        plain.close()
        ctx.set_precision("bf16x3")

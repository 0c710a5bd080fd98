"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Counts (v1 numdnu, calCR Num, v2 normalisation) must be bit-exact against
the exact integer values; accumulators are compared normwise and per cell
against the oracle with the tolerances written in each test."""
import numpy as np
import pytest

import amber_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

# relative-error bars vs the FP64 oracle (normwise ||B-B64|| / ||B64||);
# fp8 (3 mantissa bits) is a screening form with its own bar
TOL = {"fp32": 1e-5, "bf16x3": 1e-5, "fp16x3": 1e-5, "fp8": 0.15}


def _setup(ctx, N, nc, nflag, seed, lo, hi, nc10, **gen):
    gv, samp, uf = A.generate(N, nc, nflag=nflag, seed=seed, **gen)
    fk = 1 if nflag > 0 else 0
    cube = ctx.load_cube(N, nc, 1.0, fk, gv, samp)
    ctx.set_channels(cube, uf if fk else None, nc10)
    ctx.set_rings(lo, hi)
    mask = O.chan_mask(nc, fk, uf)
    return cube, gv, samp, uf, mask


def _relerr(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _exact_norm(n64, P):
    """oracle FP64 normalisation snapped to its exact value P * (k/8)"""
    return P * np.rint(n64 / P * 8.0) / 8.0


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp16x3"])
def test_small_v1_v2_vs_oracle(ctx, prec):
    N, nc, nc10 = 32, 8, 7
    lo, hi = A.log_rings(2, 10, 4)
    ctx.set_precision(prec)
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 2, 9, lo, hi, nc10)
    print("case small N=%d nc=%d prec=%s" % (N, nc, prec), flush=True)
    for (r1, r2, r3) in A.triangles(lo, hi):
        res = ctx.bispec_ex(cube, int(r1), int(r2), int(r3), want_norm=True)
        bl64, cnt64, n64 = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10, want_norm=True)
        assert np.array_equal(res["cnt"], cnt64)
        P = res["ng"] ** 2
        assert np.array_equal(res["norm64"], _exact_norm(n64, P))
        assert np.array_equal(res["norm"], _exact_norm(n64, P).astype(np.float32))
        if np.linalg.norm(bl64) > 0:
            assert _relerr(res["bl64"], bl64) <= TOL[prec], (r1, r2, r3)
    ctx.release_cube(cube)


def test_fp8_form(ctx):
    # e4m3 carries 3 mantissa bits: the contraction error is a few 1e-2 of
    # |B|; the counts and the (BF16x3) normalisation stay exact
    N, nc, nc10 = 48, 10, 9
    lo, hi = A.log_rings(2, 14, 4)
    ctx.set_precision("fp8")
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 1, 41, lo, hi, nc10)
    worst = 0.0
    for (r1, r2, r3) in A.triangles(lo, hi)[::3]:
        res = ctx.bispec_ex(cube, int(r1), int(r2), int(r3), want_norm=True)
        bl64, cnt64, n64 = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10, want_norm=True)
        assert np.array_equal(res["cnt"], cnt64)
        P = res["ng"] ** 2
        assert np.array_equal(res["norm64"], _exact_norm(n64, P))
        if np.linalg.norm(bl64) > 0:
            worst = max(worst, _relerr(res["bl64"], bl64))
    print("fp8 worst normwise rel err %.3g" % worst, flush=True)
    assert worst <= 0.15
    ctx.set_precision("bf16x3")
    ctx.release_cube(cube)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3", "fp32", "fp8"])
def test_r1_eq_r2_symmetry_is_bit_exact(ctx, prec, monkeypatch):
    """With r1 == r2 the kernel skips the row tiles below the reference
    planes and the reduction reads C[p2][p1][p3]; the A operand is the same
    product of the same two floats, so results match the full computation
    bit for bit."""
    N, nc, nc10 = 48, 300, 299
    lo, hi = A.log_rings(2, 14, 4)
    ctx.set_precision(prec)
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 5, 57, lo, hi, nc10)
    monkeypatch.setenv("AMB_NO_SYM", "1")
    full = A.Context(device=0, precision=prec)
    monkeypatch.delenv("AMB_NO_SYM")
    cube2 = full.load_cube(N, nc, 1.0, 1, gv, samp)
    full.set_channels(cube2, uf, nc10)
    full.set_rings(lo, hi)
    tris = [t for t in A.triangles(lo, hi) if t[0] == t[1]]
    assert tris
    for (r1, r2, r3) in tris[-3:]:
        a = ctx.bispec_ex(cube, int(r1), int(r2), int(r3), want_norm=(prec != "fp8"))
        b = full.bispec_ex(cube2, int(r1), int(r2), int(r3), want_norm=(prec != "fp8"))
        assert np.array_equal(a["bl64"], b["bl64"]) and np.array_equal(a["cnt"], b["cnt"])
        if prec != "fp8":
            assert np.array_equal(a["norm64"], b["norm64"])
    full.close()
    ctx.set_precision("bf16x3")
    ctx.release_cube(cube)


def test_window_and_flags(ctx):
    # nc10 < nc-1 exercises per-p1 windows; flagged channels drop out
    N, nc, nc10 = 32, 12, 3
    lo, hi = A.log_rings(2, 10, 3)
    ctx.set_precision("bf16x3")
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 3, 17, lo, hi, nc10)
    for (r1, r2, r3) in A.triangles(lo, hi)[:4]:
        res = ctx.bispec_ex(cube, int(r1), int(r2), int(r3))
        bl64, cnt64, _ = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10)
        assert np.array_equal(res["cnt"], cnt64)
        if np.linalg.norm(bl64) > 0:
            assert _relerr(res["bl64"], bl64) <= TOL["bf16x3"]
    ctx.release_cube(cube)


def test_calcr_exact(ctx):
    N, nc = 64, 8
    lo, hi = A.log_rings(2, 20, 6)
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 0, 4, lo, hi, 3, p0=0.8, u0=60.0)
    got = ctx.closed_rings(cube, plane=5)
    ref = O.calcr(N, nc, samp, lo, hi, plane=5)
    assert np.array_equal(got[:, :3], ref[:, :3])
    exact = np.rint(ref[:, 3] / N ** 2 * 8) / 8 * N ** 2
    assert np.array_equal(got[:, 3], exact)
    ctx.release_cube(cube)


@pytest.mark.parametrize("nc,rings,plane", [(40, (1, 200, 20), 5), (40, (1, 200, 24), 5)])
def test_calcr_n512(ctx, nc, rings, plane):
    """calCR on a 512^2 grid with the scaled and precision workloads' ring
    sets (20 and 24 log bins 1-200; rings past N/3 are cut by the column
    rule j < N/3): every triangle's Num equals the FP64 oracle's, snapped to
    its 1/8 grid, bit for bit."""
    N = 512
    lo, hi = A.log_rings(*rings)
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 0, 2001, lo, hi, nc - 1)
    print("case calCR N=%d rings=%d" % (N, len(lo)), flush=True)
    got = ctx.closed_rings(cube, plane=plane)
    ref = O.calcr(N, nc, samp, lo, hi, plane=plane)
    assert len(got) == len(A.triangles(lo, hi))
    assert np.array_equal(got[:, :3], ref[:, :3])
    q = ref[:, 3] / N ** 2 * 8
    assert np.max(np.abs(q - np.rint(q))) < 1e-6     # the oracle's sums sit on the grid
    assert np.array_equal(got[:, 3], np.rint(q) / 8 * N ** 2)
    ctx.release_cube(cube)


def test_deterministic_alone_and_in_batch(ctx):
    N, nc, nc10 = 32, 10, 9
    lo, hi = A.log_rings(2, 10, 4)
    ctx.set_precision("bf16x3")
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 1, 23, lo, hi, nc10)
    tris = A.triangles(lo, hi)
    alone = [ctx.bispec_ex(cube, *map(int, t))["bl"] for t in tris]
    bl, cnt, _ = ctx.bispec_batch(cube, tris[::-1])
    for k, t in enumerate(tris[::-1]):
        i = len(tris) - 1 - k
        assert np.array_equal(bl[k], alone[i])
    again = ctx.bispec_ex(cube, *map(int, tris[-1]))["bl"]
    assert np.array_equal(again, alone[-1])
    ctx.release_cube(cube)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
def test_cfg1_plain_small(ctx, prec):
    """64 ch x 128^2, 6 flagged, 8 log bins 2->60, lags to 63, all 46
    triangles; FP64 oracle on every cell of a subset of triangles."""
    N, nc, nc10 = 128, 64, 63
    lo, hi = A.log_rings(2, 60, 8)
    ctx.set_precision(prec)
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 6, 1001, lo, hi, nc10)
    tris = A.triangles(lo, hi)
    assert len(tris) == 46
    print("case cfg1 prec=%s" % prec, flush=True)
    bl, cnt, _ = ctx.bispec_batch(cube, tris)
    worst = 0.0
    for k in (0, 11, 23, 45):
        r1, r2, r3 = map(int, tris[k])
        b64, c64, _ = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10)
        assert np.array_equal(cnt[k], c64)
        e = _relerr(bl[k].astype(np.float64), b64)
        worst = max(worst, e)
    print("cfg1 worst normwise rel err %.3g" % worst, flush=True)
    assert worst <= TOL[prec]
    ctx.release_cube(cube)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp16x3"])
def test_large_grid_cells(ctx, prec):
    """A 150-200 ring bin on a 512^2 grid (Ng = 612, 375k pixels per
    plane, |B| about 1e-3 of the sum of |terms|): the FP64 oracle on a
    spread of lag cells, compared normwise over the cells and per cell
    against the sum of |terms|.  The reference's own FP32 arithmetic (the
    loop restated in oracle/orc_refalgo.c) is scored on the same cells:
    the FP32 form must do at least as well."""
    N, nc, nc10 = 512, 24, 23
    lo = np.array([150.0]); hi = np.array([200.0])
    ctx.set_precision(prec)
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 2, 77, lo, hi, nc10)
    res = ctx.bispec_ex(cube, 0, 0, 0)
    W = 2 * nc10 + 1
    rng = np.random.default_rng(5)
    i2 = rng.integers(0, W, 24)
    i3 = rng.integers(0, W, 24)
    b64, c64, _, ab = O.bispec_cells(N, nc, 1.0, gv, samp, lo, hi, 0, 0, 0, mask, nc10, i2 - nc10, i3 - nc10,
                                     want_abs=True)
    got = res["bl64"][i2, i3]
    assert np.array_equal(res["cnt"][i2, i3], c64)
    live = c64 > 0
    e = _relerr(got[live], b64[live])
    cell = np.max(np.abs(got[live] - b64[live]) / np.maximum(ab[live], 1e-300))
    rb = O.refalgo(N, nc, 1.0, gv, samp, lo, hi, 0, 0, 0, mask, nc10)[0][i2, i3].astype(np.float64)
    e_ref = _relerr(rb[live], b64[live])
    print("large grid Ng=%d prec=%s normwise %.3g per-cell/|terms| %.3g (reference FP32 loop: %.3g)"
          % (res["ng"], prec, e, cell, e_ref), flush=True)
    # per cell against the sum of |terms|: BF16x3's operands keep 16
    # significant bits (a 2^-17 residual per factor), FP32 and FP16x3's 22+
    assert cell <= (3e-8 if prec == "bf16x3" else 2e-8)
    assert e <= TOL[prec]
    if prec == "fp32":
        assert e <= e_ref
    ctx.set_precision("bf16x3")
    ctx.release_cube(cube)


@pytest.mark.parametrize("N,nc,lo,hi,ng", [(32, 6, 2.0, 9.0, 0), (64, 6, 5.0, 13.0, 0),
                                           (64, 6, 3.0, 6.0, 48), (512, 4, 150.0, 200.0, 0)])
def test_ring_image_vs_oracle(ctx, N, nc, lo, hi, ng):
    """The c2r stage alone: data and sampling images of every active plane
    against the FP64 ring gather + c2r, normwise per plane.  Ng=612
    (2^2 3^2 17) runs the radix-17 stage; ng=48 a grid above the ring's
    own."""
    lo_a, hi_a = np.array([lo]), np.array([hi])
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 1, 31, lo_a, hi_a, nc - 1)
    n = ng if ng > 0 else O.ng(hi)
    print("case ring image N=%d Ng=%d" % (N, n), flush=True)
    worst = 0.0
    for p in range(nc):
        if not mask[p]:
            with pytest.raises(RuntimeError):
                ctx.ring_image(cube, 0, p, kind=0, ng=ng)
            continue
        for kind in (0, 1):
            got = ctx.ring_image(cube, 0, p, kind=kind, ng=ng)
            ref = O.ring_c2r(N, nc, n, lo, hi, p, gv=gv if kind == 0 else None,
                             samp=samp if kind == 1 else None)
            assert got.shape == (n, n)
            if np.linalg.norm(ref) == 0:
                assert not np.any(got)
                continue
            worst = max(worst, _relerr(got.astype(np.float64), ref))
    print("ring image worst normwise rel err %.3g" % worst, flush=True)
    assert worst <= 1e-6
    ctx.release_cube(cube)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("N,nc,nflag,nc10,rings", [
    (64, 200, 7, 199, (2, 20, 5)),    # partial 128-row tiles, flags, one segment
    (256, 40, 0, 12, (20, 60, 3)),    # lag windows, Ng up to 186: several 8192-pixel segments
])
def test_launch_configurations_are_bit_identical(ctx, N, nc, nflag, nc10, rings, prec):
    """Every launch configuration the autotuner may pick runs the same
    arithmetic in the same order: the split-form kernels with direct
    global->LDS loads and pipelined k-steps (variant 2), the interleaved
    variant 1 and the plain variant 0, under any XCD-queue run length,
    wave size and grid, give the same results, data and v2 normalisation,
    per triangle and in grouped waves, bit for bit."""
    lo, hi = A.log_rings(*rings)
    gv, samp, uf = A.generate(N, nc, nflag=nflag, seed=61)
    fk = 1 if nflag else 0
    ctx.set_precision(prec)
    cb = ctx.load_cube(N, nc, 1.0, fk, gv, samp)
    ctx.set_channels(cb, uf if fk else None, nc10)
    ctx.set_rings(lo, hi)
    tris = A.triangles(lo, hi)
    print("case launch configurations %s N=%d nc=%d ntri=%d" % (prec, N, nc, len(tris)), flush=True)
    configs = [dict(variant=2), dict(variant=0), dict(variant=1, xrun_log2=3, grid_pct=50),
               dict(variant=2, xrun_log2=8, wave_gib=1, grid_pct=75)]
    try:
        for mode in ("per_triangle", "grouped"):
            ctx.set_batch_mode(mode)
            outs = []
            for cfg in configs:
                ctx.set_launch_config(**cfg)
                outs.append(ctx.bispec_batch(cb, tris, want_norm=True))
            for cfg, o in zip(configs[1:], outs[1:]):
                for k in range(3):
                    assert np.array_equal(outs[0][k], o[k]), (mode, cfg, k)
    finally:
        ctx.set_launch_config(None)
        ctx.set_batch_mode("grouped")
        ctx.set_precision("bf16x3")
        ctx.release_cube(cb)


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_uncertified_normalisation_is_recomputed_exactly(ctx, prec):
    """At Ng = 134 the v2 normalisation sums reach ~1.4e5 Ng^2, past what the
    fast path's FP32 accumulation resolves on their 1/8 grid (the FP64
    oracle's sums sit on it to 1e-9): the call counts the miss, recomputes
    the triangle in FP64 and returns the exact sums bit for bit; in strict
    mode it fails with AMB_ECERT instead."""
    N, nc, nc10 = 256, 40, 12
    lo, hi = A.log_rings(20, 60, 3)
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 0, 61, lo, hi, nc10)
    r1, r2, r3 = 1, 0, 0
    ctx.set_precision(prec)
    print("case exact normalisation %s N=%d Ng=%d" % (prec, N, A.ng(hi[r1])), flush=True)
    bl64, cnt64, n64 = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10, want_norm=True)
    P = A.ng(hi[r1]) ** 2
    assert np.max(np.abs(n64 / P * 8 - np.rint(n64 / P * 8))) < 1e-6   # the exact sums are on the grid
    ctx.reset_stats()
    res = ctx.bispec_ex(cube, r1, r2, r3, want_norm=True)
    st = ctx.stats()
    assert st["n_uncertified"] == 1 and st["n_norm_exact"] == 1 and st["cert_margin_max"] > 1.0 / 32
    assert res["cert_margin"] < 1e-6                                    # the FP64 sums' own distance
    assert np.array_equal(res["cnt"], cnt64)
    assert np.array_equal(res["norm64"], _exact_norm(n64, P))
    assert np.array_equal(res["norm"], _exact_norm(n64, P).astype(np.float32))
    # in a grouped batch with other triangles: the same exact sums
    tris = np.array([[r1, r2, r3], [2, 2, 1], [2, 1, 1]], np.int32)
    bl, cnt, norm = ctx.bispec_batch(cube, tris, want_norm=True)
    assert np.array_equal(norm[0], res["norm"])
    for k in (1, 2):
        _, _, nk = O.bispec(N, nc, 1.0, gv, samp, lo, hi, *tris[k], mask, nc10, want_norm=True)
        assert np.array_equal(norm[k], _exact_norm(nk, A.ng(hi[tris[k][0]]) ** 2).astype(np.float32))
    ctx.set_cert_strict(True)
    try:
        with pytest.raises(A.AmberError) as e:
            ctx.bispec_ex(cube, r1, r2, r3, want_norm=True)
        assert e.value.code == -8
    finally:
        ctx.set_cert_strict(False)
    ctx.set_precision("bf16x3")
    ctx.release_cube(cube)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3", "fp32", "fp8"])
def test_r2_eq_r3_symmetry(ctx, prec, monkeypatch):
    """With r2 == r3 (r1 != r2) the kernel skips the (p2, p3) tiles below the
    diagonal and the reduction reads the mirror element C[p1][p3][p2]: the
    same sum with two planes' roles exchanged, so equal to the full
    computation up to rounding; counts and the v2 normalisation stay exact,
    and the lag array comes out exactly symmetric."""
    N, nc, nc10 = 48, 300, 299
    lo, hi = A.log_rings(2, 14, 4)
    ctx.set_precision(prec)
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 5, 57, lo, hi, nc10)
    monkeypatch.setenv("AMB_NO_SYM", "1")
    full = A.Context(device=0, precision=prec)
    monkeypatch.delenv("AMB_NO_SYM")
    cube2 = full.load_cube(N, nc, 1.0, 1, gv, samp)
    full.set_channels(cube2, uf, nc10)
    full.set_rings(lo, hi)
    tris = [t for t in A.triangles(lo, hi) if t[1] == t[2] and t[0] != t[1]]
    assert tris
    print("case r2 == r3 symmetry prec=%s ntri=%d" % (prec, len(tris)), flush=True)
    for (r1, r2, r3) in tris[-3:]:
        a = ctx.bispec_ex(cube, int(r1), int(r2), int(r3), want_norm=True)
        b = full.bispec_ex(cube2, int(r1), int(r2), int(r3), want_norm=True)
        assert np.array_equal(a["cnt"], b["cnt"]) and np.array_equal(a["norm64"], b["norm64"])
        assert np.array_equal(a["bl64"], a["bl64"].T)
        # two roundings of the same sums (for fp8 two quantisations: the
        # mirror element quantises D_r1 D_r3 and D_r2 instead)
        assert _relerr(a["bl64"], b["bl64"]) <= TOL[prec]
        bl64, cnt64, _ = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10)
        assert _relerr(a["bl64"], bl64) <= TOL[prec]
    full.close()
    ctx.set_precision("bf16x3")
    ctx.release_cube(cube)


@pytest.mark.parametrize("prec", ["fp16x3", "fp32", "bf16x3", "fp8"])
def test_power_of_two_scale_is_exact(ctx, prec):
    """Scaling the cube by 2^k scales every data sum by exactly 2^(3k) in
    every form: the planes are stored normalised by a power of two and the
    exponent is undone exactly, so the operands the matrix cores see are
    the same bits for both cubes."""
    N, nc, nc10 = 48, 20, 19
    lo, hi = A.log_rings(2, 14, 4)
    ctx.set_precision(prec)
    gv, samp, uf = A.generate(N, nc, nflag=2, seed=71)
    c1 = ctx.load_cube(N, nc, 1.0, 1, gv, samp)
    c2 = ctx.load_cube(N, nc, 1.0, 1, gv * 2.0 ** -5, samp)
    for c in (c1, c2):
        ctx.set_channels(c, uf, nc10)
    ctx.set_rings(lo, hi)
    for t in A.triangles(lo, hi)[::2]:
        a = ctx.bispec_ex(c1, *map(int, t))
        b = ctx.bispec_ex(c2, *map(int, t))
        assert np.array_equal(b["bl64"], a["bl64"] * 2.0 ** -15), (prec, t)
    ctx.release_cube(c1)
    ctx.release_cube(c2)
    ctx.set_precision("bf16x3")


def test_counts_above_one_are_normalised(ctx):
    """normalize_data on upload (amber_funcs.c:1946-1978): a cube whose
    sampled cells carry counts 1..4 (V = S c, as vcg writes them) gives the
    bispectrum, counts, v2 normalisation and calCR of the oracle run on the
    normalised arrays."""
    N, nc, nc10 = 64, 10, 9
    lo, hi = A.log_rings(2, 18, 4)
    ctx.set_precision("bf16x3")
    gv, samp, uf = A.generate(N, nc, nflag=2, seed=83, count_max=4)
    assert samp.max() == 4
    gn, sn = O.normalize(gv, samp)
    cube = ctx.load_cube(N, nc, 1.0, 1, gv, samp)
    ctx.set_channels(cube, uf, nc10)
    ctx.set_rings(lo, hi)
    mask = O.chan_mask(nc, 1, uf)
    print("case counts > 1 N=%d nc=%d" % (N, nc), flush=True)
    for (r1, r2, r3) in A.triangles(lo, hi)[::2]:
        res = ctx.bispec_ex(cube, int(r1), int(r2), int(r3), want_norm=True)
        b64, c64, n64 = O.bispec(N, nc, 1.0, gn, sn, lo, hi, r1, r2, r3, mask, nc10, want_norm=True)
        assert np.array_equal(res["cnt"], c64)
        assert np.array_equal(res["norm64"], _exact_norm(n64, res["ng"] ** 2))
        if np.linalg.norm(b64) > 0:
            assert _relerr(res["bl64"], b64) <= TOL["bf16x3"]
    got = ctx.closed_rings(cube, plane=5)
    ref = O.calcr(N, nc, sn, lo, hi, plane=5)
    assert np.array_equal(got[:, 3], np.rint(ref[:, 3] / N ** 2 * 8) / 8 * N ** 2)
    ctx.release_cube(cube)


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_du_above_one(ctx, prec):
    """Du = 1.0254 (the shipped grid's, mabs_funcs.py:30): only Ng changes
    (amber.c:193-197, Ng = 3 floor((UHigh + 3) / Du) + 2 or 3), and rings
    near Ng/3 alias on the reduced grid as in the reference (3 max|k| can
    exceed Ng); bispectrum, counts and normalisation against the oracle."""
    N, nc, nc10, Du = 128, 8, 7, 1.0254
    lo, hi = A.log_rings(2, 40, 6)
    ctx.set_precision(prec)
    gv, samp, uf = A.generate(N, nc, nflag=0, seed=91)
    cube = ctx.load_cube(N, nc, Du, 0, gv, samp)
    ctx.set_channels(cube, None, nc10)
    ctx.set_rings(lo, hi)
    mask = O.chan_mask(nc, 0, uf)
    tris = A.triangles(lo, hi)
    assert any(A.ng(hi[t[0]], Du) != A.ng(hi[t[0]]) for t in tris)
    print("case Du=%g prec=%s ntri=%d" % (Du, prec, len(tris)), flush=True)
    for (r1, r2, r3) in tris[::3]:
        res = ctx.bispec_ex(cube, int(r1), int(r2), int(r3), want_norm=True)
        assert res["ng"] == O.ng(hi[r1], Du)
        b64, c64, n64 = O.bispec(N, nc, Du, gv, samp, lo, hi, r1, r2, r3, mask, nc10, want_norm=True)
        assert np.array_equal(res["cnt"], c64)
        assert np.array_equal(res["norm64"], _exact_norm(n64, res["ng"] ** 2))
        if np.linalg.norm(b64) > 0:
            assert _relerr(res["bl64"], b64) <= TOL[prec]
    ctx.release_cube(cube)
    ctx.set_precision("bf16x3")


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_cfg2_scaled(ctx, prec):
    """The scaled workload: 384 ch x 512^2, 31 flagged channels, 20 log bins
    1-200, all 383 lags.  The largest (Ng = 612) and a median triangle in
    one grouped batch: counts on sampled cells and the zero lag against the
    oracle exactly, sums normwise over 33 cells."""
    N, nc, nc10 = 512, 384, 383
    lo, hi = A.log_rings(1, 200, 20)
    ctx.set_precision(prec)
    cube, gv, samp, uf, mask = _setup(ctx, N, nc, 31, 2001, lo, hi, nc10)
    tris = A.triangles(lo, hi)
    assert len(tris) == 314
    ngs = np.array([A.ng(hi[t[0]]) for t in tris])
    pick = [int(np.argmax(ngs)), int(np.argsort(ngs)[len(ngs) // 2])]
    print("case cfg2 prec=%s triangles=%s" % (prec, [tuple(map(int, tris[k])) for k in pick]), flush=True)
    bl, cnt, _ = ctx.bispec_batch(cube, tris[pick])
    rng = np.random.default_rng(21)
    for n, k in enumerate(pick):
        r1, r2, r3 = map(int, tris[k])
        d2 = np.append(rng.integers(-nc10, nc10 + 1, 32), 0)
        d3 = np.append(rng.integers(-nc10, nc10 + 1, 32), 0)
        b64, c64, _ = O.bispec_cells(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10, d2, d3)
        got = bl[n][d2 + nc10, d3 + nc10].astype(np.float64)
        assert np.array_equal(cnt[n][d2 + nc10, d3 + nc10], c64)
        live = c64 > 0
        e = _relerr(got[live], b64[live])
        print("cfg2 %s triangle %s Ng=%d normwise %.3g" % (prec, (r1, r2, r3), ngs[k], e), flush=True)
        assert e <= TOL[prec]
    ctx.release_cube(cube)
    ctx.set_precision("bf16x3")

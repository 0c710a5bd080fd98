"""GPU: lag-space post-processing (amb_cyl_bispec) against the numpy
restatement of get_cyl_BS / cylbs_win (oracle/post.py), on random arrays
and on the pipeline's own v1 and v2 results."""
import numpy as np
import pytest

import amber_amd as A
import oracle.post as P

pytestmark = pytest.mark.gpu

K = A.ANALYSIS


def _ref(bl, den, nc10, window):
    return P.cyl_bispec(bl, den, nc10, K["ncut"], K["mabs_norm"], K["rc"], K["rc_prime"], K["dnu"], window)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("window", [False, True])
@pytest.mark.parametrize("kind", ["v1", "v2"])
def test_cyl_random(ctx, window, kind):
    rng = np.random.default_rng(11)
    nc10 = 40
    W = 2 * nc10 + 1
    bl = rng.standard_normal((3, W, W)).astype(np.float32) * 1e3
    den = (rng.integers(1, 500, (3, W, W)).astype(np.int32) if kind == "v1"
           else rng.uniform(1, 500, (3, W, W)).astype(np.float32))
    print("case cyl random %s window=%s" % (kind, window), flush=True)
    got = ctx.cyl_bispec(bl, den, window=window)
    for t in range(3):
        assert _rel(got[t], _ref(bl[t], den[t], nc10, window)) <= 1e-12


def test_cyl_on_pipeline_results(ctx):
    """v1 counts and v2 normalisation from the bispectrum path itself.  As
    in the reference's setup (nc10 = 383 of 768 channels), nc10 <= (nc-1)/2
    and a fully sampled cube keep every folded count and normalisation
    nonzero; a zero would make the quotient, and so the whole transform,
    non-finite in the reference as here."""
    N, nc, nc10 = 48, 48, 23
    lo, hi = A.log_rings(2, 14, 4)
    gv, samp, uf = A.generate(N, nc, nflag=0, seed=8, p0=1.0, u0=1e9)
    cube = ctx.load_cube(N, nc, 1.0, 0, gv, samp)
    ctx.set_channels(cube, None, nc10)
    ctx.set_rings(lo, hi)
    tris = A.triangles(lo, hi)[:4]
    bl, cnt, norm = ctx.bispec_batch(cube, tris, want_norm=True)
    for den in (cnt, norm):
        got = ctx.cyl_bispec(bl, den, window=True)
        for t in range(len(tris)):
            ref = _ref(bl[t], den[t], nc10, True)
            assert np.all(np.isfinite(ref))
            assert _rel(got[t], ref) <= 1e-12
    ctx.release_cube(cube)


@pytest.mark.parametrize("window", [False, True])
def test_cyl_full_size(ctx, window):
    """the metric workload's lag grid, W = 1535 (nc10 = 767): N = 1519 after
    the 8-lag cut, two matrix products of K = 1519 and K = 1520 on the FP64
    matrix cores, against numpy's FFT chain"""
    rng = np.random.default_rng(21)
    nc10 = 767
    W = 2 * nc10 + 1
    bl = rng.standard_normal((1, W, W)).astype(np.float32) * 1e3
    den = rng.integers(1, 767, (1, W, W)).astype(np.int32)
    print("case cyl full size window=%s" % window, flush=True)
    got = ctx.cyl_bispec(bl, den, window=window)
    assert got.shape == (1,) + (W - 2 * K["ncut"],) * 2
    assert _rel(got[0], _ref(bl[0], den[0], nc10, window)) <= 1e-12

"""The allocation-free calCR entry point (amb_closed_rings_noalloc) against
amb_closed_rings, through the C ABI: bit for bit on two calls in a row, the
second on a smaller grid with fewer rings (the static scratch keeps the
first call's contents), and a call needing more than the static scratch is
refused with AMB_EBUDGET before any device work."""
import numpy as np
import pytest

import amber_amd as A

pytestmark = pytest.mark.gpu


def _case(ctx, N, nc, rings, seed):
    lo, hi = A.log_rings(*rings)
    gv, samp, _ = A.generate(N, nc, nflag=0, seed=seed)
    cube = ctx.load_cube(N, nc, 1.0, 0, gv, samp)
    ctx.set_channels(cube, None, nc - 1)
    ctx.set_rings(lo, hi)
    return cube


def test_noalloc_closed_rings_equals_allocating():
    ctx = A.Context(device=0)
    try:
        for N, nc, rings, seed in ((512, 12, (1, 200, 20), 811), (256, 8, (2, 80, 9), 812)):
            print("case noalloc calCR N=%d rings=%d" % (N, rings[2]), flush=True)
            cube = _case(ctx, N, nc, rings, seed)
            ref = ctx.closed_rings(cube, plane=5)
            got = ctx.closed_rings(cube, plane=5, noalloc=True)
            assert len(got) == len(ref) > 0
            assert np.array_equal(got, ref)
            ctx.release_cube(cube)
        # 24 rings on a 1024^2 grid need 24 * 1024^2 * 8 B = 192 MiB of images
        cube = _case(ctx, 1024, 6, (1, 300, 24), 813)
        with pytest.raises(A.AmberError) as e:
            ctx.closed_rings(cube, plane=5, noalloc=True)
        assert e.value.code == -6  # AMB_EBUDGET
        print("oversized call refused: %s" % e.value, flush=True)
        ctx.release_cube(cube)
    finally:
        ctx.close()

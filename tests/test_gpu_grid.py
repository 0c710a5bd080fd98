"""The gridder on the GPU (amb_grid / amb_grid_cube: vcg's
group_loop_gridding, group_loop_gridding_without_flag and
collapse_channels, 3_vcg/tgefuncs.c:377-626, 816-839) against the oracle:
bit-exact, since every cell adds its groups in the reference's order."""
import numpy as np
import pytest

import amber_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def groups(n, nchan, nstokes, Umax, nu0, seed, flag_frac=0.2, outside=0.1, dense=None):
    rng = np.random.default_rng(seed)
    r = Umax * np.sqrt(rng.random(n)) * (1 + outside)
    th = rng.random(n) * 2 * np.pi
    if dense is not None:                       # many groups per cell: a few radii only
        r = rng.choice(dense, n)
    ranpar = np.zeros((n, 5), np.float32)
    ranpar[:, 0] = r * np.cos(th) / nu0
    ranpar[:, 1] = r * np.sin(th) / nu0
    ranpar[:, 2] = rng.random(n)
    data = rng.standard_normal((n, nchan, nstokes, 3)).astype(np.float32)
    data[..., 2] = np.where(rng.random((n, nchan, nstokes)) < flag_frac, rng.choice([0.0, -1.0]), 1.0)
    return ranpar, data


@pytest.mark.parametrize("flag_key", [1, 0])
@pytest.mark.parametrize("n_collapse", [1, 4])
def test_grid_bit_exact(flag_key, n_collapse):
    nu0, Umax, FWHM = np.float32(1.5e8), 60.0, 800.0      # dU 1.90, NN 64
    dU, NN = O.grid_setup(Umax, FWHM)
    nchan, nstokes, nc, chan1 = 14, 2, 10, 3
    ranpar, data = groups(3000, nchan, nstokes, Umax, nu0, seed=7 + flag_key + n_collapse,
                          dense=[3.0 * dU, 7.5 * dU, 12.2 * dU])
    print("case grid NN=%d flag=%d collapse=%d" % (NN, flag_key, n_collapse), flush=True)
    with A.Context(device=0) as ctx:
        gv, samp, n = ctx.grid(NN, nc, ranpar, data, nu0, dU, Umax, flag_key, n_collapse, chan1)
    ogv, osamp, on = O.grid(NN, nc, ranpar, data, nu0, dU, Umax, flag_key, n_collapse, chan1)
    assert n == on and n > 2000
    assert np.array_equal(samp, osamp)
    assert np.array_equal(gv.view(np.uint64), ogv.view(np.uint64))   # bit for bit, signed zeros included
    assert samp.max() > 20                                            # cells with many groups


def test_grid_sparse_four_stokes():
    nu0, Umax, FWHM = np.float32(1.8e8), 80.0, 600.0      # dU 2.53, NN 64
    dU, NN = O.grid_setup(Umax, FWHM)
    ranpar, data = groups(5000, 6, 4, Umax, nu0, seed=3)
    with A.Context(device=0) as ctx:
        gv, samp, n = ctx.grid(NN, 6, ranpar, data, nu0, dU, Umax)
    ogv, osamp, on = O.grid(NN, 6, ranpar, data, nu0, dU, Umax)
    assert n == on and np.array_equal(samp, osamp) and np.array_equal(gv.view(np.uint64), ogv.view(np.uint64))


def test_grid_cube_equals_loaded_cube():
    """the resident cube equals amb_load_cube of the same arrays: compare
    their filtered ring images bit for bit"""
    nu0, Umax, FWHM = np.float32(1.5e8), 30.0, 1500.0     # dU 1.01, NN 60
    dU, NN = O.grid_setup(Umax, FWHM)
    nc = 8
    ranpar, data = groups(4000, nc, 2, Umax, nu0, seed=11, dense=[2.0 * dU, 5.3 * dU, 9.1 * dU])
    lo, hi = A.log_rings(1, 12, 4)
    with A.Context(device=0) as ctx:
        cid, n = ctx.grid_cube(NN, nc, ranpar, data, nu0, dU, Umax, flag_key=0)
        gv, samp, _ = ctx.grid(NN, nc, ranpar, data, nu0, dU, Umax, flag_key=0)
        ref = ctx.load_cube(NN, nc, dU, 0, gv, samp)
        ctx.set_rings(lo, hi)
        for c in (cid, ref):
            ctx.set_channels(c, None, nc - 1)
        for ring in range(4):
            for plane in (0, nc - 1):
                ng = A.ng(hi[ring], dU)
                a = ctx.ring_image(cid, ring, plane, ng=ng)
                b = ctx.ring_image(ref, ring, plane, ng=ng)
                assert np.array_equal(a, b)
                assert np.abs(a).max() > 0


def test_grid_archived_size():
    """the archived 1176^2 grid of the shipped inputparameters, 64 channels,
    100k groups: bit-exact against the oracle, totals exact"""
    nu0, Umax, FWHM = np.float32(1.54e8), 602.0, 1480.73
    dU, NN = O.grid_setup(Umax, FWHM)
    assert NN == 1176
    rng = np.random.default_rng(5)
    n = 100000
    ranpar, data = groups(n, 64, 2, Umax, nu0, seed=5, flag_frac=0.0)
    data[..., :2] = rng.integers(-8, 8, data[..., :2].shape)          # integers: every sum exact
    with A.Context(device=0) as ctx:
        gv, samp, got = ctx.grid(NN, 64, ranpar, data, nu0, dU, Umax, flag_key=0)
    ogv, osamp, on = O.grid(NN, 64, ranpar, data, nu0, dU, Umax, flag_key=0)
    assert got == on
    assert np.array_equal(samp, osamp) and np.array_equal(gv, ogv)
    assert samp.sum() == got * 64
    cells, sign = A.grid_cells(NN, ranpar, nu0, dU, Umax)
    inside = cells >= 0
    want_re = data[inside][:, :, 0, 0].astype(np.float64).sum()
    want_im = (data[inside][:, :, 0, 1] * sign[inside][:, None]).astype(np.float64).sum()
    assert gv[..., 0].sum() == want_re and gv[..., 1].sum() == want_im

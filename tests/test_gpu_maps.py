"""MAPS C_l(dnu) on the GPU (amb_maps_cldnu, the maps program's pair loop
maps.c:100-229 over cl_dnu, amber_funcs.c:1418-1488) against the FP64
oracle: cldnu within 1e-12 of the sum of |terms| (the Gram sums run in
another order than the reference's cell loop), counts, numdnu and binu exact
(same cell order and expressions).  At the 768-channel 512^2 shape the
closed form of a phase cube checks the whole grid."""
import math
import os
import subprocess

import numpy as np
import pytest

import amber_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rand_cube(N, nc, seed, pmax=3):
    rng = np.random.default_rng(seed)
    npix = N * (N // 2 + 1)
    samp = rng.integers(0, pmax, size=(npix, nc)).astype(np.float64)
    gv = rng.normal(size=(npix, nc, 2)) * samp[..., None]
    return gv, samp


def abs_scale(N, nc, Du, gv, samp, NB=15, umin=1.0, umax=500.0):
    """sum over n1 of sum_x |Re1 Re2| + |Im1 Im2| over [S1>0][S2>0] / no_b(n1): the size of cldnu's terms"""
    Nh = N // 2 + 1
    on = samp > 0
    v = np.where(on[..., None], gv / np.where(on, samp, 1.0)[..., None], gv)
    out = np.zeros((NB, nc))
    bins = np.array([O.maps_bin(N, Du, x // Nh, x % Nh, NB, umin, umax)[0] for x in range(N * Nh)])
    for b in range(NB):
        xs = np.nonzero(bins == b)[0]
        if len(xs) == 0:
            continue
        a = np.abs(v[xs] * on[xs][..., None])
        w = np.abs(v[xs])
        M = a[..., 0].T @ w[..., 0] + a[..., 1].T @ w[..., 1]
        no = on[xs].sum(0)
        for d in range(nc):
            n1 = np.arange(nc - d)
            t = M[n1, n1 + d] / np.where(no[n1] > 0, no[n1], 1)
            out[b, d] = t.sum()
    return out


def check(ctx, N, nc, Du, gv, samp, **kw):
    cl, nd, bu, nm = ctx.maps_cldnu(N, nc, Du, gv, samp, **kw)
    ocl, ond, obu, onm = O.cldnu(N, nc, Du, gv, samp, **kw)
    sc = abs_scale(N, nc, Du, gv, samp, **kw)
    err = np.abs(cl - ocl)
    assert np.all(err <= 1e-12 * sc + 1e-300), float((err / np.maximum(sc, 1e-300)).max())
    assert np.array_equal(nd, ond)
    assert np.array_equal(nm, onm)
    assert np.array_equal(bu, obu)
    return cl


@pytest.mark.parametrize("N,nc,Du", [(32, 24, 9.0), (48, 7, 11.3), (64, 130, 6.0)])
def test_maps_vs_oracle(N, nc, Du):
    gv, samp = rand_cube(N, nc, seed=N + nc)
    with A.Context(device=0) as ctx:
        check(ctx, N, nc, Du, gv, samp)


def test_maps_multi_chunk_and_ragged_tiles():
    # bins above 2048 cells are cut into chunks; 140 channels leave a ragged second tile
    N, nc, Du = 128, 140, 3.0
    gv, samp = rand_cube(N, nc, seed=5)
    with A.Context(device=0) as ctx:
        check(ctx, N, nc, Du, gv, samp)


def test_maps_edges():
    with A.Context(device=0) as ctx:
        gv, samp = rand_cube(32, 1, seed=3)
        check(ctx, 32, 1, 9.0, gv, samp)                       # one channel
        gv, samp = rand_cube(32, 6, seed=4)
        check(ctx, 32, 6, 1.0, gv, samp, NB=4, umin=200.0, umax=300.0)   # no cell in range
        samp[:, 2] = 0                                          # an unsampled channel
        check(ctx, 32, 6, 9.0, gv, samp)
        check(ctx, 32, 6, 9.0, gv, samp, NB=1)
        samp[5, 1] = -1.0                                       # a sampling count is never negative
        with pytest.raises(A.AmberError):
            ctx.maps_cldnu(32, 6, 9.0, gv, samp)
        samp[5, 1] = 1.0
        check(ctx, 32, 6, 9.0, gv, samp)                        # the context still works


def test_maps_deterministic():
    gv, samp = rand_cube(64, 40, seed=8)
    with A.Context(device=0) as ctx:
        a = ctx.maps_cldnu(64, 40, 6.0, gv, samp)[0]
        b = ctx.maps_cldnu(64, 40, 6.0, gv, samp)[0]
    assert np.array_equal(a, b)


def test_maps_full_shape_closed_form():
    """768 channels on a 512^2 grid: V = S c(x) e^{i phi n} with every cell
    sampled gives cldnu[b][d] = (nc - d) cos(phi d) <|c|^2>_b"""
    N, nc, Du, phi = 512, 768, 1.0, 0.013
    Nh = N // 2 + 1
    rng = np.random.default_rng(11)
    c = (rng.normal(size=N * Nh) + 1j * rng.normal(size=N * Nh)).astype(np.complex128)
    samp = np.ones((N * Nh, nc))
    samp[::7] = 2.0
    ph = np.exp(1j * phi * np.arange(nc))
    gv = np.empty((N * Nh, nc, 2))
    v = c[:, None] * ph[None, :]
    gv[..., 0] = v.real * samp
    gv[..., 1] = v.imag * samp
    del v
    with A.Context(device=0) as ctx:
        cl, nd, bu, nm = ctx.maps_cldnu(N, nc, Du, gv, samp)
    a2 = np.abs(c) ** 2
    cells = np.array([O.maps_bin(N, Du, x // Nh, x % Nh)[0] for x in range(N * Nh)])
    d = np.arange(nc)
    for b in range(15):
        xs = cells == b
        assert nm[b] == xs.sum()
        if not xs.any():
            assert np.all(cl[b] == 0)
            continue
        mean = a2[xs].sum() / xs.sum()
        want = (nc - d) * np.cos(phi * d) * mean
        assert np.all(np.abs(cl[b] - want) <= 1e-11 * (nc - d) * mean), b
    assert np.array_equal(nd, nc - d.astype(np.float64))


def test_maps_driver(tmp_path):
    N, nc, Du, rlzn = 32, 20, 9.0, 3
    gv, samp = rand_cube(N, nc, seed=21)
    data = tmp_path / "data"
    data.mkdir()
    (tmp_path / "results").mkdir()
    A.write_cube_files(str(data), rlzn, N, nc, Du, 0, gv, samp)
    (tmp_path / "input_params").write_text("%d %f 150.0 0.04 1.0 %d %s\n" % (N, Du, nc, data))
    exe = os.path.join(ROOT, "amber_amd", "bin", "amber_maps")
    r = subprocess.run([exe, str(rlzn), "GV_data_%d" % rlzn], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "results" / "cldnu").read_text().splitlines()
    ocl, ond, obu, onm = O.cldnu(N, nc, Du, gv, samp)
    sc = abs_scale(N, nc, Du, gv, samp)
    assert len(lines) == 15 * nc
    for k, line in enumerate(lines):
        f = line.split("\t")
        i, j = k // nc, k % nc
        assert f[0] == "%d" % i and f[2] == "%d" % j
        assert f[1] == "%e" % obu[i] and f[4] == "%e" % ond[j] and f[5] == "%d" % onm[i]
        assert abs(float(f[3]) - ocl[i, j]) <= 5e-7 * abs(ocl[i, j]) + 1e-12 * sc[i, j] + 1e-300

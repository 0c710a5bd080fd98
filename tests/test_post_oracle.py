"""CPU: the lag-space post-processing oracle (oracle/post.py) pinned by its
definitions, a brute-force transform and closed forms."""
import numpy as np

import oracle.post as P


def test_fold_definition():
    rng = np.random.default_rng(1)
    nc10 = 5
    W = 2 * nc10 + 1
    B = rng.standard_normal((W, W)).astype(np.float32)
    F = P.fold(B, nc10)
    assert F.shape == (nc10 + 1, W)
    for d2 in range(0, nc10 + 1):
        for d3 in range(-nc10, nc10 + 1):
            want = B[d2 + nc10, d3 + nc10]
            if d2 >= 1:
                want = np.float32(want + B[-d2 + nc10, -d3 + nc10])
            assert F[d2, d3 + nc10] == want
    C = rng.integers(1, 9, (W, W)).astype(np.int32)
    assert P.fold(C, nc10).dtype == np.int32


def test_transform_matches_brute_force():
    """Bdnu2Bk_fft (numpy irfft2 along (1, 0) of the column-ifftshifted
    array, then fftshift) equals the explicit cosine sums the GPU evaluates"""
    rng = np.random.default_rng(2)
    for (M, N) in ((5, 9), (6, 11), (4, 7)):
        X = rng.standard_normal((M, N))
        a = P.bdnu2bk(X, 3.0, 1.7, 0.5)
        b = P.bdnu2bk_direct(X, 3.0, 1.7, 0.5)
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))


def test_transform_closed_form():
    """a unit lag at the grid centre transforms to a constant rc^4 L^2 / N^2"""
    M, N = 6, 11
    X = np.zeros((M, N))
    X[0, N // 2] = 1.0
    rc, rcp, dnu = 2.0, 3.0, 0.25
    out = P.bdnu2bk(X, rc, rcp, dnu)
    L = N * dnu * rcp
    assert np.allclose(out, rc ** 4 * L ** 2 / N ** 2, rtol=1e-13, atol=0)


def test_window():
    w = P.blackman_nuttall(21)
    assert abs(w[0] - (0.3635819 - 0.4891775 + 0.1365995 - 0.0106411)) < 1e-15
    assert np.allclose(w, w[::-1], atol=1e-15)
    assert abs(w[10] - (0.3635819 + 0.4891775 + 0.1365995 + 0.0106411)) < 1e-15
    Wd = P.window(2 * 12 + 1, 3)
    N = 2 * 12 + 1 - 6
    assert Wd.shape == (12 + 1 - 3, N)
    wn = P.blackman_nuttall(N)
    assert np.array_equal(Wd, np.outer(wn[N // 2:], wn))


def test_cyl_types_follow_numpy():
    """v1: float32 quotient by mabs_norm, then float64 division by int
    counts; v2: float32 throughout the division"""
    rng = np.random.default_rng(3)
    nc10, n = 6, 2
    W = 2 * nc10 + 1
    bl = rng.standard_normal((W, W)).astype(np.float32)
    cnt = rng.integers(1, 50, (W, W)).astype(np.int32)
    nrm = rng.uniform(1, 50, (W, W)).astype(np.float32)
    F = P.fold(bl, nc10)
    t = F / np.float32(0.37)
    assert t.dtype == np.float32
    v1 = P.cut((t / P.fold(cnt, nc10)).astype(np.float64), n)
    v2 = P.cut((t / P.fold(nrm, nc10)).astype(np.float64), n)
    assert np.array_equal(P.cyl_bispec(bl, cnt, nc10, n, 0.37, 1.0, 1.0, 1.0), P.bdnu2bk(v1, 1.0, 1.0, 1.0))
    assert np.array_equal(P.cyl_bispec(bl, nrm, nc10, n, 0.37, 1.0, 1.0, 1.0), P.bdnu2bk(v2, 1.0, 1.0, 1.0))

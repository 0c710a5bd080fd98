"""Pins the FP64 oracle (oracle/) against things that do not share its code:
direct DFTs, numpy's pocketfft, brute-force closed-triangle sums and counts,
and the reference setup's known figures.  CPU only."""
import numpy as np
import pytest

import oracle as O
import amber_amd as A

RNG = np.random.default_rng(1234)


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 14, 17, 25, 31])
def test_fft_matches_direct_dft(n):
    x = RNG.normal(size=n) + 1j * RNG.normal(size=n)
    for sign in (+1, -1):
        a = O.fft_c2c(x, sign)
        b = O.dft_direct(x, sign)
        assert np.max(np.abs(a - b)) <= 1e-12 * max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("n", [62, 86, 128, 192, 218, 282, 362, 492, 612, 698])
def test_fft_matches_numpy(n):
    x = RNG.normal(size=n) + 1j * RNG.normal(size=n)
    assert np.allclose(O.fft_c2c(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-11 * n)
    assert np.allclose(O.fft_c2c(x, -1), np.fft.fft(x), rtol=0, atol=1e-11 * n)


# ------------------------------------------------------------------ rules
def test_ng_formula_reference_values():
    # amber.c:193-197 at the shipped setup (NN=1176, Du=1.0254384902642049):
    # UHigh=270 -> Nmax=266 -> 798 -> 800 (even + 2)
    du = 1.0254384902642049
    assert O.ng(270.0, du) == 800
    assert A.ng(270.0, du) == 800
    # odd 3*Nmax takes +3: UHigh=200, Du=1 -> 609 -> 612
    assert O.ng(200.0, 1.0) == 612
    for uh in np.linspace(1, 400, 97):
        for du in (0.7, 1.0, 1.0254384902642049, 2.5):
            g = O.ng(uh, du)
            assert g % 2 == 0 and g == A.ng(uh, du)


def test_annular_rings_reference_edges():
    lo, hi = O.annular_rings(22)
    want_lo = [1] + list(range(5, 50, 5)) + list(range(50, 100, 10)) + list(range(100, 200, 20)) + [200, 235]
    assert list(lo) == want_lo
    assert list(hi[:-1]) == want_lo[1:] and hi[-1] == 270
    alo, ahi = A.annular_rings(22)
    assert np.array_equal(lo, alo) and np.array_equal(hi, ahi)


def test_triangle_rule_reference_setup():
    lo, hi = O.annular_rings(22)
    t = O.triangles(lo, hi)
    assert len(t) == 669
    # every triple ordered r1>=r2>=r3
    assert np.all(t[:, 0] >= t[:, 1]) and np.all(t[:, 1] >= t[:, 2])
    # triangles with r1 <= 15 number 365 (rp=365 is the first with r1=16)
    assert int(np.sum(t[:, 0] <= 15)) == 365 and t[365, 0] == 16
    assert np.array_equal(t, A.triangles(lo, hi))
    # equilateral bins always pass (t=1, mu=0.5 on the boundary 0.5/t)
    assert all((r, r, r) in set(map(tuple, t)) for r in range(22))


def test_triangle_rule_geometry():
    # the rule is the cosine law on bin averages: r3 closes a triangle with
    # sides U1>=U2>=U3 and angle between U1 and U2 at most 60 degrees
    lo, hi = A.log_rings(1, 200, 24)
    u = O.uavg(lo, hi)
    t = set(map(tuple, O.triangles(lo, hi)))
    for r1 in range(24):
        for r2 in range(r1 + 1):
            for r3 in range(r2 + 1):
                a, b, c = u[r1], u[r2], u[r3]
                cos_g = (a * a + b * b - c * c) / (2 * a * b)
                inside = b / a >= 0.5 and cos_g >= 0.5 * a / b - 1e-12 and cos_g <= 1.0 + 1e-12
                if abs(cos_g - 0.5 * a / b) > 1e-9 and abs(cos_g - 1.0) > 1e-9 and abs(b / a - 0.5) > 1e-9:
                    assert ((r1, r2, r3) in t) == inside


# ------------------------------------------------------------------ c2r
def _half_plane(gv, N, nc, Ng, lo, hi, plane, samp=None):
    Nh, Ngh = N // 2 + 1, Ng // 2 + 1
    H = np.zeros((Ng, Ngh), complex)
    for ia in range(-Ng // 2 + 1, Ng // 2 + 1):
        for j in range(Ng // 3):
            uu = np.sqrt(float(ia * ia + j * j))
            if lo <= uu < hi and uu > 0:
                src = (ia % N) * Nh + j
                H[ia % Ng, j] = (gv[src, plane, 0] + 1j * gv[src, plane, 1]) if samp is None else samp[src, plane]
    return H


@pytest.mark.parametrize("Ng,lo,hi", [(20, 2.0, 6.0), (26, 1.0, 5.0), (42, 4.0, 11.5), (62, 3.0, 9.0)])
def test_ring_c2r_matches_numpy_irfft2(Ng, lo, hi):
    N, nc = 32, 3
    gv, samp, _ = A.generate(N, nc, seed=7)
    for plane in range(nc):
        H = _half_plane(gv, N, nc, Ng, lo, hi, plane)
        ref = np.fft.irfft2(H, s=(Ng, Ng)) * Ng * Ng
        got = O.ring_c2r(N, nc, Ng, lo, hi, plane, gv=gv)
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1, np.max(np.abs(ref)))
        Hs = _half_plane(gv, N, nc, Ng, lo, hi, plane, samp=samp)
        got_s = O.ring_c2r(N, nc, Ng, lo, hi, plane, samp=samp)
        assert np.allclose(got_s, np.fft.irfft2(Hs, s=(Ng, Ng)) * Ng * Ng, atol=1e-10)


def test_ring_c2r_non_hermitian_dc_column_is_symmetrised():
    # a j=0 column that is not Hermitian enters as its Hermitian-symmetrised
    # version (pocketfft/FFTW c2r behaviour on such input)
    N, nc, Ng = 32, 1, 20
    gv = np.zeros((N * (N // 2 + 1), nc, 2))
    gv[3 * (N // 2 + 1), 0] = (2.0, 1.0)      # (ia=3, j=0), partner (ia=-3, 0) left empty
    got = O.ring_c2r(N, nc, Ng, 1.0, 5.0, 0, gv=gv)
    x = np.arange(Ng)[:, None]
    sym = 0.5 * ((2 + 1j) * np.exp(2j * np.pi * 3 * x / Ng) + (2 - 1j) * np.exp(-2j * np.pi * 3 * x / Ng))
    assert np.allclose(got, np.broadcast_to(sym.real, (Ng, Ng)), atol=1e-12)


def test_ring_c2r_direct_2d_sum():
    # brute-force direct 2-D sum over the Hermitian extension, no FFT anywhere
    N, nc, Ng, lo, hi = 16, 1, 14, 1.0, 4.0
    gv, _, _ = A.generate(N, nc, seed=11)
    H = _half_plane(gv, N, nc, Ng, lo, hi, 0)
    F = _full_hermitian(H, Ng)
    k = np.arange(Ng)
    E = np.exp(2j * np.pi * np.outer(k, k) / Ng)
    ref = (E @ F @ E.T).real
    assert np.allclose(O.ring_c2r(N, nc, Ng, lo, hi, 0, gv=gv), ref, atol=1e-10)


# ------------------------------------------------------------------ triple sums
def _full_hermitian(H, Ng):
    """Full Ng x Ng spectrum whose inverse DFT is the c2r of half plane H."""
    F = np.zeros((Ng, Ng), complex)
    Ngh = Ng // 2 + 1
    for kx in range(Ng):
        for ky in range(1, Ngh):
            if 2 * ky == Ng:
                continue
            F[kx, ky] += H[kx, ky]
            F[(-kx) % Ng, (-ky) % Ng] += np.conj(H[kx, ky])
        F[kx, 0] += 0.5 * (H[kx, 0] + np.conj(H[(-kx) % Ng, 0]))
    return F


def _closed_sum(F1, F2, F3, Ng):
    """sum_{k1+k2+k3 = 0 mod Ng} F1(k1) F2(k2) F3(k3), by enumeration."""
    s = 0.0
    nz1 = np.argwhere(F1 != 0)
    ky = np.arange(Ng)
    for (ax, ay) in nz1:
        # all k2 at once; k3 = -(k1 + k2)
        k3x = (-(ax + ky[:, None])) % Ng
        k3y = (-(ay + ky[None, :])) % Ng
        s += F1[ax, ay] * np.sum(F2 * F3[k3x, k3y])
    return s


def test_bispec_matches_closed_triangle_enumeration():
    N, nc, nc10 = 16, 4, 3
    lo = np.array([1.0, 2.0, 3.0]); hi = np.array([2.0, 3.0, 4.5])
    gv, samp, uf = A.generate(N, nc, seed=21, p0=0.9, u0=100.0)
    mask = np.ones(nc, np.uint8)
    r1, r2, r3 = 2, 1, 1
    Ng = O.ng(hi[r1], 1.0)
    bl, cnt, norm = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10, want_norm=True)
    F = {}
    for r in {r1, r2, r3}:
        for p in range(nc):
            F[r, p, 0] = _full_hermitian(_half_plane(gv, N, nc, Ng, lo[r], hi[r], p), Ng)
            F[r, p, 1] = _full_hermitian(_half_plane(gv, N, nc, Ng, lo[r], hi[r], p, samp=samp), Ng)
    for d2 in range(-nc10, nc10 + 1):
        for d3 in range(-nc10, nc10 + 1):
            want = 0.0
            wantn = 0.0
            n = 0
            for p1 in range(nc):
                p2, p3 = p1 + d2, p1 + d3
                if not (0 <= p2 < nc and 0 <= p3 < nc):
                    continue
                want += Ng * Ng * _closed_sum(F[r1, p1, 0], F[r2, p2, 0], F[r3, p3, 0], Ng)
                wantn += Ng * Ng * _closed_sum(F[r1, p1, 1], F[r2, p2, 1], F[r3, p3, 1], Ng)
                n += 1
            i2, i3 = d2 + nc10, d3 + nc10
            assert cnt[i2, i3] == n
            assert abs(bl[i2, i3] - want.real) <= 1e-9 * (1 + abs(want))
            assert abs(want.imag) <= 1e-9 * (1 + abs(want))
            assert abs(norm[i2, i3] - wantn.real) <= 1e-9 * (1 + abs(wantn))
            # sampling masks are 0/1: the normalisation is Ng^2 times an integer count
            assert abs(norm[i2, i3] / (Ng * Ng) - round(norm[i2, i3] / (Ng * Ng))) < 1e-9


def test_calcr_counts_closed_sampled_triangles():
    N, nc = 24, 6
    lo = np.array([1.0, 2.5, 4.0]); hi = np.array([2.5, 4.0, 6.0])
    _, samp, _ = A.generate(N, nc, seed=5, p0=0.7, u0=50.0)
    rd = O.calcr(N, nc, samp, lo, hi, plane=5)
    tris = O.triangles(lo, hi)
    assert np.array_equal(rd[:, :3].astype(int), tris)
    Nh = N // 2 + 1
    for (r1, r2, r3, num) in rd:
        Fs = []
        for r in (int(r1), int(r2), int(r3)):
            H = np.zeros((N, Nh), complex)
            for i in range(N):
                ia = N - i if i > N // 2 else i
                for j in range(N // 3):
                    uu = np.sqrt(float(ia * ia + j * j))
                    if lo[r] <= uu < hi[r] and uu > 0:
                        H[i, j] = samp[i * Nh + j, 5]
            Fs.append(_full_hermitian(H, N))
        count = _closed_sum(Fs[0], Fs[1], Fs[2], N).real
        assert abs(num / N ** 2 - count) < 1e-9
        assert abs(count * 8 - round(count * 8)) < 1e-9


def test_bispec_channel_constant_and_counts():
    # identical planes in every channel: B[d2][d3] = cnt[d2][d3] * B_single,
    # and cnt is the closed-form triplet count over the unflagged set
    N, nc, nc10 = 16, 7, 4
    gv, samp, _ = A.generate(N, 1, seed=3, p0=0.9, u0=100.0)
    gvc = np.repeat(gv, nc, axis=1)
    sc = np.repeat(samp, nc, axis=1)
    lo = np.array([1.0, 2.5]); hi = np.array([2.5, 4.0])
    uf = np.array([0, 1, 3, 4, 5])
    mask = O.chan_mask(nc, 1, uf)
    bl, cnt, _ = O.bispec(N, nc, 1.0, gvc, sc, lo, hi, 1, 1, 0, mask, nc10)
    b1, c1, _ = O.bispec(N, 1, 1.0, gv, samp, lo, hi, 1, 1, 0, np.ones(1, np.uint8), 0)
    ufs = set(uf.tolist())
    for i2 in range(2 * nc10 + 1):
        for i3 in range(2 * nc10 + 1):
            d2, d3 = i2 - nc10, i3 - nc10
            n = sum(1 for p in range(nc) if p in ufs and (p + d2) in ufs and (p + d3) in ufs)
            assert cnt[i2, i3] == n
            assert abs(bl[i2, i3] - n * b1[0, 0]) <= 1e-9 * (1 + abs(n * b1[0, 0]))


def test_refalgo_agrees_with_oracle():
    # the reference-algorithm FP32 path: same counts, FP32-level values
    N, nc, nc10 = 32, 8, 5
    gv, samp, uf = A.generate(N, nc, nflag=2, seed=9)
    lo, hi = A.log_rings(2, 10, 4)
    mask = O.chan_mask(nc, 1, uf)
    bl64, cnt64, n64 = O.bispec(N, nc, 1.0, gv, samp, lo, hi, 3, 2, 2, mask, nc10, want_norm=True)
    bl, numf, cnt, ntrip, _, _ = O.refalgo(N, nc, 1.0, gv, samp, lo, hi, 3, 2, 2, mask, nc10, v2=True)
    assert np.array_equal(cnt, cnt64)
    assert ntrip == int(cnt64.sum())
    scale = np.sqrt(np.mean(bl64 ** 2))
    assert np.max(np.abs(bl - bl64)) <= 1e-4 * scale
    assert np.max(np.abs(numf - n64)) <= 1e-4 * max(1.0, np.max(np.abs(n64)))


def test_normalize_data_pins():
    """normalize_data (amber_funcs.c:1946-1978): V = S c with integer counts
    S >= 1 gives back c exactly where S is a power of two and to one
    rounding otherwise, with S := 1; cells with S <= 0 keep V and S; the
    operation is idempotent."""
    rng = np.random.default_rng(3)
    cells, nc = 300, 7
    c = rng.standard_normal((cells, nc, 2))
    S = rng.integers(-1, 6, (cells, nc)).astype(np.float64)
    gv = c * np.where(S > 0, S, 1.0)[..., None]
    gv[S <= 0] = rng.standard_normal(((S <= 0).sum(), 2))   # untouched values
    g, s = O.normalize(gv, S)
    pos = S > 0
    assert np.array_equal(s[pos], np.ones(pos.sum())) and np.array_equal(s[~pos], S[~pos])
    assert np.array_equal(g[~pos], gv[~pos])
    assert np.allclose(g[pos], c[pos], rtol=2 ** -52, atol=0)
    p2 = pos & np.isin(S, [1, 2, 4])
    assert np.array_equal(g[p2], c[p2])                     # exact division by powers of two
    g2, s2 = O.normalize(g, s)
    assert np.array_equal(g2, g) and np.array_equal(s2, s)


def test_bispec_cells_equals_bispec():
    """orc_bispec_cells (the checker of the large and full-size GPU tests)
    against orc_bispec on every lag cell of a flagged cube with lag windows,
    data and normalisation: the same triplets and products give the same
    sums (to FP64 summation order) and identical counts."""
    N, nc, nc10 = 32, 9, 5
    lo, hi = A.log_rings(2, 10, 4)
    gv, samp, uf = A.generate(N, nc, nflag=2, seed=19)
    mask = O.chan_mask(nc, 1, uf)
    W = 2 * nc10 + 1
    d = np.arange(-nc10, nc10 + 1)
    d2, d3 = [x.ravel() for x in np.meshgrid(d, d, indexing="ij")]
    for (r1, r2, r3) in A.triangles(lo, hi)[::3]:
        b, c, n = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10, want_norm=True)
        bc, cc, nn, ab = O.bispec_cells(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10, d2, d3,
                                        want_norm=True, want_abs=True)
        assert np.array_equal(cc.reshape(W, W), c)
        assert np.all(np.abs(bc.reshape(W, W) - b) <= 1e-12 * np.maximum(ab.reshape(W, W), 1e-300))
        assert np.allclose(nn.reshape(W, W), n, rtol=1e-12, atol=1e-9)
        assert np.all(ab >= np.abs(bc) * (1 - 1e-12))


def test_uavg_closed_form_and_quadrature():
    """Uavg = (2/3)(H^3 - L^3)/(H^2 - L^2) (amber_funcs.c:2087-2093) is the
    area-weighted mean radius of the annulus [L, H): (2/3) H at L = 0, and
    the midpoint-rule integral of r over the annulus's area elsewhere."""
    lo = np.array([0.0, 1.0, 2.5, 150.0, 10.0])
    hi = np.array([3.0, 2.0, 7.5, 200.0, 10.5])
    u = O.uavg(lo, hi)
    assert u[0] == pytest.approx(2.0, rel=1e-15)
    for L, H, got in zip(lo, hi, u):
        r = L + (np.arange(200000) + 0.5) * (H - L) / 200000
        assert got == pytest.approx(np.sum(r * r) / np.sum(r), rel=1e-9)
    assert np.array_equal(A.uavg(lo, hi), u)

"""The lag-Fourier form (precision "lagft") against the FP64 oracle.

Each lag cell of the reference's sum (amber.c:233-283) is computed here as
the 2-D inverse DFT of S(k2, k3) = sum_x conj F(k2 + k3) G(k2) H(k3) over
the channel spectra, so the tests pin what the direct sum fixes: every cell
against the oracle on small cubes (flagged channels, partial lag windows,
odd grids, r1 == r2 and r2 == r3 triangles), counts bit for bit, cells
without triplets exactly zero, batches equal to single triangles bit for
bit, the v2 normalisation exact, the symmetric tilings (r1 == r2,
r2 == r3) equal to the full one, the serving queue equal to the batch path,
and graph capture (which it does not support) refused."""
import numpy as np
import pytest

import amber_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    # N, nc, nflag, nc10, rings (lo, hi, n), seed
    (64, 24, 3, 23, (2, 20, 6), 5),
    (64, 24, 3, 9, (2, 20, 6), 6),
    (96, 40, 0, 39, (2, 30, 5), 7),
    (128, 64, 6, 63, (2, 60, 8), 1001),
]


@pytest.mark.parametrize("case", CASES)
def test_cells_vs_oracle(ctx, case):
    N, nc, nflag, nc10, rings, seed = case
    lo, hi = A.log_rings(*rings)
    gv, samp, uf = A.generate(N, nc, nflag=nflag, seed=seed)
    fk = 1 if nflag else 0
    mask = O.chan_mask(nc, fk, uf)
    tris = A.triangles(lo, hi)
    pick = [0, len(tris) // 2, len(tris) - 1]
    sym = [i for i, t in enumerate(tris) if t[0] == t[1] and t[1] != t[2]][:1] + \
          [i for i, t in enumerate(tris) if t[1] == t[2] and t[0] != t[1]][:1]
    ctx.set_precision("lagft")
    try:
        ctx.set_rings(lo, hi)
        cube = ctx.load_cube(N, nc, 1.0, fk, gv, samp)
        ctx.set_channels(cube, uf if fk else None, nc10)
        for k in sorted(set(pick + sym)):
            t = tuple(int(x) for x in tris[k])
            print("case lagft N=%d nc=%d nc10=%d triangle %s" % (N, nc, nc10, t), flush=True)
            res = ctx.bispec_ex(cube, *t)
            b64, c64, _ = O.bispec(N, nc, 1.0, gv, samp, lo, hi, *t, mask, nc10)
            assert np.array_equal(res["cnt"], c64)
            assert np.all(res["bl"][c64 == 0] == 0.0) and np.all(res["bl64"][c64 == 0] == 0.0)
            e = np.linalg.norm(res["bl64"] - b64) / np.linalg.norm(b64)
            e32 = np.linalg.norm(res["bl"] - b64) / np.linalg.norm(b64)
            print("  normwise %.3g (f32 out %.3g)" % (e, e32), flush=True)
            assert e <= 2e-6 and e32 <= 2e-6, (t, e, e32)
        ctx.release_cube(cube)
    finally:
        ctx.set_precision("bf16x3")


def test_batch_equals_single_and_v2_norm_exact(ctx):
    N, nc, nc10 = 64, 24, 23
    lo, hi = A.log_rings(2, 20, 6)
    gv, samp, uf = A.generate(N, nc, nflag=3, seed=8)
    mask = O.chan_mask(nc, 1, uf)
    tris = A.triangles(lo, hi)
    ctx.set_precision("lagft")
    try:
        ctx.set_rings(lo, hi)
        cube = ctx.load_cube(N, nc, 1.0, 1, gv, samp)
        ctx.set_channels(cube, uf, nc10)
        bl, cnt, norm = ctx.bispec_batch(cube, tris, want_norm=True)
        for k in (0, len(tris) - 1):
            t = tuple(int(x) for x in tris[k])
            one = ctx.bispec_ex(cube, *t, want_norm=True)
            assert np.array_equal(one["bl"], bl[k]) and np.array_equal(one["cnt"], cnt[k])
            _, _, n64 = O.bispec(N, nc, 1.0, gv, samp, lo, hi, *t, mask, nc10, want_norm=True)
            P = one["ng"] ** 2
            snapped = P * np.rint(n64 / P * 8) / 8
            assert np.array_equal(one["norm64"], snapped)
            assert np.array_equal(norm[k], snapped.astype(np.float32))
        ctx.release_cube(cube)
    finally:
        ctx.set_precision("bf16x3")


def test_refused_in_graphs(ctx):
    N, nc = 32, 8
    lo, hi = A.log_rings(2, 10, 4)
    gv, samp, _ = A.generate(N, nc, seed=2)
    ctx.set_precision("lagft")
    try:
        ctx.set_rings(lo, hi)
        cube = ctx.load_cube(N, nc, 1.0, 0, gv, samp)
        ctx.set_channels(cube, None, nc - 1)
        t = tuple(int(x) for x in A.triangles(lo, hi)[0])
        with pytest.raises(A.AmberError):
            ctx.graph([(cube,) + t])
        ctx.bispec_ex(cube, *t)  # the context stays usable
        ctx.release_cube(cube)
    finally:
        ctx.set_precision("bf16x3")


def test_serving_equals_batch(ctx):
    """The serving queue in the lag-Fourier form: 24 one-triangle requests on
    a 512^2 x 256 cube (transform length 512, r1 == r2 triangles on pair
    tiles) and a small flagged cube, submitted at once so that fair shares
    split the larger triangles' tile items over several iterations; lag
    arrays and counts equal the batch path bit for bit, and the v2
    normalisation (the exact FP64 path) equals the batch path's."""
    lo, hi = A.log_rings(4, 200, 6)
    tris = [tuple(int(x) for x in t) for t in A.triangles(lo, hi)]
    gv, samp, _ = A.generate(512, 256, nflag=0, seed=31)
    gs, ss, us = A.generate(64, 24, nflag=3, seed=32)
    ctx.set_precision("lagft")
    try:
        ctx.set_rings(lo, hi)
        big = ctx.load_cube(512, 256, 1.0, 0, gv, samp)
        ctx.set_channels(big, None, 255)
        small = ctx.load_cube(64, 24, 1.0, 1, gs, ss)
        ctx.set_channels(small, us, 23)
        rng = np.random.default_rng(5)
        pick = [tris[i] for i in rng.choice(len(tris), size=16, replace=False)]
        pick += [t for t in tris if t[0] == t[1]][-4:]
        reqs = [(big,) + t for t in pick]
        # the small cube's N/2 = 32 bounds its rings
        reqs += [(small,) + t for t in tris if max(hi[list(t)]) <= 32][:4]
        reqs = [reqs[i] for i in rng.permutation(len(reqs))]
        print("case lagft serving: %d requests" % len(reqs), flush=True)
        ctx.serve_start(max_active=64)
        try:
            tickets = [ctx.submit([r], want_norm=(k % 5 == 2)) for k, r in enumerate(reqs)]
            got = [ctx.wait(t)[0][0] for t in tickets]
            st = ctx.stats()
        finally:
            ctx.serve_stop()
        print("iterations %d" % st["n_serve_iters"], flush=True)
        assert st["n_serve_iters"] >= 2
        for k, r in enumerate(reqs):
            want = ctx.bispec_multi([r], want_norm=(k % 5 == 2))[0]
            assert np.array_equal(got[k][0], want[0]), (k, r)
            assert np.array_equal(got[k][1], want[1]), (k, r)
            if k % 5 == 2:
                assert np.array_equal(got[k][2], want[2]), (k, r)
        ctx.release_cube(big)
        ctx.release_cube(small)
    finally:
        ctx.set_precision("bf16x3")


@pytest.mark.parametrize("sym", ["r1==r2", "r1==r2 pair tiles", "r2==r3"])
def test_symmetric_tilings(ctx, sym):
    """Triangles at a transform length of 512 (256 channels, all lags) on
    the symmetric tilings: r1 == r2 on the sheared (t, k3) tiles (default)
    or on the (i, j) pair tiles (AMB_NO_SHEAR), r2 == r3 on the (k2, k3)
    tiles without those below the diagonal of [0, L/2)^2.  Sampled cells
    against the FP64 oracle, every cell against the full (k2, k3) tiling of
    a context with the symmetries switched off (AMB_NO_SYM), counts equal."""
    import os
    N, nc, nc10 = 64, 256, 255
    lo, hi = A.log_rings(2, 20, 5)
    gv, samp, _ = A.generate(N, nc, nflag=0, seed=12)
    mask = np.ones(nc, np.uint8)
    if sym.startswith("r1==r2"):
        tris = [tuple(int(x) for x in t) for t in A.triangles(lo, hi) if t[0] == t[1]]
    else:
        tris = [tuple(int(x) for x in t) for t in A.triangles(lo, hi) if t[1] == t[2] and t[0] != t[1]]
    assert tris
    os.environ["AMB_NO_SYM"] = "1"
    try:
        plain = A.Context(device=0, precision="lagft")
    finally:
        del os.environ["AMB_NO_SYM"]
    own = None
    if sym == "r1==r2 pair tiles":
        os.environ["AMB_NO_SHEAR"] = "1"
        try:
            own = ctx = A.Context(device=0, precision="lagft")
        finally:
            del os.environ["AMB_NO_SHEAR"]
    ctx.set_precision("lagft")
    try:
        rng = np.random.default_rng(3)
        for c in (ctx, plain):
            c.set_rings(lo, hi)
        cube = ctx.load_cube(N, nc, 1.0, 0, gv, samp)
        ctx.set_channels(cube, None, nc10)
        cube2 = plain.load_cube(N, nc, 1.0, 0, gv, samp)
        plain.set_channels(cube2, None, nc10)
        for t in tris[-3:]:
            print("case lagft %s tiles triangle %s" % (sym, t), flush=True)
            a = ctx.bispec_ex(cube, *t)
            b = plain.bispec_ex(cube2, *t)
            assert np.array_equal(a["cnt"], b["cnt"])
            e = np.linalg.norm(a["bl64"] - b["bl64"]) / np.linalg.norm(b["bl64"])
            d2 = np.append(rng.integers(-nc10, nc10 + 1, 24), 0)
            d3 = np.append(rng.integers(-nc10, nc10 + 1, 24), 0)
            b64, c64, _ = O.bispec_cells(N, nc, 1.0, gv, samp, lo, hi, *t, mask, nc10, d2, d3)
            got = a["bl64"][d2 + nc10, d3 + nc10]
            eo = np.linalg.norm(got - b64) / np.linalg.norm(b64)
            print("  vs (k2,k3) tiles %.3g, vs oracle %.3g" % (e, eo), flush=True)
            assert np.array_equal(a["cnt"][d2 + nc10, d3 + nc10], c64)
            assert e <= 2e-6 and eo <= 2e-6, (t, e, eo)
        ctx.release_cube(cube)
    finally:
        plain.close()
        if own is not None:
            own.close()
        else:
            ctx.set_precision("bf16x3")


def test_compute_grids(ctx):
    """Triangles summed on their compute grids (amb_compute_grid) against the
    reference grid Ng(r1) and the FP64 oracle (which sums over Ng^2 pixels),
    at Du = 1.0254 with rings to 200 cells on a 512^2 cube: triangles with
    r1 on the last ring take the column cut j < Ng/3 = 198 of the reference
    (amber_funcs.c:582) on a smaller grid, and those whose three rings wrap
    on Ng (3 max|k| >= Ng, amber.c:193-197) stay on Ng.  Counts exact, v2
    normalisation bit for bit on both grids, sums normwise against the
    oracle and against each other."""
    N, nc, nc10, Du = 512, 12, 11, 1.0254
    lo, hi = A.log_rings(1, 200, 10)
    gv, samp, uf = A.generate(N, nc, nflag=0, seed=41)
    mask = O.chan_mask(nc, 0, uf)
    tris = [tuple(int(x) for x in t) for t in A.triangles(lo, hi)]
    last = len(lo) - 1
    pick = [t for t in tris if t[0] == last]
    kinds = [(A.compute_grid(hi[a], hi[b], hi[c], Du), A.ng(hi[a], Du)) for a, b, c in pick]
    assert any(n < g for n, g in kinds) and any(n == g for n, g in kinds)
    assert A.ng(hi[last], Du) // 3 < int(np.ceil(hi[last]))  # the cut applies
    ctx.set_precision("lagft")
    try:
        ctx.set_rings(lo, hi)
        cube = ctx.load_cube(N, nc, Du, 0, gv, samp)
        ctx.set_channels(cube, None, nc10)
        sel = pick[::max(1, len(pick) // 6)] + [pick[-1]]
        for t in sel:
            n, g = A.compute_grid(hi[t[0]], hi[t[1]], hi[t[2]], Du), A.ng(hi[t[0]], Du)
            print("case lagft Du=%g triangle %s grid %d (Ng %d)" % (Du, t, n, g), flush=True)
            ctx.set_grid(True)
            red = ctx.bispec_ex(cube, *t, want_norm=True)
            ctx.set_grid(False)
            ref = ctx.bispec_ex(cube, *t, want_norm=True)
            ctx.set_grid(True)
            b64, c64, n64 = O.bispec(N, nc, Du, gv, samp, lo, hi, *t, mask, nc10, want_norm=True)
            P = g * g
            snapped = P * np.rint(n64 / P * 8) / 8
            for res in (red, ref):
                assert res["ng"] == g
                assert np.array_equal(res["cnt"], c64)
                assert np.array_equal(res["norm64"], snapped)
            e_red = np.linalg.norm(red["bl64"] - b64) / np.linalg.norm(b64)
            e_ref = np.linalg.norm(ref["bl64"] - b64) / np.linalg.norm(b64)
            e_rr = np.linalg.norm(red["bl64"] - ref["bl64"]) / np.linalg.norm(b64)
            print("  normwise: compute grid %.3g, Ng %.3g, between %.3g" % (e_red, e_ref, e_rr), flush=True)
            # fewer pixels carry the same per-pixel FP32 rounding: the error
            # grows about as Ng/n, against the 1e-5 bar
            assert e_red <= 5e-6 and e_ref <= 2e-6 and e_rr <= 6e-6, (t, e_red, e_ref, e_rr)
            if n == g:
                assert np.array_equal(red["bl64"], ref["bl64"])
        ctx.release_cube(cube)
    finally:
        ctx.set_grid(True)
        ctx.set_precision("bf16x3")


def test_v2_norm_on_sheared_fp64_tiles(ctx):
    """The exact v2 normalisation of r1 == r2 triangles at a transform length
    of 512 (256 channels, all lags), where its FP64 sums run on the sheared
    (t, k3) tiles, and of r1 == r2 == r3 ones, which take only the sheared
    tiles of the principal domain and read every other cell from its orbit:
    equal bit for bit to the FP64 oracle's sums snapped to their 1/8 grid of
    Ng^2, and to a context with the symmetries off (AMB_NO_SYM, the full
    FP64 tiling)."""
    import os
    N, nc, nc10 = 64, 256, 255
    lo, hi = A.log_rings(2, 20, 5)
    gv, samp, _ = A.generate(N, nc, nflag=0, seed=19)
    mask = np.ones(nc, np.uint8)
    tris = [tuple(int(x) for x in t) for t in A.triangles(lo, hi) if t[0] == t[1]]
    os.environ["AMB_NO_SYM"] = "1"
    try:
        plain = A.Context(device=0, precision="lagft")
    finally:
        del os.environ["AMB_NO_SYM"]
    ctx.set_precision("lagft")
    try:
        for c in (ctx, plain):
            c.set_rings(lo, hi)
        cube = ctx.load_cube(N, nc, 1.0, 0, gv, samp)
        ctx.set_channels(cube, None, nc10)
        cube2 = plain.load_cube(N, nc, 1.0, 0, gv, samp)
        plain.set_channels(cube2, None, nc10)
        equi = [t for t in tris if t[1] == t[2]]
        assert len(equi) >= 2
        for t in [t for t in tris if t[1] != t[2]][-3:] + equi[-2:]:
            print("case lagft v2 norm on sheared FP64 tiles, triangle %s" % (t,), flush=True)
            a = ctx.bispec_ex(cube, *t, want_norm=True)
            b = plain.bispec_ex(cube2, *t, want_norm=True)
            _, _, n64 = O.bispec(N, nc, 1.0, gv, samp, lo, hi, *t, mask, nc10, want_norm=True)
            P = a["ng"] ** 2
            snapped = P * np.rint(n64 / P * 8) / 8
            assert np.array_equal(a["norm64"], snapped), t
            assert np.array_equal(a["norm64"], b["norm64"]), t
        ctx.release_cube(cube)
    finally:
        plain.close()
        ctx.set_precision("bf16x3")

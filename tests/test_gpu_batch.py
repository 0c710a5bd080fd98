"""GPU: grouped waves, multi-cube batches, captured graphs and the serving
queue through the C ABI.  Every path runs each triangle's arithmetic
unchanged, so results must equal the per-triangle path bit for bit; the
per-triangle path itself is checked against the FP64 oracle in
test_gpu_parity.py and here on a sample."""
import numpy as np
import pytest

import amber_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def _cube(ctx, N, nc, nflag, seed, nc10):
    gv, samp, uf = A.generate(N, nc, nflag=nflag, seed=seed)
    fk = 1 if nflag > 0 else 0
    cube = ctx.load_cube(N, nc, 1.0, fk, gv, samp)
    ctx.set_channels(cube, uf if fk else None, nc10)
    return cube, gv, samp, uf, fk


def _alone(ctx, reqs, want_norm=False):
    out = []
    for (cube, r1, r2, r3) in reqs:
        r = ctx.bispec_ex(cube, int(r1), int(r2), int(r3), want_norm=want_norm)
        out.append((r["bl"], r["cnt"], r.get("norm")))
    return out


def _same(a, b, want_norm=False):
    for (bl1, c1, n1), (bl2, c2, n2) in zip(a, b):
        assert np.array_equal(bl1, bl2)
        assert np.array_equal(c1, c2)
        if want_norm:
            assert np.array_equal(n1, n2)


@pytest.mark.parametrize("prec", ["bf16x3", "fp32", "fp8"])
def test_grouped_equals_per_triangle(ctx, prec):
    ctx.set_precision(prec)
    lo, hi = A.log_rings(2, 20, 6)
    ctx.set_rings(lo, hi)
    cube, gv, samp, uf, fk = _cube(ctx, 64, 24, 2, 5, 23)
    tris = A.triangles(lo, hi)
    print("case grouped vs per-triangle prec=%s ntri=%d" % (prec, len(tris)), flush=True)
    ctx.set_batch_mode("per_triangle")
    b1, c1, _ = ctx.bispec_batch(cube, tris)
    ctx.set_batch_mode("grouped")
    b2, c2, _ = ctx.bispec_batch(cube, tris)
    assert np.array_equal(b1, b2) and np.array_equal(c1, c2)
    ctx.set_precision("bf16x3")
    ctx.release_cube(cube)


def test_pipelined_waves_equal_per_triangle(monkeypatch):
    """Three triangles per grouped wave: each wave's results are copied out
    of one pinned staging buffer while the next wave runs into the other;
    data, counts and v2 normalisation must equal the per-triangle path."""
    monkeypatch.setenv("AMB_WAVE_TRIS", "3")
    c = A.Context(device=0, precision="bf16x3")
    monkeypatch.delenv("AMB_WAVE_TRIS")
    try:
        lo, hi = A.log_rings(2, 20, 6)
        c.set_rings(lo, hi)
        cube, gv, samp, uf, fk = _cube(c, 64, 24, 2, 5, 23)
        tris = A.triangles(lo, hi)
        print("case pipelined waves ntri=%d" % len(tris), flush=True)
        c.set_batch_mode("per_triangle")
        b1, c1, n1 = c.bispec_batch(cube, tris, want_norm=True)
        c.set_batch_mode("grouped")
        c.reset_stats()
        b2, c2, n2 = c.bispec_batch(cube, tris, want_norm=True)
        assert len(tris) > 6 and c.stats()["n_contract_launch"] == (len(tris) + 2) // 3
        assert np.array_equal(b1, b2) and np.array_equal(c1, c2) and np.array_equal(n1, n2)
    finally:
        c.close()


def test_multi_cube_wave_with_norm(ctx):
    """Two cubes of different shapes (and windows) in one grouped call, v2
    normalisation included; against single calls and, for a sample, the
    FP64 oracle."""
    ctx.set_precision("bf16x3")
    lo, hi = A.log_rings(2, 14, 5)
    ctx.set_rings(lo, hi)
    ca, gva, sa, ufa, fka = _cube(ctx, 32, 10, 1, 11, 9)
    cb, gvb, sb, ufb, fkb = _cube(ctx, 48, 14, 0, 12, 5)
    tris = A.triangles(lo, hi)
    reqs = [(ca,) + tuple(t) for t in tris] + [(cb,) + tuple(t) for t in tris[::-1]]
    print("case multi-cube ntri=%d" % len(reqs), flush=True)
    got = ctx.bispec_multi(reqs, want_norm=True)
    _same(got, _alone(ctx, reqs, want_norm=True), want_norm=True)
    # oracle on one triangle of each cube
    for (cube, gv, samp, uf, fk, N, nc, nc10), k in (((ca, gva, sa, ufa, fka, 32, 10, 9), 0),
                                                     ((cb, gvb, sb, ufb, fkb, 48, 14, 5), len(tris))):
        r1, r2, r3 = reqs[k][1:]
        mask = O.chan_mask(nc, fk, uf)
        b64, c64, _ = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10)
        assert np.array_equal(got[k][1], c64)
        if np.linalg.norm(b64) > 0:
            assert np.linalg.norm(got[k][0] - b64) / np.linalg.norm(b64) <= 1e-5
    ctx.release_cube(ca)
    ctx.release_cube(cb)


def test_graph_replays_follow_cube_updates(ctx):
    """A captured batch replayed three times, each after a new realisation
    is uploaded into the same cubes and eager work runs in between; every
    replay must equal the eager grouped result on the same data."""
    ctx.set_precision("bf16x3")
    lo, hi = A.log_rings(2, 14, 5)
    ctx.set_rings(lo, hi)
    ca, *_ = _cube(ctx, 32, 10, 1, 21, 9)
    cb, *_ = _cube(ctx, 48, 12, 0, 22, 11)
    tris = A.triangles(lo, hi)
    reqs = [(ca,) + tuple(t) for t in tris[:6]] + [(cb,) + tuple(t) for t in tris[-6:]]
    g = ctx.graph(reqs, want_norm=True)
    try:
        for rep in range(3):
            print("case graph replay %d" % rep, flush=True)
            gva, sa, _ = A.generate(32, 10, nflag=1, seed=100 + rep)
            gvb, sb, _ = A.generate(48, 12, nflag=0, seed=200 + rep)
            ctx.update_cube(ca, gva, sa)
            ctx.update_cube(cb, gvb, sb)
            out = g.run()
            ref = ctx.bispec_multi(reqs, want_norm=True)      # eager, between replays
            _same(out, ref, want_norm=True)
        with pytest.raises(A.AmberError):
            ctx.release_cube(ca)                              # the graph reads it
    finally:
        g.close()
    ctx.release_cube(ca)
    ctx.release_cube(cb)


def test_serving_matches_alone(ctx):
    """Requests submitted back to back are batched into shared waves; each
    result equals the same triangles computed alone."""
    ctx.set_precision("bf16x3")
    lo, hi = A.log_rings(2, 14, 5)
    ctx.set_rings(lo, hi)
    ca, *_ = _cube(ctx, 32, 10, 1, 31, 9)
    cb, *_ = _cube(ctx, 48, 12, 0, 32, 11)
    tris = A.triangles(lo, hi)
    rng = np.random.default_rng(3)
    requests = []
    for k in range(12):
        cube = ca if k % 2 else cb
        pick = rng.choice(len(tris), size=1 + k % 4, replace=False)
        requests.append([(cube,) + tuple(tris[i]) for i in pick])
    ctx.serve_start(max_active=8)
    try:
        tickets = [ctx.submit(r, want_norm=(k % 3 == 0)) for k, r in enumerate(requests)]
        results = [ctx.wait(t) for t in tickets]
    finally:
        ctx.serve_stop()
    for k, (r, (outs, lat)) in enumerate(zip(requests, results)):
        assert lat >= 0
        _same(outs, _alone(ctx, r, want_norm=(k % 3 == 0)), want_norm=(k % 3 == 0))
    ctx.release_cube(ca)
    ctx.release_cube(cb)

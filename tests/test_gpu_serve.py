"""GPU: the serving queue's continuous batching.  Requests that carry
their cubes are uploaded while other requests run; a budget far below the
working set forces filtered planes and idle uploaded cubes out (LRU)
mid-run; every request's results still equal the same triangles computed
alone on another context, bit for bit.  A small request submitted while a
large one runs finishes first (it is admitted into the running iterations
instead of waiting for the large one to end)."""
import time

import numpy as np
import pytest

import amber_amd as A

pytestmark = pytest.mark.gpu


def _same(a, b, want_norm):
    for (bl1, c1, n1), (bl2, c2, n2) in zip(a, b):
        assert np.array_equal(bl1, bl2)
        assert np.array_equal(c1, c2)
        if want_norm:
            assert np.array_equal(n1, n2)


def test_uploads_and_evictions_match_alone():
    # six carried cubes of 38 MB each cannot all stay beside the pools
    # (120 MiB of the 320 MiB budget) and the planes
    lo, hi = A.log_rings(2, 30, 6)
    tris = [tuple(int(x) for x in t) for t in A.triangles(lo, hi)]
    rng = np.random.default_rng(41)
    base = A.generate(128, 32, nflag=3, seed=400)
    carried = [A.generate(256, 96, nflag=(4 if k % 2 else 0), seed=410 + k) for k in range(6)]
    budget = 320 << 20
    srv = A.Context(device=0, precision="bf16x3", hbm_budget=budget)
    srv.set_rings(lo, hi)
    cb = srv.load_cube(128, 32, 1.0, 1, base[0], base[1])
    srv.set_channels(cb, base[2], 31)
    plan = []  # (kind, data index or cube, triangles, want_norm)
    for k in range(18):
        pick = [tris[i] for i in rng.choice(len(tris), size=int(rng.integers(1, 5)), replace=False)]
        plan.append(("up" if k % 3 == 0 else "res", k // 3, pick, k % 4 == 1))
    print("case serving uploads: %d requests, budget %d MiB" % (len(plan), budget >> 20), flush=True)
    srv.serve_start(max_active=16)
    tickets = []
    try:
        for kind, u, pick, wn in plan:
            if kind == "up":
                gv, samp, uf = carried[u]
                fk = 1 if u % 2 else 0
                t, cid = srv.submit_cube(256, 96, 1.0, fk, gv, samp, pick, uf=uf if fk else None, nc10=95,
                                         want_norm=wn)
            else:
                t = srv.submit([(cb,) + p for p in pick], want_norm=wn)
            tickets.append(t)
        results = [srv.wait(t)[0] for t in tickets]
        st = srv.stats()
    finally:
        srv.serve_stop()
        srv.close()
    print("uploads %d, evicted cubes %d, evicted plane sets %d, iterations %d" %
          (st["n_cube_uploads"], st["n_evict_cubes"], st["n_evict_planes"], st["n_serve_iters"]), flush=True)
    assert st["n_cube_uploads"] == 6
    assert st["n_evict_cubes"] > 0 and st["n_evict_planes"] > 0
    ref = A.Context(device=0, precision="bf16x3")
    try:
        ref.set_rings(lo, hi)
        rb = ref.load_cube(128, 32, 1.0, 1, base[0], base[1])
        ref.set_channels(rb, base[2], 31)
        for (kind, u, pick, wn), got in zip(plan, results):
            if kind == "up":
                gv, samp, uf = carried[u]
                fk = 1 if u % 2 else 0
                c = ref.load_cube(256, 96, 1.0, fk, gv, samp)
                ref.set_channels(c, uf if fk else None, 95)
                want = ref.bispec_multi([(c,) + p for p in pick], want_norm=wn)
                ref.release_cube(c)
            else:
                want = ref.bispec_multi([(rb,) + p for p in pick], want_norm=wn)
            _same(got, want, wn)
    finally:
        ref.close()


def test_small_request_overtakes_a_running_large_one(ctx):
    lo, hi = A.log_rings(4, 80, 6)
    tris = [tuple(int(x) for x in t) for t in A.triangles(lo, hi)]
    ctx.set_precision("bf16x3")
    ctx.set_rings(lo, hi)
    gv, samp, _ = A.generate(256, 768, nflag=0, seed=420)
    big = ctx.load_cube(256, 768, 1.0, 0, gv, samp)
    ctx.set_channels(big, None, 767)
    gs, ss, _ = A.generate(64, 16, nflag=0, seed=421)
    small = ctx.load_cube(64, 16, 1.0, 0, gs, ss)
    ctx.set_channels(small, None, 15)
    large_req = [(big,) + t for t in tris[-12:]]
    small_req = [(small,) + tris[0]]
    print("case overtaking: %d large triangles on 256^2 x 768" % len(large_req), flush=True)
    ctx.serve_start(max_active=64)
    try:
        tl = ctx.submit(large_req)
        time.sleep(0.05)
        ts = ctx.submit(small_req)
        outs_s, lat_s = ctx.wait(ts)
        large_done_first = ctx.poll(tl)
        outs_l, lat_l = ctx.wait(tl)
    finally:
        ctx.serve_stop()
    print("small %.1f ms, large %.1f ms" % (lat_s, lat_l), flush=True)
    assert not large_done_first and lat_s < lat_l
    alone = ctx.bispec_multi(small_req) + ctx.bispec_multi(large_req)
    _same(outs_s + outs_l, alone, False)
    ctx.release_cube(big)
    ctx.release_cube(small)


def test_submission_errors(ctx):
    lo, hi = A.log_rings(2, 14, 4)
    ctx.set_rings(lo, hi)
    gv, samp, _ = A.generate(32, 8, nflag=0, seed=430)
    c = ctx.load_cube(32, 8, 1.0, 0, gv, samp)
    ctx.set_channels(c, None, 7)
    ctx.serve_start(max_active=4)
    try:
        ctx._meta[c + 1000] = ctx._meta[c]              # output sizes for a cube the library does not know
        with pytest.raises(A.AmberError) as e:
            ctx.submit([(c + 1000, 0, 0, 0)])
        assert e.value.code == -7
        del ctx._meta[c + 1000]
        with pytest.raises(A.AmberError) as e:
            ctx.set_precision("fp32")                   # fixed while serving
        assert e.value.code == -7
        huge = A.generate(32, 8, nflag=0, seed=431)
        with pytest.raises(A.AmberError) as e:          # a carried cube beyond the whole budget
            ctx.submit_cube(1 << 16, 4096, 1.0, 0, huge[0], huge[1], [(0, 0, 0)])
        assert e.value.code in (-1, -6)
    finally:
        ctx.serve_stop()
        ctx.release_cube(c)

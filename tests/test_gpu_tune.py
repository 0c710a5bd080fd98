"""GPU: the autotuner of the grouped contraction's launch configuration.
A reduced mixed batch (three cubes of different grid and channel counts,
every closed triangle of five rings on each) is tuned, then run through a
captured graph under the tuned table and through per-triangle launches of
the plain form: the results must agree bit for bit, since no launch
configuration changes any triangle's arithmetic.  The table round-trips
through its file, and a table from another device is refused."""
import numpy as np
import pytest

import amber_amd as A

pytestmark = pytest.mark.gpu

SHAPES = [(64, 24), (128, 40), (96, 72)]


@pytest.fixture(scope="module")
def mix():
    lo, hi = A.log_rings(2, 20, 5)
    tris = [tuple(int(x) for x in t) for t in A.triangles(lo, hi)]
    data = [A.generate(N, nc, nflag=0, seed=700 + k) for k, (N, nc) in enumerate(SHAPES)]
    return lo, hi, tris, data


def _load(ctx, mix):
    lo, hi, tris, data = mix
    ctx.set_rings(lo, hi)
    cubes = []
    for (N, nc), (gv, samp, _) in zip(SHAPES, data):
        c = ctx.load_cube(N, nc, 1.0, 0, gv, samp)
        ctx.set_channels(c, None, nc - 1)
        cubes.append(c)
    reqs = [(c, *t) for c in cubes for t in tris]
    return cubes, reqs


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_tuned_graph_matches_per_triangle(ctx, mix, prec, tmp_path):
    ctx.set_precision(prec)
    cubes, reqs = _load(ctx, mix)
    path = str(tmp_path / "tune.txt")
    print("case autotune %s on %d triangles over %d cubes" % (prec, len(reqs), len(cubes)), flush=True)
    try:
        table, tuned_out = ctx.autotune(reqs, max_seconds=60, path=path)
        print("tuned table:", table, flush=True)
        assert len(table) >= 1 and all(e["prec"] == prec for e in table)
        assert all(e["tflops"] > 0 for e in table)
        # every bucket of the mix is tuned: (Ng, nact) of each request finds its own entry
        keys = {(e["ng_log2"], e["nact_log2"]) for e in table}
        nact = {c: nc for c, (_, nc) in zip(cubes, SHAPES)}  # no flags, all lags: every channel active
        for c, r1, _, _ in reqs:
            e = ctx.tune_lookup(A.ng(mix[1][r1]), nact[c])
            assert (e["ng_log2"], e["nact_log2"]) in keys
        # round trip through the file
        ctx.tune_clear()
        assert ctx.tune_table() == []
        ctx.tune_load(path)
        assert ctx.tune_table() == table
        # the tuned graph against per-triangle launches
        g = ctx.graph(reqs)
        gout = g.run()
        gout2 = g.run()
        g.close()
        ctx.set_batch_mode("per_triangle")
        per = ctx.bispec_multi(reqs)
        ctx.set_batch_mode("grouped")
        grp = ctx.bispec_multi(reqs)
        for i, (a, b, c2, d) in enumerate(zip(per, gout, gout2, grp)):
            for k in (0, 1):
                assert np.array_equal(a[k], b[k]), (reqs[i], k)
                assert np.array_equal(a[k], c2[k]), (reqs[i], k)
                assert np.array_equal(a[k], d[k]), (reqs[i], k)
                assert np.array_equal(a[k], tuned_out[i][k]), (reqs[i], k)
    finally:
        ctx.set_batch_mode("grouped")
        ctx.tune_clear()
        for c in cubes:
            ctx.release_cube(c)
        ctx.set_precision("bf16x3")


def test_table_from_another_device_is_refused(ctx, tmp_path):
    good = tmp_path / "t.txt"
    ctx.tune_save(str(good))
    head, *rest = good.read_text().splitlines()
    parts = head.split()
    assert parts[0] == "amber_mi355x-tune" and parts[1] == "1"
    bad = tmp_path / "bad.txt"
    bad.write_text(" ".join(parts[:2] + ["gfx000", parts[3]]) + "\n" + "\n".join(rest) + "\n")
    with pytest.raises(A.AmberError) as e:
        ctx.tune_load(str(bad))
    assert e.value.code == -3
    bad.write_text(head + "\nbf16x3 9 7 5 5 16 100 0.0\n")  # variant 5 does not exist
    with pytest.raises(A.AmberError) as e:
        ctx.tune_load(str(bad))
    assert e.value.code == -3
    with pytest.raises(A.AmberError):
        ctx.set_launch_config(variant=0, xrun_log2=13)

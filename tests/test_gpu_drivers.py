"""Drop-in drivers end to end: `amber` and `amber_v2` run in a working
directory laid out like the reference's 4_fbe/ (input_params_ess,
results/v1, results/v2) over files written in the reference formats, and
their result files are checked with the reference readers' byte layouts
(mabs_funcs.py:94-205) against the FP64 oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import amber_amd as A
import oracle as O

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(A.LIB_PATH), "bin")


def _layout(tmp, N, nc, nflag, seed, lo, hi):
    data = tmp / "data"
    data.mkdir()
    for d in ("results/v1", "results/v2"):
        (tmp / d).mkdir(parents=True)
    gv, samp, uf = A.generate(N, nc, nflag=nflag, seed=seed)
    A.write_cube_files(str(data), 1, N, nc, 1.0, 1 if nflag else 0, gv, samp)
    A.write_uf_chans(str(data / "uf_chans.txt"), uf)
    (tmp / "input_params_ess").write_text("154.255e6\t40.00e03\t385\t%s\n" % data)
    with open(tmp / "rings.txt", "w") as f:
        for a, b in zip(lo, hi):
            f.write("%.17g %.17g\n" % (a, b))
    return gv, samp, uf


def _run(tmp, prog, *args):
    r = subprocess.run([os.path.join(BIN, prog), "1", "--rings", "rings.txt"] + list(args), cwd=tmp,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    return r.stdout


@pytest.mark.parametrize("prec", ["bf16x3", "lagft"])
def test_amber_v1_and_v2_drivers(tmp_path, prec):
    N, nc, nflag, nc10 = 64, 10, 2, 6
    lo, hi = A.log_rings(2, 20, 5)
    gv, samp, uf = _layout(tmp_path, N, nc, nflag, 31, lo, hi)
    print("case drivers N=%d nc=%d precision %s" % (N, nc, prec), flush=True)
    _run(tmp_path, "amber", "--nc10", str(nc10), "--uf-nc", str(len(uf)), "--precision", prec)
    # closed_rings / rcount exactly as calCR would write them (exact integers * N^2)
    n, rd = A.read_rings(str(tmp_path / "results/v1"))
    ref = O.calcr(N, nc, samp, lo, hi, plane=5)
    assert n == len(ref) and np.array_equal(rd[:, :3], ref[:, :3])
    assert np.array_equal(rd[:, 3], np.rint(ref[:, 3] / N ** 2 * 8) / 8 * N ** 2)
    mask = O.chan_mask(nc, 1, uf)
    W = 2 * nc10 + 1
    for (r1, r2, r3, num) in rd[::3]:
        r1, r2, r3 = int(r1), int(r2), int(r3)
        f = tmp_path / "results/v1" / ("bsout_final_ra%d_rb%d_rc%d_R1" % (r1, r2, r3))
        with open(f, "rb") as fh:
            ntri = np.fromfile(fh, np.float64, 1)[0]
            Num = np.fromfile(fh, np.float64, 1)[0]
            bl = np.fromfile(fh, np.float32, W * W).reshape(W, W)
            cnt = np.fromfile(fh, np.int32, W * W).reshape(W, W)
        Ng = A.ng(hi[r1])
        gfact = (1.0 * Ng) / (1.0 * N)
        assert Num == gfact * gfact * num            # amber.c:198-200
        assert ntri == Num * (1.0 / (Ng * Ng * 1.0))  # amber.c:201,287
        b64, c64, _ = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10)
        assert np.array_equal(cnt, c64)
        assert np.linalg.norm(bl - b64) <= 1e-5 * np.linalg.norm(b64)
    # amber_v2 reads results/v2/{rcount,closed_rings}, which it never writes
    for fn in ("rcount", "closed_rings"):
        shutil.copy(tmp_path / "results/v1" / fn, tmp_path / "results/v2" / fn)
    _run(tmp_path, "amber_v2", "--nc10", str(nc10), "--uf-nc", str(len(uf)), "--precision", prec)
    for (r1, r2, r3, num) in rd[::4]:
        r1, r2, r3 = int(r1), int(r2), int(r3)
        f = tmp_path / "results/v2" / ("bsout_final2_ra%d_rb%d_rc%d_R1" % (r1, r2, r3))
        with open(f, "rb") as fh:
            inv = np.fromfile(fh, np.float64, 1)[0]
            bl = np.fromfile(fh, np.float32, W * W).reshape(W, W)
            norm = np.fromfile(fh, np.float32, W * W).reshape(W, W)
        Ng = A.ng(hi[r1])
        assert inv == 1.0 / (Ng * Ng * 1.0)
        b64, c64, n64 = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10, want_norm=True)
        P = Ng * Ng
        assert np.array_equal(norm, (P * np.rint(n64 / P * 8) / 8).astype(np.float32))
        assert np.linalg.norm(bl - b64) <= 1e-5 * np.linalg.norm(b64)


def test_driver_errors_are_reported(tmp_path):
    # missing cube file: the driver exits non-zero with a message, no crash
    (tmp_path / "results/v1").mkdir(parents=True)
    (tmp_path / "input_params_ess").write_text("1 1 1 %s\n" % (tmp_path / "nowhere"))
    r = subprocess.run([os.path.join(BIN, "amber"), "1"], cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode != 0 and "GV_data_1" in r.stdout


def test_hardened_wire_matches_legacy(tmp_path):
    """--wire both: each bsw file carries the legacy payload unchanged plus a
    header naming the triangle, nc10, Ng, nc and the precision"""
    N, nc, nflag, nc10 = 64, 10, 2, 6
    lo, hi = A.log_rings(2, 20, 5)
    gv, samp, uf = _layout(tmp_path, N, nc, nflag, 32, lo, hi)
    print("case hardened wire N=%d nc=%d" % (N, nc), flush=True)
    _run(tmp_path, "amber", "--nc10", str(nc10), "--uf-nc", str(len(uf)), "--wire", "both", "--precision", "fp32")
    for fn in ("rcount", "closed_rings"):
        shutil.copy(tmp_path / "results/v1" / fn, tmp_path / "results/v2" / fn)
    _run(tmp_path, "amber_v2", "--nc10", str(nc10), "--uf-nc", str(len(uf)), "--wire", "both")
    n, rd = A.read_rings(str(tmp_path / "results/v1"))
    W = 2 * nc10 + 1
    for (r1, r2, r3, _) in rd:
        r1, r2, r3 = int(r1), int(r2), int(r3)
        # amber_v2 ran with the drivers' default form, the lag-Fourier one
        for kind, pre, d, prec in ((1, "bsout_final", "v1", A.PREC_FP32), (2, "bsout_final2", "v2", A.PREC_LAGFT)):
            legacy = A.read_bsout(str(tmp_path / "results" / d / ("%s_ra%d_rb%d_rc%d_R1" % (pre, r1, r2, r3))))
            h, bl, den = A.read_bsw(str(tmp_path / "results" / d / ("bsw%d_ra%d_rb%d_rc%d_R1" % (kind, r1, r2, r3))))
            assert legacy[0] == kind and h["kind"] == kind
            assert (h["r1"], h["r2"], h["r3"], h["rlzn"], h["nc10"], h["W"], h["nc"]) == (r1, r2, r3, 1, nc10, W, nc)
            assert h["ng"] == A.ng(hi[r1]) and h["precision"] == prec
            assert (h["h0"], h["h1"]) == (legacy[1], legacy[2])
            assert np.array_equal(bl, legacy[3]) and np.array_equal(den, legacy[4])


def test_v2_certificate_modes(tmp_path):
    """amber_v2 on a cube whose fast normalisation sums cannot be certified
    (N = 256, Ng up to 192): by default the run recomputes them in FP64 and
    writes exact normalisations (checked against the oracle's snapped
    sums); --cert strict makes it an error exit.  Run with the split form:
    the certificate belongs to the MFMA forms."""
    N, nc, nflag, nc10 = 256, 40, 0, 12
    lo, hi = A.log_rings(20, 60, 3)
    gv, samp, _ = _layout(tmp_path, N, nc, nflag, 61, lo, hi)
    print("case v2 certificate modes N=%d nc=%d" % (N, nc), flush=True)
    _run(tmp_path, "amber", "--nc10", str(nc10), "--uf-nc", str(nc), "--precision", "bf16x3")
    for fn in ("rcount", "closed_rings"):
        shutil.copy(tmp_path / "results/v1" / fn, tmp_path / "results/v2" / fn)
    n, rd = A.read_rings(str(tmp_path / "results/v1"))
    assert n > 0 and max(A.ng(hi[int(r1)]) for r1 in rd[:, 0]) >= 134
    out = _run(tmp_path, "amber_v2", "--nc10", str(nc10), "--uf-nc", str(nc), "--precision", "bf16x3")
    assert "recomputed in FP64" in out, out[-2000:]
    mask = np.ones(nc, np.uint8)
    W = 2 * nc10 + 1
    for r1, r2, r3 in rd[:, :3].astype(int)[-2:]:
        raw = (tmp_path / "results/v2" / ("bsout_final2_ra%d_rb%d_rc%d_R1" % (r1, r2, r3))).read_bytes()
        norm = np.frombuffer(raw, np.float32, W * W, 8 + 4 * W * W).reshape(W, W)
        _, _, n64 = O.bispec(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10, want_norm=True)
        P = A.ng(hi[r1]) ** 2
        assert np.array_equal(norm, (P * np.rint(n64 / P * 8) / 8).astype(np.float32)), (r1, r2, r3)
    r = subprocess.run([os.path.join(BIN, "amber_v2"), "1", "--rings", "rings.txt", "--nc10", str(nc10), "--uf-nc",
                        str(nc), "--precision", "bf16x3", "--cert", "strict"], cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode != 0 and "1/8 grid" in r.stdout, r.stdout[-2000:]


def test_tune_option_tunes_then_loads(tmp_path):
    """--tune FILE: the first run tunes on a sample of its triangles and saves
    the table, the second loads it; both write the same result files as a
    run without a table (no launch configuration changes a result)."""
    N, nc, nflag, nc10 = 64, 10, 2, 6
    lo, hi = A.log_rings(2, 20, 5)
    _layout(tmp_path, N, nc, nflag, 33, lo, hi)
    print("case driver --tune N=%d nc=%d" % (N, nc), flush=True)
    _run(tmp_path, "amber", "--nc10", str(nc10), "--uf-nc", str(nc - nflag), "--precision", "bf16x3")
    n, rd = A.read_rings(str(tmp_path / "results/v1"))
    names = ["bsout_final_ra%d_rb%d_rc%d_R1" % tuple(int(x) for x in r[:3]) for r in rd]
    plain = {f: (tmp_path / "results/v1" / f).read_bytes() for f in names}
    table = tmp_path / "tune.txt"
    out = _run(tmp_path, "amber", "--nc10", str(nc10), "--uf-nc", str(nc - nflag), "--precision", "bf16x3", "--tune", str(table))
    assert "tuned on" in out and table.exists()
    assert table.read_text().startswith("amber_mi355x-tune 1 ")
    assert all((tmp_path / "results/v1" / f).read_bytes() == plain[f] for f in names)
    out = _run(tmp_path, "amber", "--nc10", str(nc10), "--uf-nc", str(nc - nflag), "--precision", "bf16x3", "--tune", str(table))
    assert "loaded from" in out
    assert all((tmp_path / "results/v1" / f).read_bytes() == plain[f] for f in names)

"""Host-side checks of libamber_mi355x.so (no GPU compute): every symbol the
header declares is exported, file formats round-trip with the reference's
byte layouts, rules agree with the oracle, the generator's invariants hold,
and error paths return codes instead of exiting."""
import os
import re
import struct

import numpy as np
import pytest

import amber_amd as A
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "amber_mi355x.h")).read()
    declared = set(re.findall(r"\b(amb_[a-z0-9_]+)\s*\(", hdr))
    assert len(declared) > 25
    lib = A._lib
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(A.EXPORTS) <= declared


def test_version_and_strerror():
    assert "gfx950" in A.version()
    assert A._lib.amb_strerror(-3).decode().startswith("truncated")


def test_cube_files_roundtrip_and_header_layout(tmp_path):
    N, nc = 16, 5
    gv, samp, uf = A.generate(N, nc, nflag=1, seed=2)
    A.write_cube_files(str(tmp_path), 3, N, nc, 1.25, 1, gv, samp)
    raw = open(tmp_path / "GV_data_3", "rb").read()
    # int64 NN, int64 nc, f64 Du, int32 flag_key: 28 bytes, then re/im pairs
    assert struct.unpack("<qqdi", raw[:28]) == (N, nc, 1.25, 1)
    assert len(raw) == 28 + N * (N // 2 + 1) * nc * 16
    assert np.array_equal(np.frombuffer(raw[28:], np.float64), gv.ravel())
    NN, ncc, Du, fk, g2, s2 = A.read_cube_files(str(tmp_path), 3)
    assert (NN, ncc, Du, fk) == (N, nc, 1.25, 1)
    assert np.array_equal(g2, gv) and np.array_equal(s2, samp)


def test_truncated_cube_is_an_error(tmp_path):
    N, nc = 16, 2
    gv, samp, _ = A.generate(N, nc, seed=2)
    A.write_cube_files(str(tmp_path), 1, N, nc, 1.0, 0, gv, samp)
    p = tmp_path / "GV_data_1"
    data = p.read_bytes()
    p.write_bytes(data[:-8])
    with pytest.raises(A.AmberError) as e:
        A.read_cube_files(str(tmp_path), 1)
    assert e.value.code == -3
    (tmp_path / "sampling").unlink()
    A.write_cube_files(str(tmp_path), 1, N, nc, 1.0, 0, gv, samp)
    (tmp_path / "sampling").unlink()
    with pytest.raises(A.AmberError) as e:
        A.read_cube_files(str(tmp_path), 1)
    assert e.value.code == -2


def test_uf_chans_short_and_out_of_range(tmp_path):
    p = str(tmp_path / "uf_chans.txt")
    A.write_uf_chans(p, np.array([0, 2, 5]))
    assert list(A.read_uf_chans(p, 3, 8)) == [0, 2, 5]
    with pytest.raises(A.AmberError):
        A.read_uf_chans(p, 4, 8)     # short read (the reference ignores it)
    with pytest.raises(A.AmberError):
        A.read_uf_chans(p, 3, 5)     # id 5 outside [0,5)


def test_params_ess(tmp_path):
    p = tmp_path / "input_params_ess"
    p.write_text("154.255e6\t40.00e03\t385\t/some/where/data\t\n\nnu0 deltanu chan0 data_path\n")
    nu0, dnu, c0, path = A.read_params_ess(str(p))
    assert (nu0, dnu, c0, path) == (154.255e6, 40.0e3, 385.0, "/some/where/data")


def test_rings_files_match_reference_reader(tmp_path):
    rd = np.array([[3, 2, 1, 1.5e6], [4, 4, 4, 2.0e6]])
    A.write_closed_rings(str(tmp_path), rd)
    assert (tmp_path / "rcount").read_text() == "2"
    # mabs_funcs.py:94-133 reads 4*rcount doubles
    assert np.array_equal(np.fromfile(tmp_path / "closed_rings", np.float64).reshape(-1, 4), rd)
    n, back = A.read_rings(str(tmp_path))
    assert n == 2 and np.array_equal(back, rd)


def test_bsout_layouts_match_reference_readers(tmp_path):
    W = 7
    bl = np.arange(W * W, dtype=np.float32) * 0.5
    cnt = np.arange(W * W, dtype=np.int32)
    A.write_bsout_v1(str(tmp_path), 3, 2, 1, 9, 4.0, 123.0, bl, cnt)
    f = tmp_path / "bsout_final_ra3_rb2_rc1_R9"
    raw = f.read_bytes()
    assert len(raw) == 16 + 8 * W * W
    # read_bsout_float (mabs_funcs.py:172-188)
    with open(f, "rb") as fh:
        ntri = np.fromfile(fh, np.float64, 1)[0]
        num = np.fromfile(fh, np.float64, 1)[0]
        b = np.fromfile(fh, np.float32, W * W)
        c = np.fromfile(fh, np.int32, W * W)
    assert (ntri, num) == (4.0, 123.0) and np.array_equal(b, bl) and np.array_equal(c, cnt)
    norm = np.full(W * W, 8.0, np.float32)
    A.write_bsout_v2(str(tmp_path), 3, 2, 1, 9, 1.0 / 400, bl, norm)
    f2 = tmp_path / "bsout_final2_ra3_rb2_rc1_R9"
    with open(f2, "rb") as fh:          # read_bsout_float_v2 (mabs_funcs.py:191-205)
        inv = np.fromfile(fh, np.float64, 1)[0]
        b = np.fromfile(fh, np.float32, W * W)
        nn = np.fromfile(fh, np.float32, W * W)
    assert inv == 1.0 / 400 and np.array_equal(b, bl) and np.array_equal(nn, norm)


def test_rules_agree_with_oracle():
    for (a, b, n) in [(1, 200, 24), (1, 200, 20), (2, 60, 8)]:
        lo, hi = A.log_rings(a, b, n)
        assert np.array_equal(A.triangles(lo, hi), O.triangles(lo, hi))
        assert np.array_equal(A.uavg(lo, hi), O.uavg(lo, hi))
    assert len(A.triangles(*A.log_rings(2, 60, 8))) == 46
    assert len(A.triangles(*A.log_rings(1, 200, 24))) == 506


def test_generator_invariants():
    N, nc = 32, 12
    gv, samp, uf = A.generate(N, nc, nflag=4, seed=77)
    Nh = N // 2 + 1
    g = gv.reshape(N, Nh, nc, 2)
    s = samp.reshape(N, Nh, nc)
    assert 5 in uf and len(uf) == nc - 4
    flagged = sorted(set(range(nc)) - set(uf.tolist()))
    assert np.all(gv[:, flagged] == 0) and np.all(samp[:, flagged] == 0)
    for j in (0, N // 2):
        for i in range(1, N):
            assert np.array_equal(g[i, j, :, 0], g[N - i, j, :, 0])
            assert np.array_equal(g[i, j, :, 1], -g[N - i, j, :, 1])
            assert np.array_equal(s[i, j], s[N - i, j])
    assert np.all(g[0, 0, :, 1] == 0)
    assert set(np.unique(samp)) <= {0.0, 1.0}
    gv2, samp2, uf2 = A.generate(N, nc, nflag=4, seed=77)
    assert np.array_equal(gv, gv2) and np.array_equal(uf, uf2)
    # sampling density falls off as 0.4 exp(-|U|/40)
    frac_inner = s[:4, :4, uf].mean()
    assert 0.2 < frac_inner < 0.5


def test_generator_counts_normalise_back():
    gv, samp, _ = A.generate(16, 3, seed=5, count_max=4)
    assert samp.max() > 1
    m = samp > 0
    g1, s1, _ = A.generate(16, 3, seed=5, count_max=1)
    ratio = gv[m] / samp[m][:, None]
    assert np.allclose(ratio, g1[m])


def test_context_without_gpu_is_an_error():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(A.AmberError) as e:
        A.Context()
    assert e.value.code == -4


def test_grid_cells_and_setup_match_oracle():
    """amb_grid_setup / amb_grid_cells (host side of the gridder) against the
    oracle's grid and hand cells"""
    import oracle as O
    assert A.grid_setup(602.0, 1480.73) == O.grid_setup(602.0, 1480.73) == (1.0254384902642049, 1176)
    NN, dU = 16, 2.0
    rp = np.array([[3.1, 4.9], [-3.1, 4.9], [3.1, -4.9], [30.0, 1.0]], np.float32)
    cell, sign = A.grid_cells(NN, rp, 1.0, dU, 20.0)
    assert cell.tolist() == [2 * 9 + 2, 14 * 9 + 2, 14 * 9 + 2, -1] and sign.tolist() == [1, 1, -1, 1]
    rng = np.random.default_rng(2)
    dU, NN = O.grid_setup(50.0, 700.0)
    rp = (rng.standard_normal((400, 3)) * 30 / 1.4e8).astype(np.float32)
    data = np.ones((400, 1, 1, 3), np.float32)
    cell, _ = A.grid_cells(NN, rp, np.float32(1.4e8), dU, 50.0)
    _, samp, n = O.grid(NN, 1, rp, data, np.float32(1.4e8), dU, 50.0, flag_key=0)
    assert n == (cell >= 0).sum()
    assert np.array_equal(np.bincount(cell[cell >= 0], minlength=samp.shape[0]), samp[:, 0])

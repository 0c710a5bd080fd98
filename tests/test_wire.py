"""Hardened result files (amb_write_bsw / amb_read_bsw) and the legacy
reader (amb_read_bsout): the CRC is zlib's (IEEE 802.3) CRC-32, pinned by the
standard check value and by zlib itself; the header layout is pinned by an
independent struct packing; every corruption is reported, never read."""
import os
import struct
import zlib

import numpy as np
import pytest

import amber_amd as A

HDR = struct.Struct("<8sIIIIiiiiiiqqiIddQ8s")


def test_crc32_known_answers():
    assert A.crc32(b"123456789") == 0xCBF43926
    assert A.crc32(b"") == 0
    rng = np.random.default_rng(0)
    for n in (1, 7, 1000, 65537):
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert A.crc32(b) == zlib.crc32(b)


def _arrays(W, kind, seed):
    rng = np.random.default_rng(seed)
    bl = rng.standard_normal((W, W)).astype(np.float32)
    den = rng.integers(0, 1000, (W, W)).astype(np.int32) if kind == 1 else \
        rng.standard_normal((W, W)).astype(np.float32)
    return bl, den


@pytest.mark.parametrize("kind", [1, 2])
def test_round_trip_and_layout(tmp_path, kind):
    W = 15
    bl, den = _arrays(W, kind, kind)
    A.write_bsw(str(tmp_path), kind, 5, 3, 2, 7, 114, 96, "bf16x3", 12.5, 1425.0, bl, den)
    f = tmp_path / ("bsw%d_ra5_rb3_rc2_R7" % kind)
    raw = f.read_bytes()
    payload = bl.tobytes() + den.tobytes()
    assert len(raw) == 104 + len(payload) and raw[104:] == payload
    want = HDR.pack(b"AMBRBSW\0", 1, 104, kind, 0x01020304, 7, W, 5, 3, 2, 7, 114, 96, 1,
                    zlib.crc32(payload), 12.5, 1425.0, 8 * W * W, bytes(8))
    assert raw[:104] == want
    h, b2, d2 = A.read_bsw(str(f))
    assert np.array_equal(b2, bl) and np.array_equal(d2, den) and d2.dtype == den.dtype
    assert (h["kind"], h["nc10"], h["W"], h["r1"], h["r2"], h["r3"], h["rlzn"], h["ng"], h["nc"]) == \
        (kind, 7, W, 5, 3, 2, 7, 114, 96)
    assert h["h0"] == 12.5 and h["h1"] == 1425.0 and h["precision"] == A.PREC_BF16X3


def _corrupt(path, offset, value):
    b = bytearray(path.read_bytes())
    b[offset] = value
    path.write_bytes(bytes(b))


def test_corruption_is_reported(tmp_path):
    W = 9
    bl, den = _arrays(W, 1, 3)
    A.write_bsw(str(tmp_path), 1, 2, 1, 0, 1, 20, 10, "fp32", 1.0, 2.0, bl, den)
    f = tmp_path / "bsw1_ra2_rb1_rc0_R1"
    good = f.read_bytes()
    cases = {
        "payload bit": (104 + 37, good[104 + 37] ^ 0x10),
        "magic": (0, ord("X")),
        "version": (8, 2),
        "endian": (20, 0x01),
        "kind": (16, 3),
        "nc10 vs W": (24, 5),
        "crc": (68, good[68] ^ 0xFF),
    }
    for name, (off, val) in cases.items():
        f.write_bytes(good)
        _corrupt(f, off, val)
        with pytest.raises(A.AmberError) as e:
            A.read_bsw(str(f))
        assert e.value.code == -3, name
    for cut in (1, 8 * W * W, len(good) - 50):
        f.write_bytes(good[:-cut])
        with pytest.raises(A.AmberError):
            A.read_bsw(str(f))
    f.write_bytes(good + b"\0")
    with pytest.raises(A.AmberError):
        A.read_bsw(str(f))
    f.write_bytes(good)
    assert np.array_equal(A.read_bsw(str(f))[1], bl)


@pytest.mark.parametrize("W", [1, 3, 127, 1535])
def test_legacy_reader_infers_kind_and_size(tmp_path, W):
    bl, cnt = _arrays(W, 1, W)
    _, norm = _arrays(W, 2, W + 1)
    A.write_bsout_v1(str(tmp_path), 4, 2, 1, 3, 0.25, 4096.0, bl, cnt)
    A.write_bsout_v2(str(tmp_path), 4, 2, 1, 3, 1.0 / 400, bl, norm)
    k, h0, h1, b, d = A.read_bsout(str(tmp_path / "bsout_final_ra4_rb2_rc1_R3"))
    assert (k, h0, h1) == (1, 0.25, 4096.0) and np.array_equal(b, bl) and np.array_equal(d, cnt)
    k, h0, h1, b, d = A.read_bsout(str(tmp_path / "bsout_final2_ra4_rb2_rc1_R3"))
    assert (k, h0) == (2, 1.0 / 400) and np.array_equal(b, bl) and np.array_equal(d, norm)


def test_legacy_reader_rejects_other_sizes(tmp_path):
    f = tmp_path / "bsout_final_x"
    for n in (0, 7, 16 + 8 * 4, 8 + 8 * 4, 16 + 8 * 9 + 1):   # W = 2 is even; a stray byte
        f.write_bytes(bytes(n))
        with pytest.raises(A.AmberError):
            A.read_bsout(str(f))

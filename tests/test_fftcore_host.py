"""The LDS FFT core (amber_amd/csrc/amb_fftcore.h) on the host: every
stage and butterfly of the device code, with a team's threads run one after
another per stage, against numpy's FFT and, for small n, a direct DFT.

Lengths cover each register codelet (2, 3, 4, 8 and the odd primes up to 47),
the direct stage of larger primes (109, 181, 349), the transform lengths the
ring c2r meets at the configured grids (612, 362, 218, 1176, ...), the
lag-Fourier length 1536 and the post-processing length 1519 = 7^2 31, with
team sizes that do not divide the butterfly counts; and the in-place form
(digit-reversed input) of the {2,3,5}-smooth channel-axis lengths."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "native", "fftcore_host.cpp")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fftc") / "libfftcore_host.so")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-host-only", "-O1", "-fPIC", "-shared", "-std=c++17",
                    "-I" + os.path.join(ROOT, "amber_amd", "csrc"), "-I" + os.path.join(ROOT, "include"), SRC,
                    "-o", out], check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    L = C.CDLL(out)
    L.fftc_host.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    L.fftc_host_ip.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    L.fftc_host_ip_swz.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    L.fftc_host_swz_map.argtypes = [C.c_int, C.c_void_p]
    return L


def run(lib, x, direction, dbl, team, inplace=False, maxp=47, swz=False):
    n = len(x)
    a = np.empty(2 * n)
    a[0::2], a[1::2] = x.real, x.imag
    o = np.zeros(2 * n)
    if inplace:
        f = lib.fftc_host_ip_swz if swz else lib.fftc_host_ip
        rc = f(n, direction, dbl, team, a.ctypes.data, o.ctypes.data)
    else:
        rc = lib.fftc_host(n, direction, dbl, team, maxp, a.ctypes.data, o.ctypes.data)
    assert rc > 0 or n == 1
    return o[0::2] + 1j * o[1::2]


LENGTHS = [2, 3, 4, 5, 7, 8, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 60, 64, 86, 109, 128, 218, 258, 362,
           396, 448, 470, 492, 512, 612, 698, 1176, 1519, 1536, 2048]


@pytest.mark.parametrize("n", LENGTHS)
@pytest.mark.parametrize("dbl", [1, 0])
def test_against_numpy(lib, n, dbl):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    tol = 1e-13 if dbl else 2e-6
    for team in (64, 256, 7):
        inv = run(lib, x, +1, dbl, team)  # sum x_m e^{+2 pi i m k / n}
        fwd = run(lib, x, -1, dbl, team)
        ref_inv = np.fft.ifft(x) * n
        ref_fwd = np.fft.fft(x)
        assert np.linalg.norm(inv - ref_inv) <= tol * np.linalg.norm(ref_inv), (n, team)
        assert np.linalg.norm(fwd - ref_fwd) <= tol * np.linalg.norm(ref_fwd), (n, team)


@pytest.mark.parametrize("n", [5, 7, 17, 30, 47, 53])
def test_against_direct_dft(lib, n):
    """independent of numpy: the defining sum, evaluated in FP64"""
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    m = np.arange(n)
    E = np.exp(2j * np.pi * np.outer(m, m) / n)
    assert np.allclose(run(lib, x, +1, 1, 64), E @ x, rtol=0, atol=1e-12 * np.abs(x).sum())
    assert np.allclose(run(lib, x, -1, 1, 64), np.conj(E) @ x, rtol=0, atol=1e-12 * np.abs(x).sum())


# the in-place form (digit-reversed input) on {2,3,5}-smooth lengths: the
# channel-axis transforms of the lag-Fourier spectra
IP_LENGTHS = [2, 4, 6, 10, 12, 30, 48, 60, 128, 250, 384, 768, 1000, 1152, 1536, 1600, 2048]


@pytest.mark.parametrize("n", IP_LENGTHS)
@pytest.mark.parametrize("dbl", [1, 0])
def test_in_place_against_numpy(lib, n, dbl):
    rng = np.random.default_rng(7 * n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    tol = 1e-13 if dbl else 2e-6
    for team in (64, 7):
        fwd = run(lib, x, -1, dbl, team, inplace=True)
        inv = run(lib, x, +1, dbl, team, inplace=True)
        assert np.linalg.norm(fwd - np.fft.fft(x)) <= tol * np.linalg.norm(np.fft.fft(x)), (n, team)
        assert np.linalg.norm(inv - np.fft.ifft(x) * n) <= tol * np.linalg.norm(np.fft.ifft(x) * n), (n, team)


@pytest.mark.parametrize("n", [30, 48, 256, 1000, 1536, 2048])
def test_bank_swizzle_is_a_permutation(lib, n):
    """The channel-spectra kernel's LDS index map: a permutation of [0, n)
    that moves indices only inside their aligned block of 32, and the
    identity where 32 does not divide n."""
    m = np.empty(n, dtype=np.int32)
    lib.fftc_host_swz_map(n, m.ctypes.data)
    assert sorted(m.tolist()) == list(range(n))
    if n % 32:
        assert np.array_equal(m, np.arange(n))
    else:
        assert np.array_equal(m // 32, np.arange(n) // 32)
        assert not np.array_equal(m, np.arange(n))


@pytest.mark.parametrize("n", [48, 256, 768, 1000, 1536, 2048])
@pytest.mark.parametrize("dbl", [1, 0])
def test_in_place_swizzled_against_numpy(lib, n, dbl):
    """The in-place transform with its buffer read and written through the
    bank swizzle (input scattered to swizzled digit-reversed positions,
    output read back through the swizzle) is the same DFT."""
    rng = np.random.default_rng(11 * n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    tol = 1e-13 if dbl else 2e-6
    for team in (64, 7):
        fwd = run(lib, x, -1, dbl, team, inplace=True, swz=True)
        inv = run(lib, x, +1, dbl, team, inplace=True, swz=True)
        assert np.linalg.norm(fwd - np.fft.fft(x)) <= tol * np.linalg.norm(np.fft.fft(x)), (n, team)
        assert np.linalg.norm(inv - np.fft.ifft(x) * n) <= tol * np.linalg.norm(np.fft.ifft(x) * n), (n, team)


# the direct stage of primes above a kernel's codelet limit (the ring c2r
# kernels stop at 23 in FP32, 19 in FP64, the lag-Fourier 2-D inverse at 5):
# grids 492 = 4 3 41, 258 = 2 3 43, 62, 74, the post-processing length
# 1519 = 7^2 31, larger primes, and lengths with radix-8 stages (always a
# codelet, whatever the limit: 16, 1536 = 8^3 3, 328 = 8 41)
@pytest.mark.parametrize("n", [16, 29, 62, 74, 258, 492, 1519, 1536, 109 * 2, 349, 8 * 41, 1000 - 1])
@pytest.mark.parametrize("maxp", [23, 19, 5])
def test_direct_stage_against_numpy(lib, n, maxp):
    rng = np.random.default_rng(11 * n + maxp)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for dbl, tol in ((1, 1e-13), (0, 3e-6)):
        for team in (64, 7):
            fwd = run(lib, x, -1, dbl, team, maxp=maxp)
            inv = run(lib, x, +1, dbl, team, maxp=maxp)
            assert np.linalg.norm(fwd - np.fft.fft(x)) <= tol * np.linalg.norm(np.fft.fft(x)), (n, team, dbl)
            assert np.linalg.norm(inv - np.fft.ifft(x) * n) <= tol * np.linalg.norm(np.fft.ifft(x) * n), (n, team, dbl)

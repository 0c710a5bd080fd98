"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker; the product never
imports it.  Build with ``make -C oracle``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = C.CDLL(LIB_PATH)

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_u8 = C.POINTER(C.c_uint8)

_lib.orc_fft_c2c.argtypes = [C.c_int, _dp, _dp, C.c_int]
_lib.orc_dft_direct.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int]
_lib.orc_ng.restype = C.c_long
_lib.orc_ng.argtypes = [C.c_double, C.c_double]
_lib.orc_uavg.argtypes = [_dp, _dp, C.c_int, _dp]
_lib.orc_normalize.argtypes = [C.c_long, C.c_long, _dp, _dp]
_lib.orc_annular_rings.argtypes = [C.c_int, _dp, _dp]
_lib.orc_triangles.argtypes = [_dp, C.c_int, _ip, C.c_int]
_lib.orc_ring_c2r.argtypes = [C.c_long, C.c_long, C.c_long, C.c_double, C.c_double, _dp, _dp, C.c_long, _dp]
_lib.orc_calcr.argtypes = [C.c_long, C.c_long, _dp, C.c_int, _dp, _dp, C.c_int, _dp, C.c_int]
_lib.orc_bispec.argtypes = [C.c_long, C.c_long, C.c_double, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int,
                            _u8, C.c_int, _dp, _dp, C.POINTER(C.c_int32), C.c_int]
_lib.orc_bispec_cells.argtypes = [C.c_long, C.c_long, C.c_double, _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                                  C.c_int, _u8, C.c_int, C.c_int, _ip, _ip, _dp, _dp,
                                  C.POINTER(C.c_int32), _dp, C.c_int]
_lib.orc_refalgo.restype = C.c_long
_lib.orc_refalgo.argtypes = [C.c_long, C.c_long, C.c_double, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int,
                             _u8, C.c_int, C.c_int, C.c_int, C.c_int, _fp, _fp, C.POINTER(C.c_int32),
                             C.c_int, _dp, _dp]


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def fft_c2c(x, sign=+1):
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real); im = np.ascontiguousarray(x.imag)
    _lib.orc_fft_c2c(len(x), _p(re, _dp), _p(im, _dp), sign)
    return re + 1j * im


def dft_direct(x, sign=+1):
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real); im = np.ascontiguousarray(x.imag)
    ore = np.zeros(len(x)); oim = np.zeros(len(x))
    _lib.orc_dft_direct(len(x), _p(re, _dp), _p(im, _dp), _p(ore, _dp), _p(oim, _dp), sign)
    return ore + 1j * oim


def ng(uhigh, du=1.0):
    return int(_lib.orc_ng(uhigh, du))


def uavg(lo, hi):
    lo = _f64(lo); hi = _f64(hi)
    out = np.zeros(len(lo))
    _lib.orc_uavg(_p(lo, _dp), _p(hi, _dp), len(lo), _p(out, _dp))
    return out


def annular_rings(nr=22):
    lo = np.zeros(nr); hi = np.zeros(nr)
    _lib.orc_annular_rings(nr, _p(lo, _dp), _p(hi, _dp))
    return lo, hi


def triangles(lo, hi):
    u = uavg(lo, hi)
    n = _lib.orc_triangles(_p(u, _dp), len(u), None, 0)
    t = np.zeros((max(n, 1), 3), np.int32)
    _lib.orc_triangles(_p(u, _dp), len(u), _p(t, _ip), n)
    return t[:n]


def ring_c2r(N, nc, Ng, lo, hi, plane, gv=None, samp=None):
    out = np.zeros((Ng, Ng))
    _lib.orc_ring_c2r(N, nc, Ng, lo, hi, _p(_f64(gv) if gv is not None else None, _dp),
                      _p(_f64(samp) if samp is not None else None, _dp), plane, _p(out, _dp))
    return out


def calcr(N, nc, samp, lo, hi, plane=5):
    lo = _f64(lo); hi = _f64(hi)
    n = len(triangles(lo, hi))
    rd = np.zeros((max(n, 1), 4))
    got = _lib.orc_calcr(N, nc, _p(_f64(samp), _dp), plane, _p(lo, _dp), _p(hi, _dp), len(lo), _p(rd, _dp), n)
    return rd[:got]


def normalize(gv, samp):
    """normalize_data on copies of file-layout arrays: returns (gv, samp)."""
    g = np.array(gv, np.float64, copy=True, order="C")
    s = np.array(samp, np.float64, copy=True, order="C")
    cells = s.shape[0]
    nc = s.shape[1] if s.ndim > 1 else 1
    _lib.orc_normalize(cells, nc, _p(g, _dp), _p(s, _dp))
    return g, s


def chan_mask(nc, flag_key, uf_chans):
    m = np.zeros(nc, np.uint8)
    if flag_key == 1:
        m[np.asarray(uf_chans, int)] = 1
    else:
        m[:] = 1
    return m


def bispec(N, nc, Du, gv, samp, lo, hi, r1, r2, r3, mask, nc10, want_norm=False, nthreads=0):
    W = 2 * nc10 + 1
    bl = np.zeros((W, W)); cnt = np.zeros((W, W), np.int32)
    norm = np.zeros((W, W)) if want_norm else None
    rc = _lib.orc_bispec(N, nc, Du, _p(_f64(gv), _dp), _p(_f64(samp), _dp), _p(_f64(lo), _dp),
                         _p(_f64(hi), _dp), r1, r2, r3, _p(np.ascontiguousarray(mask, np.uint8), _u8), nc10,
                         _p(bl, _dp), _p(norm, _dp), cnt.ctypes.data_as(C.POINTER(C.c_int32)), nthreads)
    if rc != 0:
        raise RuntimeError("orc_bispec failed")
    return bl, cnt, norm


def bispec_cells(N, nc, Du, gv, samp, lo, hi, r1, r2, r3, mask, nc10, d2, d3, want_bl=True,
                 want_norm=False, nthreads=0, want_abs=False):
    """Returns (bl, cnt, norm) or, with want_abs, (bl, cnt, norm, sum|terms|)."""
    d2 = np.ascontiguousarray(d2, np.int32); d3 = np.ascontiguousarray(d3, np.int32)
    n = len(d2)
    bl = np.zeros(n) if want_bl else None
    norm = np.zeros(n) if want_norm else None
    ab = np.zeros(n) if (want_abs and want_bl) else None
    cnt = np.zeros(n, np.int32)
    rc = _lib.orc_bispec_cells(N, nc, Du, _p(_f64(gv), _dp), _p(_f64(samp), _dp), _p(_f64(lo), _dp),
                               _p(_f64(hi), _dp), r1, r2, r3, _p(np.ascontiguousarray(mask, np.uint8), _u8),
                               nc10, n, _p(d2, _ip), _p(d3, _ip), _p(bl, _dp), _p(norm, _dp),
                               cnt.ctypes.data_as(C.POINTER(C.c_int32)), _p(ab, _dp), nthreads)
    if rc != 0:
        raise RuntimeError("orc_bispec_cells failed")
    if want_abs:
        return bl, cnt, norm, ab
    return bl, cnt, norm


def refalgo(N, nc, Du, gv, samp, lo, hi, r1, r2, r3, mask, nc10, p1_lo=0, p1_hi=None, v2=False, nthreads=0):
    """Reference-algorithm FP32 path; returns (bl, numf, cnt, ntriplets, t_fill, t_loop)."""
    W = 2 * nc10 + 1
    if p1_hi is None:
        p1_hi = nc
    bl = np.zeros(W * W, np.float32)
    numf = np.zeros(W * W, np.float32) if v2 else None
    cnt = np.zeros(W * W, np.int32)
    tf = C.c_double(); tl = C.c_double()
    n = _lib.orc_refalgo(N, nc, Du, _p(_f64(gv), _dp), _p(_f64(samp), _dp), _p(_f64(lo), _dp),
                         _p(_f64(hi), _dp), r1, r2, r3, _p(np.ascontiguousarray(mask, np.uint8), _u8), nc10,
                         p1_lo, p1_hi, int(bool(v2)), _p(bl, _fp), _p(numf, _fp),
                         cnt.ctypes.data_as(C.POINTER(C.c_int32)), nthreads, C.byref(tf), C.byref(tl))
    return (bl.reshape(W, W), None if numf is None else numf.reshape(W, W), cnt.reshape(W, W), n,
            tf.value, tl.value)


_lib.orc_maps_bin.restype = C.c_int
_lib.orc_maps_bin.argtypes = [C.c_long, C.c_double, C.c_long, C.c_long, C.c_int, C.c_double, C.c_double, _dp]
_lib.orc_cldnu.argtypes = [C.c_long, C.c_long, C.c_double, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp,
                           _dp, C.POINTER(C.c_int32)]


def maps_bin(N, Du, i, j, NB=15, umin=1.0, umax=500.0):
    """(bin, |U|) of half-plane cell (i, j); bin -1 outside [umin, umax)"""
    uu = C.c_double()
    b = _lib.orc_maps_bin(N, Du, i, j, NB, umin, umax, C.byref(uu))
    return b, uu.value


def cldnu(N, nc, Du, gv, samp, NB=15, umin=1.0, umax=500.0):
    """MAPS run: returns (cldnu[NB][nc], numdnu[nc], binu[NB], num[NB])"""
    cl = np.zeros((NB, nc)); nd = np.zeros(nc); bu = np.zeros(NB); nm = np.zeros(NB, np.int32)
    rc = _lib.orc_cldnu(N, nc, Du, _p(_f64(gv), _dp), _p(_f64(samp), _dp), NB, umin, umax, _p(cl, _dp),
                        _p(nd, _dp), _p(bu, _dp), nm.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != 0:
        raise RuntimeError("orc_cldnu failed")
    return cl, nd, bu, nm


_lib.orc_grid_setup.argtypes = [C.c_double, C.c_double, _dp, C.POINTER(C.c_long)]
_lib.orc_grid.restype = C.c_long
_lib.orc_grid.argtypes = [C.c_long, C.c_long, C.c_long, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                          C.c_double, C.c_double, C.c_long, _fp, _fp, _dp, _dp]


def grid_setup(Umax, FWHM):
    """(dU, NN) of initialize_gridding"""
    dU = C.c_double(); NN = C.c_long()
    _lib.orc_grid_setup(Umax, FWHM, C.byref(dU), C.byref(NN))
    return dU.value, NN.value


def grid(NN, nc, ranpar, data, nu_chan0, dU, Umax, flag_key=1, n_collapse=1, chan1=1):
    """vcg gridding: ranpar [ngroup][npar], data [ngroup][nchan][nstokes][3] float32;
    returns (gv [NN(NN/2+1)][nc/n_collapse][2], samp [..][nc/n_collapse], groups gridded)"""
    ranpar = np.ascontiguousarray(ranpar, np.float32)
    data = np.ascontiguousarray(data, np.float32)
    ngroup, npar = ranpar.shape
    nchan, nstokes = data.shape[1], data.shape[2]
    ncout = nc // n_collapse
    cells = NN * (NN // 2 + 1)
    gv = np.zeros((cells, ncout, 2)); samp = np.zeros((cells, ncout))
    n = _lib.orc_grid(NN, nc, nchan, nstokes, chan1, npar, flag_key, n_collapse, nu_chan0, dU, Umax, ngroup,
                      _p(ranpar, _fp), _p(data, _fp), _p(gv, _dp), _p(samp, _dp))
    if n < 0:
        raise RuntimeError("orc_grid failed (%d)" % n)
    return gv, samp, n

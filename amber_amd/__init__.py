"""amber_amd -- Python mirror of the amber_mi355x C ABI (include/amber_mi355x.h).

The product is the HIP library ``libamber_mi355x.so``; this module only binds
it with ctypes and converts numpy arrays.  It mirrors the reference's
in-process function set (4_fbe/mabs_funcs.h:19-60) with the same argument
meaning: ``closed_rings`` is calCR (amber_funcs.c:389-491), ``bispec_v1`` /
``bispec_v2`` are one iteration of the rp loop of amber.c:187-326 /
amber_v2.c:187-333, ``read_*``/``write_*`` are the reference's file formats.

There is no CPU fallback: importing this module raises if the library has
not been built (``python -m amber_amd.build``).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamber_mi355x.so")
if not os.path.exists(LIB_PATH):
    raise ImportError("amber_amd: %s is missing; build it with `python -m amber_amd.build`" % LIB_PATH)
_lib = C.CDLL(LIB_PATH)

PREC_FP32, PREC_BF16X3, PREC_FP8, PREC_FP16X3, PREC_LAGFT = 0, 1, 2, 3, 4
_PREC = {"fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "fp8": PREC_FP8, "fp16x3": PREC_FP16X3, "lagft": PREC_LAGFT}

EXPORTS = [
    "amb_strerror", "amb_version", "amb_read_params_ess", "amb_read_gv_header", "amb_read_cube_files",
    "amb_free", "amb_write_cube_files", "amb_read_uf_chans", "amb_write_uf_chans", "amb_write_closed_rings",
    "amb_read_rings", "amb_write_bsout_v1", "amb_write_bsout_v2", "amb_ng", "amb_annular_rings", "amb_uavg",
    "amb_triangles", "amb_generate", "amb_create", "amb_destroy", "amb_last_error", "amb_set_precision",
    "amb_load_cube", "amb_load_cube_files", "amb_release_cube", "amb_set_channels", "amb_set_rings",
    "amb_closed_rings", "amb_bispec_v1", "amb_bispec_v2", "amb_bispec_ex", "amb_bispec_batch",
    "amb_get_stats", "amb_reset_stats", "amb_drop_planes", "amb_batch_times", "amb_ring_image",
    "amb_bispec_multi", "amb_set_batch_mode", "amb_update_cube", "amb_graph_create", "amb_graph_run",
    "amb_graph_destroy", "amb_serve_start", "amb_submit", "amb_submit_cube", "amb_poll", "amb_wait",
    "amb_serve_stop",
    "amb_cyl_bispec", "amb_maps_cldnu", "amb_write_cldnu", "amb_read_params", "amb_write_bsw", "amb_read_bsw",
    "amb_read_bsout", "amb_crc32", "amb_read_vcg_inputs", "amb_grid_setup", "amb_grid_cells", "amb_grid",
    "amb_grid_cube", "amb_set_cert_strict", "amb_set_grid", "amb_compute_grid", "amb_ring_reach", "amb_closed_rings_noalloc", "amb_autotune", "amb_tune_save",
    "amb_tune_load", "amb_tune_clear", "amb_tune_entries", "amb_tune_lookup", "amb_set_launch_config",
]

BATCH_PER_TRIANGLE, BATCH_GROUPED = 0, 1

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class GenCfg(C.Structure):
    _fields_ = [("N", C.c_int64), ("nc", C.c_int64), ("nflag", C.c_int), ("shared_mask", C.c_int),
                ("count_max", C.c_int), ("seed", C.c_uint64), ("p0", C.c_double), ("u0", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("hbm_budget", C.c_uint64), ("verbose", C.c_int)]


class Result(C.Structure):
    _fields_ = [("bl", _fp), ("bl64", _dp), ("cnt", _ip), ("norm", _fp), ("norm64", _dp),
                ("inv_ng2", C.c_double), ("ng", C.c_long), ("cert_margin", C.c_double)]


class TriReq(C.Structure):
    _fields_ = [("cube_id", C.c_int), ("r1", C.c_int), ("r2", C.c_int), ("r3", C.c_int),
                ("bl", _fp), ("cnt", _ip), ("norm", _fp)]


class TuneEntry(C.Structure):
    _fields_ = [("prec", C.c_int), ("ng_log2", C.c_int), ("nact_log2", C.c_int), ("variant", C.c_int),
                ("xrun_log2", C.c_int), ("wave_gib", C.c_int), ("grid_pct", C.c_int), ("tflops", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["prec"] = {v: k for k, v in _PREC.items()}.get(self.prec, self.prec)
        return d


class CubeSrc(C.Structure):
    _fields_ = [("N", C.c_int64), ("nc", C.c_int64), ("Du", C.c_double), ("flag_key", C.c_int32),
                ("gv", C.POINTER(C.c_double)), ("samp", C.POINTER(C.c_double)), ("chans", C.POINTER(C.c_int)),
                ("uf_nc", C.c_int), ("nc10", C.c_int)]


class CylParams(C.Structure):
    _fields_ = [("ncut", C.c_int), ("mabs_norm", C.c_double), ("rc", C.c_double), ("rc_prime", C.c_double),
                ("dnu", C.c_double), ("window", C.c_int)]


def _analysis_constants():
    """the constants of the reference's analysis chain (mabs_funcs.py:7-27)"""
    import math
    theta0 = 0.6 * 24.68 * math.pi / 180
    fact = math.pi * theta0 * theta0 / 3
    nu0, k_b, c = 154.255e6, 1.38e3, 3e8
    rjfac = 2. * k_b * nu0 * nu0 / (c * c) / 1e3
    return dict(ncut=8, mabs_norm=fact * rjfac ** 3, rc=9210.11291939594, rc_prime=17., dnu=0.04)


ANALYSIS = _analysis_constants()


class Stats(C.Structure):
    _fields_ = [("ms_fft", C.c_double), ("ms_contract", C.c_double), ("ms_reduce", C.c_double),
                ("ms_norm", C.c_double), ("ms_total", C.c_double), ("flop_contract", C.c_double),
                ("bytes_fft", C.c_double), ("n_contract_launch", C.c_int64), ("n_fft_launch", C.c_int64),
                ("evals", C.c_int64), ("flop_contract_exec", C.c_double),
                ("ms_graph", C.c_double), ("n_graph_launch", C.c_int64),
                ("ms_maps", C.c_double), ("ms_maps_total", C.c_double), ("flop_maps", C.c_double),
                ("n_maps_launch", C.c_int64), ("ms_grid", C.c_double), ("n_grid_groups", C.c_int64),
                ("ms_post", C.c_double), ("flop_post", C.c_double), ("cert_margin_max", C.c_double),
                ("n_uncertified", C.c_int64), ("flop_contract_distinct", C.c_double),
                ("n_norm_exact", C.c_int64), ("n_evict_planes", C.c_int64), ("n_evict_cubes", C.c_int64),
                ("n_cube_uploads", C.c_int64), ("n_serve_iters", C.c_int64), ("flop_lagft", C.c_double),
                ("plane_pool_bytes", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib.amb_strerror.restype = C.c_char_p
_lib.amb_version.restype = C.c_char_p
_lib.amb_last_error.restype = C.c_char_p
_lib.amb_last_error.argtypes = [C.c_void_p]
_lib.amb_ng.restype = C.c_long
_lib.amb_ng.argtypes = [C.c_double, C.c_double]
_lib.amb_create.argtypes = [C.POINTER(Opts), C.POINTER(C.c_void_p)]
_lib.amb_destroy.argtypes = [C.c_void_p]
_lib.amb_generate.argtypes = [C.POINTER(GenCfg), _dp, _dp, _ip, _ip]
_lib.amb_load_cube.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_int32, _dp, _dp, _ip]
_lib.amb_load_cube_files.argtypes = [C.c_void_p, C.c_char_p, C.c_int, _ip]
_lib.amb_release_cube.argtypes = [C.c_void_p, C.c_int]
_lib.amb_set_channels.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int, C.c_int]
_lib.amb_set_rings.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
_lib.amb_set_precision.argtypes = [C.c_void_p, C.c_int]
_lib.amb_closed_rings.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, _ip]
_lib.amb_closed_rings_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, _ip]
_lib.amb_bispec_v1.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _fp, _ip, _dp, _dp]
_lib.amb_bispec_v2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _fp, _fp, _dp]
_lib.amb_bispec_ex.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Result)]
_lib.amb_bispec_batch.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int, _fp, _ip, _fp]
_lib.amb_drop_planes.argtypes = [C.c_void_p, C.c_int]
_lib.amb_ring_image.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int, _fp, C.c_int64]
_lib.amb_batch_times.argtypes = [C.c_void_p, _dp, C.c_int, _ip]
_lib.amb_bispec_multi.argtypes = [C.c_void_p, C.POINTER(TriReq), C.c_int]
_lib.amb_set_batch_mode.argtypes = [C.c_void_p, C.c_int]
_lib.amb_autotune.argtypes = [C.c_void_p, C.POINTER(TriReq), C.c_int, C.c_double, C.c_char_p]
_lib.amb_tune_save.argtypes = [C.c_void_p, C.c_char_p]
_lib.amb_tune_load.argtypes = [C.c_void_p, C.c_char_p]
_lib.amb_tune_clear.argtypes = [C.c_void_p]
_lib.amb_tune_entries.argtypes = [C.c_void_p, C.POINTER(TuneEntry), C.c_int, _ip]
_lib.amb_tune_lookup.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(TuneEntry)]
_lib.amb_set_launch_config.argtypes = [C.c_void_p, C.POINTER(TuneEntry)]
_lib.amb_update_cube.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
_lib.amb_graph_create.argtypes = [C.c_void_p, C.POINTER(TriReq), C.c_int, C.POINTER(C.c_void_p)]
_lib.amb_graph_run.argtypes = [C.c_void_p, C.c_void_p]
_lib.amb_graph_destroy.argtypes = [C.c_void_p, C.c_void_p]
_lib.amb_graph_destroy.restype = None
_lib.amb_serve_start.argtypes = [C.c_void_p, C.c_int]
_lib.amb_submit.argtypes = [C.c_void_p, C.POINTER(TriReq), C.c_int, C.POINTER(C.c_uint64)]
_lib.amb_submit_cube.argtypes = [C.c_void_p, C.POINTER(CubeSrc), C.POINTER(TriReq), C.c_int, _ip,
                                 C.POINTER(C.c_uint64)]
_lib.amb_poll.argtypes = [C.c_void_p, C.c_uint64, _ip]
_lib.amb_wait.argtypes = [C.c_void_p, C.c_uint64, _dp]
_lib.amb_serve_stop.argtypes = [C.c_void_p]
_lib.amb_cyl_bispec.argtypes = [C.c_void_p, C.c_int, C.c_int, _fp, C.c_void_p, C.c_int, C.POINTER(CylParams), _dp]
_lib.amb_maps_cldnu.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_double, _dp, _dp, C.c_int, C.c_double,
                                C.c_double, _dp, _dp, _dp, _ip]
_lib.amb_write_cldnu.argtypes = [C.c_char_p, C.c_int, C.c_int64, _dp, _dp, _dp, _ip]
_lib.amb_read_params.argtypes = [C.c_char_p, C.POINTER(C.c_int64), _dp, _dp, _dp, _dp, C.POINTER(C.c_int64),
                                 C.c_char_p, C.c_size_t]
_lib.amb_write_bsw.argtypes = [C.c_char_p, C.c_void_p, _fp, C.c_void_p]
_lib.amb_read_bsw.argtypes = [C.c_char_p, C.c_void_p, _fp, C.c_void_p, C.c_int64]
_lib.amb_read_bsout.argtypes = [C.c_char_p, _ip, _ip, _dp, _dp, _fp, C.c_void_p, C.c_int64]
_lib.amb_crc32.restype = C.c_uint32
_lib.amb_crc32.argtypes = [C.c_void_p, C.c_size_t]
class GridCfg(C.Structure):
    _fields_ = [("NN", C.c_int64), ("nc", C.c_int64), ("nchan", C.c_int64), ("nstokes", C.c_int64),
                ("chan1", C.c_int), ("npar", C.c_int), ("flag_key", C.c_int), ("n_collapse", C.c_int),
                ("nu_chan0", C.c_float), ("dU", C.c_double), ("Umax", C.c_double)]


_lib.amb_read_vcg_inputs.argtypes = [C.c_char_p, _ip, _dp, _dp, _dp]
_lib.amb_grid_setup.argtypes = [C.c_double, C.c_double, _dp, C.POINTER(C.c_int64)]
_lib.amb_grid_cells.restype = C.c_int64
_lib.amb_grid_cells.argtypes = [C.POINTER(GridCfg), C.c_int64, _fp, C.POINTER(C.c_int64), _fp]
_lib.amb_grid.argtypes = [C.c_void_p, C.POINTER(GridCfg), C.c_int64, _fp, _fp, _dp, _dp, C.POINTER(C.c_int64)]
_lib.amb_grid_cube.argtypes = [C.c_void_p, C.POINTER(GridCfg), C.c_int64, _fp, _fp, _ip, C.POINTER(C.c_int64)]
_lib.amb_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
_lib.amb_reset_stats.argtypes = [C.c_void_p]
_lib.amb_set_cert_strict.argtypes = [C.c_void_p, C.c_int]
_lib.amb_set_grid.argtypes = [C.c_void_p, C.c_int]
_lib.amb_compute_grid.restype = C.c_long
_lib.amb_compute_grid.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
_lib.amb_ring_reach.restype = C.c_long
_lib.amb_ring_reach.argtypes = [C.c_double]
_lib.amb_uavg.argtypes = [_dp, _dp, C.c_int, _dp]
_lib.amb_triangles.argtypes = [_dp, C.c_int, _ip, C.c_int]
_lib.amb_annular_rings.argtypes = [C.c_int, _dp, _dp]
_lib.amb_read_params_ess.argtypes = [C.c_char_p, _dp, _dp, _dp, C.c_char_p, C.c_size_t]
_lib.amb_read_gv_header.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _dp, C.POINTER(C.c_int32)]
_lib.amb_read_cube_files.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _dp,
                                     C.POINTER(C.c_int32), C.POINTER(_dp), C.POINTER(_dp)]
_lib.amb_free.argtypes = [C.c_void_p]
_lib.amb_write_cube_files.argtypes = [C.c_char_p, C.c_int, C.c_int64, C.c_int64, C.c_double, C.c_int32, _dp, _dp]
_lib.amb_read_uf_chans.argtypes = [C.c_char_p, C.c_int, C.c_int64, _ip]
_lib.amb_write_uf_chans.argtypes = [C.c_char_p, C.c_int, _ip]
_lib.amb_write_closed_rings.argtypes = [C.c_char_p, _dp, C.c_int]
_lib.amb_read_rings.argtypes = [C.c_char_p, _ip, C.POINTER(_dp)]
_lib.amb_write_bsout_v1.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_double, _fp, _ip]
_lib.amb_write_bsout_v2.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _fp, _fp]


class AmberError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__("%s%s" % (_lib.amb_strerror(code).decode(), (": " + msg) if msg else ""))


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _chk(rc, ctx=None):
    if rc != 0:
        msg = _lib.amb_last_error(ctx).decode() if ctx else ""
        raise AmberError(rc, msg)


def version():
    return _lib.amb_version().decode()


# ---------------------------------------------------------------- rules
def ng(uhigh, du=1.0):
    """Reduced grid side (amber.c:193-197)."""
    return int(_lib.amb_ng(float(uhigh), float(du)))


def ring_reach(uhigh):
    """Largest |component| of a cell of a ring below uhigh (amb_ring_reach)."""
    return int(_lib.amb_ring_reach(float(uhigh)))


def compute_grid(uhigh1, uhigh2, uhigh3, du=1.0):
    """The grid the lag-Fourier form computes a triangle's sums on
    (amb_compute_grid): the smallest even 7-smooth alias-free grid of its
    three rings, or ng(uhigh1)."""
    return int(_lib.amb_compute_grid(float(uhigh1), float(uhigh2), float(uhigh3), float(du)))


def annular_rings(nr=22):
    """The reference's hard-coded ring edges (amber_funcs.c:2022-2053)."""
    lo = np.zeros(nr); hi = np.zeros(nr)
    _lib.amb_annular_rings(nr, _p(lo, _dp), _p(hi, _dp))
    return lo, hi


def uavg(lo, hi):
    lo = np.ascontiguousarray(lo, np.float64); hi = np.ascontiguousarray(hi, np.float64)
    out = np.zeros(len(lo))
    _lib.amb_uavg(_p(lo, _dp), _p(hi, _dp), len(lo), _p(out, _dp))
    return out


def triangles(lo, hi):
    """Closed-triangle list r1>=r2>=r3 (amber_funcs.c:443-449)."""
    u = uavg(lo, hi)
    n = _lib.amb_triangles(_p(u, _dp), len(u), None, 0)
    t = np.zeros((max(n, 1), 3), np.int32)
    _lib.amb_triangles(_p(u, _dp), len(u), _p(t, _ip), n)
    return t[:n]


def log_rings(umin, umax, nbins):
    """Log-spaced annular bins [e_i, e_{i+1}) in cells."""
    e = np.logspace(np.log10(umin), np.log10(umax), nbins + 1)
    return e[:-1].copy(), e[1:].copy()


# ---------------------------------------------------------------- generator
def generate(N, nc, nflag=0, seed=1, p0=0.4, u0=40.0, shared_mask=False, count_max=1):
    """Synthetic cube in the file layout: gv [N*(N/2+1), nc, 2] f64,
    samp [N*(N/2+1), nc] f64, and the unflagged channel ids."""
    cells = N * (N // 2 + 1)
    gv = np.empty((cells, nc, 2), np.float64)
    samp = np.empty((cells, nc), np.float64)
    uf = np.zeros(nc, np.int32)
    nuf = C.c_int(0)
    cfg = GenCfg(N, nc, nflag, int(bool(shared_mask)), count_max, seed, p0, u0)
    _chk(_lib.amb_generate(C.byref(cfg), _p(gv, _dp), _p(samp, _dp), _p(uf, _ip), C.byref(nuf)))
    return gv, samp, uf[: nuf.value].copy()


# ---------------------------------------------------------------- files
def write_cube_files(data_path, rlzn, N, nc, Du, flag_key, gv, samp):
    _chk(_lib.amb_write_cube_files(data_path.encode(), rlzn, N, nc, Du, flag_key,
                                   _p(np.ascontiguousarray(gv, np.float64), _dp),
                                   _p(np.ascontiguousarray(samp, np.float64), _dp)))


def read_cube_files(data_path, rlzn):
    NN = C.c_int64(); nc = C.c_int64(); Du = C.c_double(); fk = C.c_int32()
    g = _dp(); s = _dp()
    _chk(_lib.amb_read_cube_files(data_path.encode(), rlzn, C.byref(NN), C.byref(nc), C.byref(Du),
                                  C.byref(fk), C.byref(g), C.byref(s)))
    cells = NN.value * (NN.value // 2 + 1)
    gv = np.ctypeslib.as_array(g, shape=(cells * nc.value * 2,)).copy().reshape(cells, nc.value, 2)
    samp = np.ctypeslib.as_array(s, shape=(cells * nc.value,)).copy().reshape(cells, nc.value)
    _lib.amb_free(g); _lib.amb_free(s)
    return NN.value, nc.value, Du.value, fk.value, gv, samp


def read_params_ess(path):
    a = C.c_double(); b = C.c_double(); c = C.c_double()
    buf = C.create_string_buffer(512)
    _chk(_lib.amb_read_params_ess(path.encode(), C.byref(a), C.byref(b), C.byref(c), buf, 512))
    return a.value, b.value, c.value, buf.value.decode()


def read_uf_chans(path, uf_nc, nc):
    out = np.zeros(max(uf_nc, 1), np.int32)
    _chk(_lib.amb_read_uf_chans(path.encode(), uf_nc, nc, _p(out, _ip)))
    return out[:uf_nc]


def write_uf_chans(path, chans):
    chans = np.ascontiguousarray(chans, np.int32)
    _chk(_lib.amb_write_uf_chans(path.encode(), len(chans), _p(chans, _ip)))


def write_closed_rings(res_path, rdata):
    rdata = np.ascontiguousarray(rdata, np.float64).reshape(-1, 4)
    _chk(_lib.amb_write_closed_rings(res_path.encode(), _p(rdata, _dp), len(rdata)))


def read_rings(res_path):
    n = C.c_int(); d = _dp()
    _chk(_lib.amb_read_rings(res_path.encode(), C.byref(n), C.byref(d)))
    out = np.ctypeslib.as_array(d, shape=(max(n.value, 1) * 4,)).copy()[: n.value * 4].reshape(n.value, 4)
    _lib.amb_free(d)
    return n.value, out


def write_bsout_v1(res_path, r1, r2, r3, rlzn, ntri, num, bl, cnt):
    W = int(round(np.sqrt(bl.size)))
    _chk(_lib.amb_write_bsout_v1(res_path.encode(), r1, r2, r3, rlzn, W, ntri, num,
                                 _p(np.ascontiguousarray(bl, np.float32), _fp),
                                 _p(np.ascontiguousarray(cnt, np.int32), _ip)))


def write_bsout_v2(res_path, r1, r2, r3, rlzn, inv, bl, norm):
    W = int(round(np.sqrt(bl.size)))
    _chk(_lib.amb_write_bsout_v2(res_path.encode(), r1, r2, r3, rlzn, W, inv,
                                 _p(np.ascontiguousarray(bl, np.float32), _fp),
                                 _p(np.ascontiguousarray(norm, np.float32), _fp)))


class BswHeader(C.Structure):
    """amb_bsw_header: the 104-byte header of a hardened result file"""
    _fields_ = [("magic", C.c_char * 8), ("version", C.c_uint32), ("header_bytes", C.c_uint32),
                ("kind", C.c_uint32), ("endian", C.c_uint32), ("nc10", C.c_int32), ("W", C.c_int32),
                ("r1", C.c_int32), ("r2", C.c_int32), ("r3", C.c_int32), ("rlzn", C.c_int32),
                ("ng", C.c_int64), ("nc", C.c_int64), ("precision", C.c_int32), ("payload_crc", C.c_uint32),
                ("h0", C.c_double), ("h1", C.c_double), ("payload_bytes", C.c_uint64),
                ("reserved", C.c_uint8 * 8)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k not in ("magic", "reserved")}


assert C.sizeof(BswHeader) == 104
BSW_KIND_V1, BSW_KIND_V2 = 1, 2


def crc32(data):
    b = bytes(data)
    return int(_lib.amb_crc32(b, len(b)))


def write_bsw(res_path, kind, r1, r2, r3, rlzn, ng, nc, precision, h0, h1, bl, den):
    """hardened result file bsw<kind>_ra.._R<rlzn>; den: int32 counts (kind 1) or float32 norm (kind 2)"""
    bl = np.ascontiguousarray(bl, np.float32)
    W = int(round(np.sqrt(bl.size)))
    den = np.ascontiguousarray(den, np.int32 if kind == BSW_KIND_V1 else np.float32)
    h = BswHeader()
    h.kind, h.nc10, h.W = kind, (W - 1) // 2, W
    h.r1, h.r2, h.r3, h.rlzn, h.ng, h.nc = r1, r2, r3, rlzn, ng, nc
    h.precision = _PREC[precision] if isinstance(precision, str) else int(precision)
    h.h0, h.h1 = h0, h1
    _chk(_lib.amb_write_bsw(res_path.encode(), C.byref(h), _p(bl, _fp), den.ctypes.data_as(C.c_void_p)))


def read_bsw(path):
    """(header dict, bl[W][W] float32, den[W][W] int32 or float32) of a hardened file"""
    h = BswHeader()
    _chk(_lib.amb_read_bsw(path.encode(), C.byref(h), None, None, 0))
    W = h.W
    bl = np.empty((W, W), np.float32)
    den = np.empty((W, W), np.int32 if h.kind == BSW_KIND_V1 else np.float32)
    _chk(_lib.amb_read_bsw(path.encode(), C.byref(h), _p(bl, _fp), den.ctypes.data_as(C.c_void_p), W * W))
    return h.as_dict(), bl, den


def read_bsout(path):
    """a legacy bsout file of either kind, W from its size:
    (kind, h0, h1, bl[W][W], den[W][W]) as the reference writes them"""
    k = C.c_int(); W = C.c_int(); h0 = C.c_double(); h1 = C.c_double()
    _chk(_lib.amb_read_bsout(path.encode(), C.byref(k), C.byref(W), C.byref(h0), C.byref(h1), None, None, 0))
    n = W.value
    bl = np.empty((n, n), np.float32)
    den = np.empty((n, n), np.int32 if k.value == 1 else np.float32)
    _chk(_lib.amb_read_bsout(path.encode(), C.byref(k), C.byref(W), C.byref(h0), C.byref(h1), _p(bl, _fp),
                             den.ctypes.data_as(C.c_void_p), n * n))
    return k.value, h0.value, h1.value, bl, den


def grid_setup(Umax, FWHM):
    """(dU, NN) of initialize_gridding (3_vcg/tgefuncs.c:305-308)"""
    dU = C.c_double(); NN = C.c_int64()
    _chk(_lib.amb_grid_setup(float(Umax), float(FWHM), C.byref(dU), C.byref(NN)))
    return dU.value, NN.value


def read_vcg_inputs(path):
    """(nc, Umax, FWHM, f) of the vcg inputparameters file"""
    nc = C.c_int(); a = C.c_double(); b = C.c_double(); f = C.c_double()
    _chk(_lib.amb_read_vcg_inputs(path.encode(), C.byref(nc), C.byref(a), C.byref(b), C.byref(f)))
    return nc.value, a.value, b.value, f.value


def _grid_cfg(NN, nc, ranpar, data, nu_chan0, dU, Umax, flag_key, n_collapse, chan1):
    ranpar = np.ascontiguousarray(ranpar, np.float32)
    data = np.ascontiguousarray(data, np.float32)
    if ranpar.ndim != 2 or data.ndim != 4 or data.shape[3] != 3 or data.shape[0] != ranpar.shape[0]:
        raise ValueError("ranpar [ngroup][npar], data [ngroup][nchan][nstokes][3]")
    cfg = GridCfg(int(NN), int(nc), data.shape[1], data.shape[2], int(chan1), ranpar.shape[1], int(flag_key),
                  int(n_collapse), float(nu_chan0), float(dU), float(Umax))
    return cfg, ranpar, data


def grid_cells(NN, ranpar, nu_chan0, dU, Umax):
    """cell index of every group (-1 beyond Umax) and its sign flip"""
    ranpar = np.ascontiguousarray(ranpar, np.float32)
    cfg = GridCfg(int(NN), 1, 1, 1, 1, ranpar.shape[1], 1, 1, float(nu_chan0), float(dU), float(Umax))
    cell = np.zeros(len(ranpar), np.int64); sign = np.zeros(len(ranpar), np.float32)
    n = _lib.amb_grid_cells(C.byref(cfg), len(ranpar), _p(ranpar, _fp),
                            cell.ctypes.data_as(C.POINTER(C.c_int64)), _p(sign, _fp))
    if n < 0:
        raise AmberError(int(n))
    return cell, sign


# ---------------------------------------------------------------- device context
class Context:
    """One device context (amb_create).  Owns resident cubes and cached
    filtered planes; all calls run on the library's own HIP stream."""

    def __init__(self, device=0, precision="bf16x3", hbm_budget=0, verbose=False):
        self._h = C.c_void_p()
        o = Opts(device, _PREC[precision], int(hbm_budget), int(bool(verbose)))
        _chk(_lib.amb_create(C.byref(o), C.byref(self._h)))
        self._meta = {}

    def close(self):
        if self._h:
            _lib.amb_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        _chk(rc, self._h)

    def set_precision(self, precision):
        self._c(_lib.amb_set_precision(self._h, _PREC[precision]))

    def load_cube(self, N, nc, Du, flag_key, gv, samp):
        gv = np.ascontiguousarray(gv, np.float64)
        samp = np.ascontiguousarray(samp, np.float64)
        cid = C.c_int()
        self._c(_lib.amb_load_cube(self._h, N, nc, Du, flag_key, _p(gv, _dp), _p(samp, _dp), C.byref(cid)))
        self._meta[cid.value] = dict(N=N, nc=nc, Du=Du, flag_key=flag_key)
        return cid.value

    def load_cube_files(self, data_path, rlzn):
        cid = C.c_int()
        self._c(_lib.amb_load_cube_files(self._h, data_path.encode(), rlzn, C.byref(cid)))
        return cid.value

    def release_cube(self, cube):
        self._c(_lib.amb_release_cube(self._h, cube))
        self._meta.pop(cube, None)

    def set_channels(self, cube, uf_chans, nc10):
        u = np.ascontiguousarray(uf_chans if uf_chans is not None else np.zeros(0), np.int32)
        self._c(_lib.amb_set_channels(self._h, cube, _p(u, _ip) if u.size else None, int(u.size), nc10))
        self._meta.setdefault(cube, {})["nc10"] = nc10

    def set_rings(self, lo, hi):
        self._lo = np.ascontiguousarray(lo, np.float64)
        self._hi = np.ascontiguousarray(hi, np.float64)
        self._c(_lib.amb_set_rings(self._h, _p(self._lo, _dp), _p(self._hi, _dp), len(self._lo)))

    def closed_rings(self, cube, plane=5, noalloc=False):
        """calCR; noalloc=True runs the entry point that allocates no device
        memory (amb_closed_rings_noalloc)."""
        n = len(triangles(self._lo, self._hi))
        rd = np.zeros((max(n, 1), 4))
        cnt = C.c_int()
        fn = _lib.amb_closed_rings_noalloc if noalloc else _lib.amb_closed_rings
        self._c(fn(self._h, cube, plane, _p(rd, _dp), n, C.byref(cnt)))
        return rd[: cnt.value]

    def _W(self, cube):
        return 2 * self._meta[cube]["nc10"] + 1

    def bispec_v1(self, cube, r1, r2, r3, rdata_num):
        W = self._W(cube)
        bl = np.zeros(W * W, np.float32)
        cnt = np.zeros(W * W, np.int32)
        ntri = C.c_double(); num = C.c_double()
        self._c(_lib.amb_bispec_v1(self._h, cube, r1, r2, r3, rdata_num, _p(bl, _fp), _p(cnt, _ip),
                                   C.byref(ntri), C.byref(num)))
        return ntri.value, num.value, bl.reshape(W, W), cnt.reshape(W, W)

    def bispec_v2(self, cube, r1, r2, r3):
        W = self._W(cube)
        bl = np.zeros(W * W, np.float32)
        norm = np.zeros(W * W, np.float32)
        inv = C.c_double()
        self._c(_lib.amb_bispec_v2(self._h, cube, r1, r2, r3, _p(bl, _fp), _p(norm, _fp), C.byref(inv)))
        return inv.value, bl.reshape(W, W), norm.reshape(W, W)

    def bispec_ex(self, cube, r1, r2, r3, want_norm=False):
        W = self._W(cube)
        out = dict(bl=np.zeros(W * W, np.float32), bl64=np.zeros(W * W), cnt=np.zeros(W * W, np.int32))
        if want_norm:
            out["norm"] = np.zeros(W * W, np.float32)
            out["norm64"] = np.zeros(W * W)
        r = Result(_p(out["bl"], _fp), _p(out["bl64"], _dp), _p(out["cnt"], _ip),
                   _p(out.get("norm"), _fp), _p(out.get("norm64"), _dp), 0.0, 0, 0.0)
        self._c(_lib.amb_bispec_ex(self._h, cube, r1, r2, r3, int(bool(want_norm)), C.byref(r)))
        res = {k: v.reshape(W, W) for k, v in out.items()}
        res.update(inv_ng2=r.inv_ng2, ng=r.ng, cert_margin=r.cert_margin)
        return res

    def ring_image(self, cube, ring, plane, kind=0, ng=0):
        """Ring-filtered image of one plane (kind 0 data, 1 sampling) on an
        ng x ng grid (ng=0: the ring's own Ng), as fill_s2d_float /
        fill_s2I (amber_funcs.c:967-1010, 1109-1131) produce it."""
        du = float(self._meta.get(cube, {}).get("Du", 1.0))
        n = int(ng) if ng > 0 else int(_lib.amb_ng(float(self._hi[ring]), du))
        out = np.zeros(n * n, np.float32)
        self._c(_lib.amb_ring_image(self._h, cube, int(ring), n, int(plane), int(kind), _p(out, _fp), n * n))
        return out.reshape(n, n)

    def drop_planes(self, cube):
        self._c(_lib.amb_drop_planes(self._h, cube))

    def bispec_batch(self, cube, tris, want_cnt=True, want_norm=False, out=None):
        """All triangles of one cube; out=(bl, cnt, norm) reuses caller arrays."""
        W = self._W(cube)
        tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        n = len(tris)
        if out is not None:
            bl, cnt, norm = out
            assert bl.shape == (n, W, W) and bl.dtype == np.float32 and bl.flags.c_contiguous
        else:
            bl = np.zeros((n, W, W), np.float32)
            cnt = np.zeros((n, W, W), np.int32) if want_cnt else None
            norm = np.zeros((n, W, W), np.float32) if want_norm else None
        self._c(_lib.amb_bispec_batch(self._h, cube, _p(tris, _ip), n, _p(bl, _fp), _p(cnt, _ip), _p(norm, _fp)))
        return bl, cnt, norm

    def set_batch_mode(self, mode):
        """"grouped" (one persistent launch per wave) or "per_triangle" """
        m = {"grouped": BATCH_GROUPED, "per_triangle": BATCH_PER_TRIANGLE}[mode]
        self._c(_lib.amb_set_batch_mode(self._h, m))

    def autotune(self, reqs, max_seconds=60.0, path=None, want_cnt=True, want_norm=False):
        """Tune the contraction's launch configuration on a sample of
        triangles [(cube, r1, r2, r3), ...] (amb_autotune); returns the
        installed table and the sample's results."""
        arr, outs = self._reqs(reqs, want_cnt, want_norm)
        self._c(_lib.amb_autotune(self._h, arr, len(reqs), float(max_seconds),
                                  None if path is None else os.fsencode(path)))
        return self.tune_table(), outs

    def tune_table(self):
        n = C.c_int()
        self._c(_lib.amb_tune_entries(self._h, None, 0, C.byref(n)))
        arr = (TuneEntry * max(n.value, 1))()
        self._c(_lib.amb_tune_entries(self._h, arr, n.value, C.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    def tune_save(self, path):
        self._c(_lib.amb_tune_save(self._h, os.fsencode(path)))

    def tune_load(self, path):
        self._c(_lib.amb_tune_load(self._h, os.fsencode(path)))

    def tune_clear(self):
        self._c(_lib.amb_tune_clear(self._h))

    def tune_lookup(self, ng, nact):
        e = TuneEntry()
        self._c(_lib.amb_tune_lookup(self._h, int(ng), int(nact), C.byref(e)))
        return e.as_dict()

    def set_launch_config(self, variant=None, xrun_log2=5, wave_gib=16, grid_pct=100):
        """One launch configuration for every shape (variant None: back to
        the tuned table)."""
        if variant is None:
            self._c(_lib.amb_set_launch_config(self._h, None))
            return
        e = TuneEntry(0, 0, 0, int(variant), int(xrun_log2), int(wave_gib), int(grid_pct), 0.0)
        self._c(_lib.amb_set_launch_config(self._h, C.byref(e)))

    def update_cube(self, cube, gv, samp):
        gv = np.ascontiguousarray(gv, np.float64)
        samp = np.ascontiguousarray(samp, np.float64)
        self._c(_lib.amb_update_cube(self._h, cube, _p(gv, _dp), _p(samp, _dp)))

    def _reqs(self, reqs, want_cnt=True, want_norm=False, W0=None):
        """reqs: (cube, r1, r2, r3) tuples -> TriReq array plus owned output
        arrays (cube 0: the carried cube of a submission, of side W0)"""
        n = len(reqs)
        arr = (TriReq * max(n, 1))()
        outs = []
        for i, (cube, r1, r2, r3) in enumerate(reqs):
            W = W0 if cube == 0 else self._W(cube)
            bl = np.zeros((W, W), np.float32)
            cnt = np.zeros((W, W), np.int32) if want_cnt else None
            norm = np.zeros((W, W), np.float32) if want_norm else None
            arr[i] = TriReq(int(cube), int(r1), int(r2), int(r3), _p(bl, _fp), _p(cnt, _ip), _p(norm, _fp))
            outs.append((bl, cnt, norm))
        return arr, outs

    def bispec_multi(self, reqs, want_cnt=True, want_norm=False):
        """Triangles of any resident cubes: reqs = [(cube, r1, r2, r3), ...];
        returns [(bl, cnt, norm), ...] in the same order."""
        arr, outs = self._reqs(reqs, want_cnt, want_norm)
        self._c(_lib.amb_bispec_multi(self._h, arr, len(reqs)))
        return outs

    def graph(self, reqs, want_cnt=True, want_norm=False):
        """A captured, replayable batch (amb_graph_create)."""
        return Graph(self, reqs, want_cnt, want_norm)

    def serve_start(self, max_active=256):
        """Starts the serving queue; max_active bounds the triangles admitted at once."""
        self._tickets = {}
        self._c(_lib.amb_serve_start(self._h, int(max_active)))

    def submit(self, reqs, want_cnt=True, want_norm=False):
        """Queue a request of triangles; returns a ticket.  The outputs are
        filled once wait(ticket) returns."""
        arr, outs = self._reqs(reqs, want_cnt, want_norm)
        t = C.c_uint64()
        self._c(_lib.amb_submit(self._h, arr, len(reqs), C.byref(t)))
        self._tickets[t.value] = (arr, outs)
        return t.value

    def submit_cube(self, N, nc, Du, flag_key, gv, samp, reqs, uf=None, nc10=None, want_cnt=True, want_norm=False):
        """Queue a request that carries its cube (amb_submit_cube): reqs are
        (r1, r2, r3) on the carried cube or (cube, r1, r2, r3) on resident
        ones (cube 0: the carried one).  Returns (ticket, cube id); the cube
        stays resident under that id (evictable once idle)."""
        gv = np.ascontiguousarray(gv, np.float64)
        samp = np.ascontiguousarray(samp, np.float64)
        nc10 = nc - 1 if nc10 is None else int(nc10)
        ch = None if uf is None else np.ascontiguousarray(uf, np.int32)
        src = CubeSrc(int(N), int(nc), float(Du), int(flag_key), _p(gv, _dp), _p(samp, _dp),
                      _p(ch, _ip) if ch is not None else None, 0 if ch is None else len(ch), nc10)
        reqs = [(0, *r) if len(r) == 3 else tuple(r) for r in reqs]
        arr, outs = self._reqs(reqs, want_cnt, want_norm, W0=2 * nc10 + 1)
        cid, t = C.c_int(), C.c_uint64()
        self._c(_lib.amb_submit_cube(self._h, C.byref(src), arr, len(reqs), C.byref(cid), C.byref(t)))
        self._meta[cid.value] = {"N": int(N), "nc": int(nc), "nc10": nc10}
        self._tickets[t.value] = (arr, outs, gv, samp, ch, src)
        return t.value, cid.value

    def poll(self, ticket):
        d = C.c_int()
        self._c(_lib.amb_poll(self._h, C.c_uint64(ticket), C.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        """-> (outputs, latency_ms)"""
        lat = C.c_double()
        rc = _lib.amb_wait(self._h, C.c_uint64(ticket), C.byref(lat))
        outs = self._tickets.pop(ticket)[1]
        self._c(rc)
        return outs, lat.value

    def serve_stop(self):
        self._c(_lib.amb_serve_stop(self._h))

    def cyl_bispec(self, bl, den, window=False, **kw):
        """Lag-space post-processing (get_cyl_BS / cylbs_win) of one or more
        triangles' W x W arrays: den is the int32 count array (v1 files) or
        the float32 normalisation (v2 files).  Keyword arguments override the
        reference analysis constants (ncut, mabs_norm, rc, rc_prime, dnu).
        Returns float64 (n, N, N) (or (N, N) for one triangle), N = W - 2 ncut."""
        bl = np.ascontiguousarray(bl, np.float32)
        one = bl.ndim == 2
        bl = bl.reshape((-1,) + bl.shape[-2:])
        den = np.asarray(den)
        is_int = np.issubdtype(den.dtype, np.integer)
        den = np.ascontiguousarray(den, np.int32 if is_int else np.float32).reshape(bl.shape)
        k = dict(ANALYSIS)
        k.update(kw)
        n, W = bl.shape[0], bl.shape[1]
        N = W - 2 * int(k["ncut"])
        out = np.zeros((n, N, N))
        p = CylParams(int(k["ncut"]), float(k["mabs_norm"]), float(k["rc"]), float(k["rc_prime"]),
                      float(k["dnu"]), int(bool(window)))
        self._c(_lib.amb_cyl_bispec(self._h, n, W, _p(bl, _fp), den.ctypes.data_as(C.c_void_p), int(is_int),
                                    C.byref(p), _p(out, _dp)))
        return out[0] if one else out

    def maps_cldnu(self, N, nc, Du, gv, samp, NB=15, umin=1.0, umax=500.0):
        """MAPS C_l(dnu) of a cube in the file layout (maps.c:100-229 over
        cl_dnu): returns (cldnu[NB][nc], numdnu[nc], binu[NB], num[NB])."""
        gv = np.ascontiguousarray(gv, np.float64)
        samp = np.ascontiguousarray(samp, np.float64)
        npix = N * (N // 2 + 1)
        if gv.size != npix * nc * 2 or samp.size != npix * nc:
            raise ValueError("cube arrays do not match N=%d nc=%d" % (N, nc))
        cl = np.zeros((NB, nc)); nd = np.zeros(nc); bu = np.zeros(NB); nm = np.zeros(NB, np.int32)
        self._c(_lib.amb_maps_cldnu(self._h, N, nc, float(Du), _p(gv, _dp), _p(samp, _dp), NB, float(umin),
                                    float(umax), _p(cl, _dp), _p(nd, _dp), _p(bu, _dp), _p(nm, _ip)))
        return cl, nd, bu, nm

    def grid(self, NN, nc, ranpar, data, nu_chan0, dU, Umax, flag_key=1, n_collapse=1, chan1=1):
        """vcg gridding on the GPU: (gv [NN(NN/2+1)][nc/n_collapse][2], samp, groups gridded)"""
        cfg, ranpar, data = _grid_cfg(NN, nc, ranpar, data, nu_chan0, dU, Umax, flag_key, n_collapse, chan1)
        cells = int(NN) * (int(NN) // 2 + 1)
        ncout = int(nc) // int(n_collapse)
        gv = np.empty((cells, ncout, 2)); samp = np.empty((cells, ncout))
        n = C.c_int64()
        self._c(_lib.amb_grid(self._h, C.byref(cfg), len(ranpar), _p(ranpar, _fp), _p(data, _fp), _p(gv, _dp),
                              _p(samp, _dp), C.byref(n)))
        return gv, samp, n.value

    def grid_cube(self, NN, nc, ranpar, data, nu_chan0, dU, Umax, flag_key=1, n_collapse=1, chan1=1):
        """vcg gridding kept resident as a cube: (cube id, groups gridded)"""
        cfg, ranpar, data = _grid_cfg(NN, nc, ranpar, data, nu_chan0, dU, Umax, flag_key, n_collapse, chan1)
        cid = C.c_int(); n = C.c_int64()
        self._c(_lib.amb_grid_cube(self._h, C.byref(cfg), len(ranpar), _p(ranpar, _fp), _p(data, _fp),
                                   C.byref(cid), C.byref(n)))
        self._meta[cid.value] = dict(N=int(NN), nc=int(nc) // int(n_collapse), Du=float(dU), flag_key=int(flag_key))
        return cid.value, n.value

    def batch_times(self):
        """completion time (ms from the call's start) of each triangle of the last batch"""
        n = C.c_int()
        self._c(_lib.amb_batch_times(self._h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1))
        self._c(_lib.amb_batch_times(self._h, _p(out, _dp), n.value, C.byref(n)))
        return out[: n.value]

    def stats(self):
        s = Stats()
        self._c(_lib.amb_get_stats(self._h, C.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        self._c(_lib.amb_reset_stats(self._h))

    def set_cert_strict(self, strict=True):
        """v2 normalisation sums not certified exact (FP32 accumulation error
        past 1/32 of their 1/8 grid): counted in stats() (default) or, when
        strict, an AmberError"""
        self._c(_lib.amb_set_cert_strict(self._h, int(bool(strict))))

    def set_grid(self, reduced=True):
        """sums on each triangle's smallest alias-free grid (compute_grid,
        default) or on the reference's Ng(r1); the split-BF16 and FP8 forms
        always use Ng(r1)"""
        self._c(_lib.amb_set_grid(self._h, int(bool(reduced))))


class Graph:
    """amb_graph_create / amb_graph_run / amb_graph_destroy.  run() replays
    the captured batch and returns the outputs [(bl, cnt, norm), ...]; the
    arrays are reused by every replay."""

    def __init__(self, ctx, reqs, want_cnt=True, want_norm=False):
        self._ctx = ctx
        self._arr, self.outputs = ctx._reqs(reqs, want_cnt, want_norm)
        self._g = C.c_void_p()
        ctx._c(_lib.amb_graph_create(ctx._h, self._arr, len(reqs), C.byref(self._g)))

    def run(self):
        self._ctx._c(_lib.amb_graph_run(self._ctx._h, self._g))
        return self.outputs

    def close(self):
        if self._g and self._ctx._h:
            _lib.amb_graph_destroy(self._ctx._h, self._g)
        self._g = C.c_void_p()

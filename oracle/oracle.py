"""ctypes front end of the FP64 CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py as the checker / timed CPU baseline.
The f_number_amd product never imports it.

The oracle restates the reference's MATLAB CPU path (das_pw.m:302-485,
das_sa.m:289-509).  Images are returned in the CPU path's scale
(``2 * sum``, das_pw.m:384,467) as complex128 arrays shaped [n_z, n_x]
(MATLAB's column-major [N_z x N_x] image).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FN_ORACLE_LIB loads another build of oracle.c (the sanitizer test's instrumented one)
_LIB_PATH = os.environ.get("FN_ORACLE_LIB") or os.path.join(_HERE, "build", "liboracle.so")

WIN = {"boxcar": 0, "hann": 1, "tukey": 2, "cosine": 3, "triangular": 4}
SEM = {"cpu": 0, "mex": 1}
NORM = {"off": 0, "on": 1, "windowed": 2}
FN = {
    "constant": 0,
    "angle_lb": 1,
    "distance_constant": 2,
    "szabo": 3,
    "directivity_soft": 4,
    "directivity_rigid": 5,
    "perrot": 6,
    "distance_none": 7,
}


class _FNumber(C.Structure):
    _fields_ = [("kind", C.c_int), ("p0", C.c_double), ("p1", C.c_double), ("p2", C.c_double)]


class _Args(C.Structure):
    _fields_ = [
        ("n_x", C.c_int), ("n_z", C.c_int),
        ("x", C.POINTER(C.c_double)), ("z", C.POINTER(C.c_double)),
        ("n_t", C.c_int), ("n_el", C.c_int), ("n_tx", C.c_int),
        ("rf", C.c_void_p), ("rf_f32", C.c_int),
        ("f_s", C.c_double), ("width", C.c_double), ("pitch", C.c_double), ("c_0", C.c_double),
        ("f_lb", C.c_double), ("f_ub", C.c_double), ("index_t0", C.c_double),
        ("angles", C.POINTER(C.c_double)),
        ("win_rx", C.c_int), ("win_rx_param", C.c_double),
        ("f_rx", _FNumber),
        ("norm_rx", C.c_int), ("norm_rx_win", C.c_int), ("norm_rx_win_param", C.c_double),
        ("win_tx", C.c_int), ("win_tx_param", C.c_double),
        ("f_tx", _FNumber),
        ("norm_tx", C.c_int), ("norm_tx_win", C.c_int), ("norm_tx_win_param", C.c_double),
        ("semantics", C.c_int), ("spectra_mode", C.c_int), ("n_threads", C.c_int),
        ("n_sel", C.c_int), ("sel_ix", C.POINTER(C.c_int)), ("sel_iz", C.POINTER(C.c_int)),
        ("image_re", C.POINTER(C.c_double)), ("image_im", C.POINTER(C.c_double)),
        ("N_dft", C.c_int), ("n_lb", C.c_int), ("n_ub", C.c_int),
        ("signal_re", C.POINTER(C.c_double)), ("signal_im", C.POINTER(C.c_double)),
        ("virtual_sources", C.POINTER(C.c_double)),
        ("n_threads_spectra", C.c_int),
    ]


_lib = None


def build() -> str:
    """Compile oracle.c with its Makefile (gcc, no GPU involved)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def build_variant(build_dir: str, plant: str) -> str:
    """Compile oracle.c with a planted mistake (``-DOR_PLANT_...``) into build_dir.

    Only the aperture tests use this, to show that their pins fail on such a
    mistake; the default library is never affected."""
    subprocess.run(["make", "-s", "-C", _HERE, f"BUILD={build_dir}", f"PLANT={plant}"], check=True)
    return os.path.join(build_dir, "liboracle.so")


@contextlib.contextmanager
def using(path: str):
    """Route every oracle call inside the block through the library at ``path``."""
    global _lib
    saved = _lib
    _lib = _configure(C.CDLL(path))
    try:
        yield
    finally:
        _lib = saved


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _configure(C.CDLL(_LIB_PATH))
    return _lib


def _configure(L):
    d = C.POINTER(C.c_double)
    i = C.POINTER(C.c_int)
    L.or_grid.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                          C.c_double, C.c_double, i, i, i]
    L.or_fnumber_values.argtypes = [C.POINTER(_FNumber), d, C.c_int, d]
    L.or_window.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int]
    L.or_window.restype = C.c_double
    L.or_spectra.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                             C.c_int, C.c_int, d, d]
    L.or_das_pw.argtypes = [C.POINTER(_Args)]
    L.or_das_sa.argtypes = [C.POINTER(_Args)]
    L.or_das_pw_c.argtypes = [C.POINTER(_Args), C.c_double, d]
    L.or_das_pw_diff.argtypes = [C.POINTER(_Args), C.c_double, d, C.POINTER(_Diff)]
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fnumber(kind: str, p0=0.0, p1=0.0, p2=0.0) -> _FNumber:
    return _FNumber(FN[kind], float(p0), float(p1), float(p2))


def grid(n_t, n_el, pitch, z_pad, f_s, c_0, f_lb, f_ub):
    N, lb, ub = C.c_int(), C.c_int(), C.c_int()
    rc = lib().or_grid(n_t, n_el, pitch, z_pad, f_s, c_0, f_lb, f_ub, C.byref(N), C.byref(lb), C.byref(ub))
    if rc:
        raise ValueError("empty frequency band")
    return N.value, lb.value, ub.value


def fnumber_values(F: _FNumber, r) -> np.ndarray:
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty_like(r)
    rc = lib().or_fnumber_values(C.byref(F), _dp(r), r.size, _dp(out))
    if rc:
        raise ValueError(f"or_fnumber_values failed ({rc})")
    return out


def window(name: str, param: float, t, semantics: str = "cpu") -> np.ndarray:
    t = np.atleast_1d(np.asarray(t, dtype=np.float64))
    L = lib()
    return np.array([L.or_window(WIN[name], float(param), float(v), SEM[semantics]) for v in t])


def spectra(rf: np.ndarray, N: int, n_lb: int, n_ub: int, mode: int = 1, n_threads: int = 0) -> np.ndarray:
    """rf: [n_ch, n_t] (C order, i.e. MATLAB column-major [n_t x n_ch]).  Returns [n_ch, n_f] complex."""
    rf = np.ascontiguousarray(rf)
    f32 = rf.dtype == np.float32
    if not f32:
        rf = np.ascontiguousarray(rf, dtype=np.float64)
    n_ch, n_t = rf.shape
    nf = n_ub - n_lb + 1
    re = np.empty((n_ch, nf))
    im = np.empty((n_ch, nf))
    rc = lib().or_spectra(rf.ctypes.data, int(f32), n_t, n_ch, N, n_lb, n_ub, mode, n_threads, _dp(re), _dp(im))
    if rc:
        raise ValueError("or_spectra failed")
    return re + 1j * im


def _args(x, z, rf, f_s, width, pitch, c_0, f_bounds, index_t0, semantics, spectra_mode, n_threads, sel):
    x = np.ascontiguousarray(x, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    rf = np.ascontiguousarray(rf)
    if rf.dtype != np.float32:
        rf = np.ascontiguousarray(rf, dtype=np.float64)
    a = _Args()
    a.n_x, a.n_z = x.size, z.size
    a.x, a.z = _dp(x), _dp(z)
    a.rf = rf.ctypes.data
    a.rf_f32 = int(rf.dtype == np.float32)
    a.f_s, a.width, a.pitch, a.c_0 = f_s, width, pitch, c_0
    a.f_lb, a.f_ub = f_bounds
    a.index_t0 = index_t0
    a.semantics = SEM[semantics]
    a.spectra_mode = spectra_mode
    a.n_threads = n_threads
    keep = [x, z, rf]
    if sel is not None:
        ix = np.ascontiguousarray(sel[0], dtype=np.int32)
        iz = np.ascontiguousarray(sel[1], dtype=np.int32)
        a.n_sel = ix.size
        a.sel_ix = ix.ctypes.data_as(C.POINTER(C.c_int))
        a.sel_iz = iz.ctypes.data_as(C.POINTER(C.c_int))
        keep += [ix, iz]
        n_out = ix.size
    else:
        a.n_sel = 0
        n_out = x.size * z.size
    re = np.zeros(n_out)
    im = np.zeros(n_out)
    a.image_re, a.image_im = _dp(re), _dp(im)
    keep += [re, im]
    return a, keep, re, im


def das_pw(x, z, rf, f_s, angles, width, pitch, c_0, f_bounds, index_t0=0.0,
           window=("tukey", 0.2), F=None, normalization="on", semantics="cpu",
           spectra_mode=1, n_threads=0, sel=None, norm_window=("boxcar", 0.0),
           c_tof=None, delay_offsets=None, want_signal=False, virtual_sources=None, n_threads_spectra=0):
    """rf: [n_tx, n_el, n_t] C-order (MATLAB [n_t x n_el x n_tx]).  Returns (image, info).

    image is [n_z, n_x] complex (or [n_sel] with ``sel=(ix, iz)``), CPU scale.
    want_signal: info['signal'] = focused RF signal of the last voxel, [N_dft] complex.
    n_threads_spectra: threads of the spectra stage (0: n_threads)."""
    rf = np.asarray(rf)
    if rf.ndim == 2:
        rf = rf[None]
    n_tx, n_el, n_t = rf.shape
    angles = np.ascontiguousarray(np.atleast_1d(angles), dtype=np.float64)
    assert angles.size == n_tx
    a, keep, re, im = _args(x, z, rf, f_s, width, pitch, c_0, f_bounds, index_t0, semantics,
                            spectra_mode, n_threads, sel)
    a.n_t, a.n_el, a.n_tx = n_t, n_el, n_tx
    a.n_threads_spectra = int(n_threads_spectra)
    a.angles = _dp(angles)
    a.win_rx, a.win_rx_param = WIN[window[0]], float(window[1])
    a.f_rx = F if F is not None else fnumber("angle_lb", 60, 1.5, 10)
    a.norm_rx = NORM[normalization]
    a.norm_rx_win, a.norm_rx_win_param = WIN[norm_window[0]], float(norm_window[1])
    sig = _signal_buffers(a, keep, n_t, want_signal and sel is None)
    _virtual_sources(a, keep, virtual_sources, n_tx)
    if c_tof is None:
        rc = lib().or_das_pw(C.byref(a))
    else:
        offs = None
        if delay_offsets is not None:
            offs = np.ascontiguousarray(delay_offsets, dtype=np.float64)
            keep.append(offs)
        rc = lib().or_das_pw_c(C.byref(a), float(c_tof), _dp(offs) if offs is not None else None)
    if rc:
        raise ValueError(f"or_das_pw failed ({rc})")
    img = re + 1j * im
    if sel is None:
        img = img.reshape(np.asarray(x).size, np.asarray(z).size).T
    info = dict(N_dft=a.N_dft, n_lb=a.n_lb, n_ub=a.n_ub)
    if sig is not None:
        info["signal"] = (sig[0] + 1j * sig[1])[:a.N_dft]
    return img, info


def _virtual_sources(a, keep, vs, n_tx):
    a.virtual_sources = None
    if vs is not None:
        vs = np.ascontiguousarray(np.asarray(vs, dtype=np.float64).reshape(n_tx, 2))
        a.virtual_sources = _dp(vs)
        keep.append(vs)


def _signal_buffers(a, keep, n_t, want):
    """signal outputs of the DFT order of the call's grid (das_pw.m:343-353)"""
    a.signal_re = a.signal_im = None
    if not want:
        return None
    z = np.ctypeslib.as_array(a.z, shape=(a.n_z,))
    z_pad = z[1] if (a.semantics == SEM["mex"] and a.n_z > 1) else z[0]
    N, _, _ = grid(int(n_t), a.n_el, a.pitch, float(z_pad), a.f_s, a.c_0, a.f_lb, a.f_ub)
    sr, si = np.zeros(N), np.zeros(N)
    a.signal_re, a.signal_im = _dp(sr), _dp(si)
    keep += [sr, si]
    return sr, si


class _Diff(C.Structure):
    _fields_ = [("dIdc_re", C.POINTER(C.c_double)), ("dIdc_im", C.POINTER(C.c_double)),
                ("g_re", C.POINTER(C.c_double)), ("g_im", C.POINTER(C.c_double)),
                ("vjp_delay", C.POINTER(C.c_double)), ("vjp_c", C.c_double),
                ("y_re", C.POINTER(C.c_double)), ("y_im", C.POINTER(C.c_double)),
                ("adj_out", C.POINTER(C.c_double))]


def das_pw_diff(x, z, rf, f_s, angles, width, pitch, c_0, f_bounds, index_t0=0.0,
                window=("tukey", 0.2), F=None, normalization="on", semantics="cpu",
                spectra_mode=1, n_threads=0, norm_window=("boxcar", 0.0),
                c_tof=None, delay_offsets=None, g=None, y=None, want_dc=False, virtual_sources=None,
                sel=None):
    """FP64 derivatives of the PW image (oracle.h or_das_pw_diff).

    g, y: complex [n_z, n_x] (image layout).  Returns a dict with 'image',
    'dIdc' (want_dc), 'vjp_delay' and 'vjp_c' (g given; vjp_c also needs
    want_dc), 'adjoint' [n_tx, n_el, n_t] (y given).
    sel=(ix, iz): only those voxels; g and y are then [n_sel] in that order (the
    cotangent is zero elsewhere) and 'image' / 'dIdc' are [n_sel]."""
    rf = np.asarray(rf)
    if rf.ndim == 2:
        rf = rf[None]
    n_tx, n_el, n_t = rf.shape
    angles = np.ascontiguousarray(np.atleast_1d(angles), dtype=np.float64)
    a, keep, re, im = _args(x, z, rf, f_s, width, pitch, c_0, f_bounds, index_t0, semantics,
                            spectra_mode, n_threads, sel)
    a.n_t, a.n_el, a.n_tx = n_t, n_el, n_tx
    a.angles = _dp(angles)
    a.win_rx, a.win_rx_param = WIN[window[0]], float(window[1])
    a.f_rx = F if F is not None else fnumber("angle_lb", 60, 1.5, 10)
    a.norm_rx = NORM[normalization]
    a.norm_rx_win, a.norm_rx_win_param = WIN[norm_window[0]], float(norm_window[1])
    nx, nz = np.asarray(x).size, np.asarray(z).size
    nv = nx * nz if sel is None else np.asarray(sel[0]).size
    a.signal_re = a.signal_im = None
    _virtual_sources(a, keep, virtual_sources, n_tx)
    d = _Diff()
    out = {}

    def vox(img):  # [n_z, n_x] -> oracle order (x-major, z fastest); [n_sel] as given
        img = np.asarray(img, dtype=np.complex128)
        img = img.ravel() if sel is not None else img.T.ravel()
        r, i = np.ascontiguousarray(img.real), np.ascontiguousarray(img.imag)
        keep.extend([r, i])
        return r, i

    if want_dc:
        dr, di = np.zeros(nv), np.zeros(nv)
        d.dIdc_re, d.dIdc_im = _dp(dr), _dp(di)
    if g is not None:
        gr, gi = vox(g)
        d.g_re, d.g_im = _dp(gr), _dp(gi)
        vj = np.zeros(n_el)
        d.vjp_delay = _dp(vj)
    if y is not None:
        yr, yi = vox(y)
        d.y_re, d.y_im = _dp(yr), _dp(yi)
        adj = np.zeros(n_tx * n_el * n_t)
        d.adj_out = _dp(adj)
    offs = None
    if delay_offsets is not None:
        offs = np.ascontiguousarray(delay_offsets, dtype=np.float64)
    rc = lib().or_das_pw_diff(C.byref(a), float(c_0 if c_tof is None else c_tof),
                              _dp(offs) if offs is not None else None, C.byref(d))
    if rc:
        raise ValueError(f"or_das_pw_diff failed ({rc})")
    shape = (lambda v: v) if sel is not None else (lambda v: v.reshape(nx, nz).T)
    out["image"] = shape(re + 1j * im)
    if want_dc:
        out["dIdc"] = shape(dr + 1j * di)
    if g is not None:
        out["vjp_delay"] = vj
        if want_dc:
            out["vjp_c"] = d.vjp_c
    if y is not None:
        out["adjoint"] = adj.reshape(n_tx, n_el, n_t)
    out["info"] = dict(N_dft=a.N_dft, n_lb=a.n_lb, n_ub=a.n_ub)
    return out


def das_sa(x, z, rf, f_s, width, pitch, c_0, f_bounds, index_t0=0.0,
           window_tx=("tukey", 0.2), window_rx=("tukey", 0.2), F_tx=None, F_rx=None,
           normalization_tx="on", normalization_rx="on", semantics="cpu", spectra_mode=1,
           n_threads=0, sel=None, norm_window=("boxcar", 0.0), want_signal=False):
    """rf: [n_tx, n_rx, n_t] C-order (MATLAB [n_t x n_rx x n_tx])."""
    rf = np.asarray(rf)
    n_tx, n_rx, n_t = rf.shape
    assert n_tx == n_rx
    a, keep, re, im = _args(x, z, rf, f_s, width, pitch, c_0, f_bounds, index_t0, semantics,
                            spectra_mode, n_threads, sel)
    a.n_t, a.n_el, a.n_tx = n_t, n_rx, n_tx
    a.win_tx, a.win_tx_param = WIN[window_tx[0]], float(window_tx[1])
    a.win_rx, a.win_rx_param = WIN[window_rx[0]], float(window_rx[1])
    a.f_tx = F_tx if F_tx is not None else fnumber("angle_lb", 60, 1.5, 10)
    a.f_rx = F_rx if F_rx is not None else fnumber("angle_lb", 60, 1.5, 10)
    a.norm_tx = NORM[normalization_tx]
    a.norm_rx = NORM[normalization_rx]
    a.norm_tx_win = a.norm_rx_win = WIN[norm_window[0]]
    a.norm_tx_win_param = a.norm_rx_win_param = float(norm_window[1])
    sig = _signal_buffers(a, keep, n_t, want_signal and sel is None)
    rc = lib().or_das_sa(C.byref(a))
    if rc:
        raise ValueError(f"or_das_sa failed ({rc})")
    img = re + 1j * im
    if sel is None:
        img = img.reshape(np.asarray(x).size, np.asarray(z).size).T
    info = dict(N_dft=a.N_dft, n_lb=a.n_lb, n_ub=a.n_ub)
    if sig is not None:
        info["signal"] = (sig[0] + 1j * sig[1])[:a.N_dft]
    return img, info

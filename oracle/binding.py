"""ctypes binding of liboracle.so — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module. It is
the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def _declare(L):
    L.orc_threads.argtypes = [C.c_int]
    L.orc_threads.restype = C.c_int
    L.orc_fft.argtypes = [_dp, C.c_int, C.c_int, C.c_int]
    L.orc_fft3d.argtypes = [_dp, C.c_int, C.c_int]
    L.orc_find_sideband.argtypes = [_fp, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip]
    L.orc_retrieve.argtypes = [_fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               _dp, _dp, _dp, _dp, _dp]
    L.orc_dct_chain.argtypes = [_dp, C.c_int, C.c_uint, C.c_int, _dp, _dp, _dp, _dp]
    L.orc_unwrap_dct.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_uint]
    for name, args in _EXTRA.items():
        if hasattr(L, name):
            getattr(L, name).argtypes = args


_EXTRA: dict = {
    "orc_rytov_spectrum": [_dp, _dp, C.c_int, C.c_int, _dp],
    "orc_ewald_scatter": [_dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp],
    "orc_ewald_gather": [_dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp],
    "orc_ewald_index": [_dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _ip],
    "orc_normalize": [_dp, _dp, _dp, C.c_int, C.c_double],
    "orc_est_n": [_dp, _dp, C.c_long, C.c_double, C.c_double],
    "orc_modify_f": [_dp, _dp, C.c_long, C.c_double, C.c_double],
    "orc_positivity": [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint],
    "orc_residues": [_fp, C.c_int, C.c_int, C.c_void_p],
    "orc_branch_cuts": [C.c_void_p, C.c_int, C.c_int, C.c_void_p],
    "orc_flood_fill": [_fp, C.c_void_p, C.c_int, C.c_int, _ip, _fp],
    "orc_goldstein": [_fp, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, _ip, _fp],
    "orc_fit_pgd": [_dp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                    C.c_double, _dp, C.c_uint, _dp],
    "orc_find_sideband_rows": [_fp, C.c_int, C.c_int],
    "orc_retrieve_rows": [_fp, _fp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp],
    "orc_median3x3": [_fp, _fp, C.c_int, C.c_int, C.c_int],
    "orc_check_maps": [C.c_void_p, C.c_void_p, C.c_long],
    "orc_recon_from_maps": [_fp, _fp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                            C.c_uint, _dp],
    "orc_reconstruct_cell": [_fp, _fp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                             C.c_double, C.c_double, C.c_double, C.c_int, C.c_uint, C.c_uint, _dp,
                             _dp, _dp, _dp],
}


def _p(a: np.ndarray | None, t):
    return None if a is None else a.ctypes.data_as(t)


def threads(n: int = 0) -> int:
    return lib().orc_threads(n)


def fft(x: np.ndarray, direction: int) -> np.ndarray:
    """Unnormalised 1-D transform along the last axis; direction -1 forward, +1 inverse."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n = a.shape[-1]
    lib().orc_fft(a.view(np.float64).ctypes.data_as(_dp), n, a.size // n, direction)
    return a


def fft3d(vol: np.ndarray, direction: int) -> np.ndarray:
    a = np.ascontiguousarray(vol, dtype=np.complex128).copy()
    lib().orc_fft3d(a.view(np.float64).ctypes.data_as(_dp), a.shape[0], direction)
    return a


def find_sideband(bg: np.ndarray, row_start: int = 0):
    b = np.ascontiguousarray(bg, dtype=np.float32)
    H, W = b.shape
    r, c, rad = C.c_int(), C.c_int(), C.c_int()
    lib().orc_find_sideband(_p(b, _fp), H, W, row_start, C.byref(r), C.byref(c), C.byref(rad))
    return r.value, c.value, rad.value


def retrieve(holo: np.ndarray, bg: np.ndarray, row: int, col: int, radius: int, M: int):
    """Returns dict(phase_ref, amp, wrapped, logamp, field) as float64 / complex128 M x M."""
    h = np.ascontiguousarray(holo, dtype=np.float32)
    b = np.ascontiguousarray(bg, dtype=np.float32)
    H, W = h.shape
    out = {k: np.zeros((M, M)) for k in ("phase_ref", "amp", "wrapped", "logamp")}
    field = np.zeros((M, M), np.complex128)
    lib().orc_retrieve(_p(h, _fp), _p(b, _fp), H, W, row, col, radius, M,
                       _p(out["phase_ref"], _dp), _p(out["amp"], _dp), _p(out["wrapped"], _dp),
                       _p(out["logamp"], _dp), field.view(np.float64).ctypes.data_as(_dp))
    out["field"] = field
    return out


SINCOS, REF_EIG, REF_IDCT, CONGRUENT = 1, 2, 4, 8


def dct_chain(x: np.ndarray, flags: int = 0, stage_in: int = 0):
    """Stages (laplace, dct, dct2, out) of the DCT unwrapper starting at `stage_in`."""
    a = np.ascontiguousarray(x, dtype=np.float64)
    N = a.shape[0]
    outs = [np.zeros((N, N)) for _ in range(4)]
    lib().orc_dct_chain(_p(a, _dp), N, flags, stage_in, *[_p(o, _dp) for o in outs])
    return dict(zip(("laplace", "dct", "dct2", "out"), outs))


def unwrap_dct(x: np.ndarray, flags: int = 0) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.ndim == 2:
        a = a[None]
    out = np.zeros_like(a)
    lib().orc_unwrap_dct(_p(a, _dp), _p(out, _dp), a.shape[0], a.shape[-1], flags)
    return out


# ---------------------------------------------------------------- reconstruction
def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _cp(a):
    return None if a is None else a.view(np.float64).ctypes.data_as(_dp)


def rytov_spectrum(phase, logamp):
    ph = np.ascontiguousarray(phase, np.float64)
    la = np.ascontiguousarray(logamp, np.float64)
    A, M = ph.shape[0], ph.shape[-1]
    out = np.zeros((A, M, M), np.complex128)
    lib().orc_rytov_spectrum(_p(ph, _dp), _p(la, _dp), A, M, _cp(out))
    return out


def _ang(angles):
    return np.ascontiguousarray(np.asarray(angles, np.float64).reshape(-1, 2))


def ewald_scatter(spec, angles, V, wavelength, n_med, dx, F=None, hits=None):
    s = _c(spec)
    a = _ang(angles)
    F = np.zeros((V, V, V), np.complex128) if F is None else _c(F).copy()
    hits = np.zeros((V, V, V)) if hits is None else np.ascontiguousarray(hits, np.float64).copy()
    lib().orc_ewald_scatter(_cp(s), _p(a, _dp), len(a), V, wavelength, n_med, dx, _cp(F), _p(hits, _dp))
    return F, hits


def ewald_gather(F, angles, wavelength, n_med, dx):
    Fc = _c(F)
    a = _ang(angles)
    V = Fc.shape[0]
    out = np.zeros((len(a), V, V), np.complex128)
    lib().orc_ewald_gather(_cp(Fc), _p(a, _dp), len(a), V, wavelength, n_med, dx, _cp(out))
    return out


def ewald_index(angles, V, wavelength, n_med, dx):
    a = _ang(angles)
    out = np.zeros((len(a), V, V), np.int32)
    lib().orc_ewald_index(_p(a, _dp), len(a), V, wavelength, n_med, dx, _p(out, _ip))
    return out


def normalize(F, hits, dx):
    Fc = _c(F).copy()
    h = np.ascontiguousarray(hits, np.float64)
    Fm = np.zeros_like(Fc)
    lib().orc_normalize(_cp(Fc), _p(h, _dp), _cp(Fm), Fc.shape[0], dx)
    return Fc, Fm


def est_n(f, wavelength, n_med):
    fc = _c(f)
    out = np.zeros_like(fc)
    lib().orc_est_n(_cp(fc), _cp(out), fc.size, wavelength, n_med)
    return out


def modify_f(n, wavelength, n_med):
    nc = _c(n)
    out = np.zeros_like(nc)
    lib().orc_modify_f(_cp(nc), _cp(out), nc.size, wavelength, n_med)
    return out


def positivity(f, Fm, hits, iters, wavelength, n_med, flags=0):
    fc = _c(f).copy()
    mc = _c(Fm)
    h = None if hits is None else np.ascontiguousarray(hits, np.float64)
    n = np.zeros_like(fc)
    lib().orc_positivity(_cp(fc), _cp(mc), _p(h, _dp), _cp(n), fc.shape[0], iters, wavelength, n_med, flags)
    return n


def reconstruct_cell(holo, bg, angles, V, crop, wavelength, n_med, dx, iters, unwrap_flags=0,
                     recon_flags=0, outputs=False):
    h = np.ascontiguousarray(holo, np.float32)
    b = np.ascontiguousarray(bg, np.float32)
    a = _ang(angles)
    A, H = h.shape[0], h.shape[-1]
    ri = np.zeros((V, V, V), np.complex128)
    extra = [np.zeros((A, V, V)) for _ in range(3)] if outputs else [None] * 3
    row, col, rad = crop[:3]
    lib().orc_reconstruct_cell(_p(h, _fp), _p(b, _fp), _p(a, _dp), A, H, V, row, col, rad, wavelength,
                               n_med, dx, iters, unwrap_flags, recon_flags, _cp(ri),
                               *[_p(e, _dp) for e in extra])
    if outputs:
        return ri, dict(zip(("wrapped", "unwrapped", "logamp"), extra))
    return ri


# ---------------------------------------------------------------- Goldstein
def residues(phi):
    p = np.ascontiguousarray(phi, np.float32)
    H, W = p.shape
    out = np.zeros((H, W), np.int8)
    lib().orc_residues(_p(p, _fp), H, W, out.ctypes.data)
    return out


def branch_cuts(res):
    r = np.ascontiguousarray(res, np.int8)
    H, W = r.shape
    out = np.zeros((H, W), np.uint8)
    lib().orc_branch_cuts(r.ctypes.data, H, W, out.ctypes.data)
    return out


def flood_fill(phi, cuts):
    p = np.ascontiguousarray(phi, np.float32)
    c = np.ascontiguousarray(cuts, np.uint8)
    H, W = p.shape
    k = np.zeros((H, W), np.int32)
    un = np.zeros((H, W), np.float32)
    lib().orc_flood_fill(_p(p, _fp), c.ctypes.data, H, W, _p(k, _ip), _p(un, _fp))
    return k, un


def goldstein(phi):
    p = np.ascontiguousarray(phi, np.float32)
    if p.ndim == 2:
        p = p[None]
    B, H, W = p.shape
    res = np.zeros((B, H, W), np.int8)
    cut = np.zeros((B, H, W), np.uint8)
    k = np.zeros((B, H, W), np.int32)
    un = np.zeros((B, H, W), np.float32)
    lib().orc_goldstein(_p(p, _fp), B, H, W, res.ctypes.data, cut.ctypes.data, _p(k, _ip), _p(un, _fp))
    return res, cut, k, un


def fit_pgd(f0, spec, angles, wavelength, n_med, dx, iters, step, flags=0, objective=False):
    """n (and, with objective=True, J before each step as an array of iters values)."""
    f = _c(f0)
    sp = _c(spec)
    a = _ang(angles)
    n = np.zeros_like(f)
    J = np.zeros(max(iters, 1))
    lib().orc_fit_pgd(_cp(f), _cp(sp), _p(a, _dp), len(a), f.shape[0], wavelength, n_med, dx, iters, step,
                      _cp(n), flags, _p(J, _dp) if objective else None)
    return (n, J[:iters]) if objective else n


def median3x3(maps):
    m = np.ascontiguousarray(maps, np.float32)
    m3 = m.reshape((-1,) + m.shape[-2:])
    out = np.empty_like(m3)
    lib().orc_median3x3(_p(m3, _fp), _p(out, _fp), m3.shape[0], m3.shape[1], m3.shape[2])
    return out.reshape(m.shape)


def check_maps(phase=None, amp=None):
    """Returns sanitised copies (phase, amp) as float32."""
    ph = None if phase is None else np.ascontiguousarray(phase, np.float32).copy()
    am = None if amp is None else np.ascontiguousarray(amp, np.float32).copy()
    n = (ph if ph is not None else am).size
    lib().orc_check_maps(None if ph is None else ph.ctypes.data, None if am is None else am.ctypes.data, n)
    return ph, am


def recon_from_maps(unwrapped, map2, angles, wavelength, n_med, dx, iters, flags=0):
    u = np.ascontiguousarray(unwrapped, np.float32)
    m = np.ascontiguousarray(map2, np.float32)
    a = _ang(angles)
    A, V = u.shape[0], u.shape[1]
    ri = np.zeros((V, V, V), np.complex128)
    lib().orc_recon_from_maps(_p(u, _fp), _p(m, _fp), _p(a, _dp), A, V, wavelength, n_med, dx, iters, flags, _cp(ri))
    return ri


def find_sideband_rows(bg):
    b = np.ascontiguousarray(bg, np.float32)
    H, W = b.shape
    L = lib()
    L.orc_find_sideband_rows.restype = C.c_int
    return int(L.orc_find_sideband_rows(_p(b, _fp), H, W))


def retrieve_rows(holo, bg, center):
    """Algorithm B maps, shape (W/4, H/4) indexed [n][y'] (the reference's transposed layout)."""
    h = np.ascontiguousarray(holo, np.float32)
    b = np.ascontiguousarray(bg, np.float32)
    H, W = h.shape
    M, N2 = W // 4, H // 4
    w = np.zeros((M, N2))
    la = np.zeros((M, N2))
    ph = np.zeros((M, N2))
    lib().orc_retrieve_rows(_p(h, _fp), _p(b, _fp), H, W, center, _p(w, _dp), _p(la, _dp), _p(ph, _dp))
    return {"wrapped": w, "logamp": la, "phase_ref": ph}


def angle_from_image(img, pixel_um, mag, wavelength_nm, n_med):
    """Returns (angle [rad], shifted column of the peak)."""
    a = np.ascontiguousarray(img, np.float32)
    H, W = a.shape
    L = lib()
    L.orc_angle_from_image.argtypes = [_fp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _ip]
    L.orc_angle_from_image.restype = C.c_double
    pos = np.zeros(1, np.int32)
    ang = L.orc_angle_from_image(_p(a, _fp), H, W, pixel_um, mag, wavelength_nm, n_med, _p(pos, _ip))
    return float(ang), int(pos[0])

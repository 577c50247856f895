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


_EXTRA: dict = {}


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

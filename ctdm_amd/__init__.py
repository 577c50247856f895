"""Python binding of libctdm (the MI355X-native cTDM reconstruction path).

The library is C-ABI (include/ctdm.h); this module loads it with ctypes and passes device pointers
taken from torch tensors (torch is used for device memory and streams only). There is no CPU
fallback: if libctdm.so is missing or fails to load, every entry point raises.

Reference-named helpers (extractQPI_AlgorithmA, DCT_UWLS_Unwrapped, FastUnwrapping,
cudaGoldsteinUnwrap2D, BatchRecon) mirror the reference functions of the same names
(HilbertTransform.cu:1546, DCT_Unwrapping.cu:10, FastUnwrapping.cu:5, Goldstein.cu:5,
HilbertTransform.cu:1825) with device tensors in place of the reference's globals.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBPATH = HERE / "libctdm.so"

OK, EINVAL, ENOMEM, EDEVICE, EUNSUPPORTED = range(5)

UNWRAP_WRAPDIFF = 0x0
UNWRAP_SINCOS = 0x1
UNWRAP_REFCOMPAT_EIG = 0x2
UNWRAP_REFCOMPAT_IDCT = 0x4
UNWRAP_CONGRUENT = 0x8
UNWRAP_REFCOMPAT_FAST = 0x10
UNWRAP_REFCOMPAT = UNWRAP_REFCOMPAT_EIG | UNWRAP_REFCOMPAT_IDCT

DC_HITMASK = 0x0
DC_REFCOMPAT = 0x1
POS_NO_SHIFT = 0x2
MAPS_MEDIAN = 0x4
MAPS_AMPLITUDE = 0x8
RETRIEVE_ROWS = 0x10
HOLO_U8 = 0x20
HOLO_U16 = 0x40
EXEC_PLAIN = 0x100

PIXEL_F32, PIXEL_U8, PIXEL_U16 = 0, 1, 2


class CtdmError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"ctdm status {status}: {msg}")
        self.status = status


class Crop(C.Structure):
    _fields_ = [("row", C.c_int), ("col", C.c_int), ("radius", C.c_int), ("M", C.c_int)]

    def astuple(self):
        return (self.row, self.col, self.radius, self.M)


class Optics(C.Structure):
    _fields_ = [("wavelength", C.c_double), ("n_med", C.c_double), ("dx", C.c_double)]


class Angle(C.Structure):
    _fields_ = [("ang_x", C.c_double), ("ang_y", C.c_double)]


class Config(C.Structure):
    _fields_ = [("device", C.c_int), ("reserved", C.c_int * 7)]


class Cell(C.Structure):
    _fields_ = [("holo", C.c_void_p), ("bg", C.c_void_p), ("angles", C.POINTER(Angle)),
                ("A", C.c_int), ("H", C.c_int), ("W", C.c_int), ("V", C.c_int),
                ("iters", C.c_int), ("unwrap_flags", C.c_uint), ("recon_flags", C.c_uint),
                ("optics", Optics), ("crop", Crop), ("ri", C.c_void_p),
                ("wrapped_opt", C.c_void_p), ("unwrapped_opt", C.c_void_p),
                ("logamp_opt", C.c_void_p)]


class BatchOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("graph", C.c_int), ("autotune", C.c_int), ("lanes", C.c_int),
                ("reserved", C.c_int * 4)]


class SynthParams(C.Structure):
    _fields_ = [("H", C.c_int), ("pixel", C.c_double), ("wavelength", C.c_double),
                ("n_med", C.c_double), ("na_ill", C.c_double), ("n_cell", C.c_double),
                ("n_nucleoli", C.c_double), ("kappa_cell", C.c_double),
                ("semi_axes", C.c_double * 3), ("n_nuc", C.c_int), ("r_nuc", C.c_double),
                ("noise", C.c_double), ("carrier", C.c_double * 2), ("field_noise", C.c_double),
                ("model", C.c_int)]


_vp = C.c_void_p
# name -> (argtypes, restype); kept in the order of include/ctdm.h
SIGNATURES = {
    "ctdm_create": ([C.POINTER(_vp), C.POINTER(Config), _vp], C.c_int),
    "ctdm_destroy": ([_vp], C.c_int),
    "ctdm_set_stream": ([_vp, _vp], C.c_int),
    "ctdm_sync": ([_vp], C.c_int),
    "ctdm_last_error": ([_vp], C.c_char_p),
    "ctdm_version": ([], C.c_char_p),
    "ctdm_find_sideband": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Crop)], C.c_int),
    "ctdm_retrieve": ([_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(Crop), _vp, _vp, _vp], C.c_int),
    "ctdm_find_sideband_rows": ([_vp, _vp, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "ctdm_retrieve_rows": ([_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp], C.c_int),
    "ctdm_angle_from_image": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.POINTER(C.c_double)], C.c_int),
    "ctdm_find_sideband_px": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Crop)], C.c_int),
    "ctdm_retrieve_px": ([_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Crop), _vp, _vp, _vp], C.c_int),
    "ctdm_find_sideband_rows_px": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "ctdm_retrieve_rows_px": ([_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp], C.c_int),
    "ctdm_angle_from_image_px": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.POINTER(C.c_double)], C.c_int),
    "ctdm_unwrap_dct": ([_vp, _vp, _vp, C.c_int, C.c_int, C.c_uint], C.c_int),
    "ctdm_unwrap_dct_noalloc": ([_vp, _vp, _vp, C.c_int, C.c_int, C.c_uint], C.c_int),
    "ctdm_unwrap_dct_adjoint": ([_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_uint], C.c_int),
    "ctdm_goldstein": ([_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp], C.c_int),
    "ctdm_rytov_spectrum": ([_vp, _vp, _vp, C.c_int, C.c_int, _vp], C.c_int),
    "ctdm_ewald_scatter": ([_vp, _vp, C.POINTER(Angle), C.c_int, C.c_int, C.c_int, C.POINTER(Optics), _vp, _vp], C.c_int),
    "ctdm_ewald_gather": ([_vp, _vp, C.POINTER(Angle), C.c_int, C.c_int, C.c_int, C.POINTER(Optics), _vp], C.c_int),
    "ctdm_normalize_measured": ([_vp, _vp, _vp, _vp, C.c_int, C.c_double], C.c_int),
    "ctdm_fft3d": ([_vp, _vp, C.c_int, C.c_int], C.c_int),
    "ctdm_positivity": ([_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.POINTER(Optics), C.c_uint], C.c_int),
    "ctdm_median3x3": ([_vp, _vp, _vp, C.c_int, C.c_int, C.c_int], C.c_int),
    "ctdm_check_maps": ([_vp, _vp, _vp, C.c_size_t], C.c_int),
    "ctdm_recon_from_maps": ([_vp, _vp, _vp, C.POINTER(Angle), C.c_int, C.c_int, C.POINTER(Optics), C.c_int, C.c_uint, _vp], C.c_int),
    "ctdm_reconstruct": ([_vp, C.POINTER(Cell), C.c_int, C.POINTER(BatchOpts)], C.c_int),
    "ctdm_autotune": ([_vp, C.POINTER(C.c_int), C.c_int, C.c_char_p, C.c_size_t], C.c_int),
    "ctdm_get_tuning": ([_vp, C.c_int, C.POINTER(C.c_int), C.c_int], C.c_int),
    "ctdm_set_tuning": ([_vp, C.c_int, C.POINTER(C.c_int), C.c_int], C.c_int),
    "ctdm_time_pass": ([_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "ctdm_time_loop": ([_vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "ctdm_fit_pgd": ([_vp, _vp, _vp, C.POINTER(Angle), C.c_int, C.c_int, C.POINTER(Optics), C.c_int, C.c_double, _vp, C.c_uint], C.c_int),
    "ctdm_fit_pgd_trace": ([_vp, _vp, _vp, C.POINTER(Angle), C.c_int, C.c_int, C.POINTER(Optics), C.c_int, C.c_double,
                            _vp, C.c_uint, C.POINTER(C.c_double)], C.c_int),
    "ctdm_phimap_write": ([C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, _vp, _vp], C.c_int),
    "ctdm_phimap_read": ([C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
                          C.POINTER(C.c_int), C.POINTER(C.c_int), _vp, _vp, C.c_size_t], C.c_int),
    "ctdm_rimap_write": ([C.c_char_p, _vp, C.c_size_t, C.c_int], C.c_int),
    "ctdm_rimap_read": ([C.c_char_p, _vp, C.c_size_t, C.c_int], C.c_int),
    "ctdm_synth_default_params": ([C.POINTER(SynthParams)], None),
    "ctdm_synth_cell": ([C.c_uint64, C.c_int, C.POINTER(SynthParams), _vp, _vp, C.POINTER(Angle)], C.c_int),
    "ctdm_synth_phantom": ([C.c_uint64, C.POINTER(SynthParams), C.c_int, C.c_double, _vp], C.c_int),
    "ctdm_synth_hologram": ([C.c_uint64, C.c_int, _vp, _vp], C.c_int),
    "ctdm_synth_wrapped": ([C.c_uint64, C.c_int, C.c_int, C.c_int, _vp], C.c_int),
}

_lib = None


def lib():
    """Load libctdm.so (raises if it is missing: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise RuntimeError(f"{LIBPATH} is missing: run `python -m ctdm_amd.build` first")
        L = C.CDLL(str(LIBPATH))
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    if not t.is_cuda:
        raise ValueError("libctdm takes device tensors")
    if not t.is_contiguous():
        raise ValueError("libctdm takes contiguous tensors")
    return t.data_ptr()


def pixel_format(t) -> int:
    """PIXEL_F32 / PIXEL_U8 / PIXEL_U16 of a hologram tensor or array (float32, uint8 or uint16)."""
    name = str(t.dtype).replace("torch.", "")
    fmt = {"float32": PIXEL_F32, "uint8": PIXEL_U8, "uint16": PIXEL_U16}.get(name)
    if fmt is None:
        raise ValueError(f"holograms must be float32, uint8 or uint16, not {t.dtype}")
    return fmt


def _pair_format(holo, bg) -> int:
    px = pixel_format(holo)
    if pixel_format(bg) != px:
        raise ValueError("holograms and backgrounds must share one pixel format")
    return px


def angles_array(angles) -> C.Array:
    """(A, 2) array of (ang_x, ang_y) radians -> ctypes Angle[A]."""
    a = np.asarray(angles, dtype=np.float64).reshape(-1, 2)
    arr = (Angle * len(a))()
    for i, (ax, ay) in enumerate(a):
        arr[i].ang_x, arr[i].ang_y = float(ax), float(ay)
    return arr


# ---------------------------------------------------------------- synthetic inputs (host)
def synth_params(**kw) -> SynthParams:
    p = SynthParams()
    lib().ctdm_synth_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("semi_axes", "carrier"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def phimap_write(path, phase, amp, ang_x: float = 0.0, ang_y: float = 0.0, status: bool = True):
    """Writes one frame in the reference's .phimap format (host arrays, sizeY x sizeX)."""
    ph = np.ascontiguousarray(phase, np.float32)
    am = np.ascontiguousarray(amp, np.float32)
    sy, sx = ph.shape
    _check_free(lib().ctdm_phimap_write(str(path).encode(), int(bool(status)), ang_x, ang_y, sx, sy,
                                        ph.ctypes.data, am.ctypes.data))


def phimap_read(path):
    """Returns dict(status, ang_x, ang_y, phase, amp) of a .phimap file."""
    st, sx, sy = C.c_int(), C.c_int(), C.c_int()
    ax, ay = C.c_double(), C.c_double()
    L = lib()
    _check_free(L.ctdm_phimap_read(str(path).encode(), C.byref(st), C.byref(ax), C.byref(ay), C.byref(sx),
                                   C.byref(sy), None, None, 0))
    ph = np.empty((sy.value, sx.value), np.float32)
    am = np.empty_like(ph)
    _check_free(L.ctdm_phimap_read(str(path).encode(), None, None, None, None, None, ph.ctypes.data,
                                   am.ctypes.data, ph.size))
    return dict(status=bool(st.value), ang_x=ax.value, ang_y=ay.value, phase=ph, amp=am)


def load_phimap_stack(paths):
    """The reference's reload path (LoadDateFromDir + RefreshStack, HilbertTransform.cu:1453-1530): frames
    whose status is 0 are dropped; returns (phase stack, amplitude stack, angles A x 2). The reference
    reads angle y from the x field (:1479); here each angle keeps its own field."""
    ph, am, ang = [], [], []
    for p in paths:
        d = phimap_read(p)
        if not d["status"]:
            continue
        ph.append(d["phase"])
        am.append(d["amp"])
        ang.append((d["ang_x"], d["ang_y"]))
    return np.stack(ph), np.stack(am), np.array(ang)


def rimap_write(path, ri, part: int = 0):
    """Writes a volume as BatchRecon exports it: part 0 complex64, 1 real, 2 imaginary."""
    v = np.ascontiguousarray(ri, np.complex64)
    _check_free(lib().ctdm_rimap_write(str(path).encode(), v.ctypes.data, v.size, part))


def rimap_read(path, shape, part: int = 0):
    v = np.zeros(shape, np.complex64)
    _check_free(lib().ctdm_rimap_read(str(path).encode(), v.ctypes.data, v.size, part))
    return v


def synth_hologram(seed: int, H: int = 1024):
    holo = np.empty((H, H), np.float32)
    bg = np.empty((H, H), np.float32)
    _check_free(lib().ctdm_synth_hologram(seed, H, holo.ctypes.data, bg.ctypes.data))
    return holo, bg


def synth_cell(seed: int, A: int, params: SynthParams | None = None):
    p = params or synth_params()
    holo = np.empty((A, p.H, p.H), np.float32)
    bg = np.empty((A, p.H, p.H), np.float32)
    ang = (Angle * A)()
    _check_free(lib().ctdm_synth_cell(seed, A, C.byref(p), holo.ctypes.data, bg.ctypes.data, ang))
    angles = np.array([(a.ang_x, a.ang_y) for a in ang], np.float64)
    return holo, bg, angles


def synth_phantom(seed: int, V: int, dx: float, params: SynthParams | None = None):
    """Refractive index of the phantom synth_cell(seed, ...) images, on the V^3 output grid ([z][y][x])."""
    p = params or synth_params()
    n = np.empty((V, V, V), np.complex64)
    _check_free(lib().ctdm_synth_phantom(seed, C.byref(p), V, dx, n.ctypes.data))
    return n


def synth_wrapped(seed: int, batch: int, H: int, W: int):
    out = np.empty((batch, H, W), np.float32)
    _check_free(lib().ctdm_synth_wrapped(seed, batch, H, W, out.ctypes.data))
    return out


def _check_free(status, what: str = "host call rejected its arguments"):
    if status != OK:
        raise CtdmError(status, what)


def optics_for(V: int, H: int = 1024, pixel: float = 0.1e-6, wavelength: float = 532e-9,
               n_med: float = 1.337) -> Optics:
    """Optics of a cell whose H x H holograms (sample-plane pixel `pixel`) are cropped to V x V:
    dx = pixel * H / V (reference BatchRecon: CameraPixelSize/Mag * 1024 / ReconZ)."""
    return Optics(wavelength, n_med, pixel * H / V)


# ---------------------------------------------------------------- context
class Context:
    """Owns a ctdm_ctx bound to a HIP device and a stream (torch's current stream by default)."""

    def __init__(self, device: int = 0, stream=None):
        import torch

        self._torch = torch
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self.stream = stream
        cfg = Config(device, (C.c_int * 7)())
        h = _vp()
        st = lib().ctdm_create(C.byref(h), C.byref(cfg), _vp(stream.cuda_stream))
        if st != OK:
            raise CtdmError(st, "ctdm_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().ctdm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        st = getattr(lib(), name)(self.h, *args)
        if st != OK:
            raise CtdmError(st, lib().ctdm_last_error(self.h).decode())

    def sync(self):
        self._call("ctdm_sync")

    def set_stream(self, stream):
        """ctdm_set_stream: later work goes to `stream` (a torch.cuda.Stream) after the current one drains."""
        self._call("ctdm_set_stream", _vp(stream.cuda_stream))
        self.stream = stream

    # -- retrieval
    def find_sideband(self, bg, M: int = 256, row_start: int = 0) -> Crop:
        H, W = bg.shape[-2:]
        crop = Crop()
        px = pixel_format(bg)
        if px == PIXEL_F32:
            self._call("ctdm_find_sideband", _ptr(bg), H, W, M, row_start, C.byref(crop))
        else:
            self._call("ctdm_find_sideband_px", _ptr(bg), px, H, W, M, row_start, C.byref(crop))
        return crop

    def angle_from_image(self, img, pixel_um: float, mag: float, wavelength_nm: float, n_med: float):
        """AngCal_GPU: illumination angle [rad] of each angle-calibration image (batch x H x W)."""
        im = img if img.dim() == 3 else img[None]
        B, H, W = im.shape
        out = (C.c_double * B)()
        px = pixel_format(im)
        if px == PIXEL_F32:
            self._call("ctdm_angle_from_image", _ptr(im), B, H, W, pixel_um, mag, wavelength_nm, n_med, out)
        else:
            self._call("ctdm_angle_from_image_px", _ptr(im), px, B, H, W, pixel_um, mag, wavelength_nm, n_med, out)
        return list(out) if img.dim() == 3 else out[0]

    def find_sideband_rows(self, bg) -> int:
        H, W = bg.shape[-2:]
        c = C.c_int()
        px = pixel_format(bg)
        if px == PIXEL_F32:
            self._call("ctdm_find_sideband_rows", _ptr(bg), H, W, C.byref(c))
        else:
            self._call("ctdm_find_sideband_rows_px", _ptr(bg), px, H, W, C.byref(c))
        return c.value

    def retrieve_rows(self, holo, bg, center: int):
        """Algorithm B maps, (batch, W/4, H/4) indexed [n][y'] (the reference's transposed layout)."""
        torch = self._torch
        if holo.dim() == 2:
            holo, bg = holo[None], bg[None]
        B, H, W = holo.shape
        wrapped = torch.empty((B, W // 4, H // 4), dtype=torch.float32, device=holo.device)
        logamp = torch.empty_like(wrapped)
        px = _pair_format(holo, bg)
        if px == PIXEL_F32:
            self._call("ctdm_retrieve_rows", _ptr(holo), _ptr(bg), B, H, W, center, _ptr(wrapped), _ptr(logamp))
        else:
            self._call("ctdm_retrieve_rows_px", _ptr(holo), _ptr(bg), px, B, H, W, center, _ptr(wrapped),
                       _ptr(logamp))
        return wrapped, logamp

    def retrieve(self, holo, bg, crop: Crop, field: bool = False):
        torch = self._torch
        if holo.dim() == 2:
            holo, bg = holo[None], bg[None]
        B, H, W = holo.shape
        M = crop.M
        wrapped = torch.empty((B, M, M), dtype=torch.float32, device=holo.device)
        logamp = torch.empty_like(wrapped)
        fld = torch.empty((B, M, M), dtype=torch.complex64, device=holo.device) if field else None
        px = _pair_format(holo, bg)
        if px == PIXEL_F32:
            self._call("ctdm_retrieve", _ptr(holo), _ptr(bg), B, H, W, C.byref(crop), _ptr(wrapped),
                       _ptr(logamp), _ptr(fld))
        else:
            self._call("ctdm_retrieve_px", _ptr(holo), _ptr(bg), px, B, H, W, C.byref(crop), _ptr(wrapped),
                       _ptr(logamp), _ptr(fld))
        return (wrapped, logamp, fld) if field else (wrapped, logamp)

    # -- unwrapping
    def unwrap_dct(self, wrapped, flags: int = 0):
        w = wrapped if wrapped.dim() == 3 else wrapped[None]
        out = self._torch.empty_like(w)
        self._call("ctdm_unwrap_dct", _ptr(w), _ptr(out), w.shape[0], w.shape[-1], flags)
        return out if wrapped.dim() == 3 else out[0]

    def unwrap_dct_noalloc(self, wrapped, out, flags: int = 0):
        """ctdm_unwrap_dct_noalloc: the same solve with the library's fixed device scratch; the caller
        passes the output buffer too, so the call allocates nothing."""
        w = wrapped if wrapped.dim() == 3 else wrapped[None]
        o = out if out.dim() == 3 else out[None]
        if o.shape != w.shape or o.dtype != w.dtype:
            raise ValueError("out must match wrapped")
        self._call("ctdm_unwrap_dct_noalloc", _ptr(w), _ptr(o), w.shape[0], w.shape[-1], flags)
        return out

    def unwrap_dct_adjoint(self, wrapped, grad_out, flags: int = 0):
        w = wrapped if wrapped.dim() == 3 else wrapped[None]
        g = grad_out if grad_out.dim() == 3 else grad_out[None]
        out = self._torch.empty_like(w)
        self._call("ctdm_unwrap_dct_adjoint", _ptr(w), _ptr(g), _ptr(out), w.shape[0],
                   w.shape[-1], flags)
        return out if wrapped.dim() == 3 else out[0]

    def goldstein(self, wrapped):
        torch = self._torch
        w = wrapped if wrapped.dim() == 3 else wrapped[None]
        B, H, W = w.shape
        res = torch.empty((B, H, W), dtype=torch.int8, device=w.device)
        cuts = torch.empty((B, H, W), dtype=torch.uint8, device=w.device)
        k = torch.empty((B, H, W), dtype=torch.int32, device=w.device)
        un = torch.empty((B, H, W), dtype=torch.float32, device=w.device)
        self._call("ctdm_goldstein", _ptr(w), B, H, W, _ptr(res), _ptr(cuts), _ptr(k), _ptr(un))
        return res, cuts, k, un

    # -- reconstruction primitives
    def rytov_spectrum(self, unwrapped, logamp):
        A, M, _ = unwrapped.shape
        spec = self._torch.empty((A, M, M), dtype=torch_c64(self._torch), device=unwrapped.device)
        self._call("ctdm_rytov_spectrum", _ptr(unwrapped), _ptr(logamp), A, M, _ptr(spec))
        return spec

    def ewald_scatter(self, spec, angles, V: int, optics: Optics, F=None, hits=None):
        torch = self._torch
        A, M, _ = spec.shape
        if F is None:
            F = torch.zeros((V, V, V), dtype=torch.complex64, device=spec.device)
        if hits is None:
            hits = torch.zeros((V, V, V), dtype=torch.float32, device=spec.device)
        arr = angles_array(angles)
        self._call("ctdm_ewald_scatter", _ptr(spec), arr, A, M, V, C.byref(optics), _ptr(F),
                   _ptr(hits))
        return F, hits

    def ewald_gather(self, F, angles, M: int, optics: Optics):
        arr = angles_array(angles)
        A = len(arr)
        V = F.shape[0]
        spec = self._torch.empty((A, M, M), dtype=self._torch.complex64, device=F.device)
        self._call("ctdm_ewald_gather", _ptr(F), arr, A, M, V, C.byref(optics), _ptr(spec))
        return spec

    def normalize_measured(self, F, hits, dx: float, with_meas: bool = True):
        V = F.shape[0]
        Fm = self._torch.empty_like(F) if with_meas else None
        self._call("ctdm_normalize_measured", _ptr(F), _ptr(hits), _ptr(Fm), V, dx)
        return Fm

    def fft3d(self, vol, direction: int):
        self._call("ctdm_fft3d", _ptr(vol), vol.shape[0], direction)
        return vol

    def positivity(self, f, F_meas, hits, iters: int, optics: Optics, flags: int = 0):
        n = self._torch.empty_like(f)
        self._call("ctdm_positivity", _ptr(f), _ptr(F_meas), _ptr(hits), _ptr(n), f.shape[0],
                   iters, C.byref(optics), flags)
        return n

    def fit_pgd(self, f, spec, angles, optics: Optics, iters: int, step: float, flags: int = 0):
        arr = angles_array(angles)
        n = self._torch.empty_like(f)
        self._call("ctdm_fit_pgd", _ptr(f), _ptr(spec), arr, len(arr), f.shape[0],
                   C.byref(optics), iters, step, _ptr(n), flags)
        return n

    def fit_pgd_trace(self, f, spec, angles, optics: Optics, iters: int, step: float, flags: int = 0):
        """(n, objective): the fit and J before each of its steps (list of iters floats)."""
        arr = angles_array(angles)
        n = self._torch.empty_like(f)
        J = (C.c_double * max(iters, 1))()
        self._call("ctdm_fit_pgd_trace", _ptr(f), _ptr(spec), arr, len(arr), f.shape[0],
                   C.byref(optics), iters, step, _ptr(n), flags, J)
        return n, list(J)[:iters]

    def median3x3(self, maps):
        """3x3 median of each map (batch x H x W float32), borders copied (MedianFilter_gpu)."""
        m = maps if maps.dim() == 3 else maps[None]
        out = self._torch.empty_like(m)
        self._call("ctdm_median3x3", _ptr(m), _ptr(out), m.shape[0], m.shape[1], m.shape[2])
        return out if maps.dim() == 3 else out[0]

    def check_maps(self, phase=None, amp=None):
        """checkMaps in place: phase NaN/Inf -> 0; amplitude NaN/Inf -> 1, 0 -> 0.01, else |A|."""
        n = (phase if phase is not None else amp).numel()
        self._call("ctdm_check_maps", _ptr(phase), _ptr(amp), n)

    def recon_from_maps(self, unwrapped, map2, angles, optics: Optics, iters: int, flags: int = 0):
        """BatchRecon: map2 is ln A, or A with MAPS_AMPLITUDE; MAPS_MEDIAN filters both maps first."""
        A, V, _ = unwrapped.shape
        arr = angles_array(angles)
        ri = self._torch.empty((V, V, V), dtype=self._torch.complex64, device=unwrapped.device)
        self._call("ctdm_recon_from_maps", _ptr(unwrapped), _ptr(map2), arr, A, V, C.byref(optics), iters,
                   flags, _ptr(ri))
        return ri

    # -- pipeline
    def reconstruct(self, cells: list[dict], mode: int = 1, graph: bool = False,
                    autotune: bool = False, lanes: int = 0):
        """cells: dicts with holo, bg (device, A x H x W; float32, uint8 or uint16 pixels), angles
        (A x 2), V, iters, optics, ri (device V^3 complex64) and optional crop / flags / wrapped /
        unwrapped / logamp.
        lanes: batched multi-cell sequences run on two concurrent lanes unless lanes == 1."""
        arr = (Cell * len(cells))()
        keep = []
        for i, d in enumerate(cells):
            A, H, W = d["holo"].shape
            ang = angles_array(d["angles"])
            keep.append(ang)
            crop = d.get("crop") or Crop(0, 0, -1, d["V"])
            holo_flag = {PIXEL_F32: 0, PIXEL_U8: HOLO_U8, PIXEL_U16: HOLO_U16}[_pair_format(d["holo"], d["bg"])]
            arr[i] = Cell(_ptr(d["holo"]), _ptr(d["bg"]), ang, A, H, W, d["V"], d["iters"],
                          d.get("unwrap_flags", 0), d.get("recon_flags", 0) | holo_flag, d["optics"], crop,
                          _ptr(d["ri"]), _ptr(d.get("wrapped")), _ptr(d.get("unwrapped")),
                          _ptr(d.get("logamp")))
        opts = BatchOpts(mode, int(graph), int(autotune), int(lanes), (C.c_int * 4)())
        self._call("ctdm_reconstruct", arr, len(cells), C.byref(opts))
        for i, d in enumerate(cells):  # report the crop the pipeline used
            d["crop_used"] = arr[i].crop
        return cells

    def get_tuning(self, V: int):
        """dict(pencils=[x, y, z+dc, grid], sweep=[x, y], yplan=0|1) of volume side V."""
        v = (C.c_int * 7)()
        self._call("ctdm_get_tuning", V, v, 7)
        return dict(pencils=list(v[:4]), sweep=list(v[4:6]), yplan=v[6])

    def set_tuning(self, V: int, pencils=(16, 16, 16, 8), sweep=(0, 0), yplan: int = 0):
        v = (C.c_int * 7)(*pencils, *sweep, yplan)
        self._call("ctdm_set_tuning", V, v, 7)

    def time_pass(self, pass_kind: int, V: int, pencils: int = 0, reps: int = 20):
        """(average ms per launch, algorithmic bytes per launch) of a fused volume pass."""
        ms, nb = C.c_double(), C.c_double()
        self._call("ctdm_time_pass", pass_kind, V, pencils, reps, C.byref(ms), C.byref(nb))
        return ms.value, nb.value

    def time_loop(self, V: int, iters: int = 10):
        """Per-pass average ms and algorithmic bytes ([x, y, z+dc]) measured inside the loop sequence."""
        ms, nb = (C.c_double * 3)(), (C.c_double * 3)()
        self._call("ctdm_time_loop", V, iters, ms, nb)
        return list(ms), list(nb)

    def autotune(self, sizes) -> str:
        s = (C.c_int * len(sizes))(*sizes)
        buf = C.create_string_buffer(8192)
        self._call("ctdm_autotune", s, len(sizes), buf, 8192)
        return buf.value.decode()

    # -- reference-named wrappers ----------------------------------------------------------
    def extractQPI_AlgorithmA(self, SP, BG, crop: Crop | None = None):
        """Reference extractQPI_AlgorithmA + calcWrapPhase (HilbertTransform.cu:1546, :3618):
        returns (wrapped phase, log-amplitude, crop); the crop is located from BG when not
        given (the reference does so on frame 1)."""
        if crop is None:
            crop = self.find_sideband(BG if BG.dim() == 2 else BG[0], M=SP.shape[-1] // 4)
        w, la = self.retrieve(SP, BG, crop)
        return w, la, crop

    def DCT_UWLS_Unwrapped(self, src, flags: int = 0):
        """Reference DCT_UWLS_Unwrapped(dst, src, nx, ny) (DCT_Unwrapping.cu:10)."""
        return self.unwrap_dct(src, flags)

    def FastUnwrapping(self, src):
        """Reference FastUnwrapping (FastUnwrapping.cu:5): the map rescaled to [0, 2 pi], sin/cos-Laplacian
        RHS, Poisson solve, congruence, minimum subtracted (the input tensor is not modified)."""
        return self.unwrap_dct(src, UNWRAP_SINCOS | UNWRAP_CONGRUENT | UNWRAP_REFCOMPAT_FAST)

    def cudaGoldsteinUnwrap2D(self, phi):
        """Reference cudaGoldsteinUnwrap2D(phi, out, nx, ny, frame) (Goldstein.cu:5)."""
        return self.goldstein(phi)[3]

    def BatchRecon(self, holo, bg, angles, V: int, iters: int, optics: Optics, **kw):
        """Per-cell pipeline ending in the reference's BatchRecon (HilbertTransform.cu:1825)."""
        mode = kw.pop("mode", 1)
        graph = kw.pop("graph", False)
        ri = self._torch.empty((V, V, V), dtype=self._torch.complex64, device=holo.device)
        cell = dict(holo=holo, bg=bg, angles=angles, V=V, iters=iters, optics=optics, ri=ri, **kw)
        self.reconstruct([cell], mode=mode, graph=graph)
        return ri


def torch_c64(torch):
    return torch.complex64

"""Edge cases and error paths through the C ABI: empty batches are no-ops, unsupported sizes and bad
arguments return a status (never a crash), a one-angle cell and the smallest volume reconstruct within
1e-5 of the oracle, and uneven tiny cells share one graph."""
import ctypes as C

import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu
LAM, NM = 532e-9, 1.337


def _st(ctx, name, *args):
    return getattr(ctdm_amd.lib(), name)(ctx.h, *args)


def _p(t):
    return C.c_void_p(t.data_ptr())


def test_empty_batches_are_noops(ctx):
    x = torch.zeros(16, device="cuda")
    crop = ctdm_amd.Crop(768, 768, 85, 256)
    assert _st(ctx, "ctdm_retrieve", _p(x), _p(x), 0, 1024, 1024, C.byref(crop), _p(x), _p(x), None) == 0
    assert _st(ctx, "ctdm_unwrap_dct", _p(x), _p(x + 1), 0, 64, 0) == 0
    assert _st(ctx, "ctdm_goldstein", _p(x), 0, 64, 64, None, None, None, _p(x + 1)) == 0
    assert _st(ctx, "ctdm_median3x3", _p(x), _p(x + 1), 0, 8, 8) == 0
    assert _st(ctx, "ctdm_retrieve_rows", _p(x), _p(x), 0, 1024, 1024, 768, _p(x), _p(x)) == 0
    assert _st(ctx, "ctdm_reconstruct", None, 0, C.byref(ctdm_amd.BatchOpts(1, 0, 0, 0, (C.c_int * 4)()))) == 0
    ctx.sync()


def test_unsupported_sizes_and_bad_arguments_return_status(ctx):
    x = torch.zeros(1 << 20, device="cuda")
    y = torch.zeros(1 << 20, device="cuda")
    crop = ctdm_amd.Crop(768, 768, 85, 256)
    cases = {
        "hologram side without a kernel": ("ctdm_retrieve", _p(x), _p(x), 1, 300, 300, C.byref(crop), _p(y), _p(y), None),
        "non-square hologram": ("ctdm_retrieve", _p(x), _p(x), 1, 512, 1024, C.byref(crop), _p(y), _p(y), None),
        "crop leaving the spectrum": ("ctdm_retrieve", _p(x), _p(x), 1, 1024, 1024,
                                      C.byref(ctdm_amd.Crop(1000, 768, 85, 256)), _p(y), _p(y), None),
        "null output": ("ctdm_retrieve", _p(x), _p(x), 1, 1024, 1024, C.byref(crop), None, _p(y), None),
        "unwrap side without a kernel": ("ctdm_unwrap_dct", _p(x), _p(y), 1, 100, 0),
        "unknown unwrap flag": ("ctdm_unwrap_dct", _p(x), _p(y), 1, 64, 0x40),
        "in-place unwrap": ("ctdm_unwrap_dct", _p(x), _p(x), 1, 64, 0),
        "volume side without a kernel": ("ctdm_fft3d", _p(x), 100, -1),
        "fft direction": ("ctdm_fft3d", _p(x), 64, 2),
        "tiny Goldstein map": ("ctdm_goldstein", _p(x), 1, 2, 2, None, None, None, _p(y)),
        "Goldstein map with 65536 rows": ("ctdm_goldstein", _p(x), 1, 65536, 3, None, None, None, _p(y)),
        "65536 Goldstein maps in one call": ("ctdm_goldstein", _p(x), 65536, 3, 3, None, None, None, _p(y)),
        "aliased median": ("ctdm_median3x3", _p(x), _p(x), 1, 8, 8),
        "rows band outside": ("ctdm_retrieve_rows", _p(x), _p(x), 1, 256, 256, 250, _p(y), _p(y)),
        "time_pass kind": ("ctdm_time_pass", 3, 64, 0, 1, C.byref(C.c_double()), C.byref(C.c_double())),
    }
    for what, (name, *args) in cases.items():
        st = _st(ctx, name, *args)
        msg = ctdm_amd.lib().ctdm_last_error(ctx.h).decode()
        print(f"{what}: status {st}, '{msg}'")
        assert st in (ctdm_amd.EINVAL, ctdm_amd.EUNSUPPORTED), what
        assert msg
    ctx.sync()  # the context stays usable
    assert _st(ctx, "ctdm_fft3d", _p(x), 64, -1) == 0
    ctx.sync()


def _cell(seed, A, H, V, iters=2):
    holo, bg, ang = ctdm_amd.synth_cell(seed, A, ctdm_amd.synth_params(H=H))
    ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    return dict(holo=torch.from_numpy(holo).cuda(), bg=torch.from_numpy(bg).cuda(), angles=ang, V=V, iters=iters,
                optics=ctdm_amd.optics_for(V, H), ri=ri), (holo, bg, ang)


def test_one_angle_and_smallest_volume(ctx, orc):
    for (A, H, V) in [(1, 256, 64), (3, 256, 32)]:
        cell, (holo, bg, ang) = _cell(60 + A, A, H, V)
        for mode in (1, 0):
            ctx.reconstruct([cell], mode=mode)
            torch.cuda.synchronize()
            ref = orc.reconstruct_cell(holo, bg, ang, V, cell["crop_used"].astuple(), LAM, NM, cell["optics"].dx, 2)
            err = np.abs(cell["ri"].cpu().numpy() - ref).max()
            print(f"A={A} V={V} mode {mode}: max|dn| {err:.2e}")
            assert err <= 1e-5


def test_tiny_uneven_cells_one_graph(ctx):
    spec = [(1, 256, 32), (2, 256, 64), (5, 256, 32)]
    cells = [_cell(70 + i, A, H, V)[0] for i, (A, H, V) in enumerate(spec)]
    ctx.reconstruct(cells, mode=1, graph=True)
    torch.cuda.synchronize()
    first = [c["ri"].clone() for c in cells]
    for c in cells:
        c["ri"].zero_()
    ctx.reconstruct(cells, mode=1, graph=True)  # replay
    torch.cuda.synchronize()
    for c, f in zip(cells, first):
        assert torch.equal(c["ri"], f)
        one = dict(c, ri=torch.empty_like(f))
        ctx.reconstruct([one], mode=1)
        torch.cuda.synchronize()
        assert torch.equal(one["ri"], f)

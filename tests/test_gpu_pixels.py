"""Integer hologram pixels (CTDM_HOLO_U8 / CTDM_HOLO_U16 and the *_px entry points).

The reference reads its frames as 8-bit BMP images and widens them to float on the host
(HilbertTransform.cu:419-420, 660-661, read_bmp :2694); here the kernels that read holograms widen
8- and 16-bit pixels as they load them. The bar is bit-identity: every output for integer pixels
equals, bit for bit, the output for the same integer values passed as float32 (the float32 path is
itself checked against the FP64 oracle by the other retrieval and cell tests), on the batched,
graph-replayed, plain and row-wise paths, the sideband searches and the angle calibration. One
retrieval of 8-bit frames is also compared with the oracle directly."""
import ctypes as C

import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu
LAM, NM = 532e-9, 1.337


def _quantise(x, top, dtype):
    """Frames of integer pixels spanning [0, top] (a camera's range) from synthetic intensities."""
    q = np.rint(np.clip(x / x.max(), 0.0, 1.0) * top).astype(dtype)
    return q, q.astype(np.float32)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _wrap(a):
    return np.angle(np.exp(1j * a))


@pytest.mark.parametrize("dtype,top", [(np.uint8, 255), (np.uint16, 4095), (np.uint16, 65535)])
def test_retrieve_and_search_px_equal_float(ctx, dtype, top):
    holo, bg = ctdm_amd.synth_hologram(5, 1024)
    hq, hf = _quantise(holo, top, dtype)
    bq, bf = _quantise(bg, top, dtype)
    c_int = ctx.find_sideband(_dev(bq), M=256)
    c_f32 = ctx.find_sideband(_dev(bf), M=256)
    assert c_int.astuple() == c_f32.astuple()
    outs = []
    for h, b in ((hq, bq), (hf, bf)):
        w, la, fld = ctx.retrieve(_dev(h), _dev(b), c_f32, field=True)
        torch.cuda.synchronize()
        outs.append((w, la, fld))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
    # row-wise (Algorithm B) search and retrieval
    center_int = ctx.find_sideband_rows(_dev(bq))
    center_f32 = ctx.find_sideband_rows(_dev(bf))
    assert center_int == center_f32
    stack_q, stack_f = np.stack([hq, bq]), np.stack([hf, bf])
    r_int = ctx.retrieve_rows(_dev(stack_q), _dev(stack_q[::-1]), center_f32)
    r_f32 = ctx.retrieve_rows(_dev(stack_f), _dev(stack_f[::-1]), center_f32)
    torch.cuda.synchronize()
    for a, b in zip(r_int, r_f32):
        assert torch.equal(a, b)


def test_retrieve_u8_matches_oracle(ctx, orc):
    holo, bg = ctdm_amd.synth_hologram(1, 1024)
    hq, hf = _quantise(holo, 255, np.uint8)
    bq, bf = _quantise(bg, 255, np.uint8)
    crop = ctx.find_sideband(_dev(bq), M=256)
    assert (crop.row, crop.col, crop.radius) == orc.find_sideband(bf)
    w, la = ctx.retrieve(_dev(hq), _dev(bq), crop)
    r = orc.retrieve(hf, bf, crop.row, crop.col, crop.radius, 256)
    dw = _wrap(w[0].cpu().numpy().astype(np.float64) - r["wrapped"])
    print("u8 frames: wrapped rms", np.sqrt(np.mean(dw ** 2)))
    assert np.sqrt(np.mean(dw ** 2)) <= 1e-4
    assert np.sqrt(np.mean((la[0].cpu().numpy() - r["logamp"]) ** 2)) <= 1e-4


def test_angle_from_image_px_equal_float(ctx):
    holo, bg = ctdm_amd.synth_hologram(3, 1024)
    for dtype, top in ((np.uint8, 255), (np.uint16, 4095)):
        q = np.stack([_quantise(holo, top, dtype)[0], _quantise(bg, top, dtype)[0]])
        a_int = ctx.angle_from_image(_dev(q), 5.5, 40.0, 532.0, NM)
        a_f32 = ctx.angle_from_image(_dev(q.astype(np.float32)), 5.5, 40.0, 532.0, NM)
        assert a_int == a_f32


def _cell(seed, A, H, V, dtype, top, flags=0):
    p = ctdm_amd.synth_params(H=H)
    holo, bg, ang = ctdm_amd.synth_cell(seed, A, p)
    hq, hf = _quantise(holo, top, dtype)
    bq, bf = _quantise(bg, top, dtype)
    base = dict(angles=ang, V=V, iters=3, optics=ctdm_amd.optics_for(V, H), recon_flags=flags)
    c_int = dict(base, holo=_dev(hq), bg=_dev(bq), ri=torch.empty((V, V, V), dtype=torch.complex64, device="cuda"),
                 wrapped=torch.empty((A, V, V), dtype=torch.float32, device="cuda"))
    c_f32 = dict(base, holo=_dev(hf), bg=_dev(bf), ri=torch.empty((V, V, V), dtype=torch.complex64, device="cuda"),
                 wrapped=torch.empty((A, V, V), dtype=torch.float32, device="cuda"))
    return c_int, c_f32


@pytest.mark.parametrize("dtype,top", [(np.uint8, 255), (np.uint16, 4095)])
def test_cell_px_equal_float_batched_and_graph(ctx, dtype, top):
    c_int, c_f32 = _cell(11, 10, 512, 128, dtype, top)
    ctx.reconstruct([c_f32], mode=1)
    torch.cuda.synchronize()
    for graph in (False, True, True):  # eager, capture, replay
        c_int["ri"].fill_(float("nan"))
        ctx.reconstruct([c_int], mode=1, graph=graph)
        torch.cuda.synchronize()
        assert c_int["crop_used"].astuple() == c_f32["crop_used"].astuple()
        assert torch.equal(c_int["wrapped"], c_f32["wrapped"])
        assert torch.equal(c_int["ri"], c_f32["ri"]), graph


def test_cells_mixed_formats_one_graph(ctx):
    """8-bit, 16-bit and float32 cells in one two-lane graph: each equals its float32 twin run alone."""
    pairs = [_cell(21, 6, 256, 64, np.uint8, 255), _cell(22, 7, 256, 64, np.uint16, 4095),
             _cell(23, 5, 256, 32, np.uint8, 255)]
    for _, f in pairs:
        ctx.reconstruct([f], mode=1)
    torch.cuda.synchronize()
    mixed = [pairs[0][0], pairs[1][0], pairs[2][1], pairs[2][0]]
    twins = [pairs[0][1], pairs[1][1], pairs[2][1], pairs[2][1]]
    refs = [t["ri"].clone() for t in twins]
    for d in mixed:
        d["ri"] = torch.empty_like(d["ri"])
    for rep in range(2):
        ctx.reconstruct(mixed, mode=1, graph=True)
        torch.cuda.synchronize()
        for d, r in zip(mixed, refs):
            assert torch.equal(d["ri"], r), rep


def test_cell_u8_plain_and_rows_equal_float(ctx):
    # plain (reference-structured) mode
    c_int, c_f32 = _cell(31, 4, 256, 64, np.uint8, 255)
    for d in (c_int, c_f32):
        ctx.reconstruct([d], mode=0)
    torch.cuda.synchronize()
    assert torch.equal(c_int["ri"], c_f32["ri"])
    # row-wise retrieval (V = W / 4), batched and plain
    for mode in (1, 0):
        c_int, c_f32 = _cell(32, 4, 256, 64, np.uint8, 255, flags=ctdm_amd.RETRIEVE_ROWS)
        for d in (c_int, c_f32):
            ctx.reconstruct([d], mode=mode)
        torch.cuda.synchronize()
        assert torch.equal(c_int["ri"], c_f32["ri"]), mode


def test_pixel_format_errors(ctx):
    c_int, c_f32 = _cell(41, 2, 256, 64, np.uint8, 255)
    with pytest.raises(ValueError, match="pixel format"):
        ctx.reconstruct([dict(c_int, bg=c_f32["bg"])], mode=1)
    with pytest.raises(ValueError, match="float32, uint8 or uint16"):
        ctx.reconstruct([dict(c_int, holo=c_int["holo"].to(torch.int16), bg=c_int["bg"].to(torch.int16))], mode=1)
    lib = ctdm_amd.lib()
    # both integer flags at once, through the C ABI
    A, H = 2, 256
    ang = ctdm_amd.angles_array(c_int["angles"])
    cell = ctdm_amd.Cell(c_int["holo"].data_ptr(), c_int["bg"].data_ptr(), ang, A, H, H, 64, 1, 0,
                         ctdm_amd.HOLO_U8 | ctdm_amd.HOLO_U16, c_int["optics"], ctdm_amd.Crop(0, 0, -1, 64),
                         c_int["ri"].data_ptr(), None, None, None)
    opts = ctdm_amd.BatchOpts(1, 0, 0, 0)
    st = lib.ctdm_reconstruct(ctx.h, C.byref(cell), 1, C.byref(opts))
    assert st == ctdm_amd.EINVAL
    assert b"HOLO_U8" in lib.ctdm_last_error(ctx.h)
    st = lib.ctdm_retrieve_px(ctx.h, c_int["holo"].data_ptr(), c_int["bg"].data_ptr(), 7, 1, H, H,
                              C.byref(ctdm_amd.Crop(192, 192, 21, 64)), c_int["wrapped"].data_ptr(),
                              c_int["wrapped"].data_ptr(), None)
    assert st == ctdm_amd.EINVAL
    assert b"pixel format" in lib.ctdm_last_error(ctx.h)
    # 8-bit frames are read in pixel pairs: an image starting at an odd byte address is refused before
    # any work is queued
    st = lib.ctdm_retrieve_px(ctx.h, c_int["holo"].data_ptr() + 1, c_int["bg"].data_ptr(), ctdm_amd.PIXEL_U8, 1,
                              H, H, C.byref(ctdm_amd.Crop(192, 192, 21, 64)), c_int["wrapped"].data_ptr(),
                              c_int["wrapped"].data_ptr(), None)
    assert st == ctdm_amd.EINVAL
    assert b"aligned to two pixels" in lib.ctdm_last_error(ctx.h)

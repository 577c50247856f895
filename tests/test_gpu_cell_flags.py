"""Whole cells with every unwrapping variant, against the FP64 oracle's cell reconstruction.

The batched pipeline takes one of two routes after the retrieval: flag sets without the reference's
IDCT or FastUnwrapping conventions run the unwrapping's last pass fused with the Rytov row transform
(k_uw_rows_inv_rytov), the others the plain unwrapping followed by the Rytov spectrum. Each variant is
run batched (eager and graph-replayed) and plain, and must give the oracle's RI within 1e-5 and the
oracle's unwrapped maps (written by the fused pass when asked for) within FP32 rounding; with the
congruence step a pixel may round to the neighbouring multiple of 2 pi where the least-squares value
sits on a half-integer, so those maps are compared outside such pixels.

FastUnwrapping's conventions (CTDM_UNWRAP_REFCOMPAT_FAST) rescale each map by its own minimum and
maximum (PhaseNormalize, FastUnwrapping.cu:119-142), which the reference takes of calcWrapPhase's value
atan2(BG) - atan2(SP) in (-2 pi, 2 pi) (HilbertTransform.cu:3630), not of the wrapped phase: every
path, both retrievals included, unwraps that value then, and the wrapped-map output stays the wrapped
phase."""
import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu
LAM, NM = 532e-9, 1.337

FLAGS = [
    ctdm_amd.UNWRAP_SINCOS,                                   # fused route
    ctdm_amd.UNWRAP_SINCOS | ctdm_amd.UNWRAP_CONGRUENT,       # fused, congruence in the last pass
    ctdm_amd.UNWRAP_REFCOMPAT_EIG,                            # fused, the reference's eigenvalues
    ctdm_amd.UNWRAP_CONGRUENT,                                # fused
    ctdm_amd.UNWRAP_REFCOMPAT,                                # unfused: the reference's IDCT
    ctdm_amd.UNWRAP_SINCOS | ctdm_amd.UNWRAP_CONGRUENT | ctdm_amd.UNWRAP_REFCOMPAT_FAST,  # unfused
]


def _cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def cell_data():
    A, H, V = 8, 256, 64
    holo, bg, ang = ctdm_amd.synth_cell(61, A, ctdm_amd.synth_params(H=H))
    return A, H, V, holo, bg, ang


CASES = [(f, False) for f in FLAGS] + [
    (0, True), (ctdm_amd.UNWRAP_SINCOS | ctdm_amd.UNWRAP_CONGRUENT, True),
    (ctdm_amd.UNWRAP_SINCOS | ctdm_amd.UNWRAP_CONGRUENT | ctdm_amd.UNWRAP_REFCOMPAT_FAST, True),
]


@pytest.mark.parametrize("flags,rows", CASES)
def test_cell_unwrap_variant_matches_oracle(ctx, orc, cell_data, flags, rows):
    A, H, V, holo, bg, ang = cell_data
    iters = 2
    o = ctdm_amd.optics_for(V, H)
    hd, bd = _cuda(holo), _cuda(bg)
    recon = ctdm_amd.RETRIEVE_ROWS if rows else 0
    outs = {}
    for label, mode, graph in (("batched", 1, False), ("graph", 1, True), ("graph replay", 1, True),
                               ("plain", 0, False)):
        maps = {k: torch.empty((A, V, V), dtype=torch.float32, device="cuda")
                for k in ("wrapped", "unwrapped", "logamp")}
        ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
        cell = dict(holo=hd, bg=bd, angles=ang, V=V, iters=iters, optics=o, ri=ri, unwrap_flags=flags,
                    recon_flags=recon, **maps)
        ctx.reconstruct([cell], mode=mode, graph=graph)
        torch.cuda.synchronize()
        outs[label] = (ri.cpu().numpy(), {k: v.cpu().numpy() for k, v in maps.items()},
                       cell["crop_used"].astuple())
    crop = outs["batched"][2]
    ref, ext = orc.reconstruct_cell(holo, bg, ang, V, crop, LAM, NM, o.dx, iters, unwrap_flags=flags,
                                    recon_flags=recon, outputs=True)
    for label, (ri, m, cr) in outs.items():
        assert cr == crop, label
        err = float(np.abs(ri - ref).max())
        dw = _wrap(m["wrapped"].astype(np.float64) - ext["wrapped"])
        d = m["unwrapped"].astype(np.float64) - ext["unwrapped"]
        if flags & ctdm_amd.UNWRAP_CONGRUENT:
            jumps = np.abs(d) > 1.0  # a 2 pi step at a half-integer least-squares value
            assert jumps.mean() <= 1e-3, (label, jumps.mean())
            d = np.where(jumps, 0.0, d)
        rms = float(np.sqrt(np.mean(d ** 2)))
        print(f"flags {flags:#x} rows {rows} {label}: max|dn| {err:.2e}, unwrapped rms {rms:.2e}, "
              f"wrapped max {np.abs(dw).max():.1e}")
        assert err <= 1e-5, label
        assert rms <= 1e-4, label
        assert np.abs(dw).max() <= 1e-3 and np.all(np.abs(m["wrapped"]) <= np.float32(np.pi)), label
    assert np.array_equal(outs["graph"][0], outs["batched"][0])
    assert np.array_equal(outs["graph replay"][0], outs["batched"][0])


def _wrap(a):
    return np.angle(np.exp(1j * a))


@pytest.mark.parametrize("recon", [ctdm_amd.DC_REFCOMPAT, ctdm_amd.POS_NO_SHIFT,
                                   ctdm_amd.DC_REFCOMPAT | ctdm_amd.POS_NO_SHIFT])
def test_cell_recon_flags_match_oracle(ctx, orc, cell_data, recon):
    """The reference's measured-voxel test (F2.re != 0 && F2.im != 0, Recon_kernel.cu:406) and the volume
    without its final octant shift, batched and plain, against the oracle's cell."""
    A, H, V, holo, bg, ang = cell_data
    iters = 2
    o = ctdm_amd.optics_for(V, H)
    got = {}
    for mode in (1, 0):
        ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
        cell = dict(holo=_cuda(holo), bg=_cuda(bg), angles=ang, V=V, iters=iters, optics=o, ri=ri,
                    recon_flags=recon)
        ctx.reconstruct([cell], mode=mode)
        torch.cuda.synchronize()
        got[mode] = (ri.cpu().numpy(), cell["crop_used"].astuple())
    ref = orc.reconstruct_cell(holo, bg, ang, V, got[1][1], LAM, NM, o.dx, iters, recon_flags=recon)
    for mode, (ri, _) in got.items():
        err = float(np.abs(ri - ref).max())
        print(f"recon flags {recon:#x} mode {mode}: max|dn| {err:.2e}")
        assert err <= 1e-5, mode
    if recon & ctdm_amd.POS_NO_SHIFT:  # the unshifted volume is the shifted one moved by half a side
        shifted = orc.reconstruct_cell(holo, bg, ang, V, got[1][1], LAM, NM, o.dx, iters,
                                       recon_flags=recon & ~ctdm_amd.POS_NO_SHIFT)
        assert np.abs(np.fft.fftshift(got[1][0]) - shifted).max() <= 1e-5


def test_mixed_flag_cells_one_graph(ctx, cell_data):
    """Cells with different retrievals and unwrapping variants in one two-lane graph (each lane's scratch,
    the fused and unfused routes and the calcWrapPhase buffer interleaved): every volume equals its cell
    run alone, on the first call and on replays."""
    A, H, V, holo, bg, ang = cell_data
    o = ctdm_amd.optics_for(V, H)
    variants = [(ctdm_amd.UNWRAP_SINCOS | ctdm_amd.UNWRAP_CONGRUENT, 0),
                (ctdm_amd.UNWRAP_SINCOS | ctdm_amd.UNWRAP_CONGRUENT | ctdm_amd.UNWRAP_REFCOMPAT_FAST, 0),
                (ctdm_amd.UNWRAP_SINCOS | ctdm_amd.UNWRAP_CONGRUENT | ctdm_amd.UNWRAP_REFCOMPAT_FAST,
                 ctdm_amd.RETRIEVE_ROWS),
                (0, ctdm_amd.RETRIEVE_ROWS),
                (ctdm_amd.UNWRAP_REFCOMPAT, ctdm_amd.MAPS_MEDIAN)]
    hd, bd = _cuda(holo), _cuda(bg)

    def cells():
        return [dict(holo=hd, bg=bd, angles=ang, V=V, iters=2, optics=o, unwrap_flags=u, recon_flags=r,
                     ri=torch.empty((V, V, V), dtype=torch.complex64, device="cuda")) for u, r in variants]

    alone = []
    for c in cells():
        ctx.reconstruct([c], mode=1)
        torch.cuda.synchronize()
        alone.append(c["ri"].clone())
    together = cells()
    for rep in range(3):
        for c in together:
            c["ri"].fill_(float("nan"))
        ctx.reconstruct(together, mode=1, graph=True)
        torch.cuda.synchronize()
        for i, (c, r) in enumerate(zip(together, alone)):
            assert torch.equal(c["ri"], r), (rep, i)


@pytest.mark.parametrize("carrier", [(0.375, 0.375), (0.25, 0.375), (0.33, 0.2)])
def test_cell_carriers_off_the_diagonal_and_at_the_edge(ctx, orc, carrier):
    """Cells whose sideband window reaches the spectrum's edge (carrier 0.375 cycles/pixel: window
    [384, 512) of a 512 spectrum, passband as wide as the window) or sits off the diagonal: the device
    crop from the search, the batched pipeline (eager and replayed) and the plain one against the
    oracle's cell with the same crop."""
    A, H, V, iters = 8, 512, 128, 2
    holo, bg, ang = ctdm_amd.synth_cell(71, A, ctdm_amd.synth_params(H=H, carrier=carrier))
    o = ctdm_amd.optics_for(V, H)
    hd, bd = _cuda(holo), _cuda(bg)
    got = {}
    for label, mode, graph in (("batched", 1, False), ("graph", 1, True), ("replay", 1, True), ("plain", 0, False)):
        ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
        cell = dict(holo=hd, bg=bd, angles=ang, V=V, iters=iters, optics=o, ri=ri)
        ctx.reconstruct([cell], mode=mode, graph=graph)
        torch.cuda.synchronize()
        got[label] = (ri.cpu().numpy(), cell["crop_used"].astuple())
    crop = got["batched"][1]
    expect = (round(H / 2 + carrier[1] * H), round(H / 2 + carrier[0] * H))
    assert crop[:2] == expect and crop == orc.find_sideband(bg[0]) + (V,)
    ref = orc.reconstruct_cell(holo, bg, ang, V, crop, LAM, NM, o.dx, iters)
    for label, (ri, cr) in got.items():
        assert cr == crop, label
        err = float(np.abs(ri - ref).max())
        print(f"carrier {carrier} {label}: crop {cr}, max|dn| {err:.2e}")
        assert err <= 1e-5, label


def test_cell_with_many_angles(ctx, orc):
    """400 angles (13 chunks of the gridding's 32-angle staging, a data-consistency list of up to 400
    entries per pencil): batched and plain against the oracle's cell."""
    A, H, V, iters = 400, 256, 64, 2
    holo, bg, ang = ctdm_amd.synth_cell(81, A, ctdm_amd.synth_params(H=H))
    o = ctdm_amd.optics_for(V, H)
    got = {}
    for mode in (1, 0):
        ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
        cell = dict(holo=_cuda(holo), bg=_cuda(bg), angles=ang, V=V, iters=iters, optics=o, ri=ri)
        ctx.reconstruct([cell], mode=mode)
        torch.cuda.synchronize()
        got[mode] = (ri.cpu().numpy(), cell["crop_used"].astuple())
    ref = orc.reconstruct_cell(holo, bg, ang, V, got[1][1], LAM, NM, o.dx, iters)
    for mode, (ri, cr) in got.items():
        err = float(np.abs(ri - ref).max())
        print(f"{A} angles mode {mode}: max|dn| {err:.2e}")
        assert cr == got[1][1] and err <= 1e-5

"""Row-wise retrieval (the reference's default Algorithm B, ctdm_retrieve_rows / ctdm_find_sideband_rows)
against the FP64 oracle: sideband column identical, wrapped phase RMS <= 1e-4 rad, log-amplitude RMS
<= 1e-4; the cell pipeline with CTDM_RETRIEVE_ROWS (batched and plain) within 1e-5 RI of the oracle."""
import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu
LAM, NM = 532e-9, 1.337


def _rms(a):
    return float(np.sqrt(np.mean(np.asarray(a, np.float64) ** 2)))


@pytest.mark.parametrize("H", [256, 512, 1024])
def test_rows_retrieval_matches_oracle(ctx, orc, H):
    holo, bg = ctdm_amd.synth_hologram(H, H)
    center = ctx.find_sideband_rows(torch.from_numpy(bg).cuda())
    assert center == orc.find_sideband_rows(bg)
    hs = np.stack([holo, bg * 0.9 + holo * 0.1])
    bs = np.stack([bg, bg])
    w, la = ctx.retrieve_rows(torch.from_numpy(hs).cuda(), torch.from_numpy(bs).cuda(), center)
    for b in range(2):
        r = orc.retrieve_rows(hs[b], bs[b], center)
        dw = np.angle(np.exp(1j * (w[b].cpu().numpy() - r["wrapped"])))
        dl = la[b].cpu().numpy() - r["logamp"]
        print(f"H={H} item {b}: center {center}, wrapped rms {_rms(dw):.2e} max {np.abs(dw).max():.2e}, "
              f"logamp rms {_rms(dl):.2e}")
        assert _rms(dw) <= 1e-4 and _rms(dl) <= 1e-4


@pytest.mark.parametrize("k0", [40, 90, 100, 120])
def test_rows_sideband_search_matches_oracle_across_carriers(ctx, orc, k0):
    """The HistogramFT search over shifted columns [0.6 W, W) for carriers across that range, including
    the reference's mirror of a peak within W/8 of the end to W - column (HilbertTransform.cu:1641): of a
    256-wide image, carrier bins 40 and 90 peak inside the range (the resampling's x scale (W-1)/W moves
    the peak by up to one bin), bins 100 and 120 within W/8 of the end and come back mirrored."""
    N = 256
    x = np.mgrid[0:N, 0:N][1].astype(np.float64)
    bg = (2 + 2 * np.cos(2 * np.pi * k0 * x / N)).astype(np.float32)
    got = ctx.find_sideband_rows(torch.from_numpy(bg).cuda())
    ref = orc.find_sideband_rows(bg)
    print(f"carrier bin {k0}: column {got} (oracle {ref})")
    assert got == ref
    peak = k0 + N // 2
    if peak >= N - N // 8:
        assert 0 < got <= N // 8
    else:
        assert abs(got - peak) <= 1


def test_rows_retrieval_rejects_band_outside_spectrum(ctx):
    holo, bg = ctdm_amd.synth_hologram(1, 256)
    with pytest.raises(RuntimeError):
        ctx.retrieve_rows(torch.from_numpy(holo).cuda(), torch.from_numpy(bg).cuda(), 250)


@pytest.mark.parametrize("mode", [1, 0])
def test_cell_pipeline_rows(ctx, orc, mode):
    A, H, V, iters = 10, 256, 64, 3
    holo, bg, ang = ctdm_amd.synth_cell(51, A, ctdm_amd.synth_params(H=H))
    ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    o = ctdm_amd.optics_for(V, H)
    cell = dict(holo=torch.from_numpy(holo).cuda(), bg=torch.from_numpy(bg).cuda(), angles=ang, V=V, iters=iters,
                optics=o, ri=ri, recon_flags=ctdm_amd.RETRIEVE_ROWS)
    ctx.reconstruct([cell], mode=mode, graph=(mode == 1))
    torch.cuda.synchronize()
    cr = cell["crop_used"]
    assert cr.col == orc.find_sideband_rows(bg[0])
    ref = orc.reconstruct_cell(holo, bg, ang, V, cr.astuple(), LAM, NM, o.dx, iters,
                               recon_flags=ctdm_amd.RETRIEVE_ROWS)
    err = np.abs(ri.cpu().numpy() - ref).max()
    print(f"rows cell mode {mode}: max|dn| vs oracle {err:.2e}, RI {ref.real.min():.4f}..{ref.real.max():.4f}")
    assert err <= 1e-5

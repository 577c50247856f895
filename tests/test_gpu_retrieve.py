"""GPU parity of the sideband retrieval (ctdm_find_sideband / ctdm_retrieve) against the FP64
oracle on the seeded 1024^2 off-axis hologram workload (sphere phase + 2nd-order background).
Bar: wrapped phase RMS error <= 1e-4 rad (compared modulo 2 pi), log-amplitude and the SP/BG field
within 1e-4 absolute RMS."""
import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu


def _wrap(a):
    return np.angle(np.exp(1j * a))


@pytest.fixture(scope="module")
def holo1():
    return ctdm_amd.synth_hologram(1, 1024)


def test_find_sideband_matches_oracle(ctx, orc, holo1):
    _, bg = holo1
    crop = ctx.find_sideband(torch.from_numpy(bg).cuda(), M=256)
    assert (crop.row, crop.col, crop.radius, crop.M) == (768, 768, 85, 256)
    assert orc.find_sideband(bg) == (768, 768, 85)


def test_retrieve_matches_oracle(ctx, orc, holo1):
    holo, bg = holo1
    crop = ctdm_amd.Crop(768, 768, 85, 256)
    w, la, fld = ctx.retrieve(torch.from_numpy(holo).cuda(), torch.from_numpy(bg).cuda(), crop,
                              field=True)
    torch.cuda.synchronize()
    r = orc.retrieve(holo, bg, 768, 768, 85, 256)
    dw = _wrap(w[0].cpu().numpy().astype(np.float64) - r["wrapped"])
    print("wrapped rms", np.sqrt(np.mean(dw ** 2)), "max", np.abs(dw).max())
    assert np.sqrt(np.mean(dw ** 2)) <= 1e-4
    assert np.abs(dw).max() <= 1e-3
    dl = la[0].cpu().numpy() - r["logamp"]
    assert np.sqrt(np.mean(dl ** 2)) <= 1e-4
    df = fld[0].cpu().numpy() - r["field"]
    assert np.sqrt(np.mean(np.abs(df) ** 2)) <= 1e-4
    assert np.all(np.abs(w[0].cpu().numpy()) <= np.float32(np.pi))


@pytest.mark.parametrize("M,H", [(128, 1024), (192, 1024), (256, 512), (64, 512)])
def test_retrieve_other_sizes_match_oracle(ctx, orc, M, H):
    holo, bg = ctdm_amd.synth_hologram(4, H)
    c = 3 * H // 4
    rad = (c - H // 2) // 3
    crop = ctdm_amd.Crop(c, c, rad, M)
    w, la = ctx.retrieve(torch.from_numpy(holo).cuda(), torch.from_numpy(bg).cuda(), crop)
    r = orc.retrieve(holo, bg, c, c, rad, M)
    dw = _wrap(w[0].cpu().numpy().astype(np.float64) - r["wrapped"])
    assert np.sqrt(np.mean(dw ** 2)) <= 1e-4
    assert np.sqrt(np.mean((la[0].cpu().numpy() - r["logamp"]) ** 2)) <= 1e-4


def test_retrieve_batch_equals_single(ctx):
    hs, bs = zip(*[ctdm_amd.synth_hologram(s, 1024) for s in (11, 12, 13)])
    H = torch.from_numpy(np.stack(hs)).cuda()
    B = torch.from_numpy(np.stack(bs)).cuda()
    crop = ctdm_amd.Crop(768, 768, 85, 256)
    w, la = ctx.retrieve(H, B, crop)
    for i in range(3):
        w1, l1 = ctx.retrieve(H[i], B[i], crop)
        assert torch.equal(w1[0], w[i]) and torch.equal(l1[0], la[i])


def test_retrieve_rejects_bad_crop(ctx):
    z = torch.zeros((1, 1024, 1024), device="cuda")
    with pytest.raises(ctdm_amd.CtdmError) as e:
        ctx.retrieve(z, z, ctdm_amd.Crop(1000, 768, 85, 256))
    assert e.value.status == ctdm_amd.EINVAL


def _carrier_hologram(N, row, col, seed):
    """Off-axis pair whose sideband sits at (row, col) of the fftshifted N x N spectrum: object with a smooth
    phase bump and amplitude dip, unit reference wave, additive noise 0.01."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    fx, fy = (col - N // 2) / N, (row - N // 2) / N
    ref = np.exp(2j * np.pi * (fx * x + fy * y))
    r2 = ((x - 0.45 * N) ** 2 + (y - 0.55 * N) ** 2) / (0.12 * N) ** 2
    obj = (1 - 0.2 * np.exp(-r2)) * np.exp(1j * 6.0 * np.exp(-r2))
    holo = np.abs(obj + ref) ** 2 + 0.01 * rng.standard_normal((N, N))
    bg = np.abs(1 + ref) ** 2 + 0.01 * rng.standard_normal((N, N))
    return holo.astype(np.float32), bg.astype(np.float32)


@pytest.mark.parametrize("row,col,radius", [(128, 896, 85), (896, 128, 85), (128, 128, 128), (384, 640, 42)])
def test_retrieve_crops_at_the_limits_match_oracle(ctx, orc, row, col, radius):
    """Sidebands whose crop window touches the spectrum's edges (the passband's rows and columns wrap
    modulo the spectrum in the kernels), a passband as wide as the window, and an off-diagonal carrier,
    against the oracle's literal crop. The sideband search (shifted rows from 600 on) finds the carrier,
    or for a carrier in the upper half its conjugate twin, like the oracle."""
    holo, bg = _carrier_hologram(1024, row, col, row + col)
    hd, bd = torch.from_numpy(holo).cuda(), torch.from_numpy(bg).cuda()
    found = ctx.find_sideband(bd, M=256)
    expect = (row, col) if row >= 600 else (1024 - row, 1024 - col)
    assert (found.row, found.col) == expect and orc.find_sideband(bg)[:2] == expect
    crop = ctdm_amd.Crop(row, col, radius, 256)
    w, la, fld = ctx.retrieve(hd, bd, crop, field=True)
    r = orc.retrieve(holo, bg, row, col, radius, 256)
    dw = _wrap(w[0].cpu().numpy().astype(np.float64) - r["wrapped"])
    df = fld[0].cpu().numpy() - r["field"]
    dl = la[0].cpu().numpy() - r["logamp"]
    print(f"crop ({row}, {col}, {radius}): wrapped rms {np.sqrt(np.mean(dw ** 2)):.1e}, "
          f"field rms {np.sqrt(np.mean(np.abs(df) ** 2)):.1e}")
    assert np.sqrt(np.mean(dw ** 2)) <= 1e-4 and np.abs(dw).max() <= 1e-3
    assert np.sqrt(np.mean(np.abs(df) ** 2)) <= 1e-4
    assert np.sqrt(np.mean(dl ** 2)) <= 1e-4

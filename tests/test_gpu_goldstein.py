"""GPU Goldstein (ctdm_goldstein) against the host specification: residue maps, branch-cut masks,
integer 2*pi counts and unwrapped phases must be bit-exact. Workload: 64 wrapped maps of 512^2
(8 Gaussians up to 40 rad + U(-0.8 pi, 0.8 pi) noise, dense residues)."""
import time

import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu


def _check(ctx, orc, w):
    wd = torch.from_numpy(w).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res, cut, k, un = ctx.goldstein(wd)
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    t0 = time.perf_counter()
    rr, rc, rk, ru = orc.goldstein(w)
    tc = time.perf_counter() - t0
    assert np.array_equal(res.cpu().numpy(), rr), "residues"
    assert np.array_equal(cut.cpu().numpy(), rc), "cuts"
    assert np.array_equal(k.cpu().numpy(), rk), "k2pi"
    ug = un.cpu().numpy()
    bad = np.argwhere(ug.view(np.uint32) != ru.view(np.uint32))
    if len(bad):
        kk = k.cpu().numpy()
        for b, y, x in bad[:5]:
            print("mismatch", (b, y, x), "gpu", repr(ug[b, y, x]), "ref", repr(ru[b, y, x]), "k", kk[b, y, x],
                  "phi", repr(w[b, y, x]))
    assert len(bad) == 0, f"unwrapped: {len(bad)} of {ug.size} differ"
    return tg, tc, rr


def test_goldstein_small_maps_bit_exact(ctx, orc):
    for (B, H, W, seed) in [(3, 37, 53, 1), (2, 64, 64, 2), (1, 8, 8, 3), (2, 130, 97, 4)]:
        w = ctdm_amd.synth_wrapped(seed, B, H, W)
        _check(ctx, orc, w)


def test_goldstein_vortex_pair(ctx, orc):
    N = 16
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    phi = np.arctan2(y - 7.5, x - 4.5) - np.arctan2(y - 7.5, x - 10.5)
    w = np.angle(np.exp(1j * phi)).astype(np.float32)[None]
    _, _, rr = _check(ctx, orc, w)
    assert (rr != 0).sum() == 2


def test_goldstein_config3_bit_exact(ctx, orc):
    w = ctdm_amd.synth_wrapped(3, 64, 512, 512)
    ctx.goldstein(torch.from_numpy(w[:2]).cuda())  # warm-up
    tg, tc, rr = _check(ctx, orc, w)
    dens = float((rr != 0).mean())
    print(f"config3: GPU {tg * 1e3:.1f} ms, oracle {tc:.2f} s, residue density {dens:.3f}")
    assert dens > 0.01


def test_goldstein_large_maps_level_path(ctx, orc):
    """1024^2 maps: the commit's presence bits (128 KiB) and the flood fill's cut bits stay in LDS, the
    fill's levels go to global memory and the parents are taken from them afterwards."""
    w = ctdm_amd.synth_wrapped(5, 2, 1024, 1024)
    tg, tc, rr = _check(ctx, orc, w)
    print(f"2 x 1024^2: GPU {tg * 1e3:.1f} ms, oracle {tc:.2f} s")


def test_goldstein_commit_global_presence(ctx, orc):
    """Maps of 1.57 M pixels, a tall one and a wide one: their presence bits and cut bits (192 KiB each)
    exceed 160 KiB of LDS, so the commit keeps the presence bits in global memory (device-scope atomics,
    cleared once per batch and before each rescan) and the flood fill runs entirely in global memory."""
    for shape, seed in (((1, 1536, 1024), 8), ((1, 1024, 1536), 9)):
        w = ctdm_amd.synth_wrapped(seed, *shape)
        tg, tc, rr = _check(ctx, orc, w)
        print(f"{shape}: GPU {tg * 1e3:.1f} ms, oracle {tc:.2f} s, residue density {(rr != 0).mean():.3f}")


def test_goldstein_wide_and_tall_maps_level_path(ctx, orc):
    """Non-square maps past the LDS budget of the flood fill's 2-bit codes (cut bits in LDS, levels in
    global memory): rows and columns must not be exchanged anywhere on those paths; a wide map and a tall
    one in separate calls."""
    for shape, seed in (((1, 640, 1536), 6), ((1, 1400, 720), 7)):
        w = ctdm_amd.synth_wrapped(seed, *shape)
        tg, tc, rr = _check(ctx, orc, w)
        print(f"{shape}: GPU {tg * 1e3:.1f} ms, oracle {tc:.2f} s, residue density {(rr != 0).mean():.3f}")


def test_goldstein_fill_queue_spill(ctx, orc):
    """A 1264 x 1020 map: its cut bits (161 160 B) leave about 2 KiB of 160 KiB of LDS for the flood fill's
    two level queues (327 entries each), so most levels keep the rest of their frontier in global memory."""
    w = ctdm_amd.synth_wrapped(10, 1, 1264, 1020)
    tg, tc, rr = _check(ctx, orc, w)
    print(f"1 x 1264 x 1020: GPU {tg * 1e3:.1f} ms, oracle {tc:.2f} s, residue density {(rr != 0).mean():.3f}")


def test_goldstein_no_residues(ctx, orc):
    """A smooth ramp has no residues: no cuts, every pixel reached, unwrapped equal to the ramp."""
    H, W = 48, 70
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    phi = 0.3 * x + 0.2 * y
    w = np.angle(np.exp(1j * phi)).astype(np.float32)[None]
    _, _, rr = _check(ctx, orc, w)
    assert (rr != 0).sum() == 0


def test_goldstein_maximum_residue_density(ctx, orc):
    """Uniform random phase (residue density about 1/3): most listed hits are consumed before a residue
    is reached, so the commit's rescans run constantly."""
    rng = np.random.default_rng(11)
    w = rng.uniform(-np.pi, np.pi, size=(4, 96, 120)).astype(np.float32)
    _, _, rr = _check(ctx, orc, w)
    assert (rr != 0).mean() > 0.2


def test_goldstein_input_beyond_one_period(ctx, orc):
    """Inputs that are not wrapped (uniform noise of +-15 rad): neighbour differences exceed two periods,
    so the residue kernel's wrap takes its general remainder path, and must still match bit for bit."""
    rng = np.random.default_rng(13)
    w = rng.uniform(-15.0, 15.0, size=(2, 96, 120)).astype(np.float32)
    _check(ctx, orc, w)


def test_goldstein_thin_and_narrow_maps(ctx, orc):
    """Maps at the border limits: 3 rows or 3 columns, and a 9-wide tall map whose box radius
    (min(901, W/2) = 4) never reaches a border from the middle column, so an isolated vortex there
    exhausts its search and stays unbalanced."""
    rng = np.random.default_rng(12)
    for shape in [(2, 3, 40), (2, 40, 3), (1, 3, 3)]:
        _check(ctx, orc, rng.uniform(-np.pi, np.pi, size=shape).astype(np.float32))
    H, W = 200, 9
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    phi = np.arctan2(y - 100.5, x - 4.5)
    w = np.angle(np.exp(1j * phi)).astype(np.float32)[None]
    _, _, rr = _check(ctx, orc, w)
    assert (rr != 0).sum() >= 1

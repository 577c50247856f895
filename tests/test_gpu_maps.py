"""Map sanitising before gridding (BatchRecon's MedianFilter + checkMaps) and the BatchRecon boundary
with it. Bars: median and checkMaps bit-exact against the oracle, NaN/Inf windows included; RI from
maps with the median (amplitude or ln A input, batched or plain) within 1e-5 of the FP64 oracle; the
cell pipeline with CTDM_MAPS_MEDIAN within 1e-5 of the oracle cell reconstruction."""
import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu
LAM, NM = 532e-9, 1.337


def _maps_with_specials(shape, seed):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal(shape).astype(np.float32)
    flat = m.reshape(-1)
    idx = rng.permutation(m.size)
    k = max(1, m.size // 100)  # about 1 % each of NaN, +inf, -inf and exact zeros
    for j, v in enumerate((np.nan, np.inf, -np.inf, 0.0)):
        flat[idx[j * k:(j + 1) * k]] = v
    return m


@pytest.mark.parametrize("shape", [(3, 64, 80), (2, 256, 256), (1, 3, 3), (1, 2, 5)])
def test_median_bit_exact(ctx, orc, shape):
    m = _maps_with_specials(shape, sum(shape))
    got = ctx.median3x3(torch.from_numpy(m).cuda()).cpu().numpy()
    ref = orc.median3x3(m)
    assert np.array_equal(got, ref, equal_nan=True)


def test_check_maps_bit_exact(ctx, orc):
    ph = _maps_with_specials((2, 32, 32), 1)
    am = _maps_with_specials((2, 32, 32), 2)
    pd, ad = torch.from_numpy(ph).cuda(), torch.from_numpy(am).cuda()
    ctx.check_maps(pd, ad)
    rp, ra = orc.check_maps(ph, am)
    assert np.array_equal(pd.cpu().numpy(), rp) and np.array_equal(ad.cpu().numpy(), ra)


def _cell(seed, A, H):
    p = ctdm_amd.synth_params(H=H)
    return ctdm_amd.synth_cell(seed, A, p)


@pytest.mark.parametrize("flags", [ctdm_amd.MAPS_MEDIAN | ctdm_amd.MAPS_AMPLITUDE, ctdm_amd.MAPS_MEDIAN,
                                   ctdm_amd.MAPS_AMPLITUDE])
@pytest.mark.parametrize("plain", [False, True])
def test_recon_from_maps_with_median(ctx, orc, flags, plain):
    A, H, V, iters = 10, 256, 64, 3
    holo, bg, ang = _cell(41, A, H)
    un = torch.empty((A, V, V), dtype=torch.float32, device="cuda")
    la = torch.empty_like(un)
    ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    o = ctdm_amd.optics_for(V, H)
    cell = dict(holo=torch.from_numpy(holo).cuda(), bg=torch.from_numpy(bg).cuda(), angles=ang, V=V, iters=iters,
                optics=o, ri=ri, unwrapped=un, logamp=la)
    ctx.reconstruct([cell], mode=1)
    torch.cuda.synchronize()
    m2 = torch.exp(la) if flags & ctdm_amd.MAPS_AMPLITUDE else la
    got = ctx.recon_from_maps(un, m2, ang, o, iters, flags | (ctdm_amd.EXEC_PLAIN if plain else 0))
    torch.cuda.synchronize()
    ref = orc.recon_from_maps(un.cpu().numpy(), m2.cpu().numpy(), ang, LAM, NM, o.dx, iters, flags)
    err = np.abs(got.cpu().numpy() - ref).max()
    print(f"recon_from_maps flags {flags} plain {plain}: max|dn| vs oracle {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("mode", [1, 0])
def test_cell_pipeline_with_median(ctx, orc, mode):
    A, H, V, iters = 12, 256, 64, 3
    holo, bg, ang = _cell(42, A, H)
    ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    o = ctdm_amd.optics_for(V, H)
    cell = dict(holo=torch.from_numpy(holo).cuda(), bg=torch.from_numpy(bg).cuda(), angles=ang, V=V, iters=iters,
                optics=o, ri=ri, recon_flags=ctdm_amd.MAPS_MEDIAN)
    ctx.reconstruct([cell], mode=mode)
    torch.cuda.synchronize()
    ref = orc.reconstruct_cell(holo, bg, ang, V, cell["crop_used"].astuple(), LAM, NM, o.dx, iters,
                               recon_flags=ctdm_amd.MAPS_MEDIAN)
    plain_ref = orc.reconstruct_cell(holo, bg, ang, V, cell["crop_used"].astuple(), LAM, NM, o.dx, iters)
    err = np.abs(ri.cpu().numpy() - ref).max()
    moved = np.abs(ref - plain_ref).max()
    print(f"cell with median, mode {mode}: max|dn| vs oracle {err:.2e} (the median moves RI by {moved:.2e})")
    assert err <= 1e-5
    assert moved > 1e-5  # the filter has an effect the comparison can see

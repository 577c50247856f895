"""GPU parity of the DCT least-squares unwrapper (ctdm_unwrap_dct) against the FP64 oracle.

Bar: RMS phase error <= 1e-4 rad against the oracle for every flag combination and size; the
reference-compatible mode reproduces the reference's own IDCT dump from its src dump; the plain
(reference-structured) and batched paths agree; the adjoint passes a dot-product test."""
from pathlib import Path

import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden" / "ref_dct_chain"


def _raw(name):
    return np.fromfile(GOLD / f"{name}.256.256.raw", "<f4").reshape(256, 256)


def _maps(N, batch, seed, noise=0.05):
    """Wrapped maps of smooth phases (|grad| well below pi) plus small noise."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    out = []
    for _ in range(batch):
        phi = np.zeros((N, N))
        for _ in range(5):
            cx, cy = rng.uniform(0, N, 2)
            s = rng.uniform(N / 8, N / 4)
            phi += rng.uniform(-25, 25) * (N / 256) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
        phi += noise * rng.standard_normal((N, N))
        out.append(np.angle(np.exp(1j * phi)))
    return np.stack(out).astype(np.float32)


def _rms(a):
    return float(np.sqrt(np.mean(np.asarray(a, np.float64) ** 2)))


FLAGS = [0, ctdm_amd.UNWRAP_SINCOS, ctdm_amd.UNWRAP_REFCOMPAT, ctdm_amd.UNWRAP_REFCOMPAT_EIG,
         ctdm_amd.UNWRAP_CONGRUENT, ctdm_amd.UNWRAP_SINCOS | ctdm_amd.UNWRAP_CONGRUENT]


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("N", [64, 128, 192, 256])
def test_unwrap_matches_oracle(ctx, orc, N, flags):
    w = _maps(N, 3, N + flags)
    got = ctx.unwrap_dct(torch.from_numpy(w).cuda(), flags).cpu().numpy()
    ref = orc.unwrap_dct(w.astype(np.float64), flags)
    err = got - ref
    if flags & ctdm_amd.UNWRAP_CONGRUENT:  # integer snapping: compare modulo exact 2 pi steps
        assert np.mean(np.abs(err) > 1e-3) < 1e-4
        err = np.where(np.abs(err) > 1e-3, 0, err)
    print(N, flags, "rms", _rms(err), "max", np.abs(err).max())
    assert _rms(err) <= 1e-4
    assert np.abs(err).max() <= 1e-3


def test_reference_mode_reproduces_reference_dumps(ctx):
    src = torch.from_numpy(_raw("src")).cuda()
    got = ctx.unwrap_dct(src, ctdm_amd.UNWRAP_REFCOMPAT).cpu().numpy()
    err = got - _raw("IDCT")
    print("vs reference IDCT dump: rms", _rms(err), "max", np.abs(err).max())
    assert _rms(err) <= 3e-5 and np.abs(err).max() <= 2e-4


@pytest.mark.parametrize("flags", [0, ctdm_amd.UNWRAP_SINCOS, ctdm_amd.UNWRAP_REFCOMPAT,
                                   ctdm_amd.UNWRAP_CONGRUENT])
@pytest.mark.parametrize("N", [128, 192, 256])
def test_plain_equals_batched(ctx, N, flags):
    w = torch.from_numpy(_maps(N, 2, 7 * N + flags)).cuda()
    a = ctx.unwrap_dct(w, flags)
    b = ctx.unwrap_dct(w, flags | ctdm_amd.EXEC_PLAIN)
    err = (a - b).cpu().numpy()
    if flags & ctdm_amd.UNWRAP_CONGRUENT:
        err = np.where(np.abs(err) > 1e-3, 0, err)
    assert _rms(err) <= 2e-5


def test_least_squares_closed_form(ctx):
    N = 256
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    phi = 20 * np.exp(-((x - 90) ** 2 + (y - 150) ** 2) / (2 * 45.0 ** 2)) + 0.03 * x
    w = np.angle(np.exp(1j * phi)).astype(np.float32)
    got = ctx.unwrap_dct(torch.from_numpy(w).cuda()).cpu().numpy()
    assert np.abs(got - (phi - phi.mean())).max() <= 1e-3


def test_config1_retrieval_and_unwrap_chain(ctx, orc):
    holo, bg = ctdm_amd.synth_hologram(1, 1024)
    crop = ctx.find_sideband(torch.from_numpy(bg).cuda(), M=256)
    w, la = ctx.retrieve(torch.from_numpy(holo).cuda(), torch.from_numpy(bg).cuda(), crop)
    u = ctx.unwrap_dct(w).cpu().numpy()[0]
    r = orc.retrieve(holo, bg, crop.row, crop.col, crop.radius, crop.M)
    ur = orc.unwrap_dct(r["wrapped"], 0)[0]
    err = u - ur
    print("config-1 unwrapped rms", _rms(err), "max", np.abs(err).max())
    assert _rms(err) <= 1e-4


@pytest.mark.parametrize("flags", [0, ctdm_amd.UNWRAP_REFCOMPAT_EIG, ctdm_amd.UNWRAP_SINCOS])
def test_adjoint_dot_product(ctx, orc, flags):
    N = 128
    rng = np.random.default_rng(11 + flags)
    phi = _maps(N, 1, 99)[0].astype(np.float64)
    v = 0.05 * rng.standard_normal((N, N))
    g = rng.standard_normal((N, N))
    # J v by central differences of the FP64 oracle at phi
    eps = 1e-4
    Jv = (orc.unwrap_dct(phi + eps * v, flags) - orc.unwrap_dct(phi - eps * v, flags))[0] / (2 * eps)
    JTg = ctx.unwrap_dct_adjoint(torch.from_numpy(phi.astype(np.float32)).cuda(),
                                 torch.from_numpy(g.astype(np.float32)).cuda(), flags).cpu().numpy()
    lhs, rhs = float(np.sum(Jv * g)), float(np.sum(v * JTg))
    print("dot", lhs, rhs)
    assert abs(lhs - rhs) <= 1e-4 * max(1.0, abs(lhs))


@pytest.mark.parametrize("N", [128, 256])
def test_fast_unwrapping_conventions_match_oracle(ctx, orc, N):
    """CTDM_UNWRAP_REFCOMPAT_FAST (FastUnwrapping's min-max rescaling of the input and minimum subtracted
    from the result) against the oracle, on full-range wrapped maps and on a map confined to [0, 0.9] rad
    that the rescaling stretches; every output map has its minimum at exactly 0; the reference-named
    wrapper and the plain path give the same maps."""
    flags = ctdm_amd.UNWRAP_SINCOS | ctdm_amd.UNWRAP_CONGRUENT | ctdm_amd.UNWRAP_REFCOMPAT_FAST
    y, x = np.mgrid[0:N, 0:N] / N
    bump = (0.9 * np.exp(-((x - 0.5) ** 2 + (y - 0.4) ** 2) / (2 * 0.1 ** 2)))[None].astype(np.float32)
    w = np.concatenate([_maps(N, 2, 3 * N), bump])
    wd = torch.from_numpy(w).cuda()
    got = ctx.unwrap_dct(wd, flags)
    ref = orc.unwrap_dct(w.astype(np.float64), flags)
    err = got.cpu().numpy() - ref
    assert np.mean(np.abs(err) > 1e-3) < 1e-4  # integer snapping: compare modulo exact 2 pi steps
    err = np.where(np.abs(err) > 1e-3, 0, err)
    print(N, "rms", _rms(err), "max", np.abs(err).max())
    assert _rms(err) <= 1e-4 and np.abs(err).max() <= 1e-3
    assert torch.equal(got.amin(dim=(1, 2)), torch.zeros(3, device="cuda"))
    assert torch.equal(ctx.FastUnwrapping(wd), got)
    assert torch.equal(wd.cpu(), torch.from_numpy(w))  # the input is left as it was
    plain = ctx.unwrap_dct(wd, flags | ctdm_amd.EXEC_PLAIN)
    d = (plain - got).cpu().numpy()
    assert _rms(np.where(np.abs(d) > 1e-3, 0, d)) <= 2e-5

"""Differentiable reconstruction: the hand-written adjoints of the Ewald interpolation
(ctdm_ewald_gather <-> ctdm_ewald_scatter) and of the 3-D FFT, wrapped in torch.autograd.Function here
(test harness only), and the projected-gradient fit ctdm_fit_pgd.

Bars: dot-product identities to FP32 rounding; the autograd gradient of
J(f) = 1/2 |G FFT3(f) - y|^2 matches FP64 central differences (oracle gather + numpy FFT) on a 32^3 /
8-angle instance; fit with step 1 == the data-consistency loop; fit == FP64 oracle fit (RI 1e-5)."""
import numpy as np
import pytest
import torch

import ctdm_amd
from phantom_metrics import check_recovery

pytestmark = pytest.mark.gpu
LAM, NM = 532e-9, 1.337


def _angles(A, na=1.0):
    s = na / NM
    az = 2 * np.pi * np.arange(A) / A
    return np.stack([np.arcsin(s * np.sin(az)), np.arcsin(s * np.cos(az))], axis=1)


class EwaldGather(torch.autograd.Function):
    """spec = G F (ctdm_ewald_gather); backward: G^H g = ctdm_ewald_scatter(g)."""

    @staticmethod
    def forward(fctx, F, lib, ang, optics):
        fctx.lib, fctx.ang, fctx.optics, fctx.V = lib, ang, optics, F.shape[0]
        return lib.ewald_gather(F.contiguous(), ang, F.shape[0], optics)

    @staticmethod
    def backward(fctx, g):
        F, _ = fctx.lib.ewald_scatter(g.contiguous(), fctx.ang, fctx.V, fctx.optics)
        return F, None, None, None


class FFT3(torch.autograd.Function):
    """Unnormalised forward 3-D DFT (ctdm_fft3d(-1)); backward: FFT3^H g = V^3 * ctdm_fft3d(g, +1)."""

    @staticmethod
    def forward(fctx, x, lib):
        fctx.lib = lib
        return lib.fft3d(x.contiguous().clone(), -1)

    @staticmethod
    def backward(fctx, g):
        V = g.shape[0]
        return fctx.lib.fft3d(g.contiguous().clone(), +1) * float(V ** 3), None


def _rand_c(shape, seed, dev="cuda"):
    g = torch.Generator(device=dev).manual_seed(seed)
    return torch.randn(shape, dtype=torch.complex64, device=dev, generator=g)


def test_adjoint_dot_products(ctx):
    V, A = 32, 8
    o = ctdm_amd.optics_for(V)
    ang = _angles(A)
    F = _rand_c((V, V, V), 1)
    s = _rand_c((A, V, V), 2)
    GF = ctx.ewald_gather(F, ang, V, o)
    Ss, _ = ctx.ewald_scatter(s, ang, V, o)
    lhs = torch.vdot(GF.flatten().to(torch.complex128), s.flatten().to(torch.complex128)).item()
    rhs = torch.vdot(F.flatten().to(torch.complex128), Ss.flatten().to(torch.complex128)).item()
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    x, y = _rand_c((V, V, V), 3), _rand_c((V, V, V), 4)
    Fx = ctx.fft3d(x.clone(), -1)
    FHy = ctx.fft3d(y.clone(), +1) * float(V ** 3)
    lhs = torch.vdot(Fx.flatten().to(torch.complex128), y.flatten().to(torch.complex128)).item()
    rhs = torch.vdot(x.flatten().to(torch.complex128), FHy.flatten().to(torch.complex128)).item()
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_gradient_matches_fp64_finite_differences(ctx, orc):
    V, A = 32, 8
    o = ctdm_amd.optics_for(V)
    ang = _angles(A)
    rng = np.random.default_rng(7)
    f = (rng.standard_normal((V, V, V)) + 1j * rng.standard_normal((V, V, V))).astype(np.complex64)
    y = (rng.standard_normal((A, V, V)) * 30 + 1j * rng.standard_normal((A, V, V)) * 30).astype(np.complex64)
    d = (rng.standard_normal((V, V, V)) + 1j * rng.standard_normal((V, V, V))).astype(np.complex64)

    ft = torch.from_numpy(f).cuda().requires_grad_(True)
    J = 0.5 * (EwaldGather.apply(FFT3.apply(ft, ctx), ctx, ang, o) - torch.from_numpy(y).cuda()).abs().pow(2).sum()
    J.backward()
    g = ft.grad.cpu().numpy().astype(np.complex128)

    def J64(x):  # FP64 oracle of the same objective; volume x + V y + V^2 z <-> numpy [z, y, x]
        r = orc.ewald_gather(np.fft.fftn(x), ang, LAM, NM, o.dx) - y.astype(np.complex128)
        return 0.5 * float(np.sum(np.abs(r) ** 2))

    f64, d64 = f.astype(np.complex128), d.astype(np.complex128)
    eps = 1e-3
    fd = (J64(f64 + eps * d64) - J64(f64 - eps * d64)) / (2 * eps)
    an = float(np.real(np.vdot(g, d64)))
    print("directional derivative: fd", fd, "adjoint", an)
    assert abs(fd - an) <= 1e-4 * abs(fd)
    assert abs(J.item() - J64(f64)) <= 1e-5 * J64(f64)


def _spectra(ctx, seed, A, H, V, field_noise=0.0):
    p = ctdm_amd.synth_params(H=H, field_noise=field_noise)
    holo, bg, ang = ctdm_amd.synth_cell(seed, A, p)
    hd, bd = torch.from_numpy(holo).cuda(), torch.from_numpy(bg).cuda()
    crop = ctx.find_sideband(bd[0], M=V)
    w, la = ctx.retrieve(hd, bd, crop)
    un = ctx.unwrap_dct(w)
    spec = ctx.rytov_spectrum(un, la)
    o = ctdm_amd.optics_for(V, H)
    F, hits = ctx.ewald_scatter(spec, ang, V, o)
    Fm = ctx.normalize_measured(F, hits, o.dx)
    f0 = ctx.fft3d(F, +1)
    return spec, ang, o, f0, Fm, hits


def test_fit_step_one_is_the_positivity_loop(ctx):
    V = 64
    spec, ang, o, f0, Fm, hits = _spectra(ctx, 31, 12, 256, V)
    n_pos = ctx.positivity(f0.clone(), Fm, hits, 4, o)
    n_fit = ctx.fit_pgd(f0.clone(), spec, ang, o, 4, 1.0)
    err = (n_pos - n_fit).abs().max().item()
    print("fit(step 1) vs positivity max|dn|", err)
    assert err <= 1e-5


@pytest.mark.parametrize("step", [0.5, 1.0])
def test_fit_matches_oracle(ctx, orc, step):
    V = 64
    spec, ang, o, f0, _, _ = _spectra(ctx, 32, 10, 256, V)
    f0n = f0.cpu().numpy()
    n = ctx.fit_pgd(f0.clone(), spec, ang, o, 6, step).cpu().numpy()
    ref = orc.fit_pgd(f0n, spec.cpu().numpy(), ang, LAM, NM, o.dx, 6, step)
    err = np.abs(n - ref).max()
    print("fit step", step, "max|dn| vs oracle", err)
    assert err <= 1e-5


@pytest.mark.parametrize("step", [0.5, 1.0])
def test_fit_objective_matches_oracle(ctx, orc, step):
    """ctdm_fit_pgd_trace: the objective before each step, summed in the fit's own passes, equals the
    oracle's sample-by-sample J; the volume is bit-identical to ctdm_fit_pgd's."""
    V, iters = 64, 6
    spec, ang, o, f0, _, _ = _spectra(ctx, 33, 10, 256, V, field_noise=0.02)
    f0n = f0.cpu().numpy()
    n_plain = ctx.fit_pgd(f0.clone(), spec, ang, o, iters, step)
    n, J = ctx.fit_pgd_trace(f0.clone(), spec, ang, o, iters, step)
    assert torch.equal(n, n_plain)
    _, Jref = orc.fit_pgd(f0n, spec.cpu().numpy(), ang, LAM, NM, o.dx, iters, step, objective=True)
    rel = np.abs(np.array(J) - Jref) / Jref
    print("step", step, "J", np.array(J), "rel err vs oracle", rel.max())
    assert rel.max() <= 1e-4
    assert np.all(np.diff(J) < 0)


def test_config5_full_size(ctx, orc):
    """120 angles of 1024^2 holograms (field noise 0.02) -> 256^3, 50 fit iterations. The step-1 fit (the
    positivity loop on the averaged measurement) matches the oracle; a step-1/2 fit reports its objective,
    which falls at every one of the 50 steps; both volumes recover the phantom, and their RMS errors
    against it are reported side by side."""
    V, A, H, iters = 256, 120, 1024, 50
    spec, ang, o, f0, _, _ = _spectra(ctx, 5, A, H, V, field_noise=0.02)
    f0n = f0.cpu().numpy()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    n = ctx.fit_pgd(f0.clone(), spec, ang, o, iters, 1.0)
    e1.record()
    torch.cuda.synchronize()
    ref = orc.fit_pgd(f0n, spec.cpu().numpy(), ang, LAM, NM, o.dx, iters, 1.0)
    err = np.abs(n.cpu().numpy() - ref).max()
    print(f"config5: fit {e0.elapsed_time(e1):.2f} ms for {iters} iterations, max|dn| vs oracle {err:.2e}, "
          f"RI {ref.real.min():.4f}..{ref.real.max():.4f}")
    assert err <= 1e-5
    n_half, J = ctx.fit_pgd_trace(f0.clone(), spec, ang, o, iters, 0.5)
    J = np.array(J)
    print(f"config5 step 1/2: J {J[0]:.4e} -> {J[-1]:.4e}, largest change {np.diff(J).max():.3e}")
    assert np.all(np.diff(J) < 0)
    p = ctdm_amd.synth_params(H=H, field_noise=0.02)
    phantom = ctdm_amd.synth_phantom(5, V, o.dx, p)
    for name, vol in (("step 1", n), ("step 1/2", n_half)):
        v = vol.cpu().numpy()
        rms = float(np.sqrt(np.mean((v.real - phantom.real) ** 2)))
        print(f"config5 {name}: RMS |Re n - phantom| {rms:.3e}")
        assert check_recovery(v, phantom, p, o.dx)

"""Pins the FP64 oracle before it is trusted as the checker.

- FFT: numpy.fft on random data, brute-force O(n^2) DFT for small n.
- DCT unwrapping: the reference's own debug dumps (DCT_Unwrapping.cu:35,47,51,55 wrote
  src/Laplace/DCT/DCT2/IDCT.256.256.raw from one run of DCT_UWLS_Unwrapped), each stage fed from
  the dump of its own input; the exact transforms against scipy dctn/idctn (ortho); the
  least-squares closed form unwrap(wrap(phi)) = phi - mean(phi) for smooth phi.
- Retrieval: a pure-carrier hologram gives zero phase and log-amplitude; a known smooth phase and
  amplitude are recovered; the sideband of a (0.25, 0.25) carrier is found at (768, 768), radius 85.
Each check is chosen so that a dropped term, a wrong sign or a transposed index fails it.
"""
from pathlib import Path

import numpy as np
import pytest
from scipy.fft import dctn, idctn

GOLD = Path(__file__).resolve().parent / "golden" / "ref_dct_chain"


def _raw(name):
    return np.fromfile(GOLD / f"{name}.256.256.raw", "<f4").reshape(256, 256).astype(np.float64)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [2, 3, 12, 64, 128, 192, 256, 384, 512, 1024])
def test_fft_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    np.testing.assert_allclose(orc.fft(x, -1), np.fft.fft(x, axis=-1), atol=1e-10 * n)
    np.testing.assert_allclose(orc.fft(x, +1), np.fft.ifft(x, axis=-1) * n, atol=1e-10 * n)


@pytest.mark.parametrize("n", [5, 16, 48, 64])
def test_fft_matches_bruteforce_dft(orc, n):
    rng = np.random.default_rng(7 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    np.testing.assert_allclose(orc.fft(x, -1), dft, atol=1e-11 * n)


def test_fft3d_matches_numpy(orc):
    rng = np.random.default_rng(3)
    v = rng.standard_normal((16, 16, 16)) + 1j * rng.standard_normal((16, 16, 16))
    # volume index x + V*y + V*V*z  <->  numpy array [z, y, x]
    np.testing.assert_allclose(orc.fft3d(v, -1), np.fft.fftn(v), atol=1e-9)
    np.testing.assert_allclose(orc.fft3d(v, +1), np.fft.ifftn(v), atol=1e-12)


# ---------------------------------------------------------------- DCT chain vs reference dumps
def test_laplacian_matches_reference_dump(orc):
    out = orc.dct_chain(_raw("src"), orc.REF_EIG | orc.REF_IDCT, 0)
    assert np.abs(out["laplace"] - _raw("Laplace")).max() < 2e-6


def test_dct_matches_reference_dump_and_scipy(orc):
    lap = _raw("Laplace")
    out = orc.dct_chain(lap, 0, 1)
    assert np.abs(out["dct"] - _raw("DCT")).max() < 1e-6
    np.testing.assert_allclose(out["dct"], dctn(lap, type=2, norm="ortho"), atol=1e-12)


def test_eigen_division_matches_reference_dump(orc):
    dct = _raw("DCT")
    out = orc.dct_chain(dct, orc.REF_EIG, 2)["dct2"]
    ref = _raw("DCT2")
    # The reference evaluated lambda in FP32 (cosf, DCT_Unwrapping.cu:310): its absolute error is
    # ~ 4 * 2^-24 * 4, i.e. a relative error ~1e-6 / |lambda| in the quotient.
    N = 256
    i = np.arange(N)
    lam = 2 * np.cos(4 * np.pi * i / (4 * N - 1))[None, :] + 2 * np.cos(4 * np.pi * i / (4 * N - 1))[:, None] - 4
    lam[0, 0] = 1.0
    bound = 1e-6 + np.abs(ref) * (1e-6 / np.abs(lam))
    assert np.all(np.abs(out - ref) <= bound)
    assert np.sqrt(np.mean((out - ref) ** 2)) < 1e-5
    # The standard pi/N eigenvalues do NOT reproduce the dump (so the 4N-1 form is what ran).
    assert np.abs(orc.dct_chain(dct, 0, 2)["dct2"] - ref).max() > 1e-3


def test_idct_matches_reference_dump(orc):
    out = orc.dct_chain(_raw("DCT2"), orc.REF_IDCT, 3)["out"]
    assert np.abs(out - _raw("IDCT")).max() < 2e-6
    # the exact inverse differs from the dump: the reference IDCT twiddle is off by one
    exact = orc.dct_chain(_raw("DCT2"), 0, 3)["out"]
    assert np.abs(exact - _raw("IDCT")).max() > 1e-3
    np.testing.assert_allclose(exact, idctn(_raw("DCT2"), type=2, norm="ortho"), atol=1e-12)


def test_dct_chain_end_to_end_reference_mode(orc):
    out = orc.dct_chain(_raw("src"), orc.REF_EIG | orc.REF_IDCT, 0)["out"]
    err = out - _raw("IDCT")
    assert np.sqrt(np.mean(err ** 2)) < 2e-5 and np.abs(err).max() < 1e-4


def _smooth_phase(N, seed=0):
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    rng = np.random.default_rng(seed)
    phi = np.zeros((N, N))
    for _ in range(4):
        cx, cy = rng.uniform(0, N, 2)
        s = rng.uniform(N / 8, N / 4)
        phi += rng.uniform(-30, 30) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    return phi + 0.02 * x - 0.01 * y


@pytest.mark.parametrize("N", [64, 128, 192, 256])
def test_unwrap_least_squares_closed_form(orc, N):
    phi = _smooth_phase(N, N)
    assert np.abs(np.diff(phi, axis=0)).max() < np.pi and np.abs(np.diff(phi, axis=1)).max() < np.pi
    w = np.angle(np.exp(1j * phi))
    u = orc.unwrap_dct(w, 0)[0]
    np.testing.assert_allclose(u, phi - phi.mean(), atol=1e-9)
    # sin/cos right-hand side: same closed form to first order in the gradients; congruence
    # snaps to the exact phase up to the constant
    uc = orc.unwrap_dct(w, orc.SINCOS | orc.CONGRUENT)[0]
    d = uc - phi
    assert np.abs(d - np.round(d.mean() / (2 * np.pi)) * 2 * np.pi).max() < 1e-9


@pytest.mark.parametrize("N", [64, 256])
def test_fast_unwrapping_conventions_closed_form(orc, N):
    """FastUnwrapping's conventions (flag 0x10; FastUnwrapping.cu:7, 108, 119-142, 346-369). A wrapped map
    spanning the whole (-pi, pi] range is only shifted by the min-max rescaling, so the result is the
    true phase minus its minimum; a map confined to [0, 0.9] rad is stretched to [0, 2 pi] first, so the
    result is the stretched map minus its minimum, not x. (The congruence step rounds
    (LS - x) / 2 pi, about -mean(x) / 2 pi since the LS solution has zero mean, so a rescaled map whose
    mean is near pi gets 2 pi jumps in the reference as well; the bump below keeps its mean near 0.4.)"""
    flags = orc.SINCOS | orc.CONGRUENT | 0x10
    phi = _smooth_phase(N, N)
    w = np.angle(np.exp(1j * phi))
    span = w.max() - w.min()
    assert span > 2 * np.pi - 1e-2
    u = orc.unwrap_dct(w, flags)[0]
    assert u.min() == 0.0
    # the rescaled map n = 2 pi (w - min w) / span jumps by 2 pi * 2 pi / span at each wrap; congruence
    # adds the true wrap counts m back in whole 2 pi steps: u = n + 2 pi m, shifted to a zero minimum
    m = np.round((phi - w) / (2 * np.pi))
    n = (2 * np.pi * (w - w.min()) / span).astype(np.float32).astype(np.float64)
    want = n + 2 * np.pi * m
    np.testing.assert_allclose(u, want - want.min(), atol=1e-9)
    y, x = np.mgrid[0:N, 0:N] / N
    small = 0.9 * np.exp(-((x - 0.5) ** 2 + (y - 0.4) ** 2) / (2 * 0.1 ** 2))  # within [0, 0.9], no wrap
    us = orc.unwrap_dct(small, flags)[0]
    stretched = 2 * np.pi * (small - small.min()) / np.ptp(small)
    np.testing.assert_allclose(us, stretched - stretched.min(), atol=1e-5)


def test_sincos_rhs_identity(orc):
    # cos(p) L[sin p] - sin(p) L[cos p] == sum over neighbours of sin(p_n - p)
    rng = np.random.default_rng(5)
    p = rng.uniform(-np.pi, np.pi, (32, 32))
    lap = orc.dct_chain(p, orc.SINCOS, 0)["laplace"]
    ref = np.zeros_like(p)
    ref[:, :-1] += np.sin(p[:, 1:] - p[:, :-1])
    ref[:, 1:] += np.sin(p[:, :-1] - p[:, 1:])
    ref[:-1, :] += np.sin(p[1:, :] - p[:-1, :])
    ref[1:, :] += np.sin(p[:-1, :] - p[1:, :])
    np.testing.assert_allclose(lap, ref, atol=1e-12)


# ---------------------------------------------------------------- retrieval
def _holo(H, phi=None, A=None, fx=0.25, fy=0.25):
    y, x = np.mgrid[0:H, 0:H].astype(np.float64)
    phi = np.zeros((H, H)) if phi is None else phi
    A = np.ones((H, H)) if A is None else A
    I = A * A + 1 + 2 * A * np.cos(2 * np.pi * (fx * x + fy * y) + phi)
    return I.astype(np.float32)


def test_find_sideband_carrier(orc):
    bg = _holo(1024)
    assert orc.find_sideband(bg) == (768, 768, 85)


def test_retrieve_pure_carrier_is_zero(orc):
    bg = _holo(256)
    r = orc.retrieve(bg, bg, 192, 192, 21, 64)
    assert np.abs(r["wrapped"]).max() < 1e-9
    assert np.abs(r["logamp"]).max() < 1e-9


def test_retrieve_recovers_smooth_phase_and_amplitude(orc):
    H, M = 512, 128
    y, x = np.mgrid[0:H, 0:H].astype(np.float64)
    # band-limited phase and amplitude (a few low harmonics, periodic on the grid)
    phi = 1.3 * np.sin(2 * np.pi * 3 * x / H) + 0.7 * np.cos(2 * np.pi * 2 * y / H + 0.4)
    la = 0.1 * np.cos(2 * np.pi * (x + 2 * y) / H)
    holo = _holo(H, phi, np.exp(la))
    bg = _holo(H)
    r = orc.retrieve(holo, bg, 384, 384, 42, M)
    sub = (slice(None, None, H // M),) * 2
    # the crop samples every 4th pixel; arg(SP) - arg(BG) = phi exactly for band-limited inputs
    np.testing.assert_allclose(r["wrapped"], np.angle(np.exp(1j * phi[sub])), atol=2e-3)
    np.testing.assert_allclose(r["logamp"], la[sub], atol=2e-3)
    np.testing.assert_allclose(r["field"], np.exp(la[sub] + 1j * phi[sub]), atol=3e-3)
    # the reference's own wrapped output lies in (-2 pi, 2 pi) and equals ours mod 2 pi
    d = r["phase_ref"] - r["wrapped"]
    np.testing.assert_allclose(np.round(d / (2 * np.pi)) * 2 * np.pi, d, atol=1e-9)


def _rows_holograms(N, k0, phase_amp=2.0, sigma=None, amp_mod=0.0):
    """Carrier along x at bin k0 (k0 / N cycles per pixel); SP carries a smooth phase and amplitude."""
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    sigma = sigma or N / 6
    phi = phase_amp * np.exp(-((x - 0.45 * N) ** 2 + (y - 0.55 * N) ** 2) / (2 * sigma ** 2))
    a = 1.0 + amp_mod * np.exp(-((x - 0.5 * N) ** 2 + (y - 0.5 * N) ** 2) / (2 * sigma ** 2))
    sp = a * a + 1 + 2 * a * np.cos(2 * np.pi * k0 * x / N + phi)
    bg = 2 + 2 * np.cos(2 * np.pi * k0 * x / N)
    return sp.astype(np.float32), bg.astype(np.float32), phi, a


def test_rows_sideband_column_is_the_carrier(orc):
    """HistogramFT peak: carrier bin k0 sits at shifted column k0 + N/2."""
    for N, k0 in [(256, 72), (512, 140)]:
        _, bg, _, _ = _rows_holograms(N, k0)
        assert orc.find_sideband_rows(bg) == k0 + N // 2


def test_rows_sideband_mirrored_near_the_end(orc):
    """sequence1DFFT mirrors a peak within W/8 of the end to W - column (HilbertTransform.cu:1641): a carrier
    at bin 120 of 256 peaks at shifted column 247-248 and comes back as 8-9; one at bin 90 (column 218) is
    kept."""
    N = 256
    for k0, lo, hi in ((120, 1, N // 8), (90, 217, 219)):
        _, bg, _, _ = _rows_holograms(N, k0)
        assert lo <= orc.find_sideband_rows(bg) <= hi, k0


def test_rows_retrieval_identical_pair_is_zero(orc):
    sp, bg, _, _ = _rows_holograms(256, 72)
    r = orc.retrieve_rows(bg, bg, 72 + 128)
    assert np.abs(r["wrapped"]).max() == 0 and np.abs(r["logamp"]).max() == 0


def test_rows_retrieval_recovers_phase_and_amplitude(orc):
    """Demodulation by the cropped band: output (n, y') samples the object at x = 4 n (N-1)/N, y = y' (N-1)/(N/4),
    and the resampling's carrier distortion cancels between SP and BG."""
    N, k0 = 256, 72
    sp, bg, phi, a = _rows_holograms(N, k0, amp_mod=0.2)
    r = orc.retrieve_rows(sp, bg, k0 + N // 2)
    M = N // 4
    n, yq = np.meshgrid(np.arange(M), np.arange(M), indexing="ij")
    xs = 4 * n * (N - 1) / N
    ys = yq * (N - 1) / (N / 4)
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    from scipy.ndimage import map_coordinates
    phi_s = map_coordinates(phi, [ys, xs], order=1)
    la_s = np.log(map_coordinates(a, [ys, xs], order=1))
    inner = (slice(4, M - 4), slice(4, M - 4))  # away from the periodic wrap of the row transform
    assert np.abs(r["wrapped"][inner] - phi_s[inner]).max() < 0.03
    assert np.abs(r["logamp"][inner] - la_s[inner]).max() < 0.03


def test_rows_phase_ref_is_calcwrapphase(orc):
    """The row-wise retrieval's calcWrapPhase value (the difference of the two atan2, in (-2 pi, 2 pi),
    HilbertTransform.cu:3630) is congruent to its wrapped phase modulo 2 pi, and takes values outside
    (-pi, pi] where the two atan2 straddle the cut."""
    N, k0 = 256, 72
    sp, bg, _, _ = _rows_holograms(N, k0, phase_amp=9.0)
    r = orc.retrieve_rows(sp, bg, k0 + N // 2)
    ph, w = r["phase_ref"], r["wrapped"]
    k = (ph - w) / (2 * np.pi)
    assert np.abs(k - np.rint(k)).max() < 1e-9
    assert np.all(np.abs(ph) < 2 * np.pi) and np.all(np.abs(w) <= np.pi)
    assert np.any(np.abs(ph) > np.pi)


def test_angle_from_image_closed_form(orc):
    """A cosine fringe at bins (kx, ky), ky > 0: the shifted peak in the searched band is (kx + N/2, ky + N/2),
    so fx = (N/2 - 1 - posX) / (N dx) = -(kx + 1) / (N dx) and the angle follows from asin."""
    N, kx, ky = 256, 37, 30
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    img = (1 + np.cos(2 * np.pi * (kx * x + ky * y) / N)).astype(np.float32)
    dx, mag, lam, nm = 5.5, 40.0, 532.0, 1.337
    ang, pos = orc.angle_from_image(img, dx, mag, lam, nm)
    assert pos == kx + N // 2
    st = mag * (-(kx + 1) / (N * dx * 1e-6)) / (nm / (lam * 1e-9))
    assert abs(ang - abs(np.arcsin(st))) < 1e-12 and ang > 0

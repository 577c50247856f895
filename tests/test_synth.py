"""Seeded synthetic generators: reproducible, shaped as the workloads need (CPU only)."""
import numpy as np

import ctdm_amd


def test_hologram_reproducible_and_carrier():
    h1, b1 = ctdm_amd.synth_hologram(1, 256)
    h2, b2 = ctdm_amd.synth_hologram(1, 256)
    assert np.array_equal(h1, h2) and np.array_equal(b1, b2)
    h3, _ = ctdm_amd.synth_hologram(2, 256)
    assert not np.array_equal(h1, h3)
    assert abs(float(b1.mean()) - 2.0) < 0.01
    F = np.abs(np.fft.fftshift(np.fft.fft2(b1)))
    F[128 - 8:128 + 8, 128 - 8:128 + 8] = 0
    r, c = np.unravel_index(F[150:].argmax(), F[150:].shape)
    assert (r + 150, c) == (192, 192)


def test_wrapped_maps_in_range():
    w = ctdm_amd.synth_wrapped(3, 2, 64, 96)
    assert w.shape == (2, 64, 96)
    assert w.min() > -np.pi - 1e-6 and w.max() <= np.float32(np.pi)
    assert np.array_equal(w, ctdm_amd.synth_wrapped(3, 2, 64, 96))


def test_cell_angles_on_cone():
    p = ctdm_amd.synth_params(H=128)
    holo, bg, ang = ctdm_amd.synth_cell(5, 6, p)
    assert holo.shape == (6, 128, 128) and ang.shape == (6, 2)
    s = np.sqrt(np.sin(ang[:, 0]) ** 2 + np.sin(ang[:, 1]) ** 2)
    np.testing.assert_allclose(s, p.na_ill / p.n_med, rtol=1e-12)
    assert np.isfinite(holo).all() and np.isfinite(bg).all()

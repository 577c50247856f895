"""Illumination-angle estimation (ctdm_angle_from_image, the reference's AngCal_GPU) against the FP64
oracle: the same spectral peak and the same angle."""
import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu
DX, MAG, LAM, NM = 5.5, 40.0, 532.0, 1.337


def test_angle_matches_oracle(ctx, orc):
    imgs = []
    for N, kx, ky in [(256, 37, 30), (512, 61, 90)]:
        y, x = np.mgrid[0:N, 0:N].astype(np.float64)
        img = (1 + np.cos(2 * np.pi * (kx * x + ky * y) / N)).astype(np.float32)
        got = ctx.angle_from_image(torch.from_numpy(img).cuda(), DX, MAG, LAM, NM)
        ref, _ = orc.angle_from_image(img, DX, MAG, LAM, NM)
        assert abs(got - ref) <= 1e-12, (N, got, ref)
    holo, bg = ctdm_amd.synth_hologram(3, 1024)
    batch = np.stack([holo, bg])
    got = ctx.angle_from_image(torch.from_numpy(batch).cuda(), DX, MAG, LAM, NM)
    for b in range(2):
        ref, pos = orc.angle_from_image(batch[b], DX, MAG, LAM, NM)
        print(f"1024^2 image {b}: peak column {pos}, angle {got[b]:.6f} (oracle {ref:.6f})")
        assert abs(got[b] - ref) <= 1e-12

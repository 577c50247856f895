"""The reference's reload path on the GPU (ReconFlag 2: LoadDateFromDir + RefreshStack, then BatchRecon;
HilbertTransform.cu:1207-1227, 1453-1544): per-angle .phimap frames, one of them marked rejected
(status 0), are read back with load_phimap_stack, which drops that frame; the phase and amplitude stacks
go through ctdm_recon_from_maps with BatchRecon's own flags (median + checkMaps on amplitude stacks,
CTDM_MAPS_MEDIAN | CTDM_MAPS_AMPLITUDE) and must give the oracle's RI to 1e-5; the volume then makes a
.rimap round trip (complex, real and imaginary parts)."""
import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu
LAM, NM = 532e-9, 1.337


def test_phimap_reload_reconstruct_and_rimap_round_trip(ctx, orc, tmp_path):
    A, H, V, iters = 10, 256, 64, 4
    holo, bg, ang = ctdm_amd.synth_cell(31, A, ctdm_amd.synth_params(H=H))
    o = ctdm_amd.optics_for(V, H)
    un = torch.empty((A, V, V), dtype=torch.float32, device="cuda")
    la = torch.empty_like(un)
    cell = dict(holo=torch.from_numpy(holo).cuda(), bg=torch.from_numpy(bg).cuda(), angles=ang, V=V,
                iters=iters, optics=o, ri=torch.empty((V, V, V), dtype=torch.complex64, device="cuda"),
                unwrapped=un, logamp=la)
    ctx.reconstruct([cell], mode=1)
    torch.cuda.synchronize()
    phase, amp = un.cpu().numpy(), np.exp(la.cpu().numpy())
    rejected = 3
    paths = []
    for a in range(A):  # outputImg writes one file per frame: status, angles, sizes, (phase, amp) pairs
        p = tmp_path / f"buffer{a + 1:03d}.phimap"
        ctdm_amd.phimap_write(p, phase[a], amp[a], ang_x=ang[a, 0], ang_y=ang[a, 1], status=(a != rejected))
        paths.append(p)
    ph, am, angs = ctdm_amd.load_phimap_stack(paths)
    keep = [a for a in range(A) if a != rejected]
    assert ph.shape == (A - 1, V, V)
    assert np.array_equal(ph, phase[keep]) and np.array_equal(am, amp[keep].astype(np.float32))
    assert np.array_equal(angs, ang[keep])
    flags = ctdm_amd.MAPS_MEDIAN | ctdm_amd.MAPS_AMPLITUDE
    results = {}
    for mode in (0, ctdm_amd.EXEC_PLAIN):
        ri = ctx.recon_from_maps(torch.from_numpy(ph).cuda(), torch.from_numpy(am).cuda(), angs, o, iters,
                                 flags | mode)
        torch.cuda.synchronize()
        results[mode] = ri.cpu().numpy()
    ref = orc.recon_from_maps(ph, am, angs, LAM, NM, o.dx, iters, flags)
    for mode, got in results.items():
        err = float(np.abs(got - ref).max())
        print(f"reload -> recon ({'plain' if mode else 'batched'}): max|dn| {err:.2e}")
        assert err <= 1e-5
    vol = results[0]
    for part in (0, 1, 2):
        f = tmp_path / f"volume{part}.rimap"
        ctdm_amd.rimap_write(f, vol, part)
        back = ctdm_amd.rimap_read(f, vol.shape, part)
        want = vol if part == 0 else (vol.real + 0j if part == 1 else 1j * vol.imag)
        assert np.array_equal(back, want.astype(np.complex64)), part

"""Batched uneven cells: several cells with different angle counts and volume sides in one
hipGraph-captured launch sequence (ctdm_reconstruct, mode 1, graph) give, cell by cell, the same RI as
running each cell alone, and stay within 1e-5 of the FP64 oracle."""
import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu
LAM, NM = 532e-9, 1.337


def _cells(spec):
    cells, host = [], []
    for i, (A, H, V) in enumerate(spec):
        p = ctdm_amd.synth_params(H=H)
        holo, bg, ang = ctdm_amd.synth_cell(100 + i, A, p)
        ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
        cells.append(dict(holo=torch.from_numpy(holo).cuda(), bg=torch.from_numpy(bg).cuda(), angles=ang, V=V,
                          iters=3, optics=ctdm_amd.optics_for(V, H), ri=ri))
        host.append((holo, bg, ang))
    return cells, host


def test_uneven_cells_one_graph(ctx, orc):
    spec = [(6, 256, 64), (9, 512, 128), (12, 256, 64), (7, 512, 192)]
    cells, host = _cells(spec)
    ctx.reconstruct(cells, mode=1, graph=True, autotune=True)
    torch.cuda.synchronize()
    batched = [c["ri"].clone() for c in cells]
    for rep in range(2):  # replays of the same multi-cell graph
        for c in cells:
            c["ri"].zero_()
        ctx.reconstruct(cells, mode=1, graph=True)
        torch.cuda.synchronize()
        for c, b in zip(cells, batched):
            assert torch.equal(c["ri"], b), rep
    for (A, H, V), c, (holo, bg, ang), b in zip(spec, cells, host, batched):
        single = torch.empty_like(b)
        one = dict(c, ri=single)
        ctx.reconstruct([one], mode=1)
        torch.cuda.synchronize()
        assert torch.equal(single, b)
        ref = orc.reconstruct_cell(holo, bg, ang, V, c["crop_used"].astuple(), LAM, NM, c["optics"].dx, 3)
        err = np.abs(b.cpu().numpy() - ref).max()
        print(f"cell A={A} V={V}: max|dn| {err:.2e}")
        assert err <= 1e-5

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


def test_two_lanes_equal_one_lane(ctx):
    """Cells alternating over two concurrent lanes (the default for several cells) give bit-identical
    RI to the one-lane sequence, eager and graph-replayed, with ordinary eager work between replays."""
    spec = [(5, 256, 64), (8, 256, 32), (6, 256, 64), (4, 256, 32), (3, 256, 64)]  # odd count: lane 0 ends
    cells, _ = _cells(spec)
    ctx.reconstruct(cells, mode=1, lanes=1)
    torch.cuda.synchronize()
    ref = [c["ri"].clone() for c in cells]
    for graph in (False, True):
        for rep in range(3):
            for c in cells:
                c["ri"].fill_(float("nan"))
            ctx.reconstruct(cells, mode=1, graph=graph, lanes=2)
            torch.cuda.synchronize()
            for i, (c, r) in enumerate(zip(cells, ref)):
                assert torch.equal(c["ri"], r), (graph, rep, i)
            scratch = torch.arange(1 << 16, device="cuda", dtype=torch.int64)
            assert scratch.sum().item() == (1 << 16) * ((1 << 16) - 1) // 2
    ctx.reconstruct(cells, mode=1, graph=True, lanes=1)  # its own graph (the key holds the lane count)
    torch.cuda.synchronize()
    for c, r in zip(cells, ref):
        assert torch.equal(c["ri"], r)
    with pytest.raises(RuntimeError, match="lanes"):
        ctx.reconstruct(cells, mode=1, lanes=3)

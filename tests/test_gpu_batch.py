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


def test_graph_follows_a_moved_carrier(ctx):
    """The batched retrieval reads its crop from device memory, written by the sideband search queued
    ahead of the cell: a replayed graph follows a carrier that moved between calls in the same buffers,
    reports the crop it used, and equals a fresh eager run."""
    A, H, V = 5, 256, 64
    srcs = [ctdm_amd.synth_cell(7, A, ctdm_amd.synth_params(H=H, carrier=cr)) for cr in ((0.25, 0.25), (0.22, 0.27))]
    assert np.array_equal(srcs[0][2], srcs[1][2])  # same angles: the second call replays the first graph
    hd, bd = torch.from_numpy(srcs[0][0]).cuda(), torch.from_numpy(srcs[0][1]).cuda()
    ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    cell = dict(holo=hd, bg=bd, angles=srcs[0][2], V=V, iters=3, optics=ctdm_amd.optics_for(V, H), ri=ri)
    ctx.reconstruct([cell], mode=1, graph=True)
    torch.cuda.synchronize()
    crop_a = cell["crop_used"].astuple()
    hd.copy_(torch.from_numpy(srcs[1][0]))
    bd.copy_(torch.from_numpy(srcs[1][1]))
    ctx.reconstruct([cell], mode=1, graph=True)
    torch.cuda.synchronize()
    crop_b = cell["crop_used"].astuple()
    print("crops", crop_a, crop_b)
    assert crop_a != crop_b
    fresh = dict(cell, ri=torch.empty_like(ri))
    ctx.reconstruct([fresh], mode=1)
    torch.cuda.synchronize()
    assert fresh["crop_used"].astuple() == crop_b
    assert torch.equal(ri, fresh["ri"])


def test_located_crop_outside_the_spectrum_is_rejected(ctx):
    """A carrier whose crop window would leave the spectrum is reported as an error once the search has
    been read back (the queued work reads the same crop modulo the spectrum size, so it stays in bounds);
    the context keeps working afterwards."""
    A, H, V = 3, 256, 64
    holo, bg, ang = ctdm_amd.synth_cell(8, A, ctdm_amd.synth_params(H=H, carrier=(0.45, 0.25)))
    ri = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    cell = dict(holo=torch.from_numpy(holo).cuda(), bg=torch.from_numpy(bg).cuda(), angles=ang, V=V, iters=2,
                optics=ctdm_amd.optics_for(V, H), ri=ri)
    with pytest.raises(RuntimeError, match="leaves the spectrum"):
        ctx.reconstruct([cell], mode=1)
    torch.cuda.synchronize()
    good = _cells([(3, 256, 64)])[0][0]
    ctx.reconstruct([good], mode=1, graph=True)
    torch.cuda.synchronize()
    assert torch.isfinite(good["ri"]).all()


def test_cells_with_different_carriers_share_one_graph(ctx):
    """Cells whose carriers differ get their own located crops (one device slot per cell) inside one
    two-lane graph; every cell equals its own eager single-cell run, replays included."""
    specs = [((0.25, 0.25), 64), ((0.22, 0.27), 32), ((0.27, 0.22), 64)]
    cells = []
    for i, (carrier, V) in enumerate(specs):
        holo, bg, ang = ctdm_amd.synth_cell(60 + i, 4, ctdm_amd.synth_params(H=256, carrier=carrier))
        cells.append(dict(holo=torch.from_numpy(holo).cuda(), bg=torch.from_numpy(bg).cuda(), angles=ang, V=V,
                          iters=2, optics=ctdm_amd.optics_for(V, 256),
                          ri=torch.empty((V, V, V), dtype=torch.complex64, device="cuda")))
    ctx.reconstruct(cells, mode=1, graph=True)
    torch.cuda.synchronize()
    crops = [c["crop_used"].astuple() for c in cells]
    assert len(set(crops)) == 3, crops
    first = [c["ri"].clone() for c in cells]
    for c in cells:
        c["ri"].zero_()
    ctx.reconstruct(cells, mode=1, graph=True)  # replay
    torch.cuda.synchronize()
    for c, f, cr in zip(cells, first, crops):
        assert torch.equal(c["ri"], f)
        one = dict(c, ri=torch.empty_like(f))
        ctx.reconstruct([one], mode=1)
        torch.cuda.synchronize()
        assert one["crop_used"].astuple() == cr
        assert torch.equal(one["ri"], f)


def test_graph_replay_follows_new_angles(ctx):
    """A replayed graph reads the angle rows written ahead of it on every call: angles alpha, then beta,
    then alpha again on the same buffers each equal an eager run with those angles (one cell, and two
    cells on two lanes where only the second cell's angles change)."""
    A, H, V = 8, 256, 64
    (cell,), _ = _cells([(A, H, V)])
    alpha = np.array(cell["angles"])
    beta = 0.8 * alpha[::-1].copy()

    def run(cells, graph):
        ctx.reconstruct(cells, mode=1, graph=graph)
        torch.cuda.synchronize()
        return [c["ri"].clone() for c in cells]

    refs = {}
    for name, ang in (("a", alpha), ("b", beta)):
        cell["angles"] = ang
        refs[name] = run([cell], False)[0]
    assert not torch.equal(refs["a"], refs["b"])
    for name in "abab":
        cell["angles"] = alpha if name == "a" else beta
        assert torch.equal(run([cell], True)[0], refs[name]), name

    pair, _ = _cells([(6, 256, 64), (7, 256, 32)])
    a2 = np.array(pair[1]["angles"])
    b2 = 0.9 * a2[::-1].copy()
    want = {}
    for name, ang in (("a", a2), ("b", b2)):
        pair[1]["angles"] = ang
        want[name] = run(pair, False)
    for name in "abba":
        pair[1]["angles"] = a2 if name == "a" else b2
        got = run(pair, True)
        for g, w in zip(got, want[name]):
            assert torch.equal(g, w), name


def test_scratch_and_graph_cache_stay_bounded(ctx):
    """Per-cell scratch is per lane, not per output buffer, and the graph cache evicts: a dozen calls that
    differ only in the output pointer (each captured as its own graph) leave device memory where the
    first call left it, and every output equals the first."""
    A, H, V = 30, 512, 128
    (cell,), _ = _cells([(A, H, V)])
    outs = [torch.empty((V, V, V), dtype=torch.complex64, device="cuda") for _ in range(12)]
    ctx.reconstruct([dict(cell, ri=outs[0])], mode=1, graph=True)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for r in outs[1:]:
        ctx.reconstruct([dict(cell, ri=r)], mode=1, graph=True)
        ctx.reconstruct([dict(cell, ri=r)], mode=1, graph=True)  # replay
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    grown = (free0 - free1) / 2**20
    print(f"device memory after 11 more outputs: {grown:+.1f} MiB")
    # the per-output scratch of a cell this size is about 10 MiB, so 11 outputs would show > 100 MiB
    assert grown < 48
    for r in outs[1:]:
        assert torch.equal(r, outs[0])


def test_context_on_its_own_streams(ctx):
    """A context bound to a caller's stream runs there (eager, captured and replayed), and after
    ctdm_set_stream replays its cached graph on the new stream: every volume equals the one from the
    session context on torch's default stream, with other work queued on the default stream meanwhile."""
    cells, _ = _cells([(6, 256, 64), (5, 256, 32)])
    ctx.reconstruct(cells, mode=1)
    torch.cuda.synchronize()
    ref = [c["ri"].clone() for c in cells]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    own = ctdm_amd.Context(0, stream=s1)
    try:
        for rep, s in enumerate((s1, s1, s1, s2, s2)):
            if rep == 3:
                own.set_stream(s2)
            for c in cells:
                c["ri"].fill_(float("nan"))
            torch.cuda.synchronize()
            own.reconstruct(cells, mode=1, graph=rep > 0)
            busy = torch.randn(1 << 20, device="cuda").sum()  # unrelated work on the default stream
            s.synchronize()
            for i, (c, r) in enumerate(zip(cells, ref)):
                assert torch.equal(c["ri"], r), (rep, i)
            assert torch.isfinite(busy)
    finally:
        own.close()

"""Config 4 at its own workload: eight cells with A = {30, 45, 60, 60, 90, 90, 120, 120} angles of 1024^2
hologram pairs (615 pairs) and volume sides cyclic over {128, 192, 256} (crop side = V), 10 positivity
iterations, all in one captured graph on two lanes. Every cell is compared with the FP64 oracle
(RI within 1e-5), the replayed graph with the eager pass, and the two-lane sequence with the one-lane
sequence (bit-identical)."""
import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu
LAM, NM = 532e-9, 1.337
H, ITERS = 1024, 10
SPEC = list(zip([30, 45, 60, 60, 90, 90, 120, 120], [128, 192, 256, 128, 192, 256, 128, 192]))


@pytest.fixture(scope="module")
def config4(ctx):
    cells, host = [], []
    for i, (A, V) in enumerate(SPEC):
        holo, bg, ang = ctdm_amd.synth_cell(40 + i, A)
        cells.append(dict(holo=torch.from_numpy(holo).cuda(), bg=torch.from_numpy(bg).cuda(), angles=ang, V=V,
                          iters=ITERS, optics=ctdm_amd.optics_for(V, H),
                          ri=torch.empty((V, V, V), dtype=torch.complex64, device="cuda")))
        host.append((holo, bg, ang))
    print(f"config4: {sum(a for a, _ in SPEC)} hologram pairs synthesised", flush=True)
    ctx.reconstruct(cells, mode=1, graph=True)  # eager pass + capture
    torch.cuda.synchronize()
    eager = [c["ri"].clone() for c in cells]
    for c in cells:
        c["ri"].fill_(float("nan"))
    ctx.reconstruct(cells, mode=1, graph=True)  # replay
    torch.cuda.synchronize()
    two = [c["ri"].clone() for c in cells]
    ctx.reconstruct(cells, mode=1, graph=True, lanes=1)
    torch.cuda.synchronize()
    one = [c["ri"].clone() for c in cells]
    crops = [c["crop_used"].astuple() for c in cells]
    yield host, eager, two, one, crops
    for c in cells:
        del c["holo"], c["bg"]
    torch.cuda.empty_cache()


def test_config4_replay_and_lanes_bit_identical(config4):
    _, eager, two, one, crops = config4
    for i, (e, t, o) in enumerate(zip(eager, two, one)):
        assert torch.equal(e, t), i
        assert torch.equal(t, o), i
    assert all(cr[:3] == (768, 768, 85) for cr in crops), crops


@pytest.mark.parametrize("i", range(len(SPEC)))
def test_config4_cell_matches_oracle(config4, orc, i):
    host, _, two, _, crops = config4
    holo, bg, ang = host[i]
    A, V = SPEC[i]
    o = ctdm_amd.optics_for(V, H)
    ref = orc.reconstruct_cell(holo, bg, ang, V, crops[i], LAM, NM, o.dx, ITERS)
    err = float(np.abs(two[i].cpu().numpy() - ref).max())
    print(f"cell {i}: A={A} V={V} max|dn| {err:.2e}")
    assert err <= 1e-5

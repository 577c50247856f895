"""Pins the oracle's Goldstein specification (residues, branch cuts, flood fill) with hand-enumerated
cases and invariants:
- a vortex pair gives exactly one +1 and one -1 residue at the loops around the vortex centres;
- residues agree with the integer circulation of rounded differences (the same quantity computed
  without the reference's float/double modulo), away from exact +-pi ties;
- the box search reproduces hand-traced cuts (opposite charge, border, same-sign skip);
- on a residue-free map the flood fill recovers the true phase up to one global 2*pi multiple;
- a cut wall isolates the far side (INT32_MIN), and cut pixels take k from a reached neighbour."""
import numpy as np

INT_MIN = np.iinfo(np.int32).min


def _vortex_pair(N=16, p1=(4.5, 7.5), p2=(10.5, 7.5)):
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    phi = np.arctan2(y - p1[1], x - p1[0]) - np.arctan2(y - p2[1], x - p2[0])
    return np.angle(np.exp(1j * phi)).astype(np.float32)


def test_vortex_pair_residues(orc):
    r = orc.residues(_vortex_pair())
    nz = np.argwhere(r != 0)
    assert len(nz) == 2
    # loop (i..i+1, j..j+1) encloses the centre (i + 0.5, j + 0.5); arrays are [y, x]
    assert set(map(tuple, nz)) == {(7, 4), (7, 10)}
    assert r[7, 4] == -r[7, 10]
    assert r[-1].sum() == 0 and r[:, -1].sum() == 0


def test_residues_equal_integer_circulation(orc):
    rng = np.random.default_rng(0)
    phi = rng.uniform(-np.pi, np.pi, (64, 64)).astype(np.float32)
    r = orc.residues(phi)
    p = phi.astype(np.float64)

    def s(a, b):  # number of 2 pi removed by wrapping a - b into [-pi, pi)
        return np.floor((a - b + np.pi) / (2 * np.pi))

    a, b, c, d = p[:-1, :-1], p[:-1, 1:], p[1:, 1:], p[1:, :-1]
    circ = -(s(a, b) + s(b, c) + s(c, d) + s(d, a))
    ref = np.zeros_like(r)
    ref[:-1, :-1] = circ
    assert np.array_equal(r, ref)


def test_branch_cut_hand_cases(orc):
    res = np.zeros((8, 8), np.int8)
    res[3, 3], res[3, 5] = 1, -1            # (x=3,y=3) + and (x=5,y=3) -
    cuts = orc.branch_cuts(res)
    exp = np.zeros((8, 8), np.uint8)
    exp[3, 3:6] = 1                          # straight segment (3,3)-(5,3)
    assert np.array_equal(cuts, exp)

    res = np.zeros((8, 8), np.int8)
    res[4, 1] = 1                            # x=1 next to the left border
    cuts = orc.branch_cuts(res)
    exp = np.zeros((8, 8), np.uint8)
    exp[4, 0:2] = 1
    assert np.array_equal(cuts, exp)

    res = np.zeros((8, 8), np.int8)
    res[0, 3] = -1                           # on the border row: one-pixel cut
    cuts = orc.branch_cuts(res)
    assert cuts.sum() == 1 and cuts[0, 3] == 1

    res = np.zeros((12, 12), np.int8)
    res[5, 5], res[5, 6], res[5, 8] = 1, 1, -1   # same sign skipped; raster order pairs (5,5)
    cuts = orc.branch_cuts(res)
    assert cuts[5, 5] == 1 and cuts[5, 8] == 1 and cuts[5, 6] == 1 and cuts[5, 7] == 1
    # (6,5) is then left unbalanced; at r = 5 the first scanned perimeter point is (1, 0), whose row
    # is on the border, so the cut runs straight up along x = 6
    assert cuts[0:6, 6].all() and not cuts[5, 9:12].any()
    assert cuts.sum() == 4 + 5


def test_flood_fill_recovers_smooth_phase(orc):
    N = 48
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    true = 9 * np.exp(-((x - 20) ** 2 + (y - 30) ** 2) / 200.0) + 0.2 * x
    w = np.angle(np.exp(1j * true)).astype(np.float32)
    assert not orc.residues(w).any()
    k, un = orc.flood_fill(w, np.zeros((N, N), np.uint8))
    assert (k != INT_MIN).all()
    d = un.astype(np.float64) - true
    c = 2 * np.pi * np.round(d.mean() / (2 * np.pi))
    assert np.abs(d - c).max() < 1e-4


def test_flood_fill_cut_wall(orc):
    N = 32
    w = np.zeros((N, N), np.float32)
    cuts = np.zeros((N, N), np.uint8)
    cuts[:, 20] = 1                          # full-height wall at x = 20
    k, un = orc.flood_fill(w, cuts)
    assert (k[:, :20] == 0).all()            # seed (16, 20) side reached
    assert (k[:, 21:] == INT_MIN).all() and (un[:, 21:] == 0).all()
    # interior wall pixels adjoin reached pixels on the left: filled from there
    assert (k[1:-1, 20] == 0).all()
    assert k[0, 20] == INT_MIN and k[-1, 20] == INT_MIN   # border rows are not filled


def test_goldstein_full_chain_balanced(orc):
    rng = np.random.default_rng(4)
    N = 64
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    true = 25 * np.exp(-((x - 30) ** 2 + (y - 34) ** 2) / 300.0)
    w = np.angle(np.exp(1j * (true + rng.uniform(-0.8 * np.pi, 0.8 * np.pi, (N, N))))).astype(np.float32)
    res, cut, k, un = orc.goldstein(w)
    reached = k[0] != INT_MIN
    assert reached.mean() > 0.3 and res.any()
    d = (un[0].astype(np.float64) - w)[reached]
    np.testing.assert_allclose(d, 2 * np.pi * k[0][reached], atol=1e-4)
    assert cut[0][res[0] != 0].all()          # every residue lies on a cut

"""Pins the scale and orientation of the oracle's Ewald mapping independently of its own formula.

1. Born quadrature. The first-order Born field of a weak Gaussian potential is computed by direct
   quadrature of the scattering integral with the outgoing Green's function e^{ikR} / (4 pi R) (no
   Fourier-diffraction formula is used), taken as the first-order Rytov field psi = u_s / u_0 on the
   detector grid, refocused to z = 0 by angular-spectrum propagation, and passed through the oracle's
   Rytov spectrum, Ewald scatter and Initial_F_F2 normalisation. Every measured voxel must then hold the
   analytic 3-D transform of the potential, -O(K) / dx^3 (the reference's potential is f = -o,
   o = k0^2 (n^2 - n_m^2)). This fixes the 4 pi f_z weight (Recon_kernel.cu:224-225), the 1/dx of
   Initial_F_F2 (:269-270) and the metre units together: a 2 pi for 4 pi slip, a micrometre dx or a
   dropped f_z moves every voxel by a large factor. It also fixes the axis convention: the reference
   transposes the maps (cuCopyArray, :162) and takes f_x0 from AngY (:175), so volume x is image y and
   AngY tilts the illumination along image y; the swapped reading is rejected.
2. A literal numpy restatement of fillEdwardSphere / cuMod / cuRound / Initial_F_F2
   (Recon_kernel.cu:187-276) on a sparse spectrum and two angles that hit one voxel twice: the oracle's
   F, hit counts and F2 equal it.
"""
import numpy as np
import pytest

LAM, NM = 532e-9, 1.337
FM = NM / LAM


def _geometry(ax, ay, V, dx):
    df = 1.0 / (V * dx)
    fx0 = FM * np.sin(ay)  # EdwardSphere: f_x0 from AngY (volume x)
    fy0 = FM * np.sin(ax)
    fz0 = np.sqrt(FM * FM - fx0 * fx0 - fy0 * fy0)
    return df, fx0, fy0, fz0


class _Blob:
    """o(r) = eps exp(-|r - r0|^2 / 2 s^2) in image coordinates (x, y, z) [m]; o in 1/m^2."""

    def __init__(self, eps, s, r0):
        self.eps, self.s, self.r0 = eps, s, np.asarray(r0, float)

    def ft(self, K):
        """Continuous transform O(K) = int o(r) exp(-2 pi i K.r) d^3r, K in cycles/m (image axes)."""
        K = np.asarray(K, float)
        k2 = float(K @ K)
        return (self.eps * (2 * np.pi * self.s ** 2) ** 1.5 * np.exp(-2 * np.pi ** 2 * self.s ** 2 * k2)
                * np.exp(-2j * np.pi * float(K @ self.r0)))

    def born_rytov(self, f_in, V, dx, zd, h):
        """psi(x, y) = u_s(x, y, zd) / u_0(x, y, zd) on the V x V detector grid ([y][x], origin at the
        centre pixel), u_s by quadrature of int G(r - r') o(r') u_0(r') d^3r' on a grid of step h."""
        g1 = np.arange(-5 * self.s, 5 * self.s + h / 2, h)
        ox, oy, oz = (a.ravel() for a in np.meshgrid(*(self.r0[i] + g1 for i in range(3)), indexing="ij"))
        o = self.eps * np.exp(-((ox - self.r0[0]) ** 2 + (oy - self.r0[1]) ** 2 + (oz - self.r0[2]) ** 2)
                              / (2 * self.s ** 2))
        src = h ** 3 * o * np.exp(2j * np.pi * (f_in[0] * ox + f_in[1] * oy + f_in[2] * oz))
        c = (np.arange(V) - V // 2) * dx
        k = 2 * np.pi * FM
        psi = np.empty((V, V), complex)
        for j, y in enumerate(c):
            R = np.sqrt((c[:, None] - ox[None]) ** 2 + (y - oy[None]) ** 2 + (zd - oz[None]) ** 2)
            us = (np.exp(1j * k * R) / (4 * np.pi * R)) @ src
            psi[j] = us / np.exp(2j * np.pi * (f_in[0] * c + f_in[1] * y + f_in[2] * zd))
        return psi


def _measured_vs_analytic(orc, blob, ax, ay, V, dx, zd, convention):
    df, fx0, fy0, fz0 = _geometry(ax, ay, V, dx)
    # incident direction in image coordinates: volume x is image y (the reference's transposition)
    f_in = np.array([fy0, fx0, fz0]) if convention == "reference" else np.array([fx0, fy0, fz0])
    psi = blob.born_rytov(f_in, V, dx, zd, h=0.2e-6)
    spec = orc.rytov_spectrum(psi.imag[None], psi.real[None])
    F, hits = orc.ewald_scatter(spec, [(ax, ay)], V, LAM, NM, dx)
    Fn, _ = orc.normalize(F, hits, dx)
    got, want = [], []
    for nz, ny, nx in zip(*np.nonzero(hits)):
        ii = nx if nx < V // 2 else nx - V
        jj = ny if ny < V // 2 else ny - V
        fz = np.sqrt(FM ** 2 - (ii * df + fx0) ** 2 - (jj * df + fy0) ** 2)
        K_img = (jj * df, ii * df, fz - fz0)
        want.append(-blob.ft(K_img) / dx ** 3)
        got.append(Fn[nz, ny, nx] * np.exp(-2j * np.pi * (fz - fz0) * zd))  # refocus zd -> 0
    return np.array(got), np.array(want)


# The 40-degree tilt moves the scattered field 4 um sideways on its way to the detector plane, so it
# needs the wider 19.2 um window (V = 48): at 12.8 um its truncated edge alone costs 4e-4 of the peak.
@pytest.mark.parametrize("ax,ay,V", [(0.0, 0.0, 32), (0.3, -0.2, 32), (-0.45, 0.5, 48)])
def test_ewald_scatter_matches_born_quadrature(orc, ax, ay, V):
    blob = _Blob(eps=1.0e12, s=0.7e-6, r0=(0.4e-6, -0.8e-6, 0.6e-6))
    got, want = _measured_vs_analytic(orc, blob, ax, ay, V=V, dx=0.4e-6, zd=5e-6, convention="reference")
    peak = np.abs(want).max()
    err = np.abs(got - want).max() / peak
    print(f"angles ({ax}, {ay}): {len(got)} measured voxels, max |F - F_born| / peak = {err:.2e}")
    assert len(got) > 300
    assert err <= 1e-4
    # a factor 2 in the weight (2 pi for 4 pi) or a micrometre dx would be off by 0.5 / 1e6 of the peak
    ratio = np.vdot(want, got).real / np.vdot(want, want).real
    assert abs(ratio - 1.0) <= 1e-4


def test_swapped_axis_convention_is_rejected(orc):
    blob = _Blob(eps=1.0e12, s=0.7e-6, r0=(0.4e-6, -0.8e-6, 0.6e-6))
    got, want = _measured_vs_analytic(orc, blob, 0.3, -0.2, V=32, dx=0.4e-6, zd=5e-6, convention="swapped")
    assert np.abs(got - want).max() / np.abs(want).max() > 0.3


def _cuMod(a, b):
    return ((a % b) + b) % b if a < 0 else a % b  # Recon_kernel.cu:232-235 (C remainder semantics)


def _cuRound(x):
    return np.floor(x + 0.5) if x > 0.0 else np.ceil(x - 0.5)  # :236-239


def _fill_edward_sphere_literal(u_sp, F, C, fx0, fy0, fz0, fm0, df, V):
    """fillEdwardSphere (Recon_kernel.cu:187-229) line by line, one (i, j) thread at a time."""
    fm02 = fm0 * fm0
    for j in range(V):
        for i in range(V):
            ii = int(i - V / 2.0)
            jj = int(j - V / 2.0)
            fx, fy = ii * df + fx0, jj * df + fy0
            err = (fm02 - fx0 * fx0 - fy0 * fy0) < 0 or (fm02 - fx * fx - fy * fy) < 0
            if err:
                continue
            fz = np.sqrt(fm02 - (fx * fx + fy * fy))
            Nx, Ny = _cuMod(ii, V), _cuMod(jj, V)
            Nz = _cuMod(int(_cuRound((fz - fz0) / df)), V)
            idx = Nx + Ny * V + Nz * V * V
            u = u_sp[Nx + Ny * V]
            F[idx] += complex(-fz * 4 * np.pi * u.imag, fz * 4 * np.pi * u.real)
            C[idx] += 1.0


def test_scatter_and_normalise_equal_literal_restatement(orc):
    V, dx = 16, 0.4e-6
    df = 1.0 / (V * dx)
    # two angles whose DC samples (ii = jj = 0) both land on voxel (0, 0, 0): tilts of one f_z step or less
    angles = [(0.0, 0.0), (0.0, 0.012)]
    rng = np.random.default_rng(3)
    spec = np.zeros((2, V, V), complex)
    for a in range(2):
        for (y, x) in ((0, 0), (1, 0), (0, 3), (V - 2, 5), (7, V - 1)):
            spec[a, y, x] = complex(*rng.standard_normal(2))
    F = np.zeros(V ** 3, complex)
    C = np.zeros(V ** 3)
    for a, (ax, ay) in enumerate(angles):
        _, fx0, fy0, fz0 = _geometry(ax, ay, V, dx)
        _fill_edward_sphere_literal(spec[a].ravel(), F, C, fx0, fy0, fz0, FM, df, V)
    assert C[0] == 2.0  # the twice-hit voxel
    # Initial_F_F2 (:253-276): F /= C where C != 0; F2 = F / dx; F = F2
    F2 = np.where(C != 0, F / np.where(C != 0, C, 1.0), F) / dx
    Fo, Ho = orc.ewald_scatter(spec, angles, V, LAM, NM, dx)
    assert np.array_equal(Ho.ravel(), C)
    assert np.abs(Fo.ravel() - F).max() <= 1e-12 * np.abs(F).max()
    Fn, Fm = orc.normalize(Fo, Ho, dx)
    assert np.abs(Fm.ravel() - F2).max() <= 1e-12 * np.abs(F2).max()
    assert np.array_equal(Fn, Fm)

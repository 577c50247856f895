"""Pins the oracle's reconstruction restatement (Rytov spectrum, Ewald mapping, normalisation,
positivity loop) with properties that a wrong sign, index or dropped term would break:
- Rytov spectrum = numpy fft2 of the fftshifted transposed field;
- every mapped sample lies on the Ewald sphere of its angle and its voxel is the nearest along z;
- scatter and gather are adjoint (complex inner product, FP64);
- est_n o modify is the identity on feasible n; the projection is idempotent;
- a fully measured spectrum of a feasible object is a fixed point of the positivity loop."""
import numpy as np
import pytest

LAM, NM = 532e-9, 1.337


def _angles(A, na=1.0):
    s = na / NM
    az = 2 * np.pi * np.arange(A) / A
    return np.stack([np.arcsin(s * np.sin(az)), np.arcsin(s * np.cos(az))], axis=1)


def test_rytov_spectrum_matches_numpy(orc):
    rng = np.random.default_rng(0)
    ph = rng.standard_normal((2, 32, 32))
    la = 0.1 * rng.standard_normal((2, 32, 32))
    got = orc.rytov_spectrum(ph, la)
    for a in range(2):
        ref = np.fft.fft2(np.fft.fftshift((la[a] + 1j * ph[a]).T))
        np.testing.assert_allclose(got[a], ref, atol=1e-10)


@pytest.mark.parametrize("V", [32, 64])
def test_cap_geometry(orc, V):
    dx = 0.1e-6 * 1024 / 256 * (256 / V)
    ang = _angles(6)
    idx = orc.ewald_index(ang, V, LAM, NM, dx)
    fm0 = NM / LAM
    df = 1.0 / (V * dx)
    f = np.fft.fftfreq(V, 1.0 / V)  # signed frequency of array index (numpy order == mod layout)
    for a, (ax, ay) in enumerate(ang):
        fx0, fy0 = fm0 * np.sin(ay), fm0 * np.sin(ax)
        fz0 = np.sqrt(fm0 ** 2 - fx0 ** 2 - fy0 ** 2)
        FX = f[None, :] * df + fx0   # index [Ny][Nx]
        FY = f[:, None] * df + fy0
        s = fm0 ** 2 - FX ** 2 - FY ** 2
        on = s >= 0
        assert np.array_equal(idx[a] >= 0, on)
        fz = np.sqrt(np.where(on, s, 0))
        nz = idx[a].astype(np.float64)
        kz = np.where(nz >= V // 2, nz - V, nz)  # signed voxel index
        # the voxel centre kz*df is the nearest grid point to fz - fz0 (mod V)
        d = (fz - fz0) / df - kz
        d = d - V * np.round(d / V)
        assert np.all(np.abs(d[on]) <= 0.5 + 1e-9)


def test_scatter_gather_adjoint(orc):
    V, A = 32, 5
    dx = 0.4e-6 * 8
    rng = np.random.default_rng(1)
    u = rng.standard_normal((A, V, V)) + 1j * rng.standard_normal((A, V, V))
    Fr = rng.standard_normal((V, V, V)) + 1j * rng.standard_normal((V, V, V))
    Su, _ = orc.ewald_scatter(u, _angles(A), V, LAM, NM, dx)
    GF = orc.ewald_gather(Fr, _angles(A), LAM, NM, dx)
    lhs = np.vdot(Su, Fr)
    rhs = np.vdot(u, GF)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_est_n_modify_identity_and_idempotence(orc):
    rng = np.random.default_rng(2)
    n = (NM + rng.uniform(0, 0.1, 1000)) + 1j * rng.uniform(0, 0.01, 1000)
    f = orc.modify_f(n, LAM, NM)
    # the literal sqrt((|z| - c)/2) form cancels when Im n is small: ~1e-11 in FP64
    np.testing.assert_allclose(orc.est_n(f, LAM, NM), n, rtol=0, atol=1e-10)
    g = rng.standard_normal(1000) * 1e13 + 1j * rng.standard_normal(1000) * 1e12
    p1 = orc.modify_f(orc.est_n(g, LAM, NM), LAM, NM)
    p2 = orc.modify_f(orc.est_n(p1, LAM, NM), LAM, NM)
    np.testing.assert_allclose(p2, p1, rtol=0, atol=1e-9 * np.abs(p1).max())


def test_positivity_fixed_point(orc):
    V = 16
    rng = np.random.default_rng(3)
    n = (NM + rng.uniform(0, 0.05, (V, V, V))) + 1j * rng.uniform(0, 0.005, (V, V, V))
    f = orc.modify_f(n, LAM, NM)
    F = np.fft.fftn(f)
    out = orc.positivity(f, F, np.ones((V, V, V)), 3, LAM, NM, flags=2)
    np.testing.assert_allclose(out, n, atol=1e-12)
    shifted = orc.positivity(f, F, np.ones((V, V, V)), 1, LAM, NM, flags=0)
    np.testing.assert_allclose(shifted, np.fft.fftshift(n), atol=1e-12)


def _fit_case(orc, V=16, A=5, seed=4):
    rng = np.random.default_rng(seed)
    dx = 0.4e-6 * 256 / V
    ang = _angles(A)
    spec = rng.standard_normal((A, V, V)) + 1j * rng.standard_normal((A, V, V))
    n = (NM + rng.uniform(0, 0.03, (V, V, V))) + 1j * rng.uniform(0, 0.003, (V, V, V))
    f0 = orc.modify_f(n, LAM, NM)
    return dx, ang, spec, f0


def test_fit_step_one_is_the_positivity_loop(orc):
    """F - 1 * G^T(G F - m) / hits replaces every measured voxel by the hit-averaged measurement, i.e.
    the data-consistency step of the positivity loop with F_meas = normalize(scatter(spec))."""
    V = 16
    dx, ang, spec, f0 = _fit_case(orc, V)
    F, hits = orc.ewald_scatter(spec, ang, V, LAM, NM, dx)
    _, Fm = orc.normalize(F, hits, dx)
    for iters in (1, 3):
        fit = orc.fit_pgd(f0, spec, ang, LAM, NM, dx, iters, 1.0)
        pos = orc.positivity(f0, Fm, hits, iters, LAM, NM)
        np.testing.assert_allclose(fit, pos, rtol=0, atol=1e-10)  # summation order differs


def test_fit_step_zero_keeps_the_feasible_start(orc):
    """With step 0 the k-space update vanishes: each iteration is FFT, inverse FFT and the projection,
    which leaves a feasible n unchanged."""
    V = 16
    dx, ang, spec, f0 = _fit_case(orc, V)
    n0 = orc.est_n(f0, LAM, NM)
    out = orc.fit_pgd(f0, spec, ang, LAM, NM, dx, 3, 0.0, flags=2)
    np.testing.assert_allclose(out, n0, rtol=0, atol=1e-12)


def test_fit_half_step_moves_halfway(orc):
    """One iteration with step 1/2: every measured voxel becomes (F0 + Fm) / 2, with F0 the spectrum of
    the projected start, so the fit equals one positivity iteration whose measured spectrum is that
    average."""
    V = 16
    dx, ang, spec, f0 = _fit_case(orc, V, seed=5)
    F, hits = orc.ewald_scatter(spec, ang, V, LAM, NM, dx)
    _, Fm = orc.normalize(F, hits, dx)
    F0 = np.fft.fftn(orc.modify_f(orc.est_n(f0, LAM, NM), LAM, NM))
    half = np.where(hits > 0, 0.5 * (F0 + Fm), F0)
    fit = orc.fit_pgd(f0, spec, ang, LAM, NM, dx, 1, 0.5)
    pos = orc.positivity(f0, half, hits, 1, LAM, NM)
    np.testing.assert_allclose(fit, pos, rtol=0, atol=1e-10)


def _windows_median(m):
    """Median of every interior 3x3 window by numpy (odd count: the exact middle element)."""
    H, W = m.shape
    out = m.copy()
    st = np.stack([m[1 + dy:H - 1 + dy, 1 + dx:W - 1 + dx] for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
    out[1:-1, 1:-1] = np.median(st, axis=0)
    return out


def test_median_matches_numpy_windows(orc):
    rng = np.random.default_rng(6)
    maps = rng.standard_normal((3, 17, 23)).astype(np.float32)
    maps[1, 5:9, 5:9] = 0.25  # ties
    got = orc.median3x3(maps)
    for b in range(3):
        np.testing.assert_array_equal(got[b], _windows_median(maps[b]))


def test_median_nan_windows_follow_selection_order(orc):
    """The reference's selection sort never moves a NaN it meets at the current slot and never picks a
    NaN as a minimum, so the fifth pick depends on where the NaN sits in the column-major gather."""
    w = np.arange(9, dtype=np.float32).reshape(3, 3).T.copy()  # gather order k = 3x + y holds value k
    a = w.copy()
    a[0, 0] = np.nan  # k = 0: slot 0 keeps the NaN, slots 1..4 take 1..4 -> 4
    assert orc.median3x3(a)[1, 1] == 4.0
    b = w.copy()
    b[1, 1] = np.nan  # k = 4 (centre): slots 0..3 take 0..3, slot 4 keeps the NaN -> NaN
    assert np.isnan(orc.median3x3(b)[1, 1])
    c = w.copy()
    c[2, 2] = np.nan  # k = 8: never reached by a minimum -> the ordinary median of 0..7, i.e. 4
    assert orc.median3x3(c)[1, 1] == 4.0


def test_check_maps_branches(orc):
    ph, am = orc.check_maps(np.array([np.nan, np.inf, -np.inf, 1.5, -2.0]),
                            np.array([np.nan, np.inf, 0.0, -3.0, 2.0]))
    np.testing.assert_array_equal(ph, np.array([0, 0, 0, 1.5, -2.0], np.float32))
    np.testing.assert_array_equal(am, np.array([1, 1, 0.01, 3.0, 2.0], np.float32))


def test_recon_from_maps_amplitude_and_log_agree(orc):
    """Median commutes with the logarithm: A maps with the amplitude flag give the RI of ln A maps;
    without the median the maps reconstruct exactly as the cell path's tail."""
    V, A = 16, 4
    rng = np.random.default_rng(7)
    ph = (0.3 * rng.standard_normal((A, V, V))).astype(np.float32)
    la = (0.05 * rng.standard_normal((A, V, V))).astype(np.float32)
    ang = _angles(A)
    dx = 0.4e-6 * 256 / V
    r_log = orc.recon_from_maps(ph, la, ang, LAM, NM, dx, 2, flags=4)
    r_amp = orc.recon_from_maps(ph, np.exp(la.astype(np.float64)).astype(np.float32), ang, LAM, NM, dx, 2, flags=12)
    np.testing.assert_allclose(r_amp, r_log, rtol=0, atol=1e-6)
    r_med = orc.recon_from_maps(orc.median3x3(ph), orc.median3x3(la), ang, LAM, NM, dx, 2, flags=0)
    np.testing.assert_allclose(r_med, r_log, rtol=0, atol=1e-12)


@pytest.mark.parametrize("step", [1.0, 0.5])
def test_fit_objective_is_the_gathered_residual(orc, step):
    """J before the first step equals 1/2 sum over samples of |F0[v] - m|^2 computed here from numpy's
    3-D FFT of the projected start, the cap voxel of every sample and m = 4 pi i f_z spec / dx
    (Recon_kernel.cu:224-225, 269-270); on consistent data (spectra gathered from a feasible object)
    J then falls at every step."""
    V = 16
    dx, ang, _, f0 = _fit_case(orc, V, seed=11)
    rng = np.random.default_rng(12)
    n_true = (NM + rng.uniform(0, 0.02, (V, V, V))) + 1j * rng.uniform(0, 0.002, (V, V, V))
    F_true = np.fft.fftn(orc.modify_f(n_true, LAM, NM))
    G = orc.ewald_gather(F_true, ang, LAM, NM, dx)  # conj(4 pi i f_z) F[v]
    fm0, df = NM / LAM, 1.0 / (V * dx)
    k = np.fft.fftfreq(V, 1.0 / V)
    II, JJ = np.meshgrid(k, k, indexing="xy")  # spectrum [y][x]: x = ii, y = jj
    idx = orc.ewald_index(ang, V, LAM, NM, dx)
    spec = np.zeros_like(G)
    w = np.zeros(G.shape)
    for a, (ax, ay) in enumerate(ang):
        fx = II * df + fm0 * np.sin(ay)
        fy = JJ * df + fm0 * np.sin(ax)
        fz = np.sqrt(np.maximum(fm0 ** 2 - fx ** 2 - fy ** 2, 0.0))
        w[a] = 4 * np.pi * fz
        on = idx[a] >= 0
        spec[a][on] = G[a][on] * dx / (w[a][on] ** 2)  # so that 4 pi i f_z spec / dx = F_true[v]
    n, J = orc.fit_pgd(f0, spec, ang, LAM, NM, dx, 8, step, objective=True)
    F0 = np.fft.fftn(orc.modify_f(orc.est_n(f0, LAM, NM), LAM, NM))
    j0 = 0.0
    for a in range(len(ang)):
        on = idx[a] >= 0
        yy, xx = np.nonzero(on)
        m = 1j * w[a][on] * spec[a][on] / dx
        j0 += 0.5 * np.sum(np.abs(F0[idx[a][on], yy, xx] - m) ** 2)
    assert abs(J[0] - j0) <= 1e-10 * j0
    assert np.all(np.diff(J) < 0), J

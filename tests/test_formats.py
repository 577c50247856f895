"""The reference's on-disk formats (host only, CPU): .phimap byte layout as outputImg writes it
(HilbertTransform.cu:2437-2457), round trips, the reload path's status filter, and .rimap parts."""
import struct

import numpy as np
import pytest

import ctdm_amd


def test_phimap_byte_layout(tmp_path):
    p = tmp_path / "buffer001.phimap"
    ph = np.arange(6, dtype=np.float32).reshape(2, 3)
    am = (10 + np.arange(6, dtype=np.float32)).reshape(2, 3)
    ctdm_amd.phimap_write(p, ph, am, ang_x=0.25, ang_y=-0.5, status=True)
    raw = p.read_bytes()
    assert len(raw) == 1 + 8 + 8 + 4 + 4 + 6 * 8
    st, ax, ay, sx, sy = struct.unpack("<?ddii", raw[:25])
    assert (st, ax, ay, sx, sy) == (True, 0.25, -0.5, 3, 2)
    pairs = np.frombuffer(raw[25:], "<f4").reshape(6, 2)
    assert np.array_equal(pairs[:, 0], ph.ravel()) and np.array_equal(pairs[:, 1], am.ravel())


def test_phimap_round_trip_and_reload_filter(tmp_path):
    rng = np.random.default_rng(1)
    paths = []
    for i in range(4):
        p = tmp_path / f"buffer{i + 1:03d}.phimap"
        ctdm_amd.phimap_write(p, rng.standard_normal((16, 16)), rng.uniform(0.5, 1.5, (16, 16)), ang_x=0.1 * i,
                              ang_y=-0.1 * i, status=(i != 2))
        paths.append(p)
    d = ctdm_amd.phimap_read(paths[1])
    assert d["status"] and d["ang_x"] == 0.1 and d["ang_y"] == -0.1 and d["phase"].shape == (16, 16)
    ph, am, ang = ctdm_amd.load_phimap_stack(paths)
    assert ph.shape == (3, 16, 16) and am.shape == (3, 16, 16)
    assert np.allclose(ang[:, 0], [0.0, 0.1, 0.3])


@pytest.mark.parametrize("part", [0, 1, 2])
def test_rimap_parts(tmp_path, part):
    rng = np.random.default_rng(2)
    v = (rng.standard_normal((8, 8, 8)) + 1j * rng.standard_normal((8, 8, 8))).astype(np.complex64)
    p = tmp_path / "3D-RI.rimap"
    ctdm_amd.rimap_write(p, v, part)
    assert p.stat().st_size == v.size * (8 if part == 0 else 4)
    back = ctdm_amd.rimap_read(p, v.shape, part)
    if part == 0:
        assert np.array_equal(back, v)
    elif part == 1:
        assert np.array_equal(back.real, v.real) and not back.imag.any()
    else:
        assert np.array_equal(back.imag, v.imag) and not back.real.any()


def test_phimap_errors(tmp_path):
    with pytest.raises(ctdm_amd.CtdmError):
        ctdm_amd.phimap_read(tmp_path / "missing.phimap")

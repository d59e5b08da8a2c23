"""NPZ output (piv-io/src/write.rs:113-165): the per-pair archive's keys, dtypes, shapes and the
derived arrays (u = dx * f32(1/dt), physical velocity = u * f32(m_per_px), grid centres plus
the ROI offset), read back with numpy.  Host-only: no GPU needed."""
import zipfile

import numpy as np
import pytest


@pytest.fixture(scope="module")
def piv():
    import piv_amd

    return piv_amd


def _field(piv, h, w, n, seed):
    g = piv.Grid.new(h, w, n, 0.5)
    rng = np.random.default_rng(seed)
    dx = rng.standard_normal((g.ny, g.nx)).astype(np.float32)
    dy = rng.standard_normal((g.ny, g.nx)).astype(np.float32)
    dx[0, 0] = np.nan
    ppr = (1 + 5 * rng.random((g.ny, g.nx))).astype(np.float32)
    fl = rng.choice(np.array([0, 1, 12, 2], np.uint8), size=(g.ny, g.nx))
    return piv.VectorField(g, dx, dy, ppr, fl)


def test_pair_archive_matches_the_reference_layout(piv, tmp_path):
    f = _field(piv, 200, 320, 32, 1)
    dt, m = 1.0 / 240.0, 2.5e-4
    path = piv.write_pair_npz(tmp_path, 12, 14, 0.0541666, dt, f, m_per_px=m, roi=(40, 8))
    assert path.endswith("pair_000012_000014.npz")
    with zipfile.ZipFile(path) as z:
        assert all(i.compress_type == zipfile.ZIP_DEFLATED for i in z.infolist())
        assert [i.filename for i in z.infolist()] == [
            k + ".npy" for k in ("x", "y", "dx", "dy", "u", "v", "u_mps", "v_mps", "peak_ratio",
                                 "flags", "frames", "t")]
    d = np.load(path)
    g = f.grid
    xs = (np.array([g.x0 + i * g.step for i in range(g.nx)], np.float32)
          + np.float32((g.window - 1.0) / 2.0)) + np.float32(40)
    assert d["x"].dtype == np.float32 and np.array_equal(d["x"], xs)
    assert d["y"].shape == (g.ny,) and d["y"][0] == np.float32(g.y0 + (g.window - 1) / 2.0) + 8
    for k, ref in (("dx", f.dx), ("dy", f.dy), ("peak_ratio", f.peak_ratio)):
        assert d[k].dtype == np.float32 and d[k].shape == (g.ny, g.nx)
        assert np.array_equal(d[k], ref, equal_nan=True)
    inv = np.float32(1.0 / dt)
    assert np.array_equal(d["u"], f.dx * inv, equal_nan=True)
    assert np.array_equal(d["v"], f.dy * inv, equal_nan=True)
    assert np.array_equal(d["u_mps"], (f.dx * inv) * np.float32(m), equal_nan=True)
    assert d["flags"].dtype == np.uint8 and np.array_equal(d["flags"], f.flags)
    assert d["frames"].dtype == np.uint64 and list(d["frames"]) == [12, 14]
    assert d["t"].dtype == np.float64 and d["t"].shape == () and float(d["t"]) == 0.0541666


def test_no_scale_no_physical_velocity(piv, tmp_path):
    f = _field(piv, 64, 64, 16, 2)
    path = piv.write_pair_npz(tmp_path, 0, 1, 0.5, 1.0, f)
    d = np.load(path)
    assert "u_mps" not in d.files and "v_mps" not in d.files
    assert np.array_equal(d["u"], f.dx, equal_nan=True)
    assert np.array_equal(d["x"], np.array(f.grid.xs(), np.float32))


def test_bad_arguments_are_errors(piv, tmp_path):
    f = _field(piv, 64, 64, 16, 3)
    with pytest.raises(piv.InvalidArgument):
        piv.write_pair_npz(tmp_path, 0, 1, 0.5, 0.0, f)
    with pytest.raises(piv.InvalidArgument):
        piv.write_pair_npz(tmp_path / "missing" / "dir", 0, 1, 0.5, 1.0, f)


def test_stats_archive_layout(piv, tmp_path):
    """piv-io/src/flow.rs:133-156: t (f64) first, then the five f32 columns in that order."""
    stats = [piv.FlowStats(0.5 * i, 1.0 + i, 2.0, 3.0e6 * i, 0.9, 0.1, 100, 90) for i in range(4)]
    stats[2].vorticity_mean = float("nan")
    t = np.array([0.1, 0.2, 0.3, 0.4])
    path = piv.write_stats_npz(tmp_path, t, stats)
    import zipfile
    with zipfile.ZipFile(path) as z:
        assert [i.filename for i in z.infolist()] == [
            k + ".npy" for k in ("t", "vorticity_mean", "vorticity_abs_mean", "ke_mean", "ke_sum",
                                 "coverage")]
    d = np.load(path)
    assert d["t"].dtype == np.float64 and np.array_equal(d["t"], t)
    for k in ("vorticity_mean", "vorticity_abs_mean", "ke_mean", "ke_sum", "coverage"):
        assert d[k].dtype == np.float32 and d[k].shape == (4,)
        assert np.array_equal(d[k], np.array([getattr(s, k) for s in stats], np.float32),
                              equal_nan=True)

"""Post-analysis stages on the GPU against the oracle (oracle/post_oracle.c):
- per-pair flow statistics (piv_flow_stats_device; flowstats.rs:31-137, flow.rs:13-34) on
  fields from the analyzer and on constructed fields with NaNs, replaced and rejected vectors,
  empty and nearly empty grids.  Counts and the noise estimate (an exact selection) are
  bit-exact; the means come from f64 sums whose association differs from the reference's
  serial loop, so they are compared at 1e-6 relative (a few f32 ulps);
- the time-series filters and metrics (smooth.rs:44-310): bit-exact (same f32 operation order,
  the same libm expf weights, total_cmp medians, hypot as libm computes it)."""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def piv():
    import piv_amd

    if piv_amd.device_count() < 1:
        pytest.skip("no GPU")
    return piv_amd


def same(a, b):
    """bitwise equality of float arrays (NaN payloads aside)"""
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return a.shape == b.shape and bool(np.all((a.view(np.uint32) == b.view(np.uint32)) |
                                              (np.isnan(a) & np.isnan(b))))


def close_stats(got, want, rel=1e-6):
    for k in ("vorticity_mean", "vorticity_abs_mean", "ke_mean", "ke_sum", "coverage"):
        g, w = getattr(got, k), want[k]
        if math.isnan(w):
            assert math.isnan(g), k
        else:
            assert g == pytest.approx(w, rel=rel, abs=1e-30), k
    assert got.n_usable == want["n_usable"] and got.n_vorticity == want["n_vorticity"]


def random_fields(rng, n, ny, nx, nan_frac=0.05, rej_frac=0.1):
    dx = (rng.standard_normal((n, ny, nx)) * 3).astype(np.float32)
    dy = (rng.standard_normal((n, ny, nx)) * 3).astype(np.float32)
    fl = np.zeros((n, ny, nx), np.uint8)
    r = rng.random((n, ny, nx))
    fl[r < rej_frac] = 4  # MEDIAN_OUTLIER
    fl[(r >= rej_frac) & (r < 2 * rej_frac)] = 4 | 8  # replaced: usable
    fl[(r >= 2 * rej_frac) & (r < 2.5 * rej_frac)] = 1  # NO_SIGNAL
    dx[rng.random((n, ny, nx)) < nan_frac] = np.nan
    dy[rng.random((n, ny, nx)) < nan_frac / 2] = np.nan
    return dx, dy, fl


def check_pairs(piv, an, g, dx, dy, fl, dt):
    D = piv.DeviceArray.from_numpy
    got = an.flow_stats_device(g, D(dx), D(dy), D(fl), dt)
    for p in range(dx.shape[0]):
        ok = O.flow_usable(dx[p], fl[p])
        close_stats(got[p], O.pair_stats(dx[p], dy[p], ok, g.step, dt))
        want_noise = O.neighbour_noise(dx[p], dy[p], ok)
        assert same(got[p].noise_px, want_noise), (p, got[p].noise_px, want_noise)
    return got


def test_flow_stats_constructed_fields(piv):
    an = piv.PivAnalyzer(piv.PivConfig(windows=[32]), piv.PreprocessConfig())
    rng = np.random.default_rng(11)
    g = piv.Grid.new(1024, 1024, 16, 0.5)  # 127 x 127
    dx, dy, fl = random_fields(rng, 5, g.ny, g.nx)
    # an all-rejected pair, and a pair with only a 3x3 patch usable (< 20 residuals)
    fl[3] = 4
    fl[4] = 4
    fl[4, 10:13, 10:13] = 0
    dx[4, 10:13, 10:13] = 1.0
    check_pairs(piv, an, g, dx, dy, fl, 1.0 / 240.0)


@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (5, 1), (2, 2), (31, 31)])
def test_flow_stats_small_grids(piv, shape):
    an = piv.PivAnalyzer(piv.PivConfig(windows=[32]), piv.PreprocessConfig())
    ny, nx = shape
    g = piv.Grid(32, 16, nx, ny, 0, 0)
    rng = np.random.default_rng(ny * 100 + nx)
    dx, dy, fl = random_fields(rng, 3, ny, nx)
    check_pairs(piv, an, g, dx, dy, fl, 0.01)


def test_flow_stats_rotation_closed_form(piv):
    """flowstats.rs:157-169 through the device: solid-body rotation gives 2 omega."""
    an = piv.PivAnalyzer(piv.PivConfig(windows=[32]), piv.PreprocessConfig())
    n, step = 21, 16
    g = piv.Grid(32, step, n, n, 0, 0)
    om, c = np.float32(3.0), np.float32((n - 1) * step / 2.0)
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float32) * np.float32(step)
    dx = (om * (yy - c))[None].astype(np.float32)
    dy = (-om * (xx - c))[None].astype(np.float32)
    fl = np.zeros((1, n, n), np.uint8)
    s = an.flow_stats_device(g, piv.DeviceArray.from_numpy(dx), piv.DeviceArray.from_numpy(dy),
                             piv.DeviceArray.from_numpy(fl), 1.0)[0]
    assert abs(s.vorticity_mean - 2 * om) <= 1e-4 and s.n_vorticity == n * n


def test_flow_stats_of_analysed_pairs(piv):
    """Fields the analyzer produced (device-resident) straight into the statistics."""
    import piv_amd.synthetic as S

    h = w = 256
    cfg = piv.PivConfig(windows=[64, 32], final_iterations=1, mask=piv.MaskConfig(0.0, 0.0))
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig())
    frames = [S.particle_frames(h, w, 40 + i, density=0.03, diameter=2.5, amp=(150, 255),
                                noise=3.0, field=S.lamb_oseen_shear(h, w, vmax=4.0, shear=0.01))
              for i in range(3)]
    g = an.output_grid(h, w)
    dxs, dys, fls = [], [], []
    for a, b in frames:
        r = an.analyze(an.prepare(a), an.prepare(b))
        dxs.append(r.field.dx)
        dys.append(r.field.dy)
        fls.append(r.field.flags)
    dx, dy, fl = (np.stack(x) for x in (dxs, dys, fls))
    got = check_pairs(piv, an, g, dx, dy, fl, 1.0 / 500.0)
    assert all(s.coverage > 0.5 and math.isfinite(s.noise_px) for s in got)


def test_flow_stats_errors(piv):
    an = piv.PivAnalyzer(piv.PivConfig(windows=[32]), piv.PreprocessConfig())
    g = piv.Grid(32, 16, 4, 4, 0, 0)
    z = piv.DeviceArray((1, 4, 4), np.float32)
    f = piv.DeviceArray((1, 4, 4), np.uint8)
    with pytest.raises(piv.PivError):
        an.flow_stats_device(g, z, z, f, 0.0)
    with pytest.raises(piv.ShapeMismatch):
        an.flow_stats_device(piv.Grid(32, 16, 5, 4, 0, 0), z, z, f, 1.0)


# ---- smoothing ---------------------------------------------------------------------------
def series(rng, nt, ny, nx, nan_frac=0.03):
    t, y, x = np.meshgrid(np.arange(nt), np.arange(ny), np.arange(nx), indexing="ij")
    du = (2.0 + 0.01 * t + 0.05 * y + 0.4 * rng.standard_normal((nt, ny, nx))).astype(np.float32)
    dv = (-1.0 + 0.02 * x + 0.4 * rng.standard_normal((nt, ny, nx))).astype(np.float32)
    du[rng.random(du.shape) < nan_frac] = np.nan
    ok = rng.random(du.shape) > 0.15
    gate = rng.random(du.shape) > 0.05
    return du, dv, ok, gate


@pytest.mark.parametrize("kernel,window,min_support,leave_out",
                         [("mean", 9, 1, ()), ("median", 7, 1, ()), ("median", 5, 3, (0,)),
                          ("mean", 11, 4, (0, 5, -5)), ("median", 3, 1, (0, 1, -1)),
                          ("median", 63, 10, ()), ("mean", 1, 1, ())])
def test_temporal_matches_oracle(piv, kernel, window, min_support, leave_out):
    from piv_amd import smooth

    rng = np.random.default_rng(window * 13 + min_support)
    du, dv, ok, gate = series(rng, 80, 12, 16)
    s = smooth.temporal(du, dv, ok, gate, kernel, window, min_support, leave_out)
    ox, oy, sup = O.temporal(du, dv, ok, gate, kernel, window, min_support, list(leave_out))
    assert same(s.dx, ox) and same(s.dy, oy) and np.array_equal(s.support, sup)


def test_temporal_video_sized_series_on_device(piv):
    """The batched-video grid (48 -> 24 px on 1280x800: 65 x 105 vectors) over 400 pairs, with
    device-resident inputs and outputs."""
    from piv_amd import smooth

    rng = np.random.default_rng(5)
    du, dv, ok, gate = series(rng, 400, 65, 105)
    dev = [piv.DeviceArray.from_numpy(a) for a in (du, dv, ok.astype(np.uint8),
                                                   gate.astype(np.uint8))]
    s = smooth.temporal(*dev, "mean", 9, 2)
    ox, oy, sup = O.temporal(du, dv, ok, gate, "mean", 9, 2)
    assert same(s.dx.numpy(), ox) and same(s.dy.numpy(), oy)
    assert np.array_equal(s.support.numpy(), sup)


@pytest.mark.parametrize("sigma,min_weight,leave_out", [(1.0, 0.25, None), (1.5, 0.5, 1),
                                                        (0.7, 0.0, None), (3.0, 0.3, 2),
                                                        (0.5, 0.0, 1), (11.9, 0.2, 3),
                                                        (12.6, 0.2, None)])
def test_spatial_matches_oracle(piv, sigma, min_weight, leave_out):
    """radius 2 .. 30 take the LDS-tiled kernel, radius 32 the untiled one"""
    from piv_amd import smooth

    rng = np.random.default_rng(int(sigma * 100))
    du, dv, _, gate = series(rng, 6, 23, 31)
    s = smooth.spatial(du, dv, gate, sigma, min_weight, leave_out)
    ox, oy, sup = O.spatial(du, dv, gate, sigma, min_weight, leave_out)
    assert same(s.dx, ox) and same(s.dy, oy) and np.array_equal(s.support, sup)


def test_metrics_match_oracle(piv):
    from piv_amd import smooth

    rng = np.random.default_rng(21)
    du, dv, ok, gate = series(rng, 30, 17, 19)
    pu, pv, _, _ = series(rng, 30, 17, 19)
    e = smooth.cross_validation_errors(du, dv, ok, pu, pv)
    want = O.cross_validation_errors(du, dv, ok, pu, pv)
    assert same(e, want)
    keep = rng.random(want.shape) > 0.3
    assert smooth.median_error(want) == O.median_error(want)
    assert smooth.median_error(want, keep) == O.median_error(want, keep)
    assert same(smooth.pair_to_pair_change(du, dv), O.pair_to_pair_change(du, dv))
    assert same(smooth.spatial_roughness(du, dv), O.spatial_roughness(du, dv))
    assert same(smooth.mask(du, ok), O.mask(du, ok))
    # even and odd counts, all-NaN, empty
    for m in (1, 2, 3, 10, 11):
        x = rng.standard_normal(m).astype(np.float32)
        assert smooth.median_error(x) == O.median_error(x)
    nan = np.full(5, np.nan, np.float32)
    r = smooth.median_error(nan)
    assert math.isnan(r[0]) and r[1] == 0
    one = du[:1]
    assert math.isnan(smooth.pair_to_pair_change(one, dv[:1]))


def test_smoothing_errors(piv):
    from piv_amd import smooth

    z = np.zeros((3, 4, 4), np.float32)
    b = np.ones((3, 4, 4), bool)
    with pytest.raises(piv.PivError):
        smooth.temporal(z, z, b, b, "mean", 4, 1)   # even window (smooth.rs:54)
    with pytest.raises(piv.PivError):
        smooth.temporal(z, z, b, b, "mean", 65, 1)  # beyond the kernel's window limit
    with pytest.raises(piv.PivError):
        smooth.spatial(z, z, b, 0.0, 0.1)
    with pytest.raises(piv.PivError):
        smooth.spatial(z, z, b, 20.0, 0.1)          # radius 50 > 32

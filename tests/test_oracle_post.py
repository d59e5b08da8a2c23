"""Pins the oracle's post-analysis stages (oracle/post_oracle.c) on the CPU:
- flow statistics against the reference's own unit tests (piv-core/src/flowstats.rs:147-265,
  piv-io/src/flow.rs:207-225): closed forms for rotation, translation, shear, opposite
  vortices, holes and edges, empty fields, and the noise estimate's recovery of sigma;
- the time-series filters against the reference's unit tests (piv-core/src/smooth.rs:334-519)
  and against brute-force numpy restatements on small random series (mean, median, Gaussian
  normalized convolution, the metrics' medians).
No GPU needed."""
import math

import numpy as np
import pytest

from oracle import oracle as O

N, STEP = 21, 16.0


def field(f):
    yy, xx = np.mgrid[0:N, 0:N].astype(np.float32) * np.float32(STEP)
    u, v = f(xx, yy)
    u = np.broadcast_to(np.asarray(u, np.float32), (N, N)).copy()
    v = np.broadcast_to(np.asarray(v, np.float32), (N, N)).copy()
    return u, v, np.ones((N, N), bool)


def close(a, b, tol):
    return abs(a - b) <= tol


# ---- flowstats.rs tests ----------------------------------------------------------------
def test_solid_body_rotation_counter_clockwise_on_screen_is_positive():
    om = np.float32(3.0)
    c = np.float32((N - 1) * STEP / 2.0)
    u, v, ok = field(lambda x, y: (om * (y - c), -om * (x - c)))
    w = O.vorticity(u, v, ok, STEP)
    assert np.all(np.abs(w - 2 * om) <= 1e-4)
    s = O.flow_stats(u, v, ok, STEP)
    assert close(s["vorticity_mean"], 2 * om, 1e-4) and close(s["vorticity_abs_mean"], 2 * om, 1e-4)
    assert s["n_vorticity"] == N * N


def test_uniform_translation():
    u, v, ok = field(lambda x, y: (3.0, -4.0))
    s = O.flow_stats(u, v, ok, STEP)
    assert s["vorticity_mean"] == 0.0 and s["vorticity_abs_mean"] == 0.0
    assert close(s["ke_mean"], 12.5, 1e-5)
    assert close(s["ke_sum"], 12.5 * N * N * STEP * STEP, 1.0)
    assert s["coverage"] == 1.0


def test_shear():
    g = np.float32(0.5)
    u, v, ok = field(lambda x, y: (g * y, 0.0))
    assert close(O.flow_stats(u, v, ok, STEP)["vorticity_mean"], g, 1e-5)


def test_opposite_vortices_cancel_in_the_signed_mean():
    L = np.float32(N * STEP)
    u, v, ok = field(lambda x, y: (0.0, np.sin(np.float32(2 * np.pi) * x / L) * 10.0))
    s = O.flow_stats(u, v, ok, STEP)
    assert abs(s["vorticity_mean"]) < 0.1 * s["vorticity_abs_mean"]
    assert close(s["vorticity_abs_mean"], 40.0 / L, 0.1 * 40.0 / L)


def test_holes_and_edges():
    u, v, ok = field(lambda x, y: (y * np.float32(0.1), x * np.float32(0.2)))
    ok[5, 5] = False
    w = O.vorticity(u, v, ok, STEP)
    assert np.isnan(w[5, 5])
    for p in [(4, 5), (6, 5), (5, 4), (5, 6), (0, 0), (N - 1, N - 1)]:
        assert close(w[p], -0.1, 1e-5), p
    lonely = np.zeros((N, N), bool)
    lonely[3, 3] = True
    s = O.flow_stats(u, v, lonely, STEP)
    assert s["n_vorticity"] == 0 and math.isnan(s["vorticity_mean"]) and s["n_usable"] == 1
    un = u.copy()
    un[0, 0] = np.nan
    assert O.flow_stats(un, v, ok, STEP)["n_usable"] == N * N - 2


def test_empty_field():
    u, v, _ = field(lambda x, y: (1.0, 1.0))
    s = O.flow_stats(u, v, np.zeros((N, N), bool), STEP)
    assert math.isnan(s["ke_mean"]) and math.isnan(s["vorticity_mean"])
    assert (s["ke_sum"], s["coverage"], s["n_usable"]) == (0.0, 0.0, 0)


def test_neighbour_noise_recovers_sigma():
    rng = np.random.default_rng(12345)
    n = 80
    u = (2.0 * rng.standard_normal((n, n))).astype(np.float32)
    v = (2.0 * rng.standard_normal((n, n))).astype(np.float32)
    s = O.neighbour_noise(u, v, np.ones((n, n), bool))
    assert abs(s - 2.0) < 0.15, s


def test_neighbour_noise_brute_force_and_edge_cases():
    """The restatement against an independent numpy reading of flowstats.rs:106-137: the
    neighbour median is element len/2 of the sorted neighbours (no averaging), at least 5
    neighbours, at least 20 residuals, MAD = element len/2."""
    rng = np.random.default_rng(3)
    ny, nx = 9, 11
    u = rng.standard_normal((ny, nx)).astype(np.float32)
    v = rng.standard_normal((ny, nx)).astype(np.float32)
    ok = rng.random((ny, nx)) > 0.2
    u[2, 3] = np.nan
    resid = []
    for a in (u, v):
        for iy in range(ny):
            for ix in range(nx):
                if not ok[iy, ix] or not np.isfinite(a[iy, ix]):
                    continue
                nb = [a[jy, jx] for jy in range(max(iy - 1, 0), min(iy + 2, ny))
                      for jx in range(max(ix - 1, 0), min(ix + 2, nx))
                      if (jy, jx) != (iy, ix) and ok[jy, jx] and np.isfinite(a[jy, jx])]
                if len(nb) < 5:
                    continue
                med = np.sort(np.array(nb, np.float32))[len(nb) // 2]
                resid.append(np.float32(abs(a[iy, ix] - med)))
    mad = np.sort(np.array(resid, np.float32))[len(resid) // 2]
    want = np.float32(1.4826) * mad / np.sqrt(np.float32(1.0) + np.float32(np.pi) / np.float32(16.0))
    assert O.neighbour_noise(u, v, ok) == pytest.approx(float(want), rel=1e-6)
    # fewer than 20 residuals: NaN
    few = np.zeros((ny, nx), bool)
    few[:3, :3] = True
    assert math.isnan(O.neighbour_noise(u, v, few))


def test_usable_includes_replaced():
    dx = np.array([[1.0, 1.0, np.nan, 1.0]], np.float32)
    fl = np.array([[0, O.MEDIAN_OUTLIER | O.REPLACED, 0, O.LOW_PPR]], np.uint8)
    assert O.flow_usable(dx, fl).tolist() == [[True, True, False, False]]


def test_pair_stats_scales_by_f32_inverse_dt():
    rng = np.random.default_rng(5)
    dx = rng.standard_normal((12, 15)).astype(np.float32)
    dy = rng.standard_normal((12, 15)).astype(np.float32)
    ok = rng.random((12, 15)) > 0.1
    dt = 1.0 / 240.0
    k = np.float32(1.0 / dt)
    assert O.pair_stats(dx, dy, ok, 8, dt) == O.flow_stats(dx * k, dy * k, ok, 8.0)


# ---- smooth.rs tests ---------------------------------------------------------------------
def series(nt, noise, seed=1):
    ny, nx = 12, 16
    t, y, x = np.meshgrid(np.arange(nt), np.arange(ny), np.arange(nx), indexing="ij")
    tu = (2.0 + 0.01 * t + 0.05 * y).astype(np.float32)
    tv = (-1.0 + 0.02 * x - 0.005 * t).astype(np.float32)
    rng = np.random.default_rng(seed)
    du = (tu + noise * rng.standard_normal(tu.shape)).astype(np.float32)
    dv = (tv + noise * rng.standard_normal(tv.shape)).astype(np.float32)
    return du, dv, tu, tv, np.ones(tu.shape, bool)


def rms(a, b):
    m = np.isfinite(a)
    return float(np.sqrt(np.mean((a[m] - b[m]) ** 2)))


def test_temporal_mean_reduces_noise_by_sqrt_window():
    du, dv, tu, _, ok = series(60, 0.5)
    ox, _, sup = O.temporal(du, dv, ok, ok, "mean", 9, 1)
    assert rms(ox, tu) < rms(du, tu) / 2.5
    assert sup[30, 5, 5] == 9 and sup[0, 5, 5] == 5


def test_temporal_median_rejects_outliers_and_respects_usable():
    du, dv, tu, _, ok = series(30, 0.0)
    du[10, 3, 3] = 50.0
    du[11, 3, 3] = -50.0
    ok[12, 3, 3] = False
    du[12, 3, 3] = 1e6
    ox, _, sup = O.temporal(du, dv, ok, np.ones_like(ok), "median", 7, 1)
    assert abs(ox[11, 3, 3] - tu[11, 3, 3]) < 0.05
    assert abs(ox[12, 3, 3] - tu[12, 3, 3]) < 0.05
    assert sup[12, 3, 3] == 6


def test_gate_and_min_support_control_output():
    du, dv, _, _, ok = series(20, 0.1)
    gate = ok.copy()
    gate[5, 0, 0] = False
    ox, _, _ = O.temporal(du, dv, ok, gate, "mean", 5, 4)
    assert np.isnan(ox[5, 0, 0]) and np.isnan(ox[0, 1, 1]) and np.isfinite(ox[1, 1, 1])


def test_spatial_smooths_noise_preserves_linear_fields_and_skips_nan():
    du, dv, tu, tv, ok = series(3, 0.0)
    ox, oy, _ = O.spatial(du, dv, ok, 1.0, 0.25)
    assert abs(ox[1, 6, 8] - tu[1, 6, 8]) < 1e-4 and abs(oy[1, 6, 8] - tv[1, 6, 8]) < 1e-4
    du[1, 6, 9] = np.nan
    ox, _, sup = O.spatial(du, dv, ok, 1.0, 0.25)
    assert np.isfinite(ox[1, 6, 8]) and sup[1, 6, 8] == 48
    nu, nv, tu, _, ok = series(3, 0.5)
    ox, _, _ = O.spatial(nu, nv, ok, 1.5, 0.25)
    assert rms(ox, tu) < rms(nu, tu) / 2.0


def test_cross_validation_prefers_moderate_smoothing():
    du, dv, _, _, ok = series(80, 0.5)

    def cv(w):
        ox, oy, _ = O.temporal(du, dv, ok, ok, "mean", w, 1, [0])
        return O.median_error(O.cross_validation_errors(du, dv, ok, ox, oy))[0]

    assert cv(9) < cv(3)


def test_leave_out_skips_the_requested_samples():
    du, dv, _, _, ok = series(20, 0.0)
    du[10, 4, 4] = 100.0
    du[15, 4, 4] = 100.0
    ox, _, sup = O.temporal(du, dv, ok, ok, "mean", 11, 1, [0, 5, -5])
    assert ox[10, 4, 4] < 3.0 and sup[10, 4, 4] == 10
    ox, _, sup = O.temporal(du, dv, ok, ok, "mean", 3, 1, [0, 1, -1])
    assert sup[10, 4, 4] == 2 and np.isfinite(ox[10, 4, 4])
    du, dv, _, _, ok = series(1, 0.0)
    du[0, 6, 7] = 100.0
    ox, _, _ = O.spatial(du, dv, ok, 1.5, 0.0, 1)
    assert ox[0, 6, 8] < 3.0
    ox, _, _ = O.spatial(du, dv, ok, 0.5, 0.0, 1)
    assert np.all(np.isnan(ox))


def test_median_error_respects_mask():
    e = np.array([1.0, np.nan, 3.0, 10.0], np.float32)
    assert O.median_error(e) == (3.0, 3)
    assert O.median_error(e, np.array([1, 1, 1, 0], bool)) == (2.0, 2)


def test_roughness_metrics():
    du, dv, tu, tv, _ = series(10, 0.3)
    assert O.pair_to_pair_change(du, dv) > 0.3
    assert O.pair_to_pair_change(tu, tv) < 0.05
    assert O.spatial_roughness(du, dv) > 0.2
    assert O.spatial_roughness(tu, tv) < 1e-4


# ---- brute force (independent numpy readings of smooth.rs) ---------------------------------
def _offsets(window, leave_out):
    half = window // 2
    off = [] if 0 in leave_out else [0]
    for sign in (-1, 1):
        k, taken = 1, 0
        while taken < half:
            if sign * k not in leave_out:
                off.append(sign * k)
                taken += 1
            k += 1
    return off


@pytest.mark.parametrize("kernel,window,min_support,leave_out",
                         [("mean", 5, 1, ()), ("median", 5, 2, ()), ("median", 6 + 1, 3, (0, 2)),
                          ("mean", 9, 4, (0, -3, 3))])
def test_temporal_brute_force(kernel, window, min_support, leave_out):
    rng = np.random.default_rng(window * 7 + min_support)
    nt, ny, nx = 13, 4, 5
    dx = rng.standard_normal((nt, ny, nx)).astype(np.float32)
    dy = rng.standard_normal((nt, ny, nx)).astype(np.float32)
    ok = rng.random(dx.shape) > 0.25
    gate = rng.random(dx.shape) > 0.1
    ox, oy, sup = O.temporal(dx, dy, ok, gate, kernel, window, min_support, list(leave_out))
    off = _offsets(window, leave_out)
    for t in range(nt):
        for y in range(ny):
            for x in range(nx):
                if not gate[t, y, x]:
                    assert np.isnan(ox[t, y, x]) and sup[t, y, x] == 0
                    continue
                s = [t + o for o in off if 0 <= t + o < nt and ok[t + o, y, x]]
                assert sup[t, y, x] == len(s)
                if len(s) < max(min_support, 1):
                    assert np.isnan(ox[t, y, x])
                    continue
                bu = np.array([dx[q, y, x] for q in s], np.float32)
                bv = np.array([dy[q, y, x] for q in s], np.float32)
                if kernel == "mean":
                    acc_u, acc_v = np.float32(-0.0), np.float32(-0.0)
                    for a, b in zip(bu, bv):
                        acc_u = np.float32(acc_u + a)
                        acc_v = np.float32(acc_v + b)
                    want = (acc_u / np.float32(len(s)), acc_v / np.float32(len(s)))
                else:
                    want = (np.float32(np.median(bu)), np.float32(np.median(bv)))
                    # numpy's median of two middles is their mean, as median() in smooth.rs
                assert ox[t, y, x] == want[0] and oy[t, y, x] == want[1]


@pytest.mark.parametrize("sigma,min_weight,leave_out", [(1.0, 0.25, None), (1.5, 0.5, 1),
                                                        (0.7, 0.0, None)])
def test_spatial_brute_force(sigma, min_weight, leave_out):
    rng = np.random.default_rng(int(sigma * 10))
    nt, ny, nx = 2, 9, 11
    dx = rng.standard_normal((nt, ny, nx)).astype(np.float32)
    dy = rng.standard_normal((nt, ny, nx)).astype(np.float32)
    dx[rng.random(dx.shape) < 0.15] = np.nan
    gate = rng.random(dx.shape) > 0.1
    ox, oy, sup = O.spatial(dx, dy, gate, sigma, min_weight, leave_out)
    r = max(int(math.ceil(2.5 * sigma)), 1)
    ii, jj = np.meshgrid(np.arange(-r, r + 1), np.arange(-r, r + 1))
    ker = np.exp(-(ii * ii + jj * jj).astype(np.float64) / (2.0 * sigma * sigma))
    excl = np.zeros_like(ker, bool) if leave_out is None else \
        (np.maximum(np.abs(ii), np.abs(jj)) <= leave_out)
    full = ker[~excl].sum()
    for t in range(nt):
        for y in range(ny):
            for x in range(nx):
                if not gate[t, y, x]:
                    assert np.isnan(ox[t, y, x])
                    continue
                su = sw = 0.0
                n = 0
                for j in range(-r, r + 1):
                    for i in range(-r, r + 1):
                        yy, xx = y + j, x + i
                        if not (0 <= yy < ny and 0 <= xx < nx) or excl[j + r, i + r]:
                            continue
                        a, b = dx[t, yy, xx], dy[t, yy, xx]
                        if np.isfinite(a) and np.isfinite(b):
                            su += ker[j + r, i + r] * a
                            sw += ker[j + r, i + r]
                            n += 1
                assert sup[t, y, x] == n
                if sw >= min_weight * full and sw > 0:
                    assert ox[t, y, x] == pytest.approx(su / sw, rel=1e-5, abs=1e-6)
                else:
                    assert np.isnan(ox[t, y, x])


def test_metric_medians_brute_force():
    rng = np.random.default_rng(9)
    dx = rng.standard_normal((5, 6, 7)).astype(np.float32)
    dy = rng.standard_normal((5, 6, 7)).astype(np.float32)
    dx[2, 3, 3] = np.nan
    h = np.hypot(dx[1:] - dx[:-1], dy[1:] - dy[:-1]).astype(np.float32)
    assert O.pair_to_pair_change(dx, dy) == np.float32(np.median(h[np.isfinite(h)]))
    d = []
    for t in range(5):
        for y in range(1, 5):
            for x in range(1, 6):
                pts = [(y, x), (y - 1, x), (y + 1, x), (y, x - 1), (y, x + 1)]
                if not all(np.isfinite(dx[t][p]) and np.isfinite(dy[t][p]) for p in pts):
                    continue
                mu = (dx[t][pts[1]] + dx[t][pts[2]] + dx[t][pts[3]] + dx[t][pts[4]]) / np.float32(4)
                mv = (dy[t][pts[1]] + dy[t][pts[2]] + dy[t][pts[3]] + dy[t][pts[4]]) / np.float32(4)
                d.append(np.float32(np.hypot(dx[t][pts[0]] - mu, dy[t][pts[0]] - mv)))
    assert O.spatial_roughness(dx, dy) == np.float32(np.median(np.array(d, np.float32)))

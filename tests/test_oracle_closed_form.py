"""The oracle against what the mathematics fixes and the reference's own unit tests.

Each check is chosen so that a plausible slip (wrong twiddle sign, transposed operand, a
dropped conj, an off-by-one neighbour, a lost tie-break) fails it:
- FFT against a direct DFT (both directions, radix 2/3/4 and generic prime sizes);
- correlation against a direct O(n^4) circular cross-correlation in float64 (numpy);
- the known-answer unit tests of piv-core: correlate.rs:310-421, validate.rs:187-248,
  preprocess.rs:223-270, grid.rs:81-105, field.rs:96-109.
"""
import numpy as np
import pytest

from piv_amd import synthetic as S


# every radix of the oracle's FFT, and lengths 2p with a large prime p (the generic-size
# correlation kernel's sizes)
@pytest.mark.parametrize("n", [8, 12, 16, 18, 24, 32, 36, 40, 48, 50, 64, 96, 128, 14, 22, 26, 34,
                               58, 62, 122, 126])
@pytest.mark.parametrize("prec", [32, 64])
def test_fft_matches_direct_dft(oracle, n, prec):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    tol = 2e-5 if prec == 32 else 1e-12
    got = oracle.fft(x, -1, prec)
    assert np.abs(got - W @ x).max() < tol * np.abs(x).sum()
    got = oracle.fft(x, +1, prec)
    assert np.abs(got - W.conj() @ x).max() < tol * np.abs(x).sum()


def direct_circular_xcorr(a, b):
    """c[dy, dx] = sum_{y,x} a[y, x] b[y+dy, x+dx] (circular) -- the plane conj(FA)*FB gives."""
    n = a.shape[0]
    c = np.zeros((n, n))
    for dy in range(n):
        for dx in range(n):
            c[dy, dx] = np.sum(a * np.roll(np.roll(b, -dy, 0), -dx, 1))
    return c


@pytest.mark.parametrize("n", [8, 12, 16])
def test_correlation_plane_matches_direct(oracle, n):
    rng = np.random.default_rng(100 + n)
    img_a = rng.random((n + 4, n + 4)).astype(np.float32)
    img_b = rng.random((n + 4, n + 4)).astype(np.float32)
    ax, ay, bx, by = 1.25, 2.5, 0.75, 1.0
    wa, ea = oracle.load_window(img_a, ax, ay, n)
    wb, eb = oracle.load_window(img_b, bx, by, n)
    ref = direct_circular_xcorr(wa.astype(np.float64), wb.astype(np.float64))
    ref /= np.sqrt(ea * eb)

    def bilinear64(img, x, y):
        x0, y0 = int(np.floor(x)), int(np.floor(y))
        tx, ty = x - x0, y - y0
        i = img.astype(np.float64)[y0:y0 + n + 1, x0:x0 + n + 1]
        w = ((1 - tx) * (1 - ty) * i[:-1, :-1] + tx * (1 - ty) * i[:-1, 1:]
             + (1 - tx) * ty * i[1:, :-1] + tx * ty * i[1:, 1:])
        return w - w.mean()

    da, db = bilinear64(img_a, ax, ay), bilinear64(img_b, bx, by)
    ref64 = direct_circular_xcorr(da, db) / np.sqrt((da * da).sum() * (db * db).sum())
    for prec, r, tol in ((32, ref, 1e-5), (64, ref64, 1e-12)):
        p, corr = oracle.correlate(img_a, img_b, n, ax, ay, bx, by, prec, want_corr=True)
        assert np.abs(corr - r).max() < tol
        # the integer peak is the direct plane's first row-major maximum
        i = int(np.argmax(r))
        assert (p.iy, p.ix) == (i // n, i % n)


def test_integer_shift_is_exact_for_circular_content(oracle):
    """correlate.rs:310-330"""
    n = 32
    a = S.uniform(7, n * n).reshape(n, n)
    b = np.roll(np.roll(a, -3, 0), 5, 1)  # +5 in x, -3 in y
    for prec in (32, 64):
        p = oracle.correlate(a, b, n, 0.0, 0.0, 0.0, 0.0, prec)
        assert abs(p.dx - 5.0) < 0.02 and abs(p.dy + 3.0) < 0.02
        assert abs(p.height - 1.0) < 1e-4
        assert p.ppr > 2.0


def test_isolated_particle_is_unbiased(oracle):
    """correlate.rs:350-382"""
    a = S.render(64, 64, [31.3], [30.7], [1.5], [1.0])
    b = S.render(64, 64, [31.7], [30.45], [1.5], [1.0])
    for prec in (32, 64):
        p = oracle.correlate(a, b, 64, 0.0, 0.0, 0.0, 0.0, prec)
        assert abs(p.dx - 0.4) < 1e-3 and abs(p.dy + 0.25) < 1e-3


def test_subpixel_shift_of_particles(oracle):
    """correlate.rs:332-348"""
    a, b = S.particle_pair(64, 64, 3, 0.05, 1.2, S.uniform_field(0.4, -0.3))
    p = oracle.correlate(a, b, 32, 16.0, 16.0, 16.0, 16.0)
    assert abs(p.dx - 0.4) < 0.08 and abs(p.dy + 0.3) < 0.08


def test_fractional_origin_interpolates(oracle):
    """correlate.rs:384-391"""
    yy, xx = np.mgrid[0:40, 0:40]
    a = (xx + 2 * yy).astype(np.float32)
    d, _ = oracle.load_window(a, 3.5, 2.25, 16)
    assert abs(d[0, 1] - d[0, 0] - 1.0) < 1e-4 and abs(d[1, 0] - d[0, 0] - 2.0) < 1e-4


def test_flat_window_has_no_peak(oracle):
    """correlate.rs:393-402"""
    a = np.full((32, 32), 0.3, np.float32)
    assert oracle.correlate(a, a, 32, 0.0, 0.0, 0.0, 0.0) is None
    assert oracle.correlate(a, a, 32, 0.0, 0.0, 0.0, 0.0, 64) is None


def test_peak_ratio_uses_local_maxima(oracle):
    """correlate.rs:404-421"""
    n = 16
    y, x = np.mgrid[0:n, 0:n]
    dy = np.minimum(y, n - y).astype(np.float32)
    dx = np.minimum(x, n - x).astype(np.float32)
    c = np.exp(-(dx * dx + dy * dy) / np.float32(8.0)).astype(np.float32)
    c[8, 8] = 0.25
    c[2, 0] += np.float32(0.05)
    c[3, 0] -= np.float32(0.05)
    p = oracle.find_peak(c)
    assert (p.dx, p.dy) == (0.0, 0.0)
    assert abs(p.ppr - 4.0) < 1e-3
    p64 = oracle.find_peak64(c.astype(np.float64))
    assert abs(p64.ppr - 4.0) < 1e-3


def test_argmax_tie_break_is_first_row_major(oracle):
    n = 16
    c = np.zeros((n, n), np.float32)
    c[5, 9] = 1.0
    c[5, 10] = 1.0
    c[11, 2] = 1.0
    p = oracle.find_peak(c)
    assert (p.iy, p.ix) == (5, 9)


def test_subpixel_gaussian_three_point_is_exact_for_a_gaussian(oracle):
    """A sampled Gaussian peak at a known sub-pixel offset is recovered exactly by the
    logarithmic three-point fit; a parabolic fallback is used when a neighbour is <= 0."""
    n = 32
    y, x = np.mgrid[0:n, 0:n]
    x0, y0 = 3.3, -2.2
    dxs = ((x - x0 + n / 2) % n) - n / 2
    dys = ((y - y0 + n / 2) % n) - n / 2
    c = np.exp(-(dxs ** 2 + dys ** 2) / 3.0)
    p = oracle.find_peak64(c)
    assert abs(p.dx - x0) < 1e-9 and abs(p.dy - y0) < 1e-9
    p32 = oracle.find_peak(c.astype(np.float32))
    assert abs(p32.dx - x0) < 1e-4 and abs(p32.dy - y0) < 1e-4
    c2 = np.zeros((n, n))
    c2[0, 0], c2[0, 1], c2[0, n - 1] = 1.0, 0.5, -0.5
    p2 = oracle.find_peak64(c2)
    # parabolic: (m - p) / (2 (m - 2c + p)) with m=-0.5, c=1, p=0.5
    assert abs(p2.dx - (-1.0) / (2 * (-0.5 - 2 + 0.5))) < 1e-12


def _uniform_field(O, u, v):
    g = O.grid_new(160, 160, 32, 0.5)
    shp = (g.ny, g.nx)
    return O.Field(g.as_tuple(), np.full(shp, u, np.float32), np.full(shp, v, np.float32),
                   np.full(shp, 5.0, np.float32), np.zeros(shp, np.uint8))


def test_validate_catches_and_replaces_injected_outliers(oracle):
    """validate.rs:187-207"""
    f = _uniform_field(oracle, 2.0, -1.0)
    i = np.arange(f.dx.size).reshape(f.dx.shape)
    f.dx += (np.float32(0.05) * ((i * 7 % 5).astype(np.float32) - 2)).astype(np.float32)
    outliers = [(2, 3), (5, 5), (8, 0)]
    for iy, ix in outliers:
        f.dx[iy, ix], f.dy[iy, ix] = 9.0, 4.0
    v, s = oracle.validate(f)
    assert s["median_outliers"] == 3 and s["replaced"] == 3
    for iy, ix in outliers:
        assert v.flags[iy, ix] & 4 and v.flags[iy, ix] & 8
        assert abs(v.dx[iy, ix] - 2.0) < 0.2 and abs(v.dy[iy, ix] + 1.0) < 1e-6


def test_validate_low_ppr_and_no_signal(oracle):
    """validate.rs:209-221"""
    f = _uniform_field(oracle, 1.0, 1.0)
    f.ppr[4, 4] = 1.05
    f.flags[0, 0] = 1
    f.dx[0, 0] = f.dy[0, 0] = np.nan
    v, s = oracle.validate(f)
    assert (s["low_ppr"], s["no_signal"], s["median_outliers"]) == (1, 1, 0)
    assert v.flags[4, 4] == 2 | 8
    assert v.flags[0, 0] == 1 and np.isnan(v.dx[0, 0])


def test_validate_out_of_range(oracle):
    """validate.rs:227-242"""
    f = _uniform_field(oracle, 1.0, 0.5)
    f.dx[3, 3] = 40.0
    v, s = oracle.validate(f, max_displacement=32.0)
    assert s["out_of_range"] == 1
    assert v.flags[3, 3] & 16 and v.flags[3, 3] & 8 and v.dx[3, 3] == 1.0


def test_median_even_and_odd(oracle):
    """validate.rs:244-248"""
    assert oracle.median([3.0, 1.0, 2.0]) == 2.0
    assert oracle.median([4.0, 1.0, 2.0, 3.0]) == 2.5


def test_integral_stats_match_direct(oracle):
    """preprocess.rs:223-235"""
    f = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 10]], np.float32)
    m, s = oracle.sat_stats(oracle.sat_build(f), 3, 1, 1, 2, 2)
    assert abs(m - 7.25) < 1e-6
    var = np.mean((np.array([5, 6, 8, 10], np.float32) - 7.25) ** 2)
    assert abs(s - np.sqrt(var)) < 1e-5


def test_blur_preserves_constant_and_mass(oracle):
    """preprocess.rs:237-248"""
    b = oracle.gaussian_blur(np.full((20, 30), 0.5, np.float32), 1.5)
    assert np.all(np.abs(b - 0.5) < 1e-6)
    d = np.zeros((21, 21), np.float32)
    d[10, 10] = 1
    b = oracle.gaussian_blur(d, 1.0)
    assert abs(b.sum() - 1) < 1e-5
    assert b[10, 10] > b[10, 11] > b[10, 12]
    k, r = oracle.gauss_kernel(1.0)
    assert r == 3 and k.size == 7 and abs(k.sum() - 1) < 1e-6
    assert oracle.gauss_kernel(0.2)[1] == 1


def test_normalize_removes_gradient(oracle):
    """preprocess.rs:250-256"""
    y, x = np.mgrid[0:40, 0:40]
    f = (0.01 * x + 0.002 * y).astype(np.float32)
    assert abs(oracle.local_normalize(f, 5, 0.0)[20, 20]) < 1e-3


def test_grid_geometry(oracle):
    """grid.rs:81-105"""
    g = oracle.grid_new(1080, 1920, 32, 0.5)
    assert g.step == 16 and (g.nx, g.ny) == (119, 66) and (g.x0, g.y0) == (0, 4)
    g = oracle.grid_new(64, 64, 32, 0.0)
    assert (g.nx, g.ny, g.step, g.x0) == (2, 2, 32, 0)
    for args in ((16, 64, 32, 0.5), (64, 64, 32, 1.0), (64, 64, 32, -0.1)):
        with pytest.raises(oracle.PivError):
            oracle.grid_new(*args)


def test_sample_interpolates_and_skips_nan(oracle):
    """field.rs:96-109"""
    g = oracle.grid_new(64, 64, 32, 0.0)
    dx = np.array([[0.0, 2.0], [0.0, 2.0]], np.float32)
    dy = np.ones((2, 2), np.float32)
    assert oracle.field_sample(g, dx, dy, 31.5, 20.0) == (1.0, 1.0)
    assert oracle.field_sample(g, dx, dy, -100.0, 15.5) == (0.0, 1.0)
    dx[0, 0] = np.nan
    assert oracle.field_sample(g, dx, dy, 31.5, 15.5) == (2.0, 1.0)
    assert oracle.field_sample(g, dx, dy, 15.5, 15.5) == (0.0, 0.0)


def test_decode_normalisation(oracle):
    """decode.rs:227-234: exact IEEE quotients."""
    v8 = np.arange(256, dtype=np.uint8)
    assert np.array_equal(oracle.normalize_pixels(v8), v8.astype(np.float32) / np.float32(255))
    v16 = np.arange(65536, dtype=np.uint16)
    assert np.array_equal(oracle.normalize_pixels(v16),
                          v16.astype(np.float32) / np.float32(65535))


def test_bad_configs_rejected(oracle):
    """multipass.rs:477-507"""
    a = np.zeros((64, 64), np.float32)
    pa = oracle.Prepared(a, a, oracle.sat_build(a))
    for cfg in (oracle.PivConfig(windows=[]), oracle.PivConfig(windows=[33]),
                oracle.PivConfig(overlap=1.0), oracle.PivConfig(final_iterations=0)):
        with pytest.raises(oracle.PivError):
            oracle.analyze(pa, pa, cfg)


def _run(O, a, b, windows, **kw):
    pa, pb = O.prepare(a), O.prepare(b)
    cfg = O.PivConfig(windows=windows, final_iterations=1 if len(windows) == 1 else 2, **kw)
    return O.analyze(pa, pb, cfg)


def _errors(f, truth, margin, size):
    g = f.grid
    xs = g[4] + np.arange(g[2]) * g[1] + (g[0] - 1) / 2
    ys = g[5] + np.arange(g[3]) * g[1] + (g[0] - 1) / 2
    X, Y = np.meshgrid(xs, ys)
    keep = (X >= margin) & (Y >= margin) & (X <= size - margin) & (Y <= size - margin)
    u, v = truth(X.astype(np.float32), Y.astype(np.float32))
    e = np.sqrt((f.dx - u) ** 2 + (f.dy - v) ** 2)[keep]
    return np.sqrt(np.mean(e ** 2)), e.max()


def test_multipass_uniform_and_large_shift(oracle):
    """multipass.rs:318-343, 378-391"""
    for i, (u, v) in enumerate([(0.25, 0.0), (3.7, -1.3), (-2.2, 2.6)]):
        a, b = S.particle_pair(256, 256, 10 + i, 0.04, 1.2, S.uniform_field(u, v))
        f, s = _run(oracle, a, b, [64, 32])
        rms, _ = _errors(f, S.uniform_field(u, v), 20, 256)
        assert rms < 0.05
        assert s["median_outliers"] * 50 <= f.dx.size
    a, b = S.particle_pair(320, 320, 5, 0.04, 1.2, S.uniform_field(20.0, -6.0))
    f, _ = _run(oracle, a, b, [64, 32])
    assert _errors(f, S.uniform_field(20.0, -6.0), 48, 320)[0] < 0.1
    f1, _ = _run(oracle, a, b, [32])
    assert _errors(f1, S.uniform_field(20.0, -6.0), 48, 320)[0] > 1.0

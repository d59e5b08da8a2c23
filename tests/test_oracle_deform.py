"""The oracle's window deformation (the library's definition; the reference offsets windows
rigidly, multipass.rs:207-231) pinned by what the mathematics fixes:
  - a uniform predictor shifts no pixel, so every deformed window and every pass result equals
    the rigid path bitwise (FP32 and FP64);
  - the predictor interpolant reproduces an affine field exactly (up to f32 rounding), and an
    affine field on a linear image resamples exactly in FP64, for both kernels;
  - the cubic (4-point Lagrange) kernel reproduces a cubic image exactly; bilinear does not;
  - where a predictor neighbour is NaN the interpolant is VectorField::sample (field.rs:56-81);
  - the sign: windows deformed by the true field of a sheared particle pair line up, so the
    normalised correlation peak rises and the residual displacement falls (the opposite sign
    makes both worse);
  - end to end, deformation lowers the error against the generating vortex."""
import numpy as np
import pytest

from piv_amd import synthetic as S


def grid_field(O, h, w, n, fn):
    g = O.grid_new(h, w, n, 0.5)
    xs = g.x0 + np.arange(g.nx) * g.step + (n - 1) / 2.0
    ys = g.y0 + np.arange(g.ny) * g.step + (n - 1) / 2.0
    X, Y = np.meshgrid(xs, ys)
    u, v = fn(X, Y)
    return g, np.asarray(u, np.float32), np.asarray(v, np.float32)


def test_uniform_predictor_is_rigid_bitwise(oracle):
    a, b = S.particle_frames(192, 224, 11, field=S.uniform_field(2.3, -1.7))
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    h, w = a.shape
    prev, pu, pv = grid_field(oracle, h, w, 64, lambda X, Y: (np.full_like(X, 2.3),
                                                               np.full_like(X, -1.7)))
    g = oracle.grid_new(h, w, 32, 0.5)
    origins = oracle.place_windows(g, h, w, prev, pu, pv)
    hx, hy = oracle.deform_shifts(prev, pu, pv, 32, *origins[0])
    assert not hx.any() and not hy.any()
    for prec in (32, 64):
        rigid = oracle.correlate_pass(pa, pb, g, origins, precision=prec)
        for mode in (1,):
            d = oracle.correlate_pass(pa, pb, g, origins, precision=prec,
                                      deform=(mode, prev, pu, pv))
            assert np.array_equal(d.dx.view(np.uint32), rigid.dx.view(np.uint32)), prec
            assert np.array_equal(d.dy.view(np.uint32), rigid.dy.view(np.uint32)), prec
            assert np.array_equal(d.flags, rigid.flags)
    # the loaded windows themselves, at a fractional origin
    img = pa.image
    for x, y in ((10.25, 7.5), (40.0, 33.0), (17.6, 50.1)):
        ref = oracle.load_window(img, x, y, 32)
        z = np.zeros((32, 32), np.float32)
        for sgn in (-1, 1):
            win, e = oracle.load_window_deformed(img, x, y, 32, 1, z, z, sgn)
            assert np.array_equal(win.view(np.uint32), ref[0].view(np.uint32))
            assert e == ref[1]


def test_affine_field_linear_image_exact(oracle):
    h, w, n = 160, 192, 32
    A = np.array([[0.02, -0.015], [0.01, 0.025]])
    b0 = np.array([1.25, -0.75])
    affine = lambda X, Y: (b0[0] + A[0, 0] * X + A[0, 1] * Y, b0[1] + A[1, 0] * X + A[1, 1] * Y)
    prev, pu, pv = grid_field(oracle, h, w, 64, affine)
    # the interpolant reproduces the affine field inside the grid (f32 rounding only)
    for x, y in ((70.3, 61.7), (95.0, 80.25), (100.5, 90.5)):
        u, v = oracle.deform_sample(prev, pu, pv, x, y)
        eu, ev = affine(np.float64(x), np.float64(y))
        assert abs(u - eu) < 2e-6 and abs(v - ev) < 2e-6
    img = (0.003 * np.arange(w)[None, :] + 0.002 * np.arange(h)[:, None] + 0.1).astype(np.float32)
    lin = lambda X, Y: 0.003 * X + 0.002 * Y + 0.1
    ax, ay, bx, by = 60.3, 50.6, 62.9, 49.4
    hx, hy = oracle.deform_shifts(prev, pu, pv, n, ax, ay, bx, by)
    # shifts: half the affine field's change from the window centre
    c = np.arange(n)
    half = (n - 1) / 2
    ehx = 0.5 * (A[0, 0] * (c[None, :] - half) + A[0, 1] * (c[:, None] - half))
    ehy = 0.5 * (A[1, 0] * (c[None, :] - half) + A[1, 1] * (c[:, None] - half))
    assert np.abs(hx - ehx).max() < 2e-6 and np.abs(hy - ehy).max() < 2e-6
    for mode in (1, 2):
        for sgn, (ox, oy) in ((-1, (ax, ay)), (1, (bx, by))):
            win, _ = oracle.load_window_deformed(img, ox, oy, n, mode, hx, hy, sgn, precision=64)
            # positions as the oracle forms them (f32), values of the linear image there
            fx = (np.float32(ox) - np.floor(np.float32(ox))).astype(np.float32)
            fy = (np.float32(oy) - np.floor(np.float32(oy))).astype(np.float32)
            px = np.floor(ox) + c[None, :] + (fx + sgn * hx.astype(np.float32)).astype(np.float32)
            py = np.floor(oy) + c[:, None] + (fy + sgn * hy.astype(np.float32)).astype(np.float32)
            expect = lin(px.astype(np.float64), py.astype(np.float64))
            expect = expect - expect.mean()
            assert np.abs(win - expect).max() < 1e-6, (mode, sgn, np.abs(win - expect).max())


def test_cubic_kernel_reproduces_cubics(oracle):
    h, w, n = 96, 96, 16
    xs, ys = np.arange(w, dtype=np.float64), np.arange(h, dtype=np.float64)
    p = lambda x: 1e-5 * x ** 3 - 2e-3 * x ** 2 + 0.05 * x + 1.0
    q = lambda y: -2e-5 * y ** 3 + 1e-3 * y ** 2 + 0.02 * y + 0.5
    img = (q(ys)[:, None] * p(xs)[None, :]).astype(np.float32)
    z = np.zeros((n, n), np.float32)
    hx = np.full((n, n), np.float32(0.125))
    hy = np.full((n, n), np.float32(-0.375))
    errs = {}
    for mode in (1, 2):
        worst = 0.0
        for ox, oy in ((30.3, 41.7), (45.55, 28.2)):
            win, _ = oracle.load_window_deformed(img, ox, oy, n, mode, hx, hy, -1, precision=64)
            fx = np.float32(ox) - np.floor(np.float32(ox)) - np.float32(0.125)
            fy = np.float32(oy) - np.floor(np.float32(oy)) + np.float32(0.375)
            X = np.floor(ox) + np.arange(n)[None, :] + np.float64(fx)
            Y = np.floor(oy) + np.arange(n)[:, None] + np.float64(fy)
            # the cubic itself at the sample positions (the samples are stored in f32)
            expect = q(Y) * p(X)
            expect = expect - expect.mean()
            worst = max(worst, float(np.abs(win - expect).max()))
        errs[mode] = worst
    assert errs[2] < 2e-6, errs  # f32 storage of the samples is the only error
    assert errs[1] > 50 * errs[2], errs  # bilinear cannot follow the curvature


def test_nan_neighbours_fall_back_to_sample(oracle):
    h, w = 160, 160
    prev, pu, pv = grid_field(oracle, h, w, 32, lambda X, Y: (0.01 * X + 0.5, -0.02 * Y))
    pu[3, 4] = np.nan
    pv[5, 2] = np.nan
    g = prev
    # points with a NaN among their four neighbours: exactly VectorField::sample
    for x, y in ((g.x0 + 4 * g.step + 15.5 + 3.0, g.y0 + 3 * g.step + 15.5 - 2.0),
                 (g.x0 + 2 * g.step + 15.5 + 7.0, g.y0 + 5 * g.step + 15.5 + 1.0)):
        assert oracle.deform_sample(g, pu, pv, x, y) == oracle.field_sample(g, pu, pv, x, y)
    # a point whose four neighbours are finite takes the lerp form, which for an affine field
    # equals the weighted form to rounding
    u, v = oracle.deform_sample(g, pu, pv, 20.0, 20.0)
    su, sv = oracle.field_sample(g, pu, pv, 20.0, 20.0)
    assert abs(u - su) < 1e-6 and abs(v - sv) < 1e-6


def test_shift_sign_aligns_sheared_windows(oracle):
    """A pure shear u = 0.1 (y - c): windows deformed by the true field hold the same
    particles, so the peak ratio rises and the measured vector matches the truth at the
    window centre; flipping the field's sign (the wrong-sign mistake) does worse."""
    h = w = 192
    c = h / 2
    shear = lambda X, Y: (np.float32(0.1) * (Y - np.float32(c)), np.zeros_like(Y))
    a, b = S.particle_frames(h, w, 5, field=shear, density=0.03)
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    prev, pu, pv = grid_field(oracle, h, w, 64, shear)
    n = 48
    g = oracle.grid_new(h, w, n, 0.5)
    origins = oracle.place_windows(g, h, w, prev, pu, pv)
    rigid = oracle.correlate_pass(pa, pb, g, origins)
    good = oracle.correlate_pass(pa, pb, g, origins, deform=(1, prev, pu, pv))
    bad = oracle.correlate_pass(pa, pb, g, origins, deform=(1, prev, -pu, -pv))
    xs = g.x0 + np.arange(g.nx) * g.step + (n - 1) / 2.0
    ys = g.y0 + np.arange(g.ny) * g.step + (n - 1) / 2.0
    X, Y = np.meshgrid(xs, ys)
    tu, tv = shear(X.astype(np.float32), Y.astype(np.float32))
    inner = np.zeros_like(X, bool)
    inner[1:-1, 1:-1] = True

    def err(f):
        return float(np.sqrt(((f.dx - tu) ** 2 + (f.dy - tv) ** 2)[inner].mean()))

    assert np.median(good.ppr[inner]) > np.median(rigid.ppr[inner])
    assert err(good) < 0.5 * err(rigid), (err(good), err(rigid))
    assert np.median(bad.ppr[inner]) < np.median(good.ppr[inner])
    assert err(bad) > err(good)


@pytest.mark.parametrize("mode", ["bilinear", "bicubic"])
def test_deformation_reduces_vortex_error(oracle, mode):
    h = w = 384
    truth = S.lamb_oseen_shear(h, w, vmax=6.0, shear=0.02)
    a, b = S.particle_frames(h, w, 2024, field=truth)
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    cfg = oracle.PivConfig(windows=[64, 32, 16], overlap=0.5, final_iterations=1)
    rigid, _ = oracle.analyze(pa, pb, cfg)
    deformed, _ = oracle.analyze(pa, pb, cfg, deformation=mode)
    g = oracle.grid_new(h, w, 16, 0.5)
    xs = g.x0 + np.arange(g.nx) * g.step + 7.5
    ys = g.y0 + np.arange(g.ny) * g.step + 7.5
    X, Y = np.meshgrid(xs.astype(np.float32), ys.astype(np.float32))
    u, v = truth(X, Y)
    keep = (X >= 32) & (Y >= 32) & (X <= w - 32) & (Y <= h - 32)

    def rms(f):
        ok = keep & ((f.flags & 23) == 0)
        return float(np.sqrt(((f.dx - u) ** 2 + (f.dy - v) ** 2)[ok].mean()))

    r0, r1 = rms(rigid), rms(deformed)
    print(f"{mode}: rms vs truth {r1:.4f} px (rigid {r0:.4f})")
    assert r1 < 0.6 * r0

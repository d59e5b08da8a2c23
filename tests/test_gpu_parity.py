"""GPU parity: the HIP path against the Rust golden fixtures and the CPU oracle.

Bars (from the reference's own GPU-port parity suite,
swift/Tests/PIVKitTests/RustParityGPUTests.swift:37-123 and PreprocessParityTests.swift):
  - sample normalisation, prepared images, window origins, validation: bit-exact;
  - raw peaks <= 1e-4 px, PPR rel <= 1e-5, NO_SIGNAL decisions identical;
  - integer peak indices identical to the oracle's (no exact ties occur in these inputs);
  - final fields RMS <= 1e-4 px with identical flags and summaries.
"""
import numpy as np
import pytest

import rust_fixtures as rf

pytestmark = pytest.mark.gpu

CASES = rf.names()


@pytest.fixture(scope="module")
def piv():
    import piv_amd

    if piv_amd.device_count() < 1:
        pytest.skip("no GPU")
    return piv_amd


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def make_cfg(piv, c):
    p, v = c.piv, c.piv["validation"]
    cfg = piv.PivConfig(windows=list(p["windows"]), overlap=rf.f32(p["overlap"]),
                        final_iterations=p["final_iterations"],
                        mask=piv.MaskConfig(rf.f32(p["mask"]["min_mean"]),
                                            rf.f32(p["mask"]["min_std"])),
                        validation=piv.ValidationConfig(rf.f32(v["ppr_threshold"]),
                                                        rf.f32(v["median_threshold"]),
                                                        rf.f32(v["median_eps"]), v["replace"]))
    pre = piv.PreprocessConfig(c.preprocess["smooth_sigma"], c.preprocess["normalize_radius"],
                               rf.f32(c.preprocess["normalize_eps"]))
    return cfg, pre


def oracle_cfg(O, c):
    p, v = c.piv, c.piv["validation"]
    return O.PivConfig(windows=list(p["windows"]), overlap=rf.f32(p["overlap"]),
                       final_iterations=p["final_iterations"],
                       min_mean=rf.f32(p["mask"]["min_mean"]), min_std=rf.f32(p["mask"]["min_std"]),
                       ppr_threshold=rf.f32(v["ppr_threshold"]),
                       median_threshold=rf.f32(v["median_threshold"]),
                       median_eps=rf.f32(v["median_eps"]), replace=v["replace"])


def test_normalisation_exhaustive(piv):
    v8 = np.arange(256, dtype=np.uint8)
    assert np.array_equal(bits(piv.normalize_pixels(v8)), bits(v8.astype(np.float32) / np.float32(255)))
    v16 = np.arange(65536, dtype=np.uint16)
    assert np.array_equal(bits(piv.normalize_pixels(v16)),
                          bits(v16.astype(np.float32) / np.float32(65535)))


@pytest.mark.parametrize("name", CASES)
def test_prepared_images_bit_exact(piv, name):
    c = rf.load(name)
    cfg, pre = make_cfg(piv, c)
    an = piv.PivAnalyzer(cfg, pre)
    for raw, ref in ((c.frame_a, c.prepared_a), (c.frame_b, c.prepared_b)):
        img, _ = an.prepare(raw, c.background).download()
        assert np.array_equal(bits(img), bits(ref)), np.abs(img - ref).max()


@pytest.mark.parametrize("name", CASES)
def test_origins_bit_exact(piv, name):
    c = rf.load(name)
    cfg, pre = make_cfg(piv, c)
    an = piv.PivAnalyzer(cfg, pre)
    prev_g, prev = None, None
    for i, p in enumerate(c.passes):
        g = piv.Grid.new(c.height, c.width, p.grid["window"], cfg.overlap)
        o = piv.place_windows(an, g, c.height, c.width, prev_g,
                              None if prev is None else prev.dx, None if prev is None else prev.dy)
        assert np.array_equal(bits(o), bits(p.origins)), f"pass {i}"
        prev_g, prev = g, p.validated


@pytest.mark.parametrize("name", CASES)
def test_raw_peaks_match_rust_and_oracle(piv, oracle, name):
    c = rf.load(name)
    cfg, pre = make_cfg(piv, c)
    an = piv.PivAnalyzer(cfg, pre)
    fa, fb = an.prepare(c.frame_a, c.background), an.prepare(c.frame_b, c.background)
    oa = oracle.prepare(c.frame_a, pre.smooth_sigma, pre.normalize_radius, pre.normalize_eps, c.background)
    ob = oracle.prepare(c.frame_b, pre.smooth_sigma, pre.normalize_radius, pre.normalize_eps, c.background)
    for i, p in enumerate(c.passes):
        n = p.grid["window"]
        dx, dy, pr, fl, pk = piv.correlate_windows(an, fa, fb, n, p.origins)
        shp = p.raw.dx.shape
        dx, dy, pr, fl, pk = (a.reshape(shp) for a in (dx, dy, pr, fl, pk))
        assert np.array_equal(fl, p.raw.flags), f"pass {i} flags"
        fin = np.isfinite(p.raw.dx)
        assert np.array_equal(fin, np.isfinite(dx))
        assert np.abs(dx - p.raw.dx)[fin].max(initial=0) < 1e-4
        assert np.abs(dy - p.raw.dy)[fin].max(initial=0) < 1e-4
        assert np.abs(pr[fin] / p.raw.ppr[fin] - 1).max(initial=0) < 1e-5
        g = oracle.grid_new(c.height, c.width, n, cfg.overlap)
        ref = oracle.correlate_pass(oa, ob, g, p.origins, cfg.mask.min_mean, cfg.mask.min_std)
        ref_pk = np.where(ref.peak_iy >= 0, ref.peak_iy * n + ref.peak_ix, -1)
        assert np.array_equal(pk, ref_pk), f"pass {i} integer peaks"
        assert np.abs(dx - ref.dx)[fin].max(initial=0) < 1e-4


@pytest.mark.parametrize("name", CASES)
def test_validation_bit_exact(piv, name):
    c = rf.load(name)
    cfg, pre = make_cfg(piv, c)
    an = piv.PivAnalyzer(cfg, pre)
    maxd = np.float32(cfg.windows[0]) / np.float32(2)
    for i, p in enumerate(c.passes):
        dx, dy, fl, s = piv.validate(an, p.raw.dx, p.raw.dy, p.raw.ppr, p.raw.flags, maxd)
        assert np.array_equal(fl, p.validated.flags), f"pass {i}"
        assert np.array_equal(bits(dx), bits(p.validated.dx))
        assert np.array_equal(bits(dy), bits(p.validated.dy))


@pytest.mark.parametrize("name", CASES)
def test_full_analysis_matches_rust(piv, name):
    c = rf.load(name)
    cfg, pre = make_cfg(piv, c)
    an = piv.PivAnalyzer(cfg, pre)
    fa, fb = an.prepare(c.frame_a, c.background), an.prepare(c.frame_b, c.background)
    r = an.analyze(fa, fb)
    ref = c.passes[-1].validated
    assert np.array_equal(r.field.flags, ref.flags)
    assert np.array_equal(np.isfinite(r.field.dx), np.isfinite(ref.dx))
    both = np.isfinite(ref.dx)
    d2 = (r.field.dx - ref.dx)[both] ** 2 + (r.field.dy - ref.dy)[both] ** 2
    if d2.size:
        assert np.sqrt(d2.mean()) < 1e-4
    assert r.summary.__dict__ == c.summary
    # traced run: every pass, and the final field is the untraced result (bitwise)
    r2, traces = an.analyze_traced(fa, fb)
    assert len(traces) == len(c.passes)
    for t, p in zip(traces, c.passes):
        assert t.raw.grid.nx == p.grid["nx"] and t.raw.grid.ny == p.grid["ny"]
    assert np.array_equal(bits(traces[-1].validated.dx), bits(r.field.dx))
    assert np.array_equal(bits(r2.field.dx), bits(r.field.dx))
    assert np.array_equal(r2.field.flags, r.field.flags)
    assert np.array_equal(bits(traces[0].origins), bits(c.passes[0].origins))


def _random_field(rng, ny, nx):
    """A noisy uniform field with every validation path present: NO_SIGNAL holes, low peak
    ratios, out-of-range and outlier vectors, blocks of invalid vectors (radius-2
    replacement, and cells with no usable neighbour at all), and +-0 ties."""
    dx = (1.5 + 0.05 * rng.standard_normal((ny, nx))).astype(np.float32)
    dy = (-0.7 + 0.05 * rng.standard_normal((ny, nx))).astype(np.float32)
    ppr = (1.0 + 4.0 * rng.random((ny, nx))).astype(np.float32)
    fl = np.zeros((ny, nx), np.uint8)
    out = rng.random((ny, nx)) < 0.04
    dx[out] += rng.choice([-6.0, 6.0, 40.0], out.sum()).astype(np.float32)
    hole = rng.random((ny, nx)) < 0.05
    dx[hole] = dy[hole] = np.nan
    ppr[hole] = 0.0
    fl[hole] = 1
    # a 3x3 block of low-PPR vectors: the centre has no usable radius-1 neighbour
    if ny > 8 and nx > 8:
        ppr[3:6, 4:7] = 1.01
        # a 5x5 block of holes: its centre has no usable neighbour within radius 2
        dx[ny - 7:ny - 2, nx - 7:nx - 2] = np.nan
        dy[ny - 7:ny - 2, nx - 7:nx - 2] = np.nan
        fl[ny - 7:ny - 2, nx - 7:nx - 2] = 1
        ppr[ny - 6:ny - 3, nx - 6:nx - 3] = 1.0
        fl[ny - 5, nx - 5] = 0
        dx[ny - 5, nx - 5], dy[ny - 5, nx - 5] = 9.0, 9.0
    z = rng.random((ny, nx)) < 0.02
    dx[z] = np.where(rng.random(z.sum()) < 0.5, np.float32(0.0), np.float32(-0.0))
    return dx, dy, ppr, fl


@pytest.mark.parametrize("seed,shape,replace", [(1, (31, 31), True), (2, (127, 127), True),
                                                (3, (17, 40), False), (4, (2, 3), True),
                                                (5, (1, 9), True), (6, (64, 65), True)])
def test_validation_random_fields_bit_exact(piv, oracle, seed, shape, replace):
    rng = np.random.default_rng(seed)
    dx, dy, ppr, fl = _random_field(rng, *shape)
    cfg = piv.PivConfig(windows=[32], final_iterations=1,
                        validation=piv.ValidationConfig(replace=replace))
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    maxd = np.float32(16.0)
    gdx, gdy, gfl, gs = piv.validate(an, dx, dy, ppr, fl, maxd)
    f = oracle.Field(None, dx, dy, ppr, fl)
    ref, rs = oracle.validate(f, 1.3, 2.0, 0.1, replace, maxd)
    assert np.array_equal(gfl, ref.flags)
    assert np.array_equal(bits(gdx), bits(ref.dx))
    assert np.array_equal(bits(gdy), bits(ref.dy))
    assert gs.__dict__ == rs

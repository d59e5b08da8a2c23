"""The oracle against the golden cases written by the Rust reference (CPU only).

Tolerances are the ones the reference's own GPU port was held to
(swift/Tests/PIVKitTests/RustParityGPUTests.swift:37-123, PreprocessParityTests.swift:29-62):
origins and validation bit-exact, raw peaks <= 1e-4 px, PPR rel <= 1e-5, identical flags,
final RMS <= 1e-4 px.
"""
import numpy as np
import pytest

import rust_fixtures as rf

CASES = rf.names()


def _cfg(O, c):
    p, v = c.piv, c.piv["validation"]
    return O.PivConfig(windows=list(p["windows"]), overlap=rf.f32(p["overlap"]),
                       final_iterations=p["final_iterations"],
                       min_mean=rf.f32(p["mask"]["min_mean"]), min_std=rf.f32(p["mask"]["min_std"]),
                       ppr_threshold=rf.f32(v["ppr_threshold"]),
                       median_threshold=rf.f32(v["median_threshold"]),
                       median_eps=rf.f32(v["median_eps"]), replace=v["replace"])


def _prep(O, c, raw):
    p = c.preprocess
    return O.prepare(raw, p["smooth_sigma"], p["normalize_radius"], rf.f32(p["normalize_eps"]),
                     c.background)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", CASES)
def test_prepared_images_bit_exact(oracle, name):
    c = rf.load(name)
    for raw, ref in ((c.frame_a, c.prepared_a), (c.frame_b, c.prepared_b)):
        assert np.array_equal(bits(_prep(oracle, c, raw).image), bits(ref))


@pytest.mark.parametrize("name", CASES)
def test_grid_and_origins_bit_exact(oracle, name):
    c = rf.load(name)
    prev_g, prev = None, None
    for i, p in enumerate(c.passes):
        g = oracle.grid_new(c.height, c.width, p.grid["window"], rf.f32(c.piv["overlap"]))
        assert g.as_tuple() == tuple(p.grid[k] for k in ("window", "step", "nx", "ny", "x0", "y0"))
        o = oracle.place_windows(g, c.height, c.width, prev_g,
                                 None if prev is None else prev.dx, None if prev is None else prev.dy)
        assert np.array_equal(bits(o), bits(p.origins)), f"pass {i}"
        prev_g, prev = g, p.validated


@pytest.mark.parametrize("name", CASES)
def test_raw_peaks_within_tolerance(oracle, name):
    c = rf.load(name)
    cfg = _cfg(oracle, c)
    pa, pb = _prep(oracle, c, c.frame_a), _prep(oracle, c, c.frame_b)
    for i, p in enumerate(c.passes):
        g = oracle.grid_new(c.height, c.width, p.grid["window"], cfg.overlap)
        raw = oracle.correlate_pass(pa, pb, g, p.origins, cfg.min_mean, cfg.min_std)
        assert np.array_equal(raw.flags, p.raw.flags), f"pass {i} flags"
        fin = np.isfinite(p.raw.dx)
        assert np.array_equal(fin, np.isfinite(raw.dx))
        assert np.abs(raw.dx - p.raw.dx)[fin].max(initial=0) < 1e-4
        assert np.abs(raw.dy - p.raw.dy)[fin].max(initial=0) < 1e-4
        assert np.abs(raw.ppr[fin] / p.raw.ppr[fin] - 1).max(initial=0) < 1e-5
        # FP64 oracle agrees with the Rust FP32 field to FP32 rounding as well
        raw64 = oracle.correlate_pass(pa, pb, g, p.origins, cfg.min_mean, cfg.min_std, precision=64)
        assert np.array_equal(raw64.flags, p.raw.flags)
        assert np.abs(raw64.dx64 - p.raw.dx)[fin].max(initial=0) < 1e-4


@pytest.mark.parametrize("name", CASES)
def test_validation_bit_exact(oracle, name):
    c = rf.load(name)
    cfg = _cfg(oracle, c)
    for i, p in enumerate(c.passes):
        f = oracle.Field(None, p.raw.dx, p.raw.dy, p.raw.ppr, p.raw.flags)
        val, _ = oracle.validate(f, cfg.ppr_threshold, cfg.median_threshold, cfg.median_eps,
                                 cfg.replace, np.float32(cfg.windows[0]) / np.float32(2))
        assert np.array_equal(val.flags, p.validated.flags), f"pass {i}"
        assert np.array_equal(bits(val.dx), bits(p.validated.dx))
        assert np.array_equal(bits(val.dy), bits(p.validated.dy))


@pytest.mark.parametrize("name", CASES)
def test_full_analysis_matches_rust(oracle, name):
    c = rf.load(name)
    cfg = _cfg(oracle, c)
    pa, pb = _prep(oracle, c, c.frame_a), _prep(oracle, c, c.frame_b)
    final, summary, traces = oracle.analyze(pa, pb, cfg, traced=True)
    assert len(traces) == len(c.passes)
    ref = c.passes[-1].validated
    assert np.array_equal(final.flags, ref.flags)
    both = np.isfinite(ref.dx) & np.isfinite(final.dx)
    assert np.array_equal(np.isfinite(ref.dx), np.isfinite(final.dx))
    d2 = (final.dx - ref.dx)[both] ** 2 + (final.dy - ref.dy)[both] ** 2
    assert np.sqrt(d2.mean()) < 1e-4 if d2.size else True
    assert summary == c.summary

"""GPU parity at the benchmark configurations and on the paths the fixtures leave open.

For each full-size run the traced GPU analysis is replayed stage by stage through the oracle:
  - the GPU's window origins of every pass go into the oracle's correlation: integer peak
    indices must be identical (no exact ties occur in these seeded inputs), displacements within
    1e-4 px (the FP32 parity bar is 1e-3 px), NO_SIGNAL decisions identical;
  - the GPU's raw field goes into the oracle's validation: flags and values bit-exact;
  - the final fields of an independent oracle run agree within 1e-3 px with identical flags.
Accuracy against the generating displacement is checked where the reference's own tests set
a bar (multipass.rs:318-443).
"""
import numpy as np
import pytest

from piv_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def piv():
    import piv_amd

    if piv_amd.device_count() < 1:
        pytest.skip("no GPU")
    return piv_amd


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def stage_parity(piv, O, an, fa, fb, pa, pb, cfg, tol=1e-4):
    """Replay every pass of a traced GPU run through the oracle."""
    res, traces = an.analyze_traced(fa, fb)
    h, w = pa.image.shape
    maxd = np.float32(cfg.windows[0]) / np.float32(2)
    v = cfg.validation
    for i, t in enumerate(traces):
        g = t.raw.grid
        og = O.grid_new(h, w, g.window, cfg.overlap)
        assert og.as_tuple() == (g.window, g.step, g.nx, g.ny, g.x0, g.y0)
        ref = O.correlate_pass(pa, pb, og, t.origins, cfg.mask.min_mean, cfg.mask.min_std)
        assert np.array_equal(t.raw.flags, ref.flags), f"pass {i} NO_SIGNAL"
        ref_pk = np.where(ref.peak_iy >= 0, ref.peak_iy * g.window + ref.peak_ix, -1)
        mism = np.count_nonzero(t.peak_index != ref_pk)
        assert mism == 0, f"pass {i}: {mism} integer peaks differ"
        fin = np.isfinite(ref.dx)
        assert np.array_equal(fin, np.isfinite(t.raw.dx))
        assert np.abs(t.raw.dx - ref.dx)[fin].max(initial=0) < tol
        assert np.abs(t.raw.dy - ref.dy)[fin].max(initial=0) < tol
        assert np.abs(t.raw.peak_ratio[fin] / ref.ppr[fin] - 1).max(initial=0) < 1e-4
        f = O.Field(None, t.raw.dx, t.raw.dy, t.raw.peak_ratio, t.raw.flags)
        val, s = O.validate(f, v.ppr_threshold, v.median_threshold, v.median_eps, v.replace, maxd)
        assert np.array_equal(val.flags, t.validated.flags), f"pass {i} validation flags"
        assert np.array_equal(bits(val.dx), bits(t.validated.dx))
        assert np.array_equal(bits(val.dy), bits(t.validated.dy))
        assert s == t.summary.__dict__
    return res, traces


def full_parity(O, res, pa, pb, cfg, tol=1e-3):
    ocfg = O.PivConfig(windows=list(cfg.windows), overlap=cfg.overlap,
                       final_iterations=cfg.final_iterations, min_mean=cfg.mask.min_mean,
                       min_std=cfg.mask.min_std, ppr_threshold=cfg.validation.ppr_threshold,
                       median_threshold=cfg.validation.median_threshold,
                       median_eps=cfg.validation.median_eps, replace=cfg.validation.replace)
    ref, s = O.analyze(pa, pb, ocfg)
    assert np.array_equal(res.field.flags, ref.flags)
    fin = np.isfinite(ref.dx)
    assert np.array_equal(fin, np.isfinite(res.field.dx))
    assert np.abs(res.field.dx - ref.dx)[fin].max(initial=0) < tol
    assert np.abs(res.field.dy - ref.dy)[fin].max(initial=0) < tol
    assert s == res.summary.__dict__
    return ref


def truth_rms(field, grid, truth, margin):
    xs = grid.x0 + np.arange(grid.nx) * grid.step + (grid.window - 1) / 2.0
    ys = grid.y0 + np.arange(grid.ny) * grid.step + (grid.window - 1) / 2.0
    X, Y = np.meshgrid(xs.astype(np.float32), ys.astype(np.float32))
    u, v = truth(X, Y)
    H = grid.y0 * 2 + (grid.ny - 1) * grid.step + grid.window
    W = grid.x0 * 2 + (grid.nx - 1) * grid.step + grid.window
    keep = (X >= margin) & (Y >= margin) & (X <= W - margin) & (Y <= H - margin)
    keep &= (field.flags & 23) == 0
    e2 = (field.dx - u) ** 2 + (field.dy - v) ** 2
    return float(np.sqrt(e2[keep].mean()))


def test_512_u8_single_pass(piv, oracle):
    """512x512 u8, uniform (2.3, -1.7), [32] at 50 % (31x31 vectors)."""
    truth = S.uniform_field(2.3, -1.7)
    a, b = S.particle_frames(512, 512, 101, field=truth)
    cfg = piv.PivConfig(windows=[32], final_iterations=1)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    fa, fb = an.prepare(a), an.prepare(b)
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    res, tr = stage_parity(piv, oracle, an, fa, fb, pa, pb, cfg)
    assert (res.field.grid.nx, res.field.grid.ny) == (31, 31)
    full_parity(oracle, res, pa, pb, cfg)
    assert truth_rms(res.field, res.field.grid, truth, 20) < 0.1


def test_1024_u8_three_pass(piv, oracle):
    """1024x1024 u8, Lamb-Oseen vortex + shear, 64 -> 32 -> 16 (127x127 vectors)."""
    truth = S.lamb_oseen_shear(1024, 1024)
    a, b = S.particle_frames(1024, 1024, 2024, field=truth)
    cfg = piv.PivConfig(windows=[64, 32, 16], final_iterations=1)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    fa, fb = an.prepare(a), an.prepare(b)
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    res, tr = stage_parity(piv, oracle, an, fa, fb, pa, pb, cfg)
    assert (res.field.grid.nx, res.field.grid.ny) == (127, 127)
    full_parity(oracle, res, pa, pb, cfg)
    assert truth_rms(res.field, res.field.grid, truth, 32) < 0.15


def test_1280x800_u16_two_pass(piv, oracle):
    """1280x800 u16 with 12-bit values, 48 -> 24 px (grids 52x32 and 105x65)."""
    swirl = S.lamb_oseen_shear(800, 1280, vmax=4.0, shear=0.0, rc=150.0)
    a, b = S.particle_frames(800, 1280, 77, density=0.015, diameter_range=(2.0, 3.0),
                             amp=(150, 255), noise=3.0, field=swirl, bits=12,
                             container=np.uint16, amp_scale=16.0)
    assert a.dtype == np.uint16 and a.max() <= 4095
    cfg = piv.PivConfig(windows=[48, 24], final_iterations=1)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    fa, fb = an.prepare(a), an.prepare(b)
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    res, tr = stage_parity(piv, oracle, an, fa, fb, pa, pb, cfg)
    assert [(t.raw.grid.nx, t.raw.grid.ny) for t in tr] == [(52, 32), (105, 65)]
    full_parity(oracle, res, pa, pb, cfg)
    assert truth_rms(res.field, res.field.grid, swirl, 24) < 0.15


def test_default_conditioning_fog(piv, oracle):
    """Library defaults end to end: blur sigma 1 + local normalisation r 4, [96, 64] with two
    final iterations, on noisy fog texture (multipass.rs:345-363)."""
    truth = S.uniform_field(2.4, -1.7)
    # fog: overlapping broad blobs (synthetic.rs:104-112 semantics)
    n = int((256 + 60) ** 2 / 40)
    u = S.uniform(4, 4 * n).reshape(n, 4)
    fx = u[:, 0] * np.float32(316) - np.float32(30)
    fy = u[:, 1] * np.float32(316) - np.float32(30)
    fs = np.float32(2.0) + np.float32(4.0) * u[:, 2]
    fa_ = np.float32(0.05) + np.float32(0.1) * u[:, 3]
    bx, by = S.displace(fx, fy, truth)
    a = S.add_noise(S.render(256, 256, fx, fy, fs, fa_), 0.01, 1)
    b = S.add_noise(S.render(256, 256, bx, by, fs, fa_), 0.01, 2)
    cfg = piv.PivConfig()
    pre = piv.PreprocessConfig()
    an = piv.PivAnalyzer(cfg, pre)
    fa, fb = an.prepare(a), an.prepare(b)
    pa = oracle.prepare(a, 1.0, 4, np.float32(0.02))
    pb = oracle.prepare(b, 1.0, 4, np.float32(0.02))
    img, base = fa.download()
    assert np.array_equal(bits(img), bits(pa.image))
    assert np.array_equal(bits(base), bits(pa.base))
    res, tr = stage_parity(piv, oracle, an, fa, fb, pa, pb, cfg)
    full_parity(oracle, res, pa, pb, cfg)
    assert truth_rms(res.field, res.field.grid, truth, 20) < 0.1


def test_background_and_cli_mask(piv, oracle):
    """Background subtraction + the CLI's mask thresholds (0.04 / 0.015, analyze.rs:117-122)
    on u8 frames whose left half is dark: NO_SIGNAL decisions from the f64 SATs."""
    truth = S.uniform_field(1.0, 0.5)
    a, b = S.particle_frames(256, 256, 9, field=truth)
    a[:, :100] = 3
    b[:, :100] = 3
    bg = np.full((256, 256), np.float32(2.0 / 255.0))
    cfg = piv.PivConfig(windows=[64, 32], final_iterations=2,
                        mask=piv.MaskConfig(np.float32(0.04), np.float32(0.015)))
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    fa, fb = an.prepare(a, bg), an.prepare(b, bg)
    pa = oracle.prepare(a, background=bg)
    pb = oracle.prepare(b, background=bg)
    res, tr = stage_parity(piv, oracle, an, fa, fb, pa, pb, cfg)
    assert res.summary.no_signal > 0
    full_parity(oracle, res, pa, pb, cfg)


@pytest.mark.parametrize("npairs", [3, 9])
def test_batch_equals_single_bitwise(piv, npairs):
    """The batched path (one launch per pass for several pairs, device outputs) gives every
    pair exactly what a single-pair call gives (multipass.rs: pairs are independent).  Nine
    pairs run as two groups on two streams."""
    cfg = piv.PivConfig(windows=[64, 32, 16], final_iterations=1)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    pairs = [S.particle_frames(256, 320, 300 + i, field=S.uniform_field(1.0 + 0.4 * i, -0.5))
             for i in range(npairs)]
    A = piv.DeviceArray.from_numpy(np.stack([p[0] for p in pairs]))
    B = piv.DeviceArray.from_numpy(np.stack([p[1] for p in pairs]))
    fa = [piv.PreparedFrame.from_device(an, A.slice(i)) for i in range(npairs)]
    fb = [piv.PreparedFrame.from_device(an, B.slice(i)) for i in range(npairs)]
    g = an.output_grid(256, 320)
    dx, dy, pr = (piv.DeviceArray((npairs, g.ny, g.nx), np.float32) for _ in range(3))
    fl = piv.DeviceArray((npairs, g.ny, g.nx), np.uint8)
    sums = an.analyze_batch_device(fa, fb, dx, dy, pr, fl)
    hdx, hdy, hfl = dx.numpy(), dy.numpy(), fl.numpy()
    for i, (a, b) in enumerate(pairs):
        r = an.analyze(an.prepare(a), an.prepare(b))
        assert np.array_equal(bits(r.field.dx), bits(hdx[i]))
        assert np.array_equal(bits(r.field.dy), bits(hdy[i]))
        assert np.array_equal(r.field.flags, hfl[i])
        assert r.summary == sums[i]


@pytest.mark.parametrize("lanes", ["1", "3", "4"])
def test_batch_lanes_bitwise(piv, monkeypatch, lanes):
    """A batch split over 1, 3 or 4 stream lanes (groups of pairs whose pass chains run side by
    side) gives bitwise the default two-lane result, eager and replayed, with uneven groups."""
    cfg = piv.PivConfig(windows=[48, 24], final_iterations=1)
    h, w, n = 200, 240, 17
    pairs = [S.particle_frames(h, w, 700 + i, field=S.uniform_field(-1.5 + 0.2 * i, 0.6),
                               bits=12, container=np.uint16, amp_scale=16.0) for i in range(n)]
    A = np.stack([p[0] for p in pairs])
    B = np.stack([p[1] for p in pairs])

    def run(env):
        if env is None:
            monkeypatch.delenv("PIV_BATCH_LANES", raising=False)
        else:
            monkeypatch.setenv("PIV_BATCH_LANES", env)
        an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
        dA, dB = piv.DeviceArray.from_numpy(A), piv.DeviceArray.from_numpy(B)
        fa = [piv.PreparedFrame.from_device(an, dA.slice(i)) for i in range(n)]
        fb = [piv.PreparedFrame.from_device(an, dB.slice(i)) for i in range(n)]
        g = an.output_grid(h, w)
        dx, dy, pr = (piv.DeviceArray((n, g.ny, g.nx), np.float32) for _ in range(3))
        fl = piv.DeviceArray((n, g.ny, g.nx), np.uint8)
        outs = []
        for _ in range(3):  # eager, captured, replayed
            for d in (dx, dy, pr):
                d.upload(np.full(d.shape, np.nan, np.float32))
            sums = an.analyze_batch_device(fa, fb, dx, dy, pr, fl)
            outs.append((bits(dx.numpy()), bits(dy.numpy()), bits(pr.numpy()), fl.numpy(), sums))
        return outs

    ref = run(None)[0]
    for k, o in enumerate(run(lanes)):
        for x, y in zip(o[:4], ref[:4]):
            assert np.array_equal(x, y), f"lanes {lanes}, call {k}"
        assert o[4] == ref[4]
    monkeypatch.delenv("PIV_BATCH_LANES", raising=False)


def test_batch_lanes_api_and_autotune(piv, monkeypatch):
    """Lane counts set through the API (batch_lanes, 1..4) give bitwise the two-lane result,
    eager and replayed, on a batch whose groups are uneven; autotune() keeps a count in 1..4
    and the result does not change; counts outside 1..4 are refused."""
    monkeypatch.delenv("PIV_BATCH_LANES", raising=False)
    cfg = piv.PivConfig(windows=[32, 16], final_iterations=1)
    h, w, n = 160, 192, 18
    pairs = [S.particle_frames(h, w, 800 + i, field=S.uniform_field(1.1 - 0.1 * i, -0.4))
             for i in range(n)]
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    assert an.batch_lanes == 2
    dA = piv.DeviceArray.from_numpy(np.stack([p[0] for p in pairs]))
    dB = piv.DeviceArray.from_numpy(np.stack([p[1] for p in pairs]))
    fa = [piv.PreparedFrame.from_device(an, dA.slice(i)) for i in range(n)]
    fb = [piv.PreparedFrame.from_device(an, dB.slice(i)) for i in range(n)]
    g = an.output_grid(h, w)
    dx, dy, pr = (piv.DeviceArray((n, g.ny, g.nx), np.float32) for _ in range(3))
    fl = piv.DeviceArray((n, g.ny, g.nx), np.uint8)

    def run():
        outs = []
        for _ in range(3):  # eager, captured, replayed
            for d in (dx, dy, pr):
                d.upload(np.full(d.shape, np.nan, np.float32))
            sums = an.analyze_batch_device(fa, fb, dx, dy, pr, fl)
            outs.append((bits(dx.numpy()), bits(dy.numpy()), bits(pr.numpy()), fl.numpy(), sums))
        return outs

    ref = run()[0]
    for lanes in (1, 3, 4, 2):
        an.batch_lanes = lanes
        assert an.batch_lanes == lanes
        for k, o in enumerate(run()):
            for x, y in zip(o[:4], ref[:4]):
                assert np.array_equal(x, y), f"lanes {lanes}, call {k}"
            assert o[4] == ref[4]
    an.autotune(fa, fb, repeats=2)
    assert 1 <= an.batch_lanes <= 4
    for k, o in enumerate(run()):
        for x, y in zip(o[:4], ref[:4]):
            assert np.array_equal(x, y), f"autotuned lanes {an.batch_lanes}, call {k}"
    for bad in (0, 5):
        with pytest.raises(piv.PivError):
            an.batch_lanes = bad


def test_async_preparation_then_caller_stream(piv):
    """Device frames are prepared asynchronously on the preparing analyzer's stream (u16 frames
    with the default mask: the tables screen runs there).  An analysis on a stream the caller
    passes, issued at once, waits for that preparation and gives the analyzer-stream result
    bitwise; so do downloads and the table query issued right after preparation."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    st = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(st), 1) == 0  # non-blocking, default priority
    try:
        h, w, n = 256, 320, 9
        frames, _ = S.video_frames(h, w, n, 23, bits=12, container=np.uint16)
        cfg = piv.PivConfig(windows=[48, 24], final_iterations=1)
        clip = piv.DeviceArray.from_numpy(np.stack(frames))
        g = None
        outs = []
        for stream in (None, st.value):
            an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
            fs = piv.PreparedFrame.from_device_batch(an, clip)
            if stream is not None:
                fs[0].tables_built()
                img, _ = fs[1].download()
                assert np.array_equal(img, frames[1].astype(np.float32) / np.float32(65535))
            g = an.output_grid(h, w)
            dx, dy, pr = (piv.DeviceArray((n - 1, g.ny, g.nx), np.float32) for _ in range(3))
            fl = piv.DeviceArray((n - 1, g.ny, g.nx), np.uint8)
            sums = an.analyze_batch_device(fs[:-1], fs[1:], dx, dy, pr, fl, stream=stream)
            outs.append((bits(dx.numpy()), bits(dy.numpy()), bits(pr.numpy()), fl.numpy(), sums))
        for x, y in zip(outs[1][:4], outs[0][:4]):
            assert np.array_equal(x, y)
        assert outs[1][4] == outs[0][4]
    finally:
        hip.hipStreamDestroy(st)


def test_batch_replay_matches_eager(piv):
    """Repeated batch calls with identical frames and outputs replay one captured graph of the
    pass chain.  Every replay, with other eager work and synchronisations between replays,
    gives the eager result bitwise; a replay reads the frames' current contents; a call on
    other buffers in between falls back to eager launches and stays correct."""
    cfg = piv.PivConfig(windows=[64, 32, 16], final_iterations=1)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    other = piv.PivAnalyzer(piv.PivConfig(windows=[32], final_iterations=1),
                            piv.PreprocessConfig.raw())
    h, w, n = 192, 256, 2
    sets = [[S.particle_frames(h, w, 500 + 10 * s + i, field=S.uniform_field(0.5 + s + i, 0.75))
             for i in range(n)] for s in range(2)]
    A = piv.DeviceArray.from_numpy(np.stack([p[0] for p in sets[0]]))
    B = piv.DeviceArray.from_numpy(np.stack([p[1] for p in sets[0]]))
    fa = [piv.PreparedFrame.from_device(an, A.slice(i)) for i in range(n)]
    fb = [piv.PreparedFrame.from_device(an, B.slice(i)) for i in range(n)]
    g = an.output_grid(h, w)
    dx, dy, pr = (piv.DeviceArray((n, g.ny, g.nx), np.float32) for _ in range(3))
    fl = piv.DeviceArray((n, g.ny, g.nx), np.uint8)

    def run():
        sums = an.analyze_batch_device(fa, fb, dx, dy, pr, fl)
        return bits(dx.numpy()), bits(dy.numpy()), bits(pr.numpy()), fl.numpy(), sums

    def single(pairs):
        ref = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
        rs = [ref.analyze(ref.prepare(a), ref.prepare(b)) for a, b in pairs]
        return (np.stack([bits(r.field.dx) for r in rs]), np.stack([r.field.flags for r in rs]),
                [r.summary for r in rs])

    eager = run()
    sdx, sfl, ssum = single(sets[0])
    assert np.array_equal(eager[0], sdx) and np.array_equal(eager[3], sfl) and eager[4] == ssum
    for k in range(4):  # capture on the second call, replays after it
        for d in (dx, dy, pr):
            d.upload(np.full(d.shape, np.nan, np.float32))
        fl.upload(np.full(fl.shape, 0xEE, np.uint8))
        o = other.analyze(other.prepare(sets[1][0][0]), other.prepare(sets[1][0][1]))
        assert np.isfinite(o.field.dx).any()
        r = run()
        for x, y in zip(r[:4], eager[:4]):
            assert np.array_equal(x, y), f"replay {k}"
        assert r[4] == eager[4]
    # new frame contents in the same buffers: the replay reads them
    A.upload(np.stack([p[0] for p in sets[1]]))
    B.upload(np.stack([p[1] for p in sets[1]]))
    tdx, tfl, tsum = single(sets[1])
    for k in range(2):
        r = run()
        assert np.array_equal(r[0], tdx) and np.array_equal(r[3], tfl) and r[4] == tsum, k
    # a single-pair call on the same analyzer in between, then the batch again
    s1 = an.analyze(an.prepare(sets[1][1][0]), an.prepare(sets[1][1][1]))
    assert np.array_equal(bits(s1.field.dx), tdx[1])
    for k in range(3):
        r = run()
        assert np.array_equal(r[0], tdx) and np.array_equal(r[3], tfl) and r[4] == tsum, k


@pytest.mark.parametrize("fmt", [np.uint8, np.uint16, np.float32])
def test_kernel_variants_bit_identical(piv, fmt):
    """Every launch variant of every window size gives the default variant's field bitwise;
    the autotuner then picks one of them and the result is unchanged."""
    h, w = 208, 224
    for n in (8, 12, 16, 24, 32, 48, 64, 96):
        a, b = S.particle_frames(h, w, 40 + n, field=S.uniform_field(1.3, -0.6))
        if fmt is np.uint16:
            a, b = a.astype(np.uint16) * 16, b.astype(np.uint16) * 16
        elif fmt is np.float32:
            a, b = a.astype(np.float32) / 255, b.astype(np.float32) / 255
        cfg = piv.PivConfig(windows=[n, n], final_iterations=1)
        an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
        fa, fb = an.prepare(a), an.prepare(b)
        ref = an.analyze(fa, fb)
        assert np.isfinite(ref.field.dx).mean() > 0.5, n
        for v in range(1, piv.kernel_variant_count(n)):
            an.set_variant(n, fmt, v)
            r = an.analyze(fa, fb)
            assert np.array_equal(bits(r.field.dx), bits(ref.field.dx)), (n, v)
            assert np.array_equal(bits(r.field.dy), bits(ref.field.dy)), (n, v)
            assert np.array_equal(r.field.flags, ref.field.flags), (n, v)
        tune = an.autotune([fa], [fb], repeats=2)
        assert [t["window"] for t in tune] == [n]
        assert 0 <= tune[0]["chosen"] < piv.kernel_variant_count(n)
        assert all(m > 0 for m in tune[0]["ms"])
        r = an.analyze(fa, fb)
        assert np.array_equal(bits(r.field.dx), bits(ref.field.dx))


@pytest.mark.parametrize("dtype,pre,n", [(np.uint16, "raw", 5), (np.uint16, "raw", 12),
                                         (np.uint8, "raw", 5), (np.uint8, "default", 5),
                                         (np.uint16, "default", 9), (np.uint8, "default", 40),
                                         (np.float32, "raw", 9)])
def test_batch_prepared_frames_equal_single(piv, dtype, pre, n):
    """Frames prepared as one batch (one launch per kernel, per chunk of 32 frames when
    conditioning; images and tables in storage the frames share) give the single-frame
    preparation's analysis and images bitwise, masks included."""
    h, w = 160, 224
    frames, _ = S.video_frames(h, w, n, 17, bits=12 if dtype is np.uint16 else 8,
                               container=np.uint8 if dtype is np.uint8 else np.uint16)
    if dtype is np.float32:
        frames = [f.astype(np.float32) / np.float32(4095) for f in frames]
    prep = piv.PreprocessConfig.raw() if pre == "raw" else piv.PreprocessConfig()
    # a mask near the frame's own mean and spread rejects about half of the windows, so the
    # tables decide outcomes (u8 frames keep the default mask, which needs no tables)
    mask = piv.MaskConfig()
    if dtype is not np.uint8:
        norm = frames[0].astype(np.float32) / (np.float32(65535) if dtype is np.uint16 else 1)
        mask = piv.MaskConfig(float(np.float32(norm.mean())), float(np.float32(0.9 * norm.std())))
    cfg = piv.PivConfig(windows=[48, 24], final_iterations=1, mask=mask)
    an = piv.PivAnalyzer(cfg, prep)
    stack = piv.DeviceArray.from_numpy(np.stack(frames))
    fs = piv.PreparedFrame.from_device_batch(an, stack)
    assert len(fs) == n
    for k in range(n - 1):
        r = an.analyze(fs[k], fs[k + 1])
        ref = an.analyze(an.prepare(frames[k]), an.prepare(frames[k + 1]))
        assert np.array_equal(bits(r.field.dx), bits(ref.field.dx)), k
        assert np.array_equal(r.field.flags, ref.field.flags), k
        assert r.summary == ref.summary
        if dtype is not np.uint8:
            assert 0 < r.summary.no_signal < r.field.flags.size, r.summary
    for k in sorted({2, n - 2}):
        img, base = fs[k].download()
        rimg, rbase = an.prepare(frames[k]).download()
        assert np.array_equal(bits(img), bits(rimg)), k
        assert np.array_equal(bits(base), bits(rbase)), k


def test_deterministic(piv):
    a, b = S.particle_frames(512, 512, 5, field=S.uniform_field(3.3, 1.1))
    an = piv.PivAnalyzer(piv.PivConfig(windows=[64, 32], final_iterations=2),
                         piv.PreprocessConfig.raw())
    fa, fb = an.prepare(a), an.prepare(b)
    r1, r2 = an.analyze(fa, fb), an.analyze(fa, fb)
    assert np.array_equal(bits(r1.field.dx), bits(r2.field.dx))
    assert np.array_equal(r1.field.flags, r2.field.flags)


def test_window_as_large_as_the_image(piv, oracle):
    """multipass.rs:365-376"""
    a, b = S.particle_pair(64, 64, 3, 0.04, 1.2, S.uniform_field(1.5, -0.5))
    for windows in ([64], [64, 64]):
        cfg = piv.PivConfig(windows=windows, final_iterations=1)
        an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
        r = an.analyze(an.prepare(a), an.prepare(b))
        assert r.field.dx.shape == (1, 1) and np.isfinite(r.field.dx[0, 0])
        assert abs(r.field.dx[0, 0] - 1.5) < 0.2 and abs(r.field.dy[0, 0] + 0.5) < 0.2
        full_parity(oracle, r, oracle.prepare(a), oracle.prepare(b), cfg)


def test_blank_regions_get_no_signal(piv, oracle):
    """multipass.rs:445-475 on float frames with a dark left half."""
    a, b = S.particle_pair(128, 128, 9, 0.04, 1.2, S.uniform_field(1.0, 0.0))
    a[:, :64] = 0
    b[:, :64] = 0
    a, b = S.add_noise(a, 0.002, 3), S.add_noise(b, 0.002, 4)
    cfg = piv.PivConfig(mask=piv.MaskConfig(0.01, 0.01))
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    r = an.analyze(an.prepare(a), an.prepare(b))
    xs = r.field.grid.xs()
    for ix, x in enumerate(xs):
        if x + 16 < 64:
            assert np.all(r.field.flags[:, ix] == 1) and np.all(np.isnan(r.field.dx[:, ix]))
        elif x - 16 >= 64:
            assert np.all((r.field.flags[:, ix] & 1) == 0)
    full_parity(oracle, r, oracle.prepare(a), oracle.prepare(b), cfg)


def test_errors(piv):
    an = piv.PivAnalyzer(piv.PivConfig(windows=[32], final_iterations=1),
                         piv.PreprocessConfig.raw())
    a = np.zeros((64, 64), np.uint8)
    b = np.zeros((64, 80), np.uint8)
    with pytest.raises(piv.ShapeMismatch):
        an.analyze(an.prepare(a), an.prepare(b))
    small = an.prepare(np.zeros((16, 64), np.uint8))
    with pytest.raises(piv.ImageTooSmall):
        an.analyze(small, small)
    with pytest.raises(piv.Unsupported):
        piv.PivAnalyzer(piv.PivConfig(windows=[130], final_iterations=1))


def test_background_estimator(piv):
    """preprocess.rs:258-270 known answer, then min / mean over u8 and float frames against
    the same f32 operation sequence in numpy (bit-exact)."""
    an = piv.PivAnalyzer(piv.PivConfig(windows=[32], final_iterations=1),
                         piv.PreprocessConfig.raw())
    for mode, want in (("min", [[1.0, 2.0]]), ("mean", [[2.0, 3.0]])):
        e = piv.BackgroundEstimator(an, mode, 1, 2)
        e.add(np.array([[1.0, 4.0]], np.float32))
        e.add(np.array([[3.0, 2.0]], np.float32))
        assert np.array_equal(e.finish(), np.array(want, np.float32))
    rng = np.random.default_rng(3)
    frames = [rng.integers(0, 256, (37, 53), dtype=np.uint8) for _ in range(5)]
    norm = [f.astype(np.float32) / np.float32(255) for f in frames]
    e = piv.BackgroundEstimator(an, "min", 37, 53)
    for f in frames:
        e.add(f)
    assert np.array_equal(bits(e.finish()), bits(np.minimum.reduce(norm)))
    e = piv.BackgroundEstimator(an, "mean", 37, 53)
    acc = norm[0].copy()
    for f in frames:
        e.add(f)
    for f in norm[1:]:
        acc = (acc + f).astype(np.float32)
    assert np.array_equal(bits(e.finish()), bits(acc / np.float32(len(frames))))

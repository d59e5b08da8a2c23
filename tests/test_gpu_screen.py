"""Screened signal-mask tables (u16 frames without conditioning, default mask min_mean =
min_std = 0).

With that mask only windows whose box is all zero can depend on a u16 frame's f64 tables
(IntegralImage, preprocess.rs:84-124; the mask test, multipass.rs:233-241): any other box sums
to at least 1/65535, more than the tables' rounding error, so its mean is positive.  The
preparation therefore screens each frame for aligned all-zero tiles of half the analyzer's
smallest window (any window holds such a tile) and builds the tables only where it finds
one.  Checked here: which frames get tables, re-screening when a frame is analysed with
smaller windows than it was prepared for, and, on frames with all-zero regions (origins
reaching into and out of them on every pass), the single-pair, batched and served paths
against the oracle, whose tables are always built, and bitwise against the same analysis
with screening switched off (PIV_NO_TABLE_SCREEN=1)."""
import numpy as np
import pytest

from piv_amd import synthetic as S
from test_gpu_configs import bits, full_parity, stage_parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def piv():
    import piv_amd

    if piv_amd.device_count() < 1:
        pytest.skip("no GPU")
    return piv_amd


def _pair(h, w, seed, holes_a=(), holes_b=(), shift=(1.37, -0.62)):
    a, b = S.particle_frames(h, w, seed, field=S.uniform_field(*shift), bits=12,
                             container=np.uint16, amp_scale=16.0)
    for y0, y1, x0, x1 in holes_a:
        a[y0:y1, x0:x1] = 0
    for y0, y1, x0, x1 in holes_b:
        b[y0:y1, x0:x1] = 0
    return a, b


HOLES_A = [(60, 150, 40, 170), (180, 197, 200, 215)]
HOLES_B = [(100, 200, 150, 290)]


def _same(r, ref):
    assert np.array_equal(bits(r.field.dx), bits(ref.field.dx))
    assert np.array_equal(bits(r.field.dy), bits(ref.field.dy))
    assert np.array_equal(bits(r.field.peak_ratio), bits(ref.field.peak_ratio))
    assert np.array_equal(r.field.flags, ref.field.flags)
    assert r.summary == ref.summary


def _eager(piv, monkeypatch, cfg, a, b):
    """The analysis with the tables always built."""
    monkeypatch.setenv("PIV_NO_TABLE_SCREEN", "1")
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    fa, fb = an.prepare(a), an.prepare(b)
    monkeypatch.delenv("PIV_NO_TABLE_SCREEN")
    assert fa.tables_built() and fb.tables_built()
    return an.analyze(fa, fb)


def test_screen_decides_which_frames_get_tables(piv, monkeypatch):
    """Windows 32 -> 16 screen with 8 x 8 tiles: particle frames (noise leaves no all-zero
    tile) get no tables; any aligned all-zero 8 x 8 tile, alone or inside a larger unaligned
    zero patch, gets them; a 14 x 14 zero patch that holds no aligned tile does not."""
    an = piv.PivAnalyzer(piv.PivConfig(windows=[32, 16], final_iterations=1),
                         piv.PreprocessConfig.raw())
    a, _ = _pair(200, 264, 3)
    assert not an.prepare(a).tables_built()
    z = a.copy()
    z[40:48, 64:72] = 0
    assert an.prepare(z).tables_built()
    z = a.copy()
    z[33:48, 57:72] = 0  # holds rows 40-47 x columns 64-71
    assert an.prepare(z).tables_built()
    z = a.copy()
    z[41:55, 57:71] = 0  # rows 41-54, columns 57-70: no aligned 8-row x 8-column tile
    assert not an.prepare(z).tables_built()
    # a frame whose size is not a multiple of 8: its last full tile row and column count,
    # the partial ones past them do not
    e, _ = _pair(203, 270, 4)
    assert not an.prepare(e).tables_built()
    z = e.copy()
    z[192:200, 256:264] = 0
    assert an.prepare(z).tables_built()
    z = e.copy()
    z[200:203, :] = 0
    z[:, 264:270] = 0
    assert not an.prepare(z).tables_built()
    # masks with a positive threshold read every window's statistics: always built
    strict = piv.PivAnalyzer(piv.PivConfig(windows=[32, 16], final_iterations=1,
                                           mask=piv.MaskConfig(0.001, 0.0)),
                             piv.PreprocessConfig.raw())
    assert strict.prepare(a).tables_built()
    # u8 frames without background need no tables at all
    a8, _ = S.particle_frames(200, 264, 3, field=S.uniform_field(1.0, 0.5))
    assert not an.prepare(a8).tables_built()
    # screening switched off
    monkeypatch.setenv("PIV_NO_TABLE_SCREEN", "1")
    assert an.prepare(a).tables_built()


@pytest.mark.parametrize("windows,t", [([16, 8], 4), ([24, 12], 4), ([48, 24], 8), ([32], 16),
                                       ([64], 32), ([128, 64], 32)])
@pytest.mark.parametrize("w", [512, 520, 517])
def test_screen_tile_sizes(piv, windows, t, w):
    """Tile t (the largest power of two <= n / 2 of the smallest window n) on frames whose
    width is a multiple of 8 samples (16 B loads) or not: an aligned all-zero t x t tile gets
    tables, the same patch moved by one pixel (no aligned t x t tile inside it) does not, at
    the frame's last full tile column too."""
    an = piv.PivAnalyzer(piv.PivConfig(windows=windows, final_iterations=1),
                         piv.PreprocessConfig.raw())
    h = 136
    a, _ = _pair(h, w, 7)
    assert not an.prepare(a).tables_built()
    last = (w // t - 1) * t
    for x0 in (t, 3 * t, last):
        z = a.copy()
        z[2 * t:3 * t, x0:x0 + t] = 0
        assert an.prepare(z).tables_built(), (x0, "aligned")
        z = a.copy()
        z[2 * t + 1:3 * t + 1, x0 + 1:x0 + t + 1] = 0
        assert not an.prepare(z).tables_built(), (x0, "shifted")


@pytest.mark.parametrize("windows", [[32, 16], [48, 24], [64, 32, 16]])
def test_blank_regions_against_oracle(piv, oracle, monkeypatch, windows):
    """Frames with all-zero regions: every pass replayed through the oracle (integer peaks,
    NO_SIGNAL decisions and validation identical), the final field within 1e-3 px with
    identical flags, and the whole result bitwise equal to the analysis with the tables
    always built."""
    cfg = piv.PivConfig(windows=windows, final_iterations=1)
    a, b = _pair(256, 320, 40, HOLES_A, HOLES_B)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    fa, fb = an.prepare(a), an.prepare(b)
    assert fa.tables_built() and fb.tables_built()
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    res, traces = stage_parity(piv, oracle, an, fa, fb, pa, pb, cfg)
    assert any((t.raw.flags & 1).any() for t in traces)  # the zero regions give NO_SIGNAL
    full_parity(oracle, res, pa, pb, cfg)
    _same(res, _eager(piv, monkeypatch, cfg, a, b))


def test_rescreen_for_smaller_windows(piv, oracle, monkeypatch):
    """Frames prepared for windows 64 (screened with 32 x 32 tiles) whose only zero patch holds
    aligned 8 x 8 tiles but no aligned 32 x 32 one: no tables at first; analysed with windows
    32 -> 16 they are re-screened with 8 x 8 tiles and get them, with the results of the
    tables built from the start."""
    holes = [(40, 60, 64, 84)]
    a, b = _pair(224, 288, 41, holes, holes)
    big = piv.PivAnalyzer(piv.PivConfig(windows=[64], final_iterations=1),
                          piv.PreprocessConfig.raw())
    fa, fb = big.prepare(a), big.prepare(b)
    assert not fa.tables_built() and not fb.tables_built()
    cfg = piv.PivConfig(windows=[32, 16], final_iterations=1)
    small = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    r = small.analyze(fa, fb)
    assert fa.tables_built() and fb.tables_built()
    _same(r, _eager(piv, monkeypatch, cfg, a, b))
    full_parity(oracle, r, oracle.prepare(a), oracle.prepare(b), cfg)


def test_batch_with_some_blank_frames(piv, monkeypatch):
    """A batch prepared in one call (frames of one clip, each frame the B of one pair and the
    A of the next) where only some frames hold zero regions: per-frame table decisions, and
    every pair bitwise equal to the pair analysed alone with the tables always built."""
    h, w, n = 192, 256, 7
    cfg = piv.PivConfig(windows=[48, 24], final_iterations=1)
    frames = []
    for i in range(n + 1):
        f, _ = _pair(h, w, 900 + i)
        if i in (2, 3, 6):
            f[30 + 10 * i:90 + 10 * i, 20:120] = 0
        frames.append(f)
    clip = np.stack(frames)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    dev = piv.DeviceArray.from_numpy(clip)
    fr = piv.PreparedFrame.from_device_batch(an, dev)
    assert [f.tables_built() for f in fr] == [i in (2, 3, 6) for i in range(n + 1)]
    g = an.output_grid(h, w)
    dx, dy, pr = (piv.DeviceArray((n, g.ny, g.nx), np.float32) for _ in range(3))
    fl = piv.DeviceArray((n, g.ny, g.nx), np.uint8)
    sums = an.analyze_batch_device(fr[:-1], fr[1:], dx, dy, pr, fl)
    hdx, hdy, hfl = dx.numpy(), dy.numpy(), fl.numpy()
    for i in range(n):
        ref = _eager(piv, monkeypatch, cfg, frames[i], frames[i + 1])
        assert np.array_equal(bits(hdx[i]), bits(ref.field.dx)), i
        assert np.array_equal(bits(hdy[i]), bits(ref.field.dy)), i
        assert np.array_equal(hfl[i], ref.field.flags), i
        assert sums[i] == ref.summary


def test_served_blank_frames(piv, oracle, monkeypatch):
    """Served u16 pairs (frames prepared together per stream, at their own addresses), some
    with zero regions: each result equals the oracle's analysis and, bitwise, the pair analysed
    alone with the tables always built."""
    cfg = piv.PivConfig(windows=[48, 24], final_iterations=1)
    srv = piv.PivServer(max_batch_pairs=8, max_inflight=1)
    sid = srv.add_stream(cfg, 256, 320, np.uint16, piv.PreprocessConfig.raw())
    pairs = {}
    srv.pause(True)
    for k in range(6):
        holes = (HOLES_A, HOLES_B) if k % 2 else ((), ())
        a, b = _pair(256, 320, 950 + k, *holes)
        pairs[srv.submit(sid, a, b)] = (a, b)
    srv.pause(False)
    got = {}
    while len(got) < len(pairs):
        res = srv.poll(8, wait_ms=10000)
        assert res, "no result within 10 s"
        for r in res:
            assert r.error is None, r.error
            got[r.request_id] = r
    srv.close()
    for rid, (a, b) in pairs.items():
        _same(got[rid], _eager(piv, monkeypatch, cfg, a, b))
        full_parity(oracle, got[rid], oracle.prepare(a), oracle.prepare(b), cfg)

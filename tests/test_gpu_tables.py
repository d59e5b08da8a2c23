"""Summed-area tables on the device (preprocess.rs:84-124): the signal mask's window statistics
come out of the device tables exactly as out of the oracle's f64 tables.

Mask thresholds are set to window means and standard deviations the oracle computes, so a
window exactly at a threshold passes only if the device computed the identical f32 value: a
one-ulp difference in any table entry a window reads flips a decision.  Frame shapes cover
partial 32-column strips and 32-row bands, frames no wider than one strip and frames an exact
multiple of both; the single-frame, batched and served (grouped) preparations all run, and
frames prepared in memory that released frames left in the analyzer's pool."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(37, 29), (31, 31), (33, 95), (70, 33), (64, 64), (97, 130), (129, 257)]
N = 8  # window size: the tables, not the window, are under test
NO_SIGNAL = 1


@pytest.fixture(scope="module")
def piv():
    import piv_amd

    if piv_amd.device_count() < 1:
        pytest.skip("no GPU")
    return piv_amd


def _frames(h, w, dtype, seed, count=2):
    rng = np.random.default_rng(seed)
    if dtype is np.uint16:
        return [rng.integers(0, 4096, (h, w)).astype(np.uint16) for _ in range(count)]
    # f32 samples of mixed magnitude: the f64 sums are inexact and their order matters
    return [(rng.random((h, w)) ** 3 * 4.0).astype(np.float32) for _ in range(count)]


def _origins(h, w, seed, count=300):
    rng = np.random.default_rng(seed)
    o = np.empty((count, 4), np.float32)
    o[:, 0] = rng.integers(0, w - N + 1, count)
    o[:, 1] = rng.integers(0, h - N + 1, count)
    o[:, 2] = rng.integers(0, w - N + 1, count)
    o[:, 3] = rng.integers(0, h - N + 1, count)
    return o


def _oracle_stats(oracle, frame, origins, cols):
    p = oracle.prepare(frame)
    w = frame.shape[1]
    return np.array([oracle.sat_stats(p.sat, w, int(x), int(y), N, N)
                     for x, y in origins[:, cols]], np.float32)


def _thresholds(sa, sb):
    """(min_mean, min_std) pairs at window statistics: quantiles of the means (std test
    open), then of the standard deviations (mean test open)."""
    means = np.sort(np.concatenate([sa[:, 0], sb[:, 0]]))
    stds = np.sort(np.concatenate([sa[:, 1], sb[:, 1]]))
    out = []
    for q in (0.1, 0.5, 0.9):
        out.append((float(means[int(q * (means.size - 1))]), 0.0))
        out.append((-1e30, float(stds[int(q * (stds.size - 1))])))
    return out


def _expected(sa, sb, mm, ms):
    ok_a = (sa[:, 0] >= np.float32(mm)) & (sa[:, 1] >= np.float32(ms))
    ok_b = (sb[:, 0] >= np.float32(mm)) & (sb[:, 1] >= np.float32(ms))
    return ~(ok_a & ok_b)


@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
@pytest.mark.parametrize("shape", SHAPES)
def test_single_frame_tables_decide_the_mask_like_the_oracle(piv, oracle, shape, dtype):
    h, w = shape
    fa, fb = _frames(h, w, dtype, 11 * h + w)
    o = _origins(h, w, 7 * h + w)
    sa = _oracle_stats(oracle, fa, o, [0, 1])
    sb = _oracle_stats(oracle, fb, o, [2, 3])
    for mm, ms in _thresholds(sa, sb):
        cfg = piv.PivConfig(windows=[N], final_iterations=1, mask=piv.MaskConfig(mm, ms))
        an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
        *_, fl, _ = piv.correlate_windows(an, an.prepare(fa), an.prepare(fb), N, o)
        want = _expected(sa, sb, mm, ms)
        assert 0 < want.sum() < want.size
        assert np.array_equal((fl & NO_SIGNAL) != 0, want), (shape, mm, ms)


@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
@pytest.mark.parametrize("shape", [(37, 29), (97, 130), (129, 257)])
def test_batch_tables_decide_the_mask_like_the_oracle(piv, oracle, shape, dtype):
    """Frames prepared as one batch (the table kernels over every frame in one launch)."""
    h, w = shape
    frames = _frames(h, w, dtype, 5 * h + w, count=9)
    o = _origins(h, w, 3 * h + w)
    stack = piv.DeviceArray.from_numpy(np.stack(frames))
    for k in (0, 4, 7):
        sa = _oracle_stats(oracle, frames[k], o, [0, 1])
        sb = _oracle_stats(oracle, frames[k + 1], o, [2, 3])
        mm, ms = _thresholds(sa, sb)[1]
        cfg = piv.PivConfig(windows=[N], final_iterations=1, mask=piv.MaskConfig(mm, ms))
        an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
        fs = piv.PreparedFrame.from_device_batch(an, stack)
        *_, fl, _ = piv.correlate_windows(an, fs[k], fs[k + 1], N, o)
        assert np.array_equal((fl & NO_SIGNAL) != 0, _expected(sa, sb, mm, ms)), (shape, k)


def test_served_pair_tables_match_single_frames(piv, oracle):
    """The requests a server prepares together get their tables from one launch per kernel:
    with a mask at the median window mean, every served field equals the pair analysed alone,
    and the mask splits the windows."""
    h, w = 97, 130
    frames = _frames(h, w, np.uint16, 99, count=6)
    norm = np.stack(frames).astype(np.float32) / np.float32(65535)
    cfg = piv.PivConfig(windows=[24], final_iterations=1,
                        mask=piv.MaskConfig(float(np.float32(np.median(norm))), 0.0))
    raw = piv.PreprocessConfig.raw()
    srv = piv.PivServer(max_batch_pairs=4)
    sid = srv.add_stream(cfg, h, w, np.uint16, raw)
    ids = {srv.submit(sid, frames[2 * i], frames[2 * i + 1]): i for i in range(3)}
    got = {}
    while len(got) < len(ids):
        for r in srv.poll(4, wait_ms=10000):
            assert r.error is None, r.error
            got[r.request_id] = r
    srv.close()
    for rid, i in ids.items():
        an = piv.PivAnalyzer(cfg, raw)
        ref = an.analyze(an.prepare(frames[2 * i]), an.prepare(frames[2 * i + 1]))
        f = got[rid].field
        assert np.array_equal(f.flags, ref.field.flags)
        assert np.array_equal(np.ascontiguousarray(f.dx).view(np.uint32),
                              np.ascontiguousarray(ref.field.dx).view(np.uint32))
        masked = (ref.field.flags & NO_SIGNAL) != 0
        assert 0 < masked.sum() < masked.size


@pytest.mark.parametrize("pre", ["raw", "default"])
def test_reused_frame_memory_gives_fresh_results(piv, pre):
    """Frames released to the analyzer's pool leave nothing behind: frames prepared afterwards
    in the reused memory (single and batched, tables and conditioned images) analyse bitwise
    as with a fresh analyzer."""
    h, w = 97, 130
    first = _frames(h, w, np.uint16, 1, count=6)
    second = _frames(h, w, np.uint16, 2, count=6)
    norm = np.stack(second).astype(np.float32) / np.float32(65535)
    cfg = piv.PivConfig(windows=[24], final_iterations=1,
                        mask=piv.MaskConfig(float(np.float32(np.median(norm))), 0.0))
    prep = piv.PreprocessConfig.raw() if pre == "raw" else piv.PreprocessConfig()
    an = piv.PivAnalyzer(cfg, prep)
    # fill the pool: single frames and a batch, analysed, then released
    old = [an.prepare(f) for f in first]
    an.analyze(old[0], old[1])
    del old
    stack1 = piv.DeviceArray.from_numpy(np.stack(first))
    old = piv.PreparedFrame.from_device_batch(an, stack1)
    an.analyze(old[2], old[3])
    del old
    fresh = piv.PivAnalyzer(cfg, prep)
    stack2 = piv.DeviceArray.from_numpy(np.stack(second))
    got_single = [an.prepare(f) for f in second[:2]]
    got_batch = piv.PreparedFrame.from_device_batch(an, stack2)
    ref = fresh.analyze(fresh.prepare(second[0]), fresh.prepare(second[1]))
    for r in (an.analyze(got_single[0], got_single[1]), an.analyze(got_batch[0], got_batch[1])):
        assert np.array_equal(r.field.flags, ref.field.flags)
        assert np.array_equal(np.ascontiguousarray(r.field.dx).view(np.uint32),
                              np.ascontiguousarray(ref.field.dx).view(np.uint32))
    masked = (ref.field.flags & NO_SIGNAL) != 0
    assert 0 < masked.sum() < masked.size

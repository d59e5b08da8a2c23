"""piv-analyze (the `piv analyze` host loop over the C ABI) on a synthetic clip: frame range,
stride and pair step pairing, background sampled across the whole clip, ROI, NPZ per pair and
run.json, each pair's field equal (bitwise) to the library analysing the same pair alone
(piv-cli/src/analyze.rs:243-523, piv-io/src/write.rs:113-165)."""
import json
import os
import subprocess

import numpy as np
import pytest

import piv_amd.synthetic as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "piv_amd", "bin", "piv-analyze")


@pytest.fixture(scope="module")
def piv():
    import piv_amd

    if piv_amd.device_count() < 1:
        pytest.skip("no GPU")
    if not os.path.exists(CLI):
        pytest.fail("piv-analyze is not built (make -C piv_amd)")
    return piv_amd


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def run_step(d):
    """grid step of a pair archive from its x coordinates"""
    x = d["x"]
    return int(round(float(x[1] - x[0]))) if x.size > 1 else 1


def _run(args):
    p = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout + p.stderr


@pytest.mark.parametrize("dtype,bits_", [(np.uint8, 8), (np.uint16, 12)])
def test_cli_pairs_background_roi_npz(piv, tmp_path, dtype, bits_):
    frames, _ = S.video_frames(224, 320, 9, 23, bits=bits_, container=dtype)
    clip = np.stack(frames)
    src = tmp_path / "clip.npy"
    np.save(src, clip)
    out = tmp_path / "out"
    roi = (16, 8, 288, 208)
    fps, stride, step, start = 200.0, 2, 3, 1
    rc, log = _run([str(src), "-o", str(out), "--start", str(start), "--stride", str(stride),
                    "--pair-step", str(step), "--capture-fps", str(fps),
                    "--roi", ",".join(map(str, roi)), "--background", "min:4",
                    "--windows", "64,32", "--final-iterations", "1", "--min-mean", "0",
                    "--min-std", "0", "--px-per-mm", "12.5", "--batch", "2"])
    assert rc == 0, log
    # pairs (a, a + stride) for a in range(start, end - stride, pair_step)
    starts = list(range(start, len(frames) - stride, step))
    names = sorted(f for f in os.listdir(out) if f.startswith("pair_") and f.endswith(".npz"))
    assert names == [f"pair_{a:06d}_{a + stride:06d}.npz" for a in starts]
    run = json.loads((out / "run.json").read_text())
    assert run["stride"] == stride and run["pair_step"] == step and run["start_frame"] == start
    assert run["end_frame"] == len(frames) and run["background"] == "min:4"
    assert run["piv"]["windows"] == [64, 32] and abs(run["dt_s"] - stride / fps) < 1e-15
    # the same analysis through the library: background of 4 frames sampled across the clip
    x, y, w, h = roi
    crop = [f[y:y + h, x:x + w] for f in frames]
    cfg = piv.PivConfig(windows=[64, 32], final_iterations=1, mask=piv.MaskConfig(0.0, 0.0))
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig())
    est = piv.BackgroundEstimator(an, "min", h, w)
    for k in range(4):
        est.add(crop[k * (len(frames) - 1) // 3])
    bg = est.finish()
    for a in starts:
        d = np.load(out / f"pair_{a:06d}_{a + stride:06d}.npz")
        r = an.analyze(an.prepare(crop[a], bg), an.prepare(crop[a + stride], bg))
        assert np.array_equal(bits(d["dx"]), bits(r.field.dx)), a
        assert np.array_equal(bits(d["dy"]), bits(r.field.dy)), a
        assert np.array_equal(d["flags"], r.field.flags), a
        assert np.array_equal(bits(d["peak_ratio"]), bits(r.field.peak_ratio)), a
        inv = np.float32(1.0 / (stride / fps))
        assert np.array_equal(bits(d["u"]), bits(r.field.dx * inv))
        assert np.array_equal(bits(d["v_mps"]), bits((r.field.dy * inv) * np.float32(1e-3 / 12.5)))
        assert np.array_equal(d["x"], np.array(r.field.grid.xs(), np.float32) + np.float32(x))
        assert list(d["frames"]) == [a, a + stride]
        assert float(d["t"]) == (a + stride / 2.0) / fps
        assert ((r.field.flags & 23) == 0).mean() > 0.5
    # stats.npz (analyze.rs:483-489, 549): pair_stats of every pair's field, in pair order
    from oracle import oracle as O
    st = np.load(out / "stats.npz")
    assert st["t"].shape == (len(starts),)
    for i, a in enumerate(starts):
        d = np.load(out / f"pair_{a:06d}_{a + stride:06d}.npz")
        assert float(st["t"][i]) == float(d["t"])
        ok = O.flow_usable(d["dx"], d["flags"])
        want = O.pair_stats(d["dx"], d["dy"], ok, run_step(d), stride / fps)
        for k in ("vorticity_mean", "vorticity_abs_mean", "ke_mean", "ke_sum", "coverage"):
            assert st[k][i] == pytest.approx(want[k], rel=1e-6, nan_ok=True), (a, k)


def test_cli_green_plane_of_rgb_clip(piv, tmp_path):
    """decode.rs:124-133: --channel green of an RGB(A) clip analyses the green plane, exactly
    as the one-plane clip of those samples."""
    frames, _ = S.video_frames(128, 160, 5, 31, bits=8, container=np.uint8)
    gray = np.stack(frames)
    rng = np.random.default_rng(2)
    for c in (3, 4):
        rgb = rng.integers(0, 256, size=gray.shape + (c,), dtype=np.uint8)
        rgb[..., 1] = gray
        np.save(tmp_path / f"rgb{c}.npy", rgb)
    np.save(tmp_path / "gray.npy", gray)
    common = ["--windows", "32", "--final-iterations", "1", "--min-mean", "0", "--min-std", "0",
              "--background", "none"]
    rc, log = _run([str(tmp_path / "gray.npy"), "-o", str(tmp_path / "g")] + common)
    assert rc == 0, log
    for c in (3, 4):
        out = tmp_path / f"c{c}"
        rc, log = _run([str(tmp_path / f"rgb{c}.npy"), "-o", str(out), "--channel", "green"] + common)
        assert rc == 0, log
        assert json.loads((out / "run.json").read_text())["channel"] == "green"
        for a in range(len(frames) - 1):
            name = f"pair_{a:06d}_{a + 1:06d}.npz"
            d, e = np.load(out / name), np.load(tmp_path / "g" / name)
            for k in ("dx", "dy", "peak_ratio"):
                assert np.array_equal(bits(d[k]), bits(e[k])), (c, a, k)
            assert np.array_equal(d["flags"], e["flags"])
    rc, log = _run([str(tmp_path / "gray.npy"), "-o", str(tmp_path / "x"), "--channel", "green"])
    assert rc != 0 and "colour clip" in log
    rc, log = _run([str(tmp_path / "rgb3.npy"), "-o", str(tmp_path / "y"), "--channel", "luma"])
    assert rc != 0 and "ffmpeg" in log


def test_cli_errors(piv, tmp_path):
    clip = np.zeros((3, 64, 64), np.uint8)
    src = tmp_path / "c.npy"
    np.save(src, clip)
    rc, log = _run([str(src), "-o", str(tmp_path / "o"), "--stride", "3"])
    assert rc != 0 and "too short" in log
    (tmp_path / "exists").mkdir()
    rc, log = _run([str(src), "-o", str(tmp_path / "exists")])
    assert rc != 0 and "--force" in log
    rc, log = _run([str(src), "-o", str(tmp_path / "o2"), "--format", "hdf5"])
    assert rc != 0 and "HDF5" in log


def test_cli_streamed_upload_multi_piece_clip(piv, tmp_path):
    """A one-plane clip of several 32 MiB upload pieces and no background: the loop starts
    while the rest of the clip is still uploading and each batch waits only for its own
    frames.  Pairs spread over the whole clip (the last ones in the last piece) equal the
    library analysing the same pair alone, bitwise."""
    h, w, n = 1024, 1536, 36  # 3 MiB per u16 frame: 108 MiB, four pieces
    base, _ = S.video_frames(256, 384, 6, 29, bits=12, container=np.uint16)
    clip = np.stack([np.tile(base[k % len(base)], (4, 4)) for k in range(n)])
    assert clip.shape == (n, h, w) and clip.nbytes > 3 * (32 << 20)
    src = tmp_path / "clip.npy"
    np.save(src, clip)
    out = tmp_path / "out"
    rc, log = _run([str(src), "-o", str(out), "--stride", "1", "--pair-step", "5",
                    "--capture-fps", "100", "--background", "none", "--windows", "64",
                    "--final-iterations", "1", "--min-mean", "0", "--min-std", "0",
                    "--smooth-sigma", "0", "--normalize-radius", "0", "--batch", "2"])
    assert rc == 0, log
    assert "waiting for the upload" in log and "set-up:" in log
    cfg = piv.PivConfig(windows=[64], final_iterations=1, mask=piv.MaskConfig(0.0, 0.0))
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    starts = list(range(0, n - 1, 5))
    assert starts[-1] == 30
    for a in (0, 15, 30):
        d = np.load(out / f"pair_{a:06d}_{a + 1:06d}.npz")
        r = an.analyze(an.prepare(clip[a]), an.prepare(clip[a + 1]))
        assert np.array_equal(bits(d["dx"]), bits(r.field.dx)), a
        assert np.array_equal(bits(d["dy"]), bits(r.field.dy)), a
        assert np.array_equal(d["flags"], r.field.flags), a

"""Frame conditioning (PreparedFrame::new, preprocess.rs:137-216) on the GPU against the oracle,
bit for bit: the base image (normalised samples minus the background, clamped at 0), the
Gaussian blur (preprocess.rs:161-195) and the local normalisation (preprocess.rs:199-216), for
every combination of the three steps, u8 / u16 / f32 samples, single frames and batches (one
launch per kernel and chunk of frames), frame sizes that are not multiples of the fused
kernel's 64 x 16 tile, and blur radii inside (1..16) and beyond its halo (the two-pass blur)."""
import numpy as np
import pytest

from piv_amd import synthetic as S

pytestmark = pytest.mark.gpu

COMBOS = [(0.0, 0, True), (1.0, 0, False), (1.0, 0, True), (0.0, 4, False), (0.0, 4, True),
          (1.0, 4, False), (1.0, 4, True), (2.5, 3, False), (0.3, 2, True), (6.0, 0, True),
          (5.5, 2, False)]


@pytest.fixture(scope="module")
def piv():
    import piv_amd

    if piv_amd.device_count() < 1:
        pytest.skip("no GPU")
    return piv_amd


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _frames(dtype, h, w, n, seed):
    bits_ = 12 if dtype is np.uint16 else 8
    cont = np.uint8 if dtype is np.uint8 else np.uint16
    frames, _ = S.video_frames(h, w, n, seed, bits=bits_, container=cont)
    if dtype is np.float32:
        frames = [f.astype(np.float32) / np.float32(4095) for f in frames]
    # a background of the order of the samples, with structure along both axes
    yy, xx = np.mgrid[0:h, 0:w]
    bg = (0.02 * (1 + np.sin(xx / 17.0) * np.cos(yy / 11.0))).astype(np.float32)
    if dtype is np.uint16:
        bg *= np.float32(4095 / 65535)
    return frames, bg


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("sigma,radius,with_bg", COMBOS)
def test_conditioning_bit_exact(piv, oracle, dtype, sigma, radius, with_bg):
    h, w, n = 150, 203, 3
    frames, bg = _frames(dtype, h, w, n, 31)
    bg = bg if with_bg else None
    pre = piv.PreprocessConfig(smooth_sigma=sigma, normalize_radius=radius, normalize_eps=0.02)
    an = piv.PivAnalyzer(piv.PivConfig(windows=[32], final_iterations=1), pre)
    refs = [oracle.prepare(f, sigma, radius, 0.02, bg) for f in frames]
    for f, ref in zip(frames, refs):
        img, base = an.prepare(f, bg).download()
        assert np.array_equal(bits(base), bits(ref.base))
        assert np.array_equal(bits(img), bits(ref.image)), np.abs(img - ref.image).max()
    dev = piv.DeviceArray.from_numpy(np.stack(frames))
    dbg = piv.DeviceArray.from_numpy(bg) if bg is not None else None
    fs = piv.PreparedFrame.from_device_batch(an, dev, dbg)
    for fr, ref in zip(fs, refs):
        img, base = fr.download()
        assert np.array_equal(bits(base), bits(ref.base))
        assert np.array_equal(bits(img), bits(ref.image))


def test_conditioned_analysis_equals_oracle(piv, oracle):
    """The CLI's default conditioning and mask (blur 1, normalisation radius 4, min_mean 0.04,
    min_std 0.015) with a background, through the whole analysis: flags identical, final field
    within 1e-3 px of the oracle's."""
    from test_gpu_configs import full_parity

    h, w = 256, 320
    frames, bg = _frames(np.uint16, h, w, 2, 5)
    pre = piv.PreprocessConfig()
    cfg = piv.PivConfig(windows=[48, 24], final_iterations=1, mask=piv.MaskConfig(0.04, 0.015))
    an = piv.PivAnalyzer(cfg, pre)
    r = an.analyze(an.prepare(frames[0], bg), an.prepare(frames[1], bg))
    pa = oracle.prepare(frames[0], 1.0, 4, 0.02, bg)
    pb = oracle.prepare(frames[1], 1.0, 4, 0.02, bg)
    full_parity(oracle, r, pa, pb, cfg)

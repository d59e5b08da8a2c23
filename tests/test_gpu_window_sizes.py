"""Every window size the reference accepts (any even n >= 8, multipass.rs:96-100,
correlate.rs:53-65), up to the library's largest (128): the sizes with a register-resident
FFT kernel (every even size up to 64, and 70, 72, 80, 84, 90, 96, 98, 100, 108, 112, 120, 126,
128) run it, every other size the generic-size kernel, and deformed passes above 64 px the
generic kernel too.  Each is replayed pass by pass through the
oracle (integer peaks identical, displacements within 1e-4 px, NO_SIGNAL identical, validation
bit-exact) and compared end to end (flags identical, within 1e-3 px)."""
import numpy as np
import pytest

from piv_amd import synthetic as S
from test_gpu_configs import full_parity, stage_parity
from test_gpu_deform import deformed_stage_parity

pytestmark = pytest.mark.gpu

FFT_NEW = [10, 14, 18, 20, 22, 26, 28, 30, 34, 36, 38, 40, 42, 44, 46, 50, 52, 54, 56, 58, 60, 62,
           70, 72, 80, 84, 90, 98, 100, 108, 112, 120, 126, 128]
GENERIC = [66, 74, 78, 88, 94, 104, 110, 118, 122]


@pytest.fixture(scope="module")
def piv():
    import piv_amd

    if piv_amd.device_count() < 1:
        pytest.skip("no GPU")
    return piv_amd


def frames_for(n, seed, dtype=np.uint8):
    h = w = max(6 * n, 192)
    field = S.lamb_oseen_shear(h, w, vmax=min(4.0, n / 8.0), shear=0.01)
    a, b = S.particle_frames(h, w, seed, field=field)
    if dtype is np.uint16:
        a, b = a.astype(np.uint16) * 16, b.astype(np.uint16) * 16
    elif dtype is np.float32:
        a, b = a.astype(np.float32) / 255, b.astype(np.float32) / 255
    return a, b


@pytest.mark.parametrize("n", FFT_NEW + GENERIC)
def test_window_size_parity(piv, oracle, n):
    """[n] with two final iterations: a first pass at integer origins, then a pass offset by its
    predictor (fractional origins)."""
    a, b = frames_for(n, 100 + n)
    cfg = piv.PivConfig(windows=[n], overlap=0.5, final_iterations=2)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    res, tr = stage_parity(piv, oracle, an, an.prepare(a), an.prepare(b), pa, pb, cfg)
    full_parity(oracle, res, pa, pb, cfg)
    assert np.isfinite(res.field.dx).mean() > 0.8, n


@pytest.mark.parametrize("n,dtype", [(22, np.uint16), (22, np.float32), (66, np.uint16),
                                     (66, np.float32), (40, np.uint16),
                                     (80, np.uint16), (128, np.uint16), (128, np.float32)])
def test_window_size_formats(piv, oracle, n, dtype):
    a, b = frames_for(n, 7 + n, dtype)
    cfg = piv.PivConfig(windows=[n], overlap=0.5, final_iterations=2)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    res, _ = stage_parity(piv, oracle, an, an.prepare(a), an.prepare(b), pa, pb, cfg)
    full_parity(oracle, res, pa, pb, cfg)


@pytest.mark.parametrize("windows", [[128, 64, 32], [100, 50, 26], [80, 40, 20]])
def test_mixed_sizes_multipass(piv, oracle, windows):
    a, b = frames_for(windows[0], 31)
    cfg = piv.PivConfig(windows=windows, overlap=0.5, final_iterations=1)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    res, _ = stage_parity(piv, oracle, an, an.prepare(a), an.prepare(b), pa, pb, cfg)
    full_parity(oracle, res, pa, pb, cfg)


@pytest.mark.parametrize("windows,mode", [([128, 96], "bilinear"), ([128, 96], "bicubic"),
                                          ([96, 80, 44], "bilinear"), ([100, 64, 30], "bicubic")])
def test_deformed_large_and_generic_windows(piv, oracle, windows, mode):
    a, b = frames_for(windows[0], 53)
    cfg = piv.PivConfig(windows=windows, overlap=0.5, final_iterations=1, deformation=mode)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    deformed_stage_parity(piv, oracle, an, cfg, a, b, mode)


def test_window_size_limits(piv):
    for n in (7, 9, 6):
        with pytest.raises(piv.ConfigError):
            piv.PivAnalyzer(piv.PivConfig(windows=[n], final_iterations=1))
    for n in (130, 256):
        with pytest.raises(piv.Unsupported):
            piv.PivAnalyzer(piv.PivConfig(windows=[n], final_iterations=1))
    for n in range(8, 129, 2):
        piv.PivAnalyzer(piv.PivConfig(windows=[n], final_iterations=1)).close()

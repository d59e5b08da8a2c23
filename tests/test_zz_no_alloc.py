"""piv_analyze_batch_noalloc (workspace in the library's static per-GPU array, nothing
allocated per call) against piv_analyze_batch_device on the same pairs: bitwise equal fields,
flags and summaries, on two calls in a row, the second with a smaller input, for u8 frames
(no mask tables) and u16 frames (mask tables built in the workspace)."""
import ctypes as C

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


def batch_reference(piv, an, clip, npairs):
    """piv_analyze_batch_device on frames 0..npairs of the clip (pairs k, k+1)."""
    fs = piv.PreparedFrame.from_device_batch(an, clip)
    h, w = clip.shape[1:]
    g = an.output_grid(h, w)
    dx, dy, pr = (piv.DeviceArray((npairs, g.ny, g.nx), np.float32) for _ in range(3))
    fl = piv.DeviceArray((npairs, g.ny, g.nx), np.uint8)
    sums = an.analyze_batch_device(fs[:npairs], fs[1:npairs + 1], dx, dy, pr, fl)
    return dx.numpy(), dy.numpy(), pr.numpy(), fl.numpy(), sums


def noalloc(piv, an, clip, npairs, outs):
    """pairs k, k+1 of the clip through the C entry point into the preallocated outputs."""
    from piv_amd import _lib as L

    lib = L.load()
    h, w = clip.shape[1:]
    item = clip.dtype.itemsize
    fmt = {1: L.PIXEL_U8, 2: L.PIXEL_U16}[item]
    dx, dy, pr, fl, sm = outs
    fstride = h * w * item
    base = clip.device_ptr
    rc = lib.piv_analyze_batch_noalloc(an._h, C.c_void_p(base), C.c_void_p(base + fstride), fmt,
                                       npairs, h, w, w * item, fstride, L.ptr(dx), L.ptr(dy),
                                       L.ptr(pr), L.ptr(fl), L.ptr(sm), None)
    assert rc == 0, lib.piv_last_error().decode()
    an.synchronize()
    g = an.output_grid(h, w)
    n = g.ny * g.nx
    cut = lambda a, dt: a.numpy().reshape(-1)[:npairs * n].view(dt).reshape(npairs, g.ny, g.nx)
    return (cut(dx, np.float32), cut(dy, np.float32), cut(pr, np.float32), cut(fl, np.uint8),
            sm.numpy().reshape(-1)[:npairs * 5].reshape(npairs, 5))


@pytest.mark.parametrize("dtype,windows", [(np.uint8, [64, 32, 16]), (np.uint16, [48, 24])])
def test_noalloc_equals_batch_device(piv, dtype, windows):
    h, w = 256, 320
    frames, _ = S.video_frames(h, w, 9, 5, bits=8 if dtype is np.uint8 else 12,
                               container=dtype)
    cfg = piv.PivConfig(windows=windows, overlap=0.5, final_iterations=2)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    g = an.output_grid(h, w)
    big = 8
    outs = (piv.DeviceArray((big, g.ny, g.nx), np.float32),
            piv.DeviceArray((big, g.ny, g.nx), np.float32),
            piv.DeviceArray((big, g.ny, g.nx), np.float32),
            piv.DeviceArray((big, g.ny, g.nx), np.uint8),
            piv.DeviceArray((big, 5), np.uint32))
    for npairs in (big, 3):  # the second call smaller than the first
        clip = piv.DeviceArray.from_numpy(np.stack(frames[:npairs + 1]))
        rdx, rdy, rpr, rfl, rsum = batch_reference(piv, an, clip, npairs)
        ndx, ndy, npr, nfl, nsum = noalloc(piv, an, clip, npairs, outs)
        assert np.array_equal(bits(ndx), bits(rdx)), npairs
        assert np.array_equal(bits(ndy), bits(rdy)), npairs
        assert np.array_equal(bits(npr), bits(rpr)), npairs
        assert np.array_equal(nfl, rfl), npairs
        for k in range(npairs):
            s = rsum[k]
            assert list(nsum[k]) == [s.no_signal, s.low_ppr, s.median_outliers,
                                     s.out_of_range, s.replaced], (npairs, k)
        assert np.isfinite(ndx).mean() > 0.5


def test_noalloc_rejects_oversized_and_conditioned(piv):
    from piv_amd import _lib as L

    lib = L.load()
    cfg = piv.PivConfig(windows=[16], overlap=0.75, final_iterations=1)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    frame = piv.DeviceArray((2048, 2048), np.uint16)
    out = piv.DeviceArray((4,), np.float32)
    # 2048^2 u16 needs mask tables of 64 MiB per frame: beyond the workspace
    rc = lib.piv_analyze_batch_noalloc(an._h, L.ptr(frame), L.ptr(frame), L.PIXEL_U16, 1, 2048,
                                       2048, 4096, 0, L.ptr(out), L.ptr(out), L.ptr(out),
                                       L.ptr(out), None, None)
    assert rc == L.ERR_INVALID_ARGUMENT and b"workspace" in lib.piv_last_error()
    an2 = piv.PivAnalyzer(cfg, piv.PreprocessConfig())
    rc = lib.piv_analyze_batch_noalloc(an2._h, L.ptr(frame), L.ptr(frame), L.PIXEL_U16, 1, 256,
                                       256, 4096, 0, L.ptr(out), L.ptr(out), L.ptr(out),
                                       L.ptr(out), None, None)
    assert rc == L.ERR_UNSUPPORTED

"""Time-series filters of a run on the GPU, mirroring piv-core's `smooth` module
(piv-core/src/smooth.rs:1-320): temporal(), spatial(), cross_validation_errors(),
median_error(), pair_to_pair_change(), spatial_roughness() and mask().

Series are stacked [pair, ny, nx].  Inputs may be host numpy arrays (uploaded, results returned
as numpy) or device arrays (torch tensors on the GPU or DeviceArray; results returned as
DeviceArray).  Every function runs HIP kernels through the C ABI (include/piv.h)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import DeviceArray, _check

MEAN, MEDIAN = "mean", "median"  # TemporalKernel (smooth.rs:20-24)


@dataclass
class Smoothed:
    """smooth.rs:27-35: displacement plus how many raw vectors supported each output."""
    dx: object
    dy: object
    support: object


def _device(arrays, device, dtypes):
    """Device views of the inputs: numpy arrays are uploaded (host=True), device arrays used
    as they are."""
    host = all(isinstance(a, np.ndarray) for a in arrays if a is not None)
    out = []
    for a, dt in zip(arrays, dtypes):
        if a is None:
            out.append(None)
        elif isinstance(a, np.ndarray):
            out.append(DeviceArray.from_numpy(np.ascontiguousarray(a, dt), device))
        else:
            out.append(a)
    return host, out


def _shape3(a):
    s = tuple(a.shape)
    if len(s) != 3:
        raise ValueError(f"series must be [pair, ny, nx], got {s}")
    return s


def _result(host, arrs):
    return [a.numpy() for a in arrs] if host else arrs


def _outputs(shape, device, out):
    """caller's (dx, dy, support) device buffers, or new ones"""
    if out is not None:
        return out
    return (DeviceArray(shape, np.float32, device), DeviceArray(shape, np.float32, device),
            DeviceArray(shape, np.uint16, device))


def temporal(dx, dy, usable, gate, kernel=MEAN, window=5, min_support=1, leave_out=(),
             device=0, out=None) -> Smoothed:
    """smooth.rs:44-130: moving mean or median over `window` (odd) consecutive pairs.
    out: optional (dx, dy, support) device buffers [nt, ny, nx] (f32, f32, u16) to write."""
    nt, ny, nx = _shape3(dx)
    host, (ddx, ddy, dok, dg) = _device([dx, dy, usable, gate], device,
                                        [np.float32, np.float32, np.uint8, np.uint8])
    odx, ody, sup = _outputs((nt, ny, nx), device, out)
    lo = np.ascontiguousarray(list(leave_out) or [0], np.int32)
    _check(L.load().piv_smooth_temporal_device(
        device, nt, ny, nx, L.ptr(ddx), L.ptr(ddy), L.ptr(dok), L.ptr(dg),
        0 if kernel == MEAN else 1, window, min_support, L.ptr(lo), len(leave_out),
        L.ptr(odx), L.ptr(ody), L.ptr(sup)))
    return Smoothed(*_result(host and out is None, [odx, ody, sup]))


def spatial(dx, dy, gate, sigma, min_weight, leave_out=None, device=0, out=None) -> Smoothed:
    """smooth.rs:137-208: Gaussian low-pass of each pair's field by normalized convolution.
    out: optional (dx, dy, support) device buffers to write."""
    nt, ny, nx = _shape3(dx)
    host, (ddx, ddy, dg) = _device([dx, dy, gate], device, [np.float32, np.float32, np.uint8])
    odx, ody, sup = _outputs((nt, ny, nx), device, out)
    _check(L.load().piv_smooth_spatial_device(
        device, nt, ny, nx, L.ptr(ddx), L.ptr(ddy), L.ptr(dg), float(sigma), float(min_weight),
        -1 if leave_out is None else int(leave_out), L.ptr(odx), L.ptr(ody), L.ptr(sup)))
    return Smoothed(*_result(host and out is None, [odx, ody, sup]))


def mask(values, usable, device=0):
    """smooth.rs:303-310: NaN where not usable."""
    host, (dv, dok) = _device([values, usable], device, [np.float32, np.uint8])
    out = DeviceArray(tuple(values.shape), np.float32, device)
    n = int(np.prod(values.shape))
    _check(L.load().piv_mask_device(device, n, L.ptr(dv), L.ptr(dok), L.ptr(out)))
    return out.numpy() if host else out


def cross_validation_errors(raw_dx, raw_dy, usable, pred_dx, pred_dy, device=0):
    """smooth.rs:213-232: distance of each usable raw vector to its prediction, in order."""
    host, arrs = _device([raw_dx, raw_dy, usable, pred_dx, pred_dy], device,
                         [np.float32, np.float32, np.uint8, np.float32, np.float32])
    n = int(np.prod(raw_dx.shape))
    out = DeviceArray((max(n, 1),), np.float32, device)
    cnt = C.c_uint64()
    _check(L.load().piv_cross_validation_errors_device(device, n, *[L.ptr(a) for a in arrs],
                                                       L.ptr(out), C.byref(cnt)))
    m = int(cnt.value)
    if host:
        return out.numpy()[:m]
    view = DeviceArray.__new__(DeviceArray)
    view.shape, view.dtype, view.device = (m,), out.dtype, device
    view.nbytes, view.device_ptr, view._base = m * 4, out.device_ptr, out
    return view


def median_error(errors, keep=None, device=0):
    """smooth.rs:236-249 -> (median, count)."""
    _, (de, dk) = _device([errors, keep], device, [np.float32, np.uint8])
    n = int(np.prod(errors.shape))
    med, cnt = C.c_float(), C.c_uint64()
    _check(L.load().piv_median_error_device(device, L.ptr(de), n, L.ptr(dk), C.byref(med),
                                            C.byref(cnt)))
    return float(med.value), int(cnt.value)


def pair_to_pair_change(dx, dy, device=0):
    """smooth.rs:252-274: median |d(t+1) - d(t)| (px)."""
    nt, ny, nx = _shape3(dx)
    _, (ddx, ddy) = _device([dx, dy], device, [np.float32, np.float32])
    r = C.c_float()
    _check(L.load().piv_pair_to_pair_change_device(device, nt, ny, nx, L.ptr(ddx), L.ptr(ddy),
                                                   C.byref(r)))
    return float(r.value)


def spatial_roughness(dx, dy, device=0):
    """smooth.rs:277-300: median |d - mean of its 4 neighbours| (px)."""
    nt, ny, nx = _shape3(dx)
    _, (ddx, ddy) = _device([dx, dy], device, [np.float32, np.float32])
    r = C.c_float()
    _check(L.load().piv_spatial_roughness_device(device, nt, ny, nx, L.ptr(ddx), L.ptr(ddy),
                                                 C.byref(r)))
    return float(r.value)

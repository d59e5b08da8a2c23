"""MI355X-native PIV analysis engine: Python mirror of the piv-core API over the C ABI.

Names, argument meaning and error behaviour follow the Rust crate piv-core (agoodman/piv):

    PivConfig / MaskConfig / ValidationConfig / PreprocessConfig   (multipass.rs:22-63,
                                                                    validate.rs:7-37,
                                                                    preprocess.rs:8-29)
    PivAnalyzer.new(cfg).analyze(a, b) -> PairResult             (multipass.rs:91-150)
    PivAnalyzer.analyze_traced(a, b) -> (PairResult, [PassTrace])
    PivAnalyzer.output_grid(h, w) -> Grid                         (multipass.rs:126-133)
    PreparedFrame.new(raw, preprocess, background, analyzer)      (preprocess.rs:137-153)
    Grid.new(h, w, window, overlap)                               (grid.rs:20-44)
    validate(field, cfg, max_displacement)                        (validate.rs:51-130)

Every computation runs in the HIP library (libpiv_amd.so) on the GPU; errors come back as
PivError subclasses carrying the library's message.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib as L

__all__ = ["PivConfig", "MaskConfig", "ValidationConfig", "PreprocessConfig", "PivAnalyzer",
           "PreparedFrame", "Grid", "VectorField", "PairResult", "PassTrace",
           "ValidationSummary", "PivError", "ConfigError", "ShapeMismatch", "ImageTooSmall",
           "DeviceError", "flags", "validate", "place_windows", "correlate_windows",
           "normalize_pixels", "BackgroundEstimator", "DeviceArray", "FlowStats"]


@dataclass
class FlowStats:
    """FlowStats (flowstats.rs:9-25) plus the neighbour noise estimate (px)."""
    vorticity_mean: float
    vorticity_abs_mean: float
    ke_mean: float
    ke_sum: float
    coverage: float
    noise_px: float
    n_usable: int
    n_vorticity: int


class flags:  # field.rs:6-20
    NO_SIGNAL = 1
    LOW_PPR = 2
    MEDIAN_OUTLIER = 4
    REPLACED = 8
    OUT_OF_RANGE = 16
    INVALID = NO_SIGNAL | LOW_PPR | MEDIAN_OUTLIER | OUT_OF_RANGE


class PivError(Exception):
    """PivError (error.rs:3-20) and device-side failures."""
    status = None


class ConfigError(PivError):
    status = L.ERR_CONFIG


class ShapeMismatch(PivError):
    status = L.ERR_SHAPE_MISMATCH


class ImageTooSmall(PivError):
    status = L.ERR_IMAGE_TOO_SMALL


class DeviceError(PivError):
    status = L.ERR_DEVICE


class Unsupported(PivError):
    status = L.ERR_UNSUPPORTED


class InvalidArgument(PivError):
    status = L.ERR_INVALID_ARGUMENT


class AdmissionError(PivError):
    """Serving: the request was refused by the HBM budget or a full queue."""
    status = L.ERR_ADMISSION


class OutOfMemory(PivError):
    status = L.ERR_OOM


_ERRORS = {c.status: c for c in (ConfigError, ShapeMismatch, ImageTooSmall, DeviceError,
                                 Unsupported, InvalidArgument, AdmissionError, OutOfMemory)}


def _check(rc):
    if rc != L.OK:
        msg = L.load().piv_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, PivError)(msg)


@dataclass
class MaskConfig:
    min_mean: float = 0.0
    min_std: float = 0.0


@dataclass
class ValidationConfig:
    ppr_threshold: float = 1.3
    median_threshold: float = 2.0
    median_eps: float = 0.1
    replace: bool = True


@dataclass
class PreprocessConfig:
    smooth_sigma: float = 1.0
    normalize_radius: int = 4
    normalize_eps: float = 0.02

    @staticmethod
    def raw():
        return PreprocessConfig(0.0, 0, 0.0)


@dataclass
class PivConfig:
    windows: List[int] = field(default_factory=lambda: [96, 64])
    overlap: float = 0.5
    final_iterations: int = 2
    mask: MaskConfig = field(default_factory=MaskConfig)
    validation: ValidationConfig = field(default_factory=ValidationConfig)
    # window deformation between passes: "none" (the reference's rigid offset), "bilinear"
    # or "bicubic" (include/piv.h, piv_deformation)
    deformation: str = "none"

    def to_c(self, pre: PreprocessConfig) -> L.piv_config:
        c = L.piv_config()
        if len(self.windows) > L.MAX_PASSES:
            raise ConfigError(f"invalid configuration: at most {L.MAX_PASSES} window sizes")
        for i, w in enumerate(self.windows):
            if not (0 <= int(w) < 2 ** 32):
                raise ConfigError(f"invalid configuration: window size {w}")
            c.windows[i] = int(w)
        c.num_windows = len(self.windows)
        c.overlap = self.overlap
        c.final_iterations = max(int(self.final_iterations), 0)
        c.min_mean = self.mask.min_mean
        c.min_std = self.mask.min_std
        c.ppr_threshold = self.validation.ppr_threshold
        c.median_threshold = self.validation.median_threshold
        c.median_eps = self.validation.median_eps
        c.replace = 1 if self.validation.replace else 0
        c.smooth_sigma = pre.smooth_sigma
        c.normalize_radius = int(pre.normalize_radius)
        c.normalize_eps = pre.normalize_eps
        if self.deformation not in L.DEFORM:
            raise ConfigError(f"invalid configuration: deformation {self.deformation!r}")
        c.deformation = L.DEFORM[self.deformation]
        return c


@dataclass(frozen=True)
class Grid:
    window: int
    step: int
    nx: int
    ny: int
    x0: int
    y0: int

    @staticmethod
    def new(height, width, window, overlap) -> "Grid":
        g = L.piv_grid()
        _check(L.load().piv_grid_new(height, width, window, overlap, C.byref(g)))
        return Grid._from_c(g)

    @staticmethod
    def _from_c(g) -> "Grid":
        return Grid(g.window, g.step, g.nx, g.ny, g.x0, g.y0)

    def _to_c(self):
        return L.piv_grid(self.window, self.step, self.nx, self.ny, self.x0, self.y0)

    def __len__(self):
        return self.nx * self.ny

    def center_x(self, ix):
        return np.float32(self.x0 + ix * self.step) + (np.float32(self.window) - 1) / 2

    def center_y(self, iy):
        return np.float32(self.y0 + iy * self.step) + (np.float32(self.window) - 1) / 2

    def xs(self):
        return np.array([self.center_x(i) for i in range(self.nx)], np.float32)

    def ys(self):
        return np.array([self.center_y(i) for i in range(self.ny)], np.float32)


@dataclass
class ValidationSummary:
    no_signal: int = 0
    low_ppr: int = 0
    median_outliers: int = 0
    out_of_range: int = 0
    replaced: int = 0

    @staticmethod
    def _from_c(s):
        return ValidationSummary(int(s.no_signal), int(s.low_ppr), int(s.median_outliers),
                                 int(s.out_of_range), int(s.replaced))


@dataclass
class VectorField:
    grid: Grid
    dx: np.ndarray
    dy: np.ndarray
    peak_ratio: np.ndarray
    flags: np.ndarray

    def valid_count(self):
        return int(np.count_nonzero((self.flags & flags.INVALID) == 0))


@dataclass
class PairResult:
    field: VectorField
    summary: ValidationSummary


@dataclass
class PassTrace:
    origins: np.ndarray     # [ny*nx, 4]
    raw: VectorField
    validated: VectorField
    peak_index: np.ndarray  # [ny, nx] iy*n + ix, -1 where no peak
    summary: ValidationSummary


_FMT = {np.dtype(np.uint8): L.PIXEL_U8, np.dtype(np.uint16): L.PIXEL_U16,
        np.dtype(np.float32): L.PIXEL_F32}


class PivAnalyzer:
    """PivAnalyzer (multipass.rs:85-150) on one GPU."""

    def __init__(self, cfg: Optional[PivConfig] = None,
                 preprocess: Optional[PreprocessConfig] = None, device: int = 0):
        self.cfg = cfg if cfg is not None else PivConfig()
        self.preprocess = preprocess if preprocess is not None else PreprocessConfig()
        self._lib = L.load()
        self._c = self.cfg.to_c(self.preprocess)
        h = C.c_void_p()
        _check(self._lib.piv_analyzer_create(C.byref(self._c), device, C.byref(h)))
        self._h = h
        self.device = device

    @staticmethod
    def new(cfg: PivConfig, preprocess: Optional[PreprocessConfig] = None, device: int = 0):
        return PivAnalyzer(cfg, preprocess, device)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.piv_analyzer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self):
        return self.cfg

    @property
    def num_passes(self):
        return int(self._lib.piv_analyzer_num_passes(self._h))

    def output_grid(self, height, width) -> Grid:
        g = L.piv_grid()
        _check(self._lib.piv_output_grid(self._h, height, width, C.byref(g)))
        return Grid._from_c(g)

    def prepare(self, raw, background=None) -> "PreparedFrame":
        return PreparedFrame._prepare(self, raw, background)

    def _pass_grids(self, h, w):
        wins = list(self.cfg.windows) + [self.cfg.windows[-1]] * (self.cfg.final_iterations - 1)
        return [Grid.new(h, w, n, self.cfg.overlap) for n in wins]

    def analyze(self, a: "PreparedFrame", b: "PreparedFrame") -> PairResult:
        g = self.output_grid(a.height, a.width) if a.shape == b.shape else None
        if g is None:
            raise ShapeMismatch(f"frames have different shapes: {a.shape} vs {b.shape}")
        n = len(g)
        dx, dy, pr = (np.empty(n, np.float32) for _ in range(3))
        fl = np.empty(n, np.uint8)
        out = L.piv_field_out(L.ptr(dx), L.ptr(dy), L.ptr(pr), L.ptr(fl), n)
        _check(self._lib.piv_analyze_pair(self._h, a._h, b._h, C.byref(out)))
        g = Grid._from_c(out.grid)
        shp = (g.ny, g.nx)
        return PairResult(VectorField(g, dx.reshape(shp), dy.reshape(shp), pr.reshape(shp),
                                      fl.reshape(shp)), ValidationSummary._from_c(out.summary))

    def analyze_traced(self, a: "PreparedFrame", b: "PreparedFrame"):
        if a.shape != b.shape:
            raise ShapeMismatch(f"frames have different shapes: {a.shape} vs {b.shape}")
        grids = self._pass_grids(a.height, a.width)
        n = len(grids[-1])
        dx, dy, pr = (np.empty(n, np.float32) for _ in range(3))
        fl = np.empty(n, np.uint8)
        out = L.piv_field_out(L.ptr(dx), L.ptr(dy), L.ptr(pr), L.ptr(fl), n)
        bufs = []
        traces = (L.piv_pass_trace * len(grids))()
        for i, g in enumerate(grids):
            c = len(g)
            b_ = dict(origins=np.empty(c * 4, np.float32), raw_dx=np.empty(c, np.float32),
                      raw_dy=np.empty(c, np.float32), raw_ppr=np.empty(c, np.float32),
                      raw_flags=np.empty(c, np.uint8), peak_index=np.empty(c, np.int32),
                      dx=np.empty(c, np.float32), dy=np.empty(c, np.float32),
                      ppr=np.empty(c, np.float32), flags=np.empty(c, np.uint8))
            bufs.append(b_)
            t = traces[i]
            t.capacity = c
            for k, v in b_.items():
                setattr(t, k, L.ptr(v))
        _check(self._lib.piv_analyze_traced(self._h, a._h, b._h, C.byref(out), traces,
                                            len(grids)))
        res = []
        for i, g in enumerate(grids):
            shp = (g.ny, g.nx)
            b_ = bufs[i]
            raw = VectorField(g, b_["raw_dx"].reshape(shp), b_["raw_dy"].reshape(shp),
                              b_["raw_ppr"].reshape(shp), b_["raw_flags"].reshape(shp))
            val = VectorField(g, b_["dx"].reshape(shp), b_["dy"].reshape(shp),
                              b_["ppr"].reshape(shp), b_["flags"].reshape(shp))
            res.append(PassTrace(b_["origins"].reshape(-1, 4), raw, val,
                                 b_["peak_index"].reshape(shp),
                                 ValidationSummary._from_c(traces[i].summary)))
        g = Grid._from_c(out.grid)
        shp = (g.ny, g.nx)
        pair = PairResult(VectorField(g, dx.reshape(shp), dy.reshape(shp), pr.reshape(shp),
                                      fl.reshape(shp)), ValidationSummary._from_c(out.summary))
        return pair, res

    def synchronize(self):
        """Wait for all work queued by this analyzer."""
        _check(self._lib.piv_analyzer_synchronize(self._h))

    def profile(self, enable=True):
        """Turn per-kernel HIP-event timing on (clearing counters) or off."""
        _check(self._lib.piv_profile_enable(self._h, 1 if enable else 0))

    def profile_read(self):
        """[{name, window, launches, units, total_ms, bytes}] since profile(True)."""
        cnt = C.c_uint32(0)
        _check(self._lib.piv_profile_read(self._h, None, 0, C.byref(cnt)))
        arr = (L.piv_kernel_stat * max(cnt.value, 1))()
        _check(self._lib.piv_profile_read(self._h, C.cast(arr, C.c_void_p), cnt.value,
                                          C.byref(cnt)))
        return [dict(name=arr[i].name.decode(), window=int(arr[i].window),
                     launches=int(arr[i].launches), units=int(arr[i].units),
                     total_ms=float(arr[i].total_ms), bytes=float(arr[i].bytes))
                for i in range(cnt.value)]

    def set_precision(self, precision):
        """'fp32' (FFT, the reference's arithmetic), 'fp16' or 'bf16' (DFT-matrix GEMMs on the
        matrix cores for 16/32/64 px windows, FP64 peak fit)."""
        code = {"fp32": 0, "fp16": 1, "bf16": 2}[precision]
        _check(self._lib.piv_analyzer_set_precision(self._h, code))
        self.precision = precision

    def set_variant(self, window, dtype, variant):
        """Pin the correlation-kernel variant for (window, sample dtype)."""
        _check(self._lib.piv_kernel_variant_set(self._h, window, _FMT[np.dtype(dtype)], variant))

    def flow_stats_device(self, grid, dx, dy, flags_, dt_s):
        """Per-pair statistics of validated fields on the device, as `piv analyze` gathers them
        (analyze.rs:483-489): usable vectors (flow.rs:13-17), pair_stats (flow.rs:20-34) and
        neighbour_noise (flowstats.rs:106-137).  dx, dy, flags_: device arrays (torch tensors or
        DeviceArray) [npairs, ny, nx] or [ny, nx] on this analyzer's GPU; returns one FlowStats
        per pair."""
        shape = tuple(dx.shape)
        n = 1 if len(shape) == 2 else shape[0]
        if shape[-2:] != (grid.ny, grid.nx):
            raise ShapeMismatch(f"fields {shape} vs grid {grid.ny}x{grid.nx}")
        out = (L.piv_flow_stats * max(n, 1))()
        _check(self._lib.piv_flow_stats_device(self._h, C.byref(grid._to_c()), n, L.ptr(dx),
                                               L.ptr(dy), L.ptr(flags_), float(dt_s), out))
        return [FlowStats(**{k: (int(getattr(o, k)) if k.startswith("n_") else
                                 float(getattr(o, k))) for k, _ in o._fields_})
                for o in out[:n]]

    @property
    def batch_lanes(self):
        """Groups of pairs whose pass chains a batch runs side by side (1..4; results do not
        depend on it): 2, PIV_BATCH_LANES, or what the caller (or autotune() with
        PIV_TUNE_LANES=1) last chose."""
        v = C.c_uint32(0)
        _check(self._lib.piv_batch_lanes_get(self._h, C.byref(v)))
        return int(v.value)

    @batch_lanes.setter
    def batch_lanes(self, lanes):
        _check(self._lib.piv_batch_lanes_set(self._h, int(lanes)))

    def autotune(self, frames_a, frames_b, repeats=3):
        """Time every correlation-kernel variant on this batch and keep the fastest per window
        size (variants give bit-identical results); with PIV_TUNE_LANES=1 at the analyzer's
        creation, then the batch lanes (batch_lanes).  Returns one dict per window size."""
        n = len(frames_a)
        A = (C.c_void_p * n)(*[f._h.value for f in frames_a])
        B = (C.c_void_p * n)(*[f._h.value for f in frames_b])
        arr = (L.piv_tune_entry * 16)()
        cnt = C.c_uint32(0)
        _check(self._lib.piv_autotune(self._h, A, B, n, repeats, C.cast(arr, C.c_void_p), 16,
                                      C.byref(cnt)))
        out = []
        for i in range(min(cnt.value, 16)):
            e = arr[i]
            k = e.num_variants
            out.append(dict(window=int(e.window), fmt=int(e.fmt), chosen=int(e.chosen),
                            occupancy=[int(x) for x in e.occupancy[:k]],
                            ms=[float(x) for x in e.ms[:k]]))
        return out

    def analyze_batch_device(self, frames_a, frames_b, dx, dy, ppr, flags_out,
                             summaries=True, stream=None, sync=True):
        """Analyse len(frames_a) pairs; outputs are device tensors [npairs, ny, nx]."""
        n = len(frames_a)
        A = (C.c_void_p * n)(*[f._h.value for f in frames_a])
        B = (C.c_void_p * n)(*[f._h.value for f in frames_b])
        sums = (L.piv_summary * n)() if summaries else None
        _check(self._lib.piv_analyze_batch_device(
            self._h, A, B, n, L.ptr(dx), L.ptr(dy), L.ptr(ppr), L.ptr(flags_out),
            C.cast(sums, C.c_void_p) if sums is not None else None,
            None if stream is None else C.c_void_p(stream), 0 if sync else 1))
        if sums is None:
            return None
        return [ValidationSummary._from_c(s) for s in sums]


class PreparedFrame:
    """PreparedFrame (preprocess.rs:130-153): conditioned image + signal statistics, on the
    analyzer's device."""

    def __init__(self):
        self._h = None

    @staticmethod
    def new(raw, preprocess: PreprocessConfig, background, analyzer: PivAnalyzer):
        if (preprocess.smooth_sigma, preprocess.normalize_radius, preprocess.normalize_eps) != (
                analyzer.preprocess.smooth_sigma, analyzer.preprocess.normalize_radius,
                analyzer.preprocess.normalize_eps):
            raise ConfigError("preprocess config differs from the analyzer's")
        return PreparedFrame._prepare(analyzer, raw, background)

    @staticmethod
    def _prepare(an: PivAnalyzer, raw, background=None):
        raw = np.ascontiguousarray(raw)
        if raw.ndim != 2 or raw.dtype not in _FMT:
            raise InvalidArgument(f"expected a 2-D uint8/uint16/float32 image, got {raw.dtype} "
                                  f"{raw.shape}")
        bg = None if background is None else np.ascontiguousarray(background, np.float32)
        if bg is not None and bg.shape != raw.shape:
            raise ShapeMismatch(f"background {bg.shape} vs frame {raw.shape}")
        h = C.c_void_p()
        _check(an._lib.piv_frame_prepare(an._h, L.ptr(raw), _FMT[raw.dtype], raw.shape[0],
                                         raw.shape[1], raw.strides[0], L.ptr(bg), C.byref(h)))
        f = PreparedFrame()
        f._h, f._lib, f.shape, f._owner = h, an._lib, raw.shape, an
        return f

    @staticmethod
    def from_device(an: PivAnalyzer, arr, background=None):
        """Wrap device-resident samples: a 2-D DeviceArray, or a row-major device tensor of
        a torch that shares this process's HIP runtime (u8, u16 or f32)."""
        if isinstance(arr, DeviceArray):
            if len(arr.shape) != 2 or arr.dtype not in _FMT:
                raise InvalidArgument("expected a 2-D uint8/uint16/float32 DeviceArray")
            fmt, shape, ptr_, stride = _FMT[arr.dtype], arr.shape, arr.device_ptr, arr.shape[1] * arr.dtype.itemsize
        else:
            name = str(arr.dtype).replace("torch.", "")
            fmt = {"uint8": L.PIXEL_U8, "uint16": L.PIXEL_U16, "int16": L.PIXEL_U16,
                   "float32": L.PIXEL_F32}[name]
            if arr.dim() != 2 or arr.stride(1) != 1:
                raise InvalidArgument("expected a 2-D row-major device tensor")
            shape, ptr_ = tuple(arr.shape), arr.data_ptr()
            stride = arr.stride(0) * arr.element_size()
        h = C.c_void_p()
        _check(an._lib.piv_frame_prepare_device(
            an._h, C.c_void_p(ptr_), fmt, shape[0], shape[1], stride,
            None if background is None else L.ptr(background), C.byref(h)))
        f = PreparedFrame()
        f._h, f._lib, f.shape, f._owner = h, an._lib, tuple(shape), an
        f._keep = (arr, background)
        return f

    @staticmethod
    def from_device_batch(an: PivAnalyzer, arr, background=None):
        """Wrap a device-resident stack of frames [count, h, w] (DeviceArray or row-major
        torch tensor), prepared together (one launch per conditioning kernel where the
        configuration allows).  Returns a list of frames; keep `arr` unchanged meanwhile."""
        if isinstance(arr, DeviceArray):
            if len(arr.shape) != 3 or arr.dtype not in _FMT:
                raise InvalidArgument("expected a 3-D uint8/uint16/float32 DeviceArray")
            fmt, shape, ptr_ = _FMT[arr.dtype], arr.shape, arr.device_ptr
            stride = shape[2] * arr.dtype.itemsize
            fstride = shape[1] * stride
        else:
            name = str(arr.dtype).replace("torch.", "")
            fmt = {"uint8": L.PIXEL_U8, "uint16": L.PIXEL_U16, "int16": L.PIXEL_U16,
                   "float32": L.PIXEL_F32}[name]
            if arr.dim() != 3 or arr.stride(2) != 1:
                raise InvalidArgument("expected a 3-D device tensor with unit column stride")
            shape, ptr_ = tuple(arr.shape), arr.data_ptr()
            stride = arr.stride(1) * arr.element_size()
            fstride = arr.stride(0) * arr.element_size()
        n = shape[0]
        hs = (C.c_void_p * n)()
        _check(an._lib.piv_frame_prepare_device_batch(
            an._h, C.c_void_p(ptr_), fmt, n, shape[1], shape[2], stride, fstride,
            None if background is None else L.ptr(background), hs))
        # one shared attribute set per frame, copied into each object (a clip can hold
        # thousands of frames: this loop is a visible share of a video's preparation time)
        base = {"_lib": an._lib, "shape": tuple(shape[1:]), "_owner": an,
                "_keep": (arr, background)}
        new, vp = PreparedFrame.__new__, C.c_void_p
        out = []
        for hv in hs:
            f = new(PreparedFrame)
            d = f.__dict__
            d.update(base)
            d["_h"] = vp(hv)
            out.append(f)
        return out

    @property
    def height(self):
        return self.shape[0]

    @property
    def width(self):
        return self.shape[1]

    def dim(self):
        return self.shape

    def download(self):
        """(image, base): the conditioned image and the background-subtracted samples."""
        img = np.empty(self.shape, np.float32)
        base = np.empty(self.shape, np.float32)
        _check(self._lib.piv_frame_download(self._h, L.ptr(img), L.ptr(base)))
        return img, base

    @property
    def image(self):
        return self.download()[0]

    def tables_built(self):
        """Whether the frame's signal-mask tables exist on the device (False when the mask
        needs none or when screening found that no window can depend on them)."""
        v = C.c_int(0)
        _check(self._lib.piv_frame_tables_built(self._h, C.byref(v)))
        return bool(v.value)

    def close(self):
        if self._h:
            self._lib.piv_frame_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A plain device buffer with a numpy dtype and shape (no torch needed)."""

    def __init__(self, shape, dtype, device=0):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.device = device
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(L.load().piv_device_alloc(device, self.nbytes, C.byref(p)))
        self.device_ptr = p.value

    @staticmethod
    def from_numpy(a, device=0):
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.shape, a.dtype, device)
        _check(L.load().piv_copy_to_device(device, C.c_void_p(d.device_ptr), L.ptr(a), d.nbytes))
        return d

    def upload(self, a):
        """Overwrite the buffer with the bytes of a host array of the same size."""
        a = np.ascontiguousarray(a, self.dtype)
        if a.nbytes != self.nbytes:
            raise ShapeMismatch(f"{a.shape} {a.dtype} into device array {self.shape} {self.dtype}")
        _check(L.load().piv_copy_to_device(self.device, C.c_void_p(self.device_ptr), L.ptr(a),
                                           self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(L.load().piv_copy_to_host(self.device, L.ptr(out), C.c_void_p(self.device_ptr),
                                         self.nbytes))
        return out

    def slice(self, i):
        """Device view of a[i] along the first axis (shares memory)."""
        v = DeviceArray.__new__(DeviceArray)
        v.shape, v.dtype, v.device = self.shape[1:], self.dtype, self.device
        v.nbytes = int(np.prod(v.shape)) * v.dtype.itemsize
        v.device_ptr = self.device_ptr + i * v.nbytes
        v._base = self
        return v

    def close(self):
        if getattr(self, "_base", None) is None and getattr(self, "device_ptr", None):
            L.load().piv_device_free(self.device, C.c_void_p(self.device_ptr))
            self.device_ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BackgroundEstimator:
    """BackgroundEstimator (preprocess.rs:39-72): per-pixel min or mean over sample frames,
    accumulated on the analyzer's GPU."""

    MIN, MEAN = 0, 1

    def __init__(self, an: PivAnalyzer, mode, height, width):
        self._an = an
        self._lib = an._lib
        self.shape = (height, width)
        h = C.c_void_p()
        m = {"min": 0, "mean": 1}.get(mode, mode)
        _check(self._lib.piv_background_create(an._h, int(m), height, width, C.byref(h)))
        self._h = h

    def add(self, frame):
        f = np.ascontiguousarray(frame)
        if f.shape != self.shape:
            raise ShapeMismatch(f"frame {f.shape} vs background {self.shape}")
        _check(self._lib.piv_background_add(self._h, L.ptr(f), _FMT[f.dtype], f.strides[0], 0))

    def finish(self):
        out = np.empty(self.shape, np.float32)
        _check(self._lib.piv_background_finish(self._h, L.ptr(out)))
        return out

    def close(self):
        if self._h:
            self._lib.piv_background_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- stage-level entry points ---------------------------------------------------------

def place_windows(an: PivAnalyzer, grid: Grid, height, width, prev: Optional[Grid] = None,
                  prev_dx=None, prev_dy=None):
    """Window origins [ax, ay, bx, by] per cell (multipass.rs:207-231)."""
    o = np.empty((len(grid), 4), np.float32)
    pdx = None if prev is None else np.ascontiguousarray(prev_dx, np.float32)
    pdy = None if prev is None else np.ascontiguousarray(prev_dy, np.float32)
    pg = None if prev is None else C.byref(prev._to_c())
    g = grid._to_c()
    _check(an._lib.piv_place_windows(an._h, C.byref(g), height, width, pg, L.ptr(pdx),
                                     L.ptr(pdy), L.ptr(o)))
    return o


def correlate_windows(an: PivAnalyzer, a: PreparedFrame, b: PreparedFrame, window, origins):
    """Raw field cells for given origins: (dx, dy, ppr, flags, peak_index)."""
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 4)
    n = o.shape[0]
    dx, dy, pr = (np.empty(n, np.float32) for _ in range(3))
    fl = np.empty(n, np.uint8)
    pk = np.empty(n, np.int32)
    _check(an._lib.piv_correlate_windows(an._h, a._h, b._h, window, L.ptr(o), n, L.ptr(dx),
                                         L.ptr(dy), L.ptr(pr), L.ptr(fl), L.ptr(pk)))
    return dx, dy, pr, fl, pk


def validate(an: PivAnalyzer, dx, dy, ppr, flags_, max_displacement):
    """validate (validate.rs:51-130) on copies; returns (dx, dy, flags, ValidationSummary)."""
    dx = np.array(dx, np.float32, copy=True, order="C")
    dy = np.array(dy, np.float32, copy=True, order="C")
    fl = np.array(flags_, np.uint8, copy=True, order="C")
    pr = np.ascontiguousarray(ppr, np.float32)
    ny, nx = dx.shape
    s = L.piv_summary()
    _check(an._lib.piv_validate_field(an._h, ny, nx, L.ptr(dx), L.ptr(dy), L.ptr(pr), L.ptr(fl),
                                      max_displacement, C.byref(s)))
    return dx, dy, fl, ValidationSummary._from_c(s)


def normalize_pixels(pixels, device=0):
    """decode.rs:227-234 on the device."""
    p = np.ascontiguousarray(pixels)
    out = np.empty(p.shape, np.float32)
    _check(L.load().piv_normalize_pixels(device, L.ptr(p), _FMT[p.dtype], p.size, L.ptr(out)))
    return out


@dataclass
class ServedResult:
    """One finished request of a PivServer (arrays are copies owned by Python)."""
    request_id: int
    stream_id: int
    field: Optional["VectorField"]
    summary: Optional["ValidationSummary"]
    submit_ms: float
    done_ms: float
    batch_pairs: int
    error: Optional[str] = None

    @property
    def latency_ms(self):
        return self.done_ms - self.submit_ms


class PivServer:
    """Continuous-batching request loop on one GPU (include/piv.h, serving).  Streams have
    their own frame size, sample format, pass schedule, conditioning and background; every
    result equals PivAnalyzer.analyze on the same pair and configuration."""

    def __init__(self, device=0, hbm_budget=0, max_batch_pairs=0, max_inflight=0):
        self._lib = L.load()
        cfg = L.piv_server_config(hbm_budget, max_batch_pairs, max_inflight)
        h = C.c_void_p()
        _check(self._lib.piv_server_create(device, C.byref(cfg), C.byref(h)))
        self._h = h
        self._streams = {}

    def add_stream(self, cfg: "PivConfig", height, width, dtype,
                   preprocess: "PreprocessConfig" = None, background=None):
        pre = preprocess if preprocess is not None else PreprocessConfig()
        c = cfg.to_c(pre)
        bg = None
        if background is not None:
            bg = np.ascontiguousarray(background, np.float32)
            if bg.shape != (height, width):
                raise ShapeMismatch(f"background {bg.shape} vs stream {(height, width)}")
        sid = C.c_uint32()
        _check(self._lib.piv_server_add_stream(self._h, C.byref(c), height, width,
                                               _FMT[np.dtype(dtype)],
                                               None if bg is None else L.ptr(bg), C.byref(sid)))
        self._streams[sid.value] = (height, width, np.dtype(dtype), bg)
        return sid.value

    def submit(self, stream_id, frame_a, frame_b):
        h, w, dt, _ = self._streams[stream_id]
        a = np.ascontiguousarray(frame_a, dt)
        b = np.ascontiguousarray(frame_b, dt)
        if a.shape != (h, w) or b.shape != (h, w):
            raise ShapeMismatch(f"frames {a.shape}/{b.shape} vs stream {(h, w)}")
        rid = C.c_uint64()
        _check(self._lib.piv_submit(self._h, stream_id, L.ptr(a), L.ptr(b), 0, C.byref(rid)))
        return rid.value

    def submit_device(self, stream_id, frame_a, frame_b):
        """Frames already in device memory (DeviceArray or torch tensor, [h, w] of the stream's
        dtype), read in place: keep them unchanged until the result is polled."""
        h, w, dt, _ = self._streams[stream_id]
        for f in (frame_a, frame_b):
            shp = tuple(f.shape)
            if shp != (h, w):
                raise ShapeMismatch(f"frame {shp} vs stream {(h, w)}")
        rid = C.c_uint64()
        _check(self._lib.piv_submit_device(self._h, stream_id, L.ptr(frame_a), L.ptr(frame_b), 0,
                                           C.byref(rid)))
        return rid.value

    def submit_device_many(self, requests):
        """Several (stream_id, frame_a, frame_b) requests of device frames in one call
        (piv_submit_device_many), in order.  Returns the ids of the requests admitted: fewer
        than given when the server could not admit the next one (retry the rest after
        polling)."""
        n = len(requests)
        if n == 0:
            return []
        sids = (C.c_uint32 * n)()
        pa, pb = (C.c_void_p * n)(), (C.c_void_p * n)()
        for i, (sid, fa, fb) in enumerate(requests):
            h, w, dt, _ = self._streams[sid]
            for f in (fa, fb):
                if tuple(f.shape) != (h, w):
                    raise ShapeMismatch(f"frame {tuple(f.shape)} vs stream {(h, w)}")
            sids[i] = sid
            pa[i] = L.ptr(fa).value
            pb[i] = L.ptr(fb).value
        ids = (C.c_uint64 * n)()
        got = C.c_uint32()
        st = self._lib.piv_submit_device_many(self._h, n, sids, pa, pb, 0, ids, C.byref(got))
        if st != 0 and not (st == L.ERR_ADMISSION and got.value < n):
            _check(st)
        return list(ids[:got.value])

    def poll(self, max_results=64, wait_ms=0, copy=True):
        """Finished requests, oldest first, released after reading.  copy=False skips the
        field arrays (field is None): ids, status, timings and summaries only."""
        if getattr(self, "_poll_cap", 0) < max_results:
            self._poll_arr = (L.piv_result * max_results)()
            self._poll_cap = max_results
        arr = self._poll_arr
        cnt = C.c_uint32()
        _check(self._lib.piv_poll(self._h, arr, max_results, C.byref(cnt), int(wait_ms)))
        out = []
        release = self._lib.piv_result_release
        ids = [int(arr[i].request_id) for i in range(cnt.value)]
        try:
            for i in range(cnt.value):
                r = arr[i]
                field, summ, err = None, None, None
                if r.status == L.OK:
                    summ = ValidationSummary._from_c(r.summary)
                    if copy:
                        g = Grid._from_c(r.grid)
                        n = g.nx * g.ny
                        shp = (g.ny, g.nx)
                        field = VectorField(
                            g, np.ctypeslib.as_array(r.dx, (n,)).reshape(shp).copy(),
                            np.ctypeslib.as_array(r.dy, (n,)).reshape(shp).copy(),
                            np.ctypeslib.as_array(r.peak_ratio, (n,)).reshape(shp).copy(),
                            np.ctypeslib.as_array(r.flags, (n,)).reshape(shp).copy())
                else:
                    err = f"status {r.status}"
                out.append(ServedResult(int(r.request_id), int(r.stream_id), field, summ,
                                        float(r.submit_ms), float(r.done_ms),
                                        int(r.batch_pairs), err))
        finally:
            # every polled result goes back to the server, also when a conversion failed, so
            # none keeps its request slot
            rcs = [release(self._h, rid) for rid in ids]
        for rc in rcs:
            _check(rc)
        return out

    def pause(self, paused=True):
        """Hold (True) or release (False) the launch of admitted requests."""
        _check(self._lib.piv_server_pause(self._h, 1 if paused else 0))

    def stats(self):
        s = L.piv_server_stats()
        _check(self._lib.piv_server_stats_read(self._h, C.byref(s)))
        return {k: int(getattr(s, k)) for k, _ in L.piv_server_stats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.piv_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_pair_npz(directory, frame_a, frame_b, time_s, dt_s, field: "VectorField",
                   m_per_px=None, roi=None):
    """NpzWriter::write (write.rs:129-165): <directory>/pair_AAAAAA_BBBBBB.npz."""
    g = field.grid
    cg = L.piv_grid(g.window, g.step, g.nx, g.ny, g.x0, g.y0)
    arrs = [np.ascontiguousarray(field.dx, np.float32), np.ascontiguousarray(field.dy, np.float32),
            np.ascontiguousarray(field.peak_ratio, np.float32),
            np.ascontiguousarray(field.flags, np.uint8)]
    rx, ry = (roi[0], roi[1]) if roi is not None else (0, 0)
    _check(L.load().piv_write_pair_npz(str(directory).encode(), frame_a, frame_b, time_s, dt_s,
                                       m_per_px if m_per_px else 0.0, rx, ry, C.byref(cg),
                                       *[L.ptr(x) for x in arrs]))
    return os.path.join(str(directory), f"pair_{frame_a:06d}_{frame_b:06d}.npz")


def write_stats_npz(directory, t, stats):
    """piv-io/src/flow.rs:133-145: `directory`/stats.npz with t (f64) and the FlowStats columns
    (f32) of a run, one entry per pair.  Returns the path."""
    t = np.ascontiguousarray(t, np.float64)
    if len(stats) != t.size:
        raise ShapeMismatch(f"{t.size} times vs {len(stats)} statistics")
    arr = (L.piv_flow_stats * max(len(stats), 1))()
    for i, s in enumerate(stats):
        for k, _ in L.piv_flow_stats._fields_:
            setattr(arr[i], k, getattr(s, k))
    _check(L.load().piv_write_stats_npz(str(directory).encode(), t.size, L.ptr(t), arr))
    return os.path.join(str(directory), "stats.npz")


def kernel_variant_count(window):
    """Number of launch variants of the correlation kernel for a window size."""
    return int(L.load().piv_kernel_variant_count(window))


def device_count():
    n = C.c_int(0)
    _check(L.load().piv_device_count(C.byref(n)))
    return n.value

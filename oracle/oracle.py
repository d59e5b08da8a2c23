"""ctypes front end of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP product path is compared
against; only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import it.  ``analyze`` below mirrors ``PivAnalyzer::run`` (piv-core/src/multipass.rs:152-192):
it strings the C restatements of place / mask / correlate / validate together pass by pass.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

NO_SIGNAL, LOW_PPR, MEDIAN_OUTLIER, REPLACED, OUT_OF_RANGE = 1, 2, 4, 8, 16
INVALID = NO_SIGNAL | LOW_PPR | MEDIAN_OUTLIER | OUT_OF_RANGE


class Grid(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("window", "step", "nx", "ny", "x0", "y0")]

    def as_tuple(self):
        return (self.window, self.step, self.nx, self.ny, self.x0, self.y0)

    def __len__(self):
        return self.nx * self.ny


class VCfg(C.Structure):
    _fields_ = [("ppr_threshold", C.c_float), ("median_threshold", C.c_float),
                ("median_eps", C.c_float), ("replace", C.c_int32)]


class Summary(C.Structure):
    _fields_ = [(n, C.c_int64) for n in
                ("no_signal", "low_ppr", "median_outliers", "out_of_range", "replaced")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class Peak(C.Structure):
    _fields_ = [("dx", C.c_float), ("dy", C.c_float), ("height", C.c_float),
                ("ppr", C.c_float), ("iy", C.c_int32), ("ix", C.c_int32)]


class Peak64(C.Structure):
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("height", C.c_double),
                ("ppr", C.c_double), ("iy", C.c_int32), ("ix", C.c_int32)]


class FlowStat(C.Structure):
    _fields_ = [("vorticity_mean", C.c_float), ("vorticity_abs_mean", C.c_float),
                ("ke_mean", C.c_float), ("ke_sum", C.c_float), ("coverage", C.c_float),
                ("n_usable", C.c_int64), ("n_vorticity", C.c_int64)]

    def as_dict(self):
        return {n: (int(getattr(self, n)) if n.startswith("n_") else float(getattr(self, n)))
                for n, _ in self._fields_}


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        P = np.ctypeslib.ndpointer
        f32 = P(np.float32, flags="C")
        f64 = P(np.float64, flags="C")
        u8 = P(np.uint8, flags="C")
        i32 = P(np.int32, flags="C")
        L = _lib
        L.or_grid_new.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_float, C.POINTER(Grid)]
        L.or_field_sample.argtypes = [C.POINTER(Grid), f32, f32, C.c_float, C.c_float,
                                      C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.or_place_windows.argtypes = [C.POINTER(Grid), C.c_int32, C.c_int32, C.POINTER(Grid),
                                       C.c_void_p, C.c_void_p, f32]
        L.or_sat_build.argtypes = [f32, C.c_int32, C.c_int32, f64, f64]
        L.or_sat_stats.argtypes = [f64, f64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.or_gauss_kernel.argtypes = [C.c_float, f32]
        L.or_gauss_kernel.restype = C.c_int32
        L.or_gaussian_blur.argtypes = [f32, C.c_int32, C.c_int32, C.c_float, f32]
        L.or_local_normalize.argtypes = [f32, C.c_int32, C.c_int32, C.c_int32, C.c_float, f32]
        L.or_prepare.argtypes = [f32, C.c_int32, C.c_int32, C.c_void_p, C.c_float, C.c_int32,
                                 C.c_float, f32, f32]
        L.or_normalize_u8.argtypes = [u8, C.c_int64, f32]
        L.or_normalize_u16.argtypes = [P(np.uint16, flags="C"), C.c_int64, f32]
        L.or_load_window.argtypes = [f32, f32, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                     C.c_int32]
        L.or_load_window.restype = C.c_double
        L.or_find_peak.argtypes = [f32, C.c_int32, C.POINTER(Peak)]
        L.or_find_peak64.argtypes = [f64, C.c_int32, C.POINTER(Peak64)]
        L.or_fft_f32.argtypes = [f32, C.c_int32, C.c_int32]
        L.or_fft_f64.argtypes = [f64, C.c_int32, C.c_int32]
        L.or_correlate_f32.argtypes = [f32, f32, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                       C.c_float, C.c_float, C.c_float, C.POINTER(Peak),
                                       C.c_void_p]
        L.or_correlate_f64.argtypes = [f32, f32, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                       C.c_float, C.c_float, C.c_float, C.POINTER(Peak64),
                                       C.c_void_p]
        L.or_correlate_pass.argtypes = [f32, f32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int32, C.c_int32, C.POINTER(Grid), f32, C.c_float,
                                        C.c_float, C.c_int32, C.c_int32, f32, f32, f32, u8,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_deform_sample.argtypes = [C.POINTER(Grid), f32, f32, C.c_float, C.c_float,
                                       C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.or_deform_shifts.argtypes = [C.POINTER(Grid), f32, f32, C.c_int32, C.c_float,
                                       C.c_float, C.c_float, C.c_float, f32, f32]
        L.or_load_window_deformed.argtypes = [f32, f32, C.c_int32, C.c_int32, C.c_float,
                                              C.c_float, C.c_int32, C.c_int32, f32, f32,
                                              C.c_int32]
        L.or_load_window_deformed.restype = C.c_double
        L.or_load_window_deformed64.argtypes = [f64, f32, C.c_int32, C.c_int32, C.c_float,
                                                C.c_float, C.c_int32, C.c_int32, f32, f32,
                                                C.c_int32]
        L.or_load_window_deformed64.restype = C.c_double
        L.or_validate.argtypes = [C.c_int32, C.c_int32, f32, f32, f32, u8, C.POINTER(VCfg),
                                  C.c_float, C.POINTER(Summary)]
        L.or_median.argtypes = [f32, C.c_int32]
        L.or_median.restype = C.c_float
        u16 = P(np.uint16, flags="C")
        L.or_flow_usable.argtypes = [C.c_int64, f32, u8, u8]
        L.or_vorticity.argtypes = [C.c_int32, C.c_int32, f32, f32, u8, C.c_float, f32]
        L.or_flow_stats.argtypes = [C.c_int32, C.c_int32, f32, f32, u8, C.c_float,
                                    C.POINTER(FlowStat)]
        L.or_pair_stats.argtypes = [C.c_int32, C.c_int32, f32, f32, u8, C.c_int32, C.c_double,
                                    C.POINTER(FlowStat)]
        L.or_neighbour_noise.argtypes = [C.c_int32, C.c_int32, f32, f32, u8]
        L.or_neighbour_noise.restype = C.c_float
        L.or_temporal.argtypes = [C.c_int32, C.c_int32, C.c_int32, f32, f32, u8, u8, C.c_int32,
                                  C.c_int32, C.c_int32, i32, C.c_int32, f32, f32, u16]
        L.or_spatial.argtypes = [C.c_int32, C.c_int32, C.c_int32, f32, f32, u8, C.c_float,
                                 C.c_float, C.c_int32, f32, f32, u16]
        L.or_cross_validation_errors.argtypes = [C.c_int64, f32, f32, u8, f32, f32, f32]
        L.or_cross_validation_errors.restype = C.c_int64
        L.or_median_error.argtypes = [f32, C.c_int64, C.c_void_p, C.POINTER(C.c_int64)]
        L.or_median_error.restype = C.c_float
        L.or_pair_to_pair_change.argtypes = [C.c_int32, C.c_int32, C.c_int32, f32, f32]
        L.or_pair_to_pair_change.restype = C.c_float
        L.or_spatial_roughness.argtypes = [C.c_int32, C.c_int32, C.c_int32, f32, f32]
        L.or_spatial_roughness.restype = C.c_float
        L.or_mask.argtypes = [C.c_int64, f32, u8, f32]
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------------------
class PivError(Exception):
    pass


def grid_new(h, w, window, overlap) -> Grid:
    g = Grid()
    rc = lib().or_grid_new(h, w, window, overlap, C.byref(g))
    if rc == 1:
        raise PivError(f"invalid configuration: overlap {overlap} must be in [0, 1)")
    if rc == 2:
        raise PivError(f"image ({h}x{w}) is smaller than the {window} px interrogation window")
    return g


def field_sample(g: Grid, dx, dy, x, y):
    u, v = C.c_float(), C.c_float()
    lib().or_field_sample(C.byref(g), _f32(dx).ravel(), _f32(dy).ravel(), x, y, C.byref(u),
                          C.byref(v))
    return u.value, v.value


def place_windows(g: Grid, h, w, prev: Grid | None = None, prev_dx=None, prev_dy=None):
    out = np.zeros(len(g) * 4, np.float32)
    pdx = _f32(prev_dx).ravel() if prev is not None else None
    pdy = _f32(prev_dy).ravel() if prev is not None else None
    lib().or_place_windows(C.byref(g), h, w, C.byref(prev) if prev is not None else None,
                           _ptr(pdx), _ptr(pdy), out)
    return out.reshape(-1, 4)


def sat_build(img):
    img = _f32(img)
    h, w = img.shape
    s = np.zeros((h + 1) * (w + 1), np.float64)
    q = np.zeros_like(s)
    lib().or_sat_build(img.ravel(), h, w, s, q)
    return s, q


def sat_stats(sat, w, x, y, bw, bh):
    m, sd = C.c_float(), C.c_float()
    lib().or_sat_stats(sat[0], sat[1], w, x, y, bw, bh, C.byref(m), C.byref(sd))
    return m.value, sd.value


def gauss_kernel(sigma):
    k = np.zeros(512, np.float32)
    r = lib().or_gauss_kernel(sigma, k)
    return k[: 2 * r + 1].copy(), r


def gaussian_blur(img, sigma):
    img = _f32(img)
    out = np.zeros_like(img)
    lib().or_gaussian_blur(img.ravel(), img.shape[0], img.shape[1], sigma, out.ravel())
    return out


def local_normalize(img, radius, eps):
    img = _f32(img)
    out = np.zeros_like(img)
    lib().or_local_normalize(img.ravel(), img.shape[0], img.shape[1], radius, eps, out.ravel())
    return out


def normalize_pixels(raw):
    """u8 -> v/255, u16 -> v/65535, float32 passes through (decode.rs:227-234)."""
    raw = np.ascontiguousarray(raw)
    if raw.dtype == np.float32:
        return raw
    out = np.zeros(raw.shape, np.float32)
    if raw.dtype == np.uint8:
        lib().or_normalize_u8(raw.ravel(), raw.size, out.ravel())
    elif raw.dtype == np.uint16:
        lib().or_normalize_u16(raw.ravel(), raw.size, out.ravel())
    else:
        raise TypeError(raw.dtype)
    return out


@dataclass
class Prepared:
    image: np.ndarray
    base: np.ndarray
    sat: tuple | None = None


def prepare(raw, smooth_sigma=0.0, normalize_radius=0, normalize_eps=0.0, background=None):
    """PreparedFrame::new (preprocess.rs:137-153) on pixels normalised as decode.rs does."""
    raw = normalize_pixels(raw)
    h, w = raw.shape
    img = np.zeros_like(raw)
    base = np.zeros_like(raw)
    bg = _f32(background).ravel() if background is not None else None
    lib().or_prepare(raw.ravel(), h, w, _ptr(bg), smooth_sigma, normalize_radius, normalize_eps,
                     img.ravel(), base.ravel())
    return Prepared(img, base, sat_build(base))


def load_window(img, x, y, n):
    img = _f32(img)
    dst = np.zeros(n * n, np.float32)
    e = lib().or_load_window(dst, img.ravel(), img.shape[0], img.shape[1], x, y, n)
    return dst.reshape(n, n), e


def find_peak(corr):
    corr = _f32(corr)
    p = Peak()
    lib().or_find_peak(corr.ravel(), corr.shape[0], C.byref(p))
    return p


def find_peak64(corr):
    corr = np.ascontiguousarray(corr, np.float64)
    p = Peak64()
    lib().or_find_peak64(corr.ravel(), corr.shape[0], C.byref(p))
    return p


def fft(x, sign=-1, precision=32):
    dt = np.float32 if precision == 32 else np.float64
    x = np.asarray(x)
    buf = np.zeros(2 * x.size, dt)
    buf[0::2] = x.real
    buf[1::2] = x.imag
    (lib().or_fft_f32 if precision == 32 else lib().or_fft_f64)(buf, x.size, sign)
    return buf[0::2] + 1j * buf[1::2]


def correlate(a, b, n, ax, ay, bx, by, precision=32, want_corr=False):
    a, b = _f32(a), _f32(b)
    h, w = a.shape
    if precision == 32:
        p = Peak()
        corr = np.zeros(n * n, np.float32) if want_corr else None
        ok = lib().or_correlate_f32(a.ravel(), b.ravel(), h, w, n, ax, ay, bx, by, C.byref(p),
                                    _ptr(corr))
    else:
        p = Peak64()
        corr = np.zeros(n * n, np.float64) if want_corr else None
        ok = lib().or_correlate_f64(a.ravel(), b.ravel(), h, w, n, ax, ay, bx, by, C.byref(p),
                                    _ptr(corr))
    if not ok:
        return None
    return (p, corr.reshape(n, n)) if want_corr else p


@dataclass
class Field:
    grid: tuple
    dx: np.ndarray
    dy: np.ndarray
    ppr: np.ndarray
    flags: np.ndarray
    peak_iy: np.ndarray | None = None
    peak_ix: np.ndarray | None = None
    dx64: np.ndarray | None = None
    dy64: np.ndarray | None = None

    def copy(self):
        return Field(self.grid, self.dx.copy(), self.dy.copy(), self.ppr.copy(),
                     self.flags.copy(), self.peak_iy, self.peak_ix, self.dx64, self.dy64)


DEFORM = {"none": 0, "bilinear": 1, "bicubic": 2}


def correlate_pass(pa: Prepared, pb: Prepared, g: Grid, origins, min_mean=0.0, min_std=0.0,
                   precision=32, threads=0, use_mask=True, deform=None) -> Field:
    """PivAnalyzer::pass given origins (multipass.rs:233-259).  deform: None (rigid windows,
    the reference) or (mode, prev_grid, prev_dx, prev_dy), mode 1 bilinear / 2 bicubic: the
    library's window deformation from the previous pass's validated field."""
    h, w = pa.image.shape
    cnt = len(g)
    dx = np.zeros(cnt, np.float32)
    dy = np.zeros_like(dx)
    ppr = np.zeros_like(dx)
    fl = np.zeros(cnt, np.uint8)
    piy = np.zeros(cnt, np.int32)
    pix = np.zeros(cnt, np.int32)
    d64 = np.zeros(cnt, np.float64) if precision == 64 else None
    e64 = np.zeros(cnt, np.float64) if precision == 64 else None
    sa = pa.sat if use_mask else (None, None)
    sb = pb.sat if use_mask else (None, None)
    lib().or_correlate_pass(pa.image.ravel(), pb.image.ravel(), _ptr(sa[0]), _ptr(sa[1]),
                            _ptr(sb[0]), _ptr(sb[1]), h, w, C.byref(g),
                            _f32(origins).ravel(), min_mean, min_std, precision, threads, dx, dy,
                            ppr, fl, _ptr(piy), _ptr(pix), _ptr(d64), _ptr(e64),
                            *_deform_args(deform))
    shp = (g.ny, g.nx)
    return Field(g.as_tuple(), dx.reshape(shp), dy.reshape(shp), ppr.reshape(shp),
                 fl.reshape(shp), piy.reshape(shp), pix.reshape(shp),
                 None if d64 is None else d64.reshape(shp),
                 None if e64 is None else e64.reshape(shp))


def _deform_args(deform):
    if deform is None or deform[0] == 0 or deform[1] is None:
        return 0, None, None, None
    mode, pg, pdx, pdy = deform
    pdx, pdy = _f32(pdx).ravel(), _f32(pdy).ravel()
    # the arrays must outlive the call: keep them on the function object for its duration
    _deform_args.keep = (pg, pdx, pdy)
    return (mode, C.cast(C.pointer(pg), C.c_void_p), pdx.ctypes.data_as(C.c_void_p),
            pdy.ctypes.data_as(C.c_void_p))


def deform_sample(g: Grid, dx, dy, x, y):
    u, v = C.c_float(), C.c_float()
    lib().or_deform_sample(C.byref(g), _f32(dx).ravel(), _f32(dy).ravel(), x, y, C.byref(u),
                           C.byref(v))
    return u.value, v.value


def deform_shifts(prev: Grid, pdx, pdy, n, ax, ay, bx, by):
    hx = np.zeros(n * n, np.float32)
    hy = np.zeros_like(hx)
    lib().or_deform_shifts(C.byref(prev), _f32(pdx).ravel(), _f32(pdy).ravel(), n, ax, ay, bx, by,
                           hx, hy)
    return hx.reshape(n, n), hy.reshape(n, n)


def load_window_deformed(img, x, y, n, mode, hx, hy, sgn, precision=32):
    img = _f32(img)
    h, w = img.shape
    dst = np.zeros(n * n, np.float64 if precision == 64 else np.float32)
    fn = lib().or_load_window_deformed64 if precision == 64 else lib().or_load_window_deformed
    e = fn(dst, img.ravel(), h, w, x, y, n, mode, _f32(hx).ravel(), _f32(hy).ravel(), sgn)
    return dst.reshape(n, n), e


def validate(f: Field, ppr_threshold=1.3, median_threshold=2.0, median_eps=0.1, replace=True,
             max_displacement=float("inf")):
    """validate.rs:51-130, in place on a copy; returns (field, summary dict)."""
    out = f.copy()
    ny, nx = out.dx.shape
    dx = np.ascontiguousarray(out.dx, np.float32).ravel().copy()
    dy = np.ascontiguousarray(out.dy, np.float32).ravel().copy()
    fl = np.ascontiguousarray(out.flags, np.uint8).ravel().copy()
    s = Summary()
    cfg = VCfg(ppr_threshold, median_threshold, median_eps, 1 if replace else 0)
    lib().or_validate(ny, nx, dx, dy, _f32(out.ppr).ravel(), fl, C.byref(cfg),
                      max_displacement, C.byref(s))
    out.dx, out.dy, out.flags = dx.reshape(ny, nx), dy.reshape(ny, nx), fl.reshape(ny, nx)
    return out, s.as_dict()


def median(vals):
    v = _f32(vals).copy()
    return lib().or_median(v, v.size)


@dataclass
class PassTrace:
    grid: tuple
    origins: np.ndarray
    raw: Field
    validated: Field


@dataclass
class PivConfig:
    windows: list = field(default_factory=lambda: [96, 64])
    overlap: float = 0.5
    final_iterations: int = 2
    min_mean: float = 0.0
    min_std: float = 0.0
    ppr_threshold: float = 1.3
    median_threshold: float = 2.0
    median_eps: float = 0.1
    replace: bool = True


def check_config(cfg: PivConfig):
    """PivAnalyzer::new checks (multipass.rs:91-119)."""
    if not cfg.windows:
        raise PivError("invalid configuration: at least one window size is required")
    for w in cfg.windows:
        if w < 8 or w % 2:
            raise PivError(f"invalid configuration: window size {w} must be even and >= 8")
    if not (0.0 <= np.float32(cfg.overlap) < 1.0):
        raise PivError(f"invalid configuration: overlap {cfg.overlap} must be in [0, 1)")
    if cfg.final_iterations < 1:
        raise PivError("invalid configuration: final_iterations must be at least 1")


def analyze(pa: Prepared, pb: Prepared, cfg: PivConfig, precision=32, threads=0,
            traced=False, deformation="none"):
    """PivAnalyzer::run (multipass.rs:152-192); deformation "bilinear" / "bicubic" deforms
    the windows of every pass after the first by the previous validated field (the library's
    definition, no counterpart in the reference)."""
    mode = DEFORM[deformation]
    check_config(cfg)
    if pa.image.shape != pb.image.shape:
        raise PivError(f"frames have different shapes: {pa.image.shape} vs {pb.image.shape}")
    h, w = pa.image.shape
    passes = list(cfg.windows) + [cfg.windows[-1]] * (cfg.final_iterations - 1)
    prev_grid, prev = None, None
    traces, summary = [], None
    for win in passes:
        g = grid_new(h, w, win, cfg.overlap)
        origins = place_windows(g, h, w, prev_grid, None if prev is None else prev.dx,
                                None if prev is None else prev.dy)
        dfm = None if prev is None else (mode, prev_grid, prev.dx, prev.dy)
        raw = correlate_pass(pa, pb, g, origins, cfg.min_mean, cfg.min_std, precision, threads,
                             deform=dfm)
        val, summary = validate(raw, cfg.ppr_threshold, cfg.median_threshold, cfg.median_eps,
                                cfg.replace, np.float32(cfg.windows[0]) / np.float32(2.0))
        if traced:
            traces.append(PassTrace(g.as_tuple(), origins, raw, val))
        prev_grid, prev = g, val
    return (prev, summary, traces) if traced else (prev, summary)


# ---- post-analysis stages (oracle/post_oracle.c) ----------------------------------------
def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def flow_usable(dx, flags):
    """flow.rs:13-17"""
    dx, fl = _f32(dx), _u8(flags)
    out = np.empty(dx.shape, np.uint8)
    lib().or_flow_usable(dx.size, dx, fl, out)
    return out.astype(bool)


def vorticity(u, v, usable, step):
    """flowstats.rs:48-66"""
    u, v = _f32(u), _f32(v)
    w = np.empty(u.shape, np.float32)
    lib().or_vorticity(u.shape[0], u.shape[1], u, v, _u8(usable), step, w)
    return w


def flow_stats(u, v, usable, step):
    """flowstats.rs:69-101"""
    u, v = _f32(u), _f32(v)
    s = FlowStat()
    lib().or_flow_stats(u.shape[0], u.shape[1], u, v, _u8(usable), step, C.byref(s))
    return s.as_dict()


def pair_stats(dx, dy, usable, step, dt_s):
    """flow.rs:20-34"""
    dx, dy = _f32(dx), _f32(dy)
    s = FlowStat()
    lib().or_pair_stats(dx.shape[0], dx.shape[1], dx, dy, _u8(usable), int(step), dt_s,
                        C.byref(s))
    return s.as_dict()


def neighbour_noise(u, v, usable):
    """flowstats.rs:106-137"""
    u, v = _f32(u), _f32(v)
    return float(lib().or_neighbour_noise(u.shape[0], u.shape[1], u, v, _u8(usable)))


def temporal(dx, dy, usable, gate, kernel, window, min_support, leave_out=()):
    """smooth.rs:44-130; kernel "mean" or "median" -> (dx, dy, support)"""
    dx, dy = _f32(dx), _f32(dy)
    nt, ny, nx = dx.shape
    lo = np.ascontiguousarray(list(leave_out) or [0], dtype=np.int32)
    ox, oy = np.empty_like(dx), np.empty_like(dy)
    sup = np.empty(dx.shape, np.uint16)
    lib().or_temporal(nt, ny, nx, dx, dy, _u8(usable), _u8(gate), 0 if kernel == "mean" else 1,
                      window, min_support, lo, len(leave_out), ox, oy, sup)
    return ox, oy, sup


def spatial(dx, dy, gate, sigma, min_weight, leave_out=None):
    """smooth.rs:137-208 -> (dx, dy, support)"""
    dx, dy = _f32(dx), _f32(dy)
    nt, ny, nx = dx.shape
    ox, oy = np.empty_like(dx), np.empty_like(dy)
    sup = np.empty(dx.shape, np.uint16)
    lib().or_spatial(nt, ny, nx, dx, dy, _u8(gate), sigma, min_weight,
                     -1 if leave_out is None else int(leave_out), ox, oy, sup)
    return ox, oy, sup


def cross_validation_errors(raw_dx, raw_dy, usable, pred_dx, pred_dy):
    """smooth.rs:213-232"""
    a = [_f32(x) for x in (raw_dx, raw_dy, pred_dx, pred_dy)]
    out = np.empty(a[0].size, np.float32)
    m = lib().or_cross_validation_errors(a[0].size, a[0], a[1], _u8(usable), a[2], a[3], out)
    return out[:m]


def median_error(errors, keep=None):
    """smooth.rs:236-249 -> (median, count)"""
    e = _f32(errors)
    k = None if keep is None else _u8(keep)
    n = C.c_int64()
    r = lib().or_median_error(e, e.size, _ptr(k), C.byref(n))
    return float(r), int(n.value)


def pair_to_pair_change(dx, dy):
    """smooth.rs:252-274"""
    dx, dy = _f32(dx), _f32(dy)
    return float(lib().or_pair_to_pair_change(*dx.shape, dx, dy))


def spatial_roughness(dx, dy):
    """smooth.rs:277-300"""
    dx, dy = _f32(dx), _f32(dy)
    return float(lib().or_spatial_roughness(*dx.shape, dx, dy))


def mask(values, usable):
    """smooth.rs:303-310"""
    v = _f32(values)
    out = np.empty_like(v)
    lib().or_mask(v.size, v, _u8(usable), out)
    return out

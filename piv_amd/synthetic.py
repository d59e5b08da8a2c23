"""Seeded synthetic particle / fog image pairs with a known displacement field.

Follows the semantics of piv-core/src/synthetic.rs (agoodman/piv): a SplitMix64 generator
(:10-36), Gaussian blobs rendered analytically out to 3.5 sigma (:48-62), displacement
evaluated at the blob's midpoint by three fixed-point iterations (:66-82), particles placed
over the image plus a 20 px margin (:85-101).  Vectorised with numpy; float32 throughout
like the reference, except that numpy's exp/log/cos may differ from libm by an ulp, so
images agree with the Rust generator to float32 rounding, not bit for bit.

Also provides the benchmark workloads (particle diameter model, u8/u16 quantisation, noise,
Lamb-Oseen vortex + shear, swirl) used by the benchmark and the tests.
"""
from __future__ import annotations

import math

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Outputs start..start+count of SplitMix64 seeded with `seed` (synthetic.rs:17-24)."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Uniform in [0, 1): top 24 bits / 2^24 (synthetic.rs:27-29)."""
    return ((splitmix64(seed, count, start) >> np.uint64(40)).astype(np.float32)
            / np.float32(1 << 24))


def normal(seed: int, count: int) -> np.ndarray:
    """Box-Muller from consecutive uniform pairs (synthetic.rs:32-36)."""
    u = uniform(seed, 2 * count)
    u1 = np.maximum(u[0::2], np.float32(1e-7))
    u2 = u[1::2]
    return (np.sqrt(np.float32(-2.0) * np.log(u1)) *
            np.cos(np.float32(2.0 * math.pi) * u2)).astype(np.float32)


def render(h: int, w: int, x, y, sigma, amp) -> np.ndarray:
    """Sum of Gaussian blobs rendered out to ceil(3.5 sigma) (synthetic.rs:48-62)."""
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    sigma = np.asarray(sigma, np.float32)
    amp = np.asarray(amp, np.float32)
    out = np.zeros((h, w), np.float32)
    if x.size == 0:
        return out
    r = np.ceil(np.float32(3.5) * sigma).astype(np.int64)
    cx = np.round(x).astype(np.int64)
    cy = np.round(y).astype(np.int64)
    k = np.float32(1.0) / (np.float32(2.0) * sigma * sigma)
    rmax = int(r.max())
    for oy in range(-rmax, rmax + 1):
        for ox in range(-rmax, rmax + 1):
            px, py = cx + ox, cy + oy
            ok = (abs(ox) <= r) & (abs(oy) <= r) & (px >= 0) & (px < w) & (py >= 0) & (py < h)
            if not ok.any():
                continue
            dx = px[ok].astype(np.float32) - x[ok]
            dy = py[ok].astype(np.float32) - y[ok]
            v = amp[ok] * np.exp(-(dx * dx + dy * dy) * k[ok])
            np.add.at(out, (py[ok], px[ok]), v.astype(np.float32))
    return out


def displace(x, y, field):
    """Midpoint displacement p_b = p_a + f((p_a + p_b)/2), 3 iterations (synthetic.rs:66-82)."""
    dx, dy = field(x, y)
    for _ in range(3):
        dx, dy = field(x + dx / np.float32(2.0), y + dy / np.float32(2.0))
    return (x + dx).astype(np.float32), (y + dy).astype(np.float32)


def particles(h, w, seed, density, sigma, margin=20.0):
    """Particles over the image plus margin (synthetic.rs:85-101): x, y, sigma, amp."""
    area = (np.float32(h) + np.float32(2 * margin)) * (np.float32(w) + np.float32(2 * margin))
    count = int(np.float32(density) * area)
    u = uniform(seed, 4 * count).reshape(count, 4)
    x = u[:, 0] * (np.float32(w) + np.float32(2 * margin)) - np.float32(margin)
    y = u[:, 1] * (np.float32(h) + np.float32(2 * margin)) - np.float32(margin)
    s = np.float32(sigma) * (np.float32(0.8) + np.float32(0.4) * u[:, 2])
    a = np.float32(0.3) + np.float32(0.7) * u[:, 3]
    return x, y, s, a


def particle_pair(h, w, seed, density, sigma, field):
    """synthetic.rs:120-131: float frames (values not clipped)."""
    x, y, s, a = particles(h, w, seed, density, sigma)
    xb, yb = displace(x, y, field)
    return render(h, w, x, y, s, a), render(h, w, xb, yb, s, a)


def add_noise(frame, sigma, seed):
    return (frame + np.float32(sigma) * normal(seed, frame.size).reshape(frame.shape)).astype(
        np.float32)


# ---- workloads -------------------------------------------------------------------------

def uniform_field(u, v):
    return lambda x, y: (np.full_like(x, np.float32(u)), np.full_like(y, np.float32(v)))


def lamb_oseen_shear(h, w, vmax=6.0, shear=0.02, rc=None):
    """Lamb-Oseen vortex centred in the image, scaled so max |u| = vmax px (at r = 1.1209 rc),
    plus a linear shear u_x += shear * (y - cy)."""
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    rc = rc if rc is not None else min(h, w) / 8.0
    rpk = 1.1209 * rc
    peak = (1.0 - math.exp(-(rpk / rc) ** 2)) / rpk
    gam = vmax / peak

    def f(x, y):
        xd = x.astype(np.float64) - cx
        yd = y.astype(np.float64) - cy
        r2 = xd * xd + yd * yd
        with np.errstate(invalid="ignore", divide="ignore"):
            ut = np.where(r2 > 0, gam * (1.0 - np.exp(-r2 / rc ** 2)) / np.maximum(r2, 1e-30), 0.0)
        u = -ut * yd + shear * yd
        v = ut * xd
        return u.astype(np.float32), v.astype(np.float32)

    return f


def quantize(frame_dn, bits=8, container=np.uint8):
    """Round to integer DN and clip to the sensor range."""
    hi = (1 << bits) - 1
    return np.clip(np.rint(frame_dn), 0, hi).astype(container)


def particle_frames(h, w, seed, density=0.02, diameter=2.5, amp=(150.0, 255.0), noise=3.0,
                    field=None, bits=8, container=np.uint8, amp_scale=1.0, diameter_range=None):
    """The benchmark particle model: Gaussian particles of image diameter d (sigma = d/4, e^-2
    convention), peak intensities uniform in `amp` DN, additive N(0, noise) DN, rounded and
    clipped to `bits` in `container`.  Frame B has the particles moved by `field`."""
    margin = 20.0
    area = (h + 2 * margin) * (w + 2 * margin)
    count = int(density * area)
    u = uniform(seed, 4 * count).reshape(count, 4)
    x = u[:, 0] * np.float32(w + 2 * margin) - np.float32(margin)
    y = u[:, 1] * np.float32(h + 2 * margin) - np.float32(margin)
    if diameter_range is not None:
        d = np.float32(diameter_range[0]) + np.float32(diameter_range[1] - diameter_range[0]) * u[:, 2]
    else:
        d = np.full(count, np.float32(diameter))
    s = (d / np.float32(4.0)).astype(np.float32)
    a = (np.float32(amp[0]) + np.float32(amp[1] - amp[0]) * u[:, 3]) * np.float32(amp_scale)
    fa = render(h, w, x, y, s, a)
    xb, yb = displace(x, y, field) if field is not None else (x, y)
    fb = render(h, w, xb, yb, s, a)
    fa = fa + np.float32(noise * amp_scale) * normal(seed ^ 0x5A5A, h * w).reshape(h, w)
    fb = fb + np.float32(noise * amp_scale) * normal(seed ^ 0xA5A5, h * w).reshape(h, w)
    return quantize(fa, bits, container), quantize(fb, bits, container)


def swirl_field(h, w, t, period=64.0, vmax=5.0):
    """Time-varying swirl at frame time t: a Lamb-Oseen vortex whose centre circles the image
    centre (radius h/8, one turn per `period` frames) and whose strength oscillates between
    0.5 vmax and vmax."""
    ang = 2.0 * math.pi * t / period
    cx = (w - 1) / 2.0 + (h / 8.0) * math.cos(ang)
    cy = (h - 1) / 2.0 + (h / 8.0) * math.sin(ang)
    rc = min(h, w) / 6.0
    rpk = 1.1209 * rc
    peak = (1.0 - math.exp(-(rpk / rc) ** 2)) / rpk
    gam = vmax * (0.75 + 0.25 * math.sin(ang)) / peak

    def f(x, y):
        xd = x.astype(np.float64) - cx
        yd = y.astype(np.float64) - cy
        r2 = xd * xd + yd * yd
        with np.errstate(invalid="ignore", divide="ignore"):
            ut = np.where(r2 > 0, gam * (1.0 - np.exp(-r2 / rc ** 2)) / np.maximum(r2, 1e-30), 0.0)
        return (-ut * yd).astype(np.float32), (ut * xd).astype(np.float32)

    return f


def video_frames(h, w, count, seed, density=0.015, diameter_range=(2.0, 3.0), noise=3.0,
                 bits=12, container=np.uint16, period=64.0, vmax=5.0):
    """A high-speed video: `count` frames of one particle population, each frame advected from
    the previous one by swirl_field at that frame's time (particles leaving the view are
    re-seeded on the far side, so the density stays constant).  Returns the frames and the
    field of each consecutive pair (pair k: frame k -> k+1 moves by swirl_field(t = k))."""
    margin = 20.0
    area = (h + 2 * margin) * (w + 2 * margin)
    n = int(density * area)
    u = uniform(seed, 4 * n).reshape(n, 4)
    x = u[:, 0] * np.float32(w + 2 * margin) - np.float32(margin)
    y = u[:, 1] * np.float32(h + 2 * margin) - np.float32(margin)
    d = np.float32(diameter_range[0]) + np.float32(diameter_range[1] - diameter_range[0]) * u[:, 2]
    s = (d / np.float32(4.0)).astype(np.float32)
    scale = np.float32((1 << bits) - 1) / np.float32(255.0)
    a = (np.float32(150.0) + np.float32(105.0) * u[:, 3]) * scale
    frames, fields = [], []
    for k in range(count):
        f = render(h, w, x, y, s, a)
        f = f + np.float32(noise) * scale * normal(seed ^ (0x9E37 + k), h * w).reshape(h, w)
        frames.append(quantize(f, bits, container))
        fld = swirl_field(h, w, float(k), period, vmax)
        fields.append(fld)
        x, y = displace(x, y, fld)
        wrapx = np.float32(w + 2 * margin)
        wrapy = np.float32(h + 2 * margin)
        x = ((x + np.float32(margin)) % wrapx - np.float32(margin)).astype(np.float32)
        y = ((y + np.float32(margin)) % wrapy - np.float32(margin)).astype(np.float32)
    return frames, fields[:-1]

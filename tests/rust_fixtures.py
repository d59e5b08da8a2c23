"""Loader for the golden cases written by the Rust reference.

The cases under tests/golden/rust_cases were written by piv-core's fixture example
(rust/crates/piv-core/examples/fixtures.rs:140-201 in agoodman/piv, MIT licence) and are the
only Rust-produced numbers available: frames, prepared images and, per pass, the window
origins and the raw and validated fields.  Layout (fixtures.rs:118-190): raw little-endian
row-major arrays named in meta.json.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass

import numpy as np

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rust_cases")


def names():
    with open(os.path.join(ROOT, "index.json")) as f:
        return json.load(f)


@dataclass
class FieldArrays:
    dx: np.ndarray
    dy: np.ndarray
    ppr: np.ndarray
    flags: np.ndarray


@dataclass
class PassData:
    grid: dict
    origins: np.ndarray  # [N, 4] ax, ay, bx, by
    raw: FieldArrays
    validated: FieldArrays


@dataclass
class Case:
    name: str
    height: int
    width: int
    frame_a: np.ndarray
    frame_b: np.ndarray
    background: np.ndarray | None
    prepared_a: np.ndarray
    prepared_b: np.ndarray
    preprocess: dict
    piv: dict
    passes: list
    summary: dict


def _load(d, fname, dtype, shape=None):
    a = np.fromfile(os.path.join(d, fname), dtype=np.dtype(dtype).newbyteorder("<"))
    a = a.astype(dtype)
    return a.reshape(shape) if shape is not None else a


def load(name) -> Case:
    d = os.path.join(ROOT, name)
    with open(os.path.join(d, "meta.json")) as f:
        meta = json.load(f)
    h, w = meta["height"], meta["width"]

    def field(spec, g):
        shp = (g["ny"], g["nx"])
        return FieldArrays(_load(d, spec["dx"], np.float32, shp), _load(d, spec["dy"], np.float32, shp),
                           _load(d, spec["peak_ratio"], np.float32, shp),
                           _load(d, spec["flags"], np.uint8, shp))

    passes = []
    for p in meta["passes"]:
        g = p["grid"]
        passes.append(PassData(g, _load(d, p["origins"], np.float32, (-1, 4)),
                               field(p["raw"], g), field(p["validated"], g)))
    bg = _load(d, meta["background"], np.float32, (h, w)) if meta.get("background") else None
    return Case(name, h, w, _load(d, meta["frame_a"], np.float32, (h, w)),
                _load(d, meta["frame_b"], np.float32, (h, w)), bg,
                _load(d, meta["prepared_a"], np.float32, (h, w)),
                _load(d, meta["prepared_b"], np.float32, (h, w)), meta["preprocess"], meta["piv"],
                passes, meta["summary"])


def f32(x):
    """JSON stores f32 config values widened to f64; narrow them back."""
    return float(np.float32(x))

"""The C-ABI library on the CPU: it loads, exports every entry point include/piv.h declares,
and its host-only calls (config defaults, Grid::new, error reporting) behave like piv-core.
No GPU compute here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import piv_amd
from piv_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "piv.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(piv_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L._SIGS), set(names) ^ set(L._SIGS)
    assert lib.piv_abi_version() == 2


def test_config_defaults_match_reference():
    """PivConfig/ValidationConfig/PreprocessConfig defaults (multipass.rs:53-63,
    validate.rs:28-37, preprocess.rs:21-29)."""
    c = L.piv_config()
    L.load().piv_config_default(C.byref(c))
    assert list(c.windows[:c.num_windows]) == [96, 64]
    assert c.overlap == np.float32(0.5) and c.final_iterations == 2
    assert (c.min_mean, c.min_std) == (0.0, 0.0)
    assert c.ppr_threshold == np.float32(1.3) and c.median_threshold == 2.0
    assert c.median_eps == np.float32(0.1) and c.replace == 1
    assert c.smooth_sigma == 1.0 and c.normalize_radius == 4
    assert c.normalize_eps == np.float32(0.02)
    d = piv_amd.PivConfig()
    assert d.windows == [96, 64] and d.final_iterations == 2


def test_grid_new_matches_oracle(oracle):
    rng = np.random.default_rng(0)
    for _ in range(300):
        h, w = rng.integers(8, 3000, 2)
        n = int(rng.choice([8, 12, 16, 24, 32, 48, 64, 96, 128]))
        ov = float(np.float32(rng.choice([0.0, 0.25, 0.5, 0.75, 0.6, 0.9])))
        try:
            ref = oracle.grid_new(int(h), int(w), n, ov).as_tuple()
        except oracle.PivError:
            with pytest.raises(piv_amd.PivError):
                piv_amd.Grid.new(int(h), int(w), n, ov)
            continue
        g = piv_amd.Grid.new(int(h), int(w), n, ov)
        assert (g.window, g.step, g.nx, g.ny, g.x0, g.y0) == ref


def test_grid_errors_are_typed():
    with pytest.raises(piv_amd.ImageTooSmall):
        piv_amd.Grid.new(16, 64, 32, 0.5)
    with pytest.raises(piv_amd.ConfigError):
        piv_amd.Grid.new(64, 64, 32, 1.0)
    with pytest.raises(piv_amd.ConfigError):
        piv_amd.Grid.new(64, 64, 32, -0.1)


def test_bad_configs_rejected_before_device():
    """multipass.rs:477-507: every bad config is refused (ConfigError), GPU or not."""
    for cfg in (piv_amd.PivConfig(windows=[]), piv_amd.PivConfig(windows=[33]),
                piv_amd.PivConfig(windows=[6]), piv_amd.PivConfig(overlap=1.0),
                piv_amd.PivConfig(final_iterations=0)):
        with pytest.raises(piv_amd.ConfigError):
            piv_amd.PivAnalyzer(cfg)


def test_no_device_is_a_device_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(piv_amd.DeviceError):
        piv_amd.PivAnalyzer(piv_amd.PivConfig(windows=[32]))


def test_kernel_variant_counts():
    """Every even window size from 8 to 128 has at least one correlation-kernel variant (the
    generic-size kernel has one); other sizes none."""
    import piv_amd
    for n in range(8, 129, 2):
        assert 1 <= piv_amd.kernel_variant_count(n) <= 4, n
    for n in (10, 14, 20, 28, 40, 56):  # sizes with radix-5 / radix-7 stages
        assert piv_amd.kernel_variant_count(n) >= 2, n
    for n in (0, 6, 9, 130, 256):
        assert piv_amd.kernel_variant_count(n) == 0, n

"""Window deformation between passes on the GPU (include/piv.h, piv_deformation) against the
oracle's restatement (tests/test_oracle_deform.py pins the oracle):
  - a uniform predictor deforms nothing: bilinear deformation equals the rigid pass bitwise;
  - the 1024x1024 three-pass vortex + shear workload (64 -> 32 -> 16 px), bilinear and
    bicubic: every pass replayed through the oracle from the GPU's own origins and predictor
    (integer peaks identical, displacements within 1e-4 px, validation bit-exact), then the
    final field against an independent oracle run (flags identical, within 1e-3 px); the
    error against the generating field falls below the rigid method's;
  - packed 48 / 24 px windows (u8 and u16), every pass replayed through the oracle;
  - the batched, allocation-free and served paths give the single-pair result bitwise;
  - DFT-GEMM correlation (fp16 / bf16) with deformation holds RMS <= 0.02 px against the FP64
    oracle with the same deformation."""
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


MODES = {"bilinear": 1, "bicubic": 2}


@pytest.fixture(scope="module")
def vortex(oracle):
    import bench

    pairs, field = bench.make_pairs(2)
    return bench, pairs, field


def truth_rms(bench, f, g, field):
    return bench.truth_rms(f.dx, f.dy, f.flags, g, field)[0]


def test_uniform_predictor_is_rigid_bitwise(piv):
    """A 64x64 pair with windows [64, 32]: the first pass is one vector, so the predictor of the
    second pass is uniform and bilinear deformation must give the rigid result bitwise."""
    for seed in (3, 4, 5):
        a, b = S.particle_frames(64, 64, seed, density=0.04, field=S.uniform_field(1.7, -0.6))
        out = []
        for mode in ("none", "bilinear"):
            an = piv.PivAnalyzer(piv.PivConfig(windows=[64, 32], final_iterations=1,
                                               deformation=mode), piv.PreprocessConfig.raw())
            out.append(an.analyze(an.prepare(a), an.prepare(b)).field)
        assert out[0].dx.shape == (3, 3) and np.isfinite(out[0].dx).all()
        assert np.array_equal(bits(out[1].dx), bits(out[0].dx)), seed
        assert np.array_equal(bits(out[1].dy), bits(out[0].dy)), seed
        assert np.array_equal(bits(out[1].peak_ratio), bits(out[0].peak_ratio)), seed


def deformed_stage_parity(piv, oracle, an, cfg, a, b, mode, tol=1e-4):
    """Replay every pass of a traced deformed GPU run through the oracle (from the GPU's own
    origins and previous validated field), then compare the final field with an independent
    oracle run.  Returns the GPU result."""
    res, traces = an.analyze_traced(an.prepare(a), an.prepare(b))
    pa, pb = oracle.prepare(a), oracle.prepare(b)
    h, w = a.shape
    maxd = np.float32(cfg.windows[0]) / np.float32(2)
    v = cfg.validation
    prev_g = None
    for i, t in enumerate(traces):
        g = t.raw.grid
        og = oracle.grid_new(h, w, g.window, cfg.overlap)
        deform = None
        if i > 0:
            deform = (MODES[mode], prev_g, traces[i - 1].validated.dx, traces[i - 1].validated.dy)
        ref = oracle.correlate_pass(pa, pb, og, t.origins, cfg.mask.min_mean, cfg.mask.min_std,
                                    deform=deform)
        assert np.array_equal(t.raw.flags, ref.flags), f"pass {i} NO_SIGNAL"
        ref_pk = np.where(ref.peak_iy >= 0, ref.peak_iy * g.window + ref.peak_ix, -1)
        mism = np.count_nonzero(t.peak_index != ref_pk)
        assert mism == 0, f"pass {i}: {mism} integer peaks differ"
        fin = np.isfinite(ref.dx)
        e = max(np.abs(t.raw.dx - ref.dx)[fin].max(initial=0),
                np.abs(t.raw.dy - ref.dy)[fin].max(initial=0))
        assert e < tol, (i, e)
        f = oracle.Field(None, t.raw.dx, t.raw.dy, t.raw.peak_ratio, t.raw.flags)
        val, s = oracle.validate(f, v.ppr_threshold, v.median_threshold, v.median_eps, v.replace,
                                 maxd)
        assert np.array_equal(val.flags, t.validated.flags)
        assert np.array_equal(bits(val.dx), bits(t.validated.dx))
        print(f"{mode} pass {i} (n = {g.window}): max |raw d - oracle| = {e:.2e} px")
        prev_g = og
    ocfg = oracle.PivConfig(windows=list(cfg.windows), overlap=cfg.overlap,
                            final_iterations=cfg.final_iterations, min_mean=cfg.mask.min_mean,
                            min_std=cfg.mask.min_std)
    ref, s = oracle.analyze(pa, pb, ocfg, deformation=mode)
    assert np.array_equal(res.field.flags, ref.flags)
    fin = np.isfinite(ref.dx)
    assert np.abs(res.field.dx - ref.dx)[fin].max(initial=0) < 1e-3
    assert np.abs(res.field.dy - ref.dy)[fin].max(initial=0) < 1e-3
    assert s == res.summary.__dict__
    return res


@pytest.mark.parametrize("mode", ["bilinear", "bicubic"])
def test_vortex_1024_stage_parity(piv, oracle, vortex, mode):
    bench, pairs, field = vortex
    a, b = pairs[0]
    cfg = piv.PivConfig(windows=[64, 32, 16], overlap=0.5, final_iterations=1, deformation=mode)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    res = deformed_stage_parity(piv, oracle, an, cfg, a, b, mode)
    rigid = piv.PivAnalyzer(piv.PivConfig(windows=[64, 32, 16], final_iterations=1),
                            piv.PreprocessConfig.raw())
    r0 = rigid.analyze(rigid.prepare(a), rigid.prepare(b))
    g = res.field.grid
    e_def, e_rig = truth_rms(bench, res.field, g, field), truth_rms(bench, r0.field, g, field)
    print(f"{mode}: rms vs truth {e_def:.4f} px, rigid {e_rig:.4f} px")
    assert e_def < 0.9 * e_rig


@pytest.mark.parametrize("mode", ["bilinear", "bicubic"])
@pytest.mark.parametrize("container", [np.uint8, np.uint16])
def test_packed_windows_stage_parity(piv, oracle, mode, container):
    """Windows of 48 and 24 px run in packed three-wave workgroups (four and eight windows,
    some across a wave boundary): their deformed passes against the oracle, u8 and 12-bit
    u16 samples."""
    h, w = 256, 320
    frames, _ = S.video_frames(h, w, 2, 13, bits=12 if container is np.uint16 else 8,
                               container=container)
    cfg = piv.PivConfig(windows=[48, 24], overlap=0.5, final_iterations=2, deformation=mode)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    deformed_stage_parity(piv, oracle, an, cfg, frames[0], frames[1], mode)


@pytest.mark.parametrize("mode", ["bilinear", "bicubic"])
def test_batch_noalloc_serving_equal_single(piv, mode):
    from piv_amd import _lib as L
    import ctypes as C

    h, w, n = 256, 320, 4
    frames, _ = S.video_frames(h, w, n + 1, 9, bits=8, container=np.uint8)
    cfg = piv.PivConfig(windows=[48, 24], overlap=0.5, final_iterations=2, deformation=mode)
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    singles = [an.analyze(an.prepare(frames[k]), an.prepare(frames[k + 1])) for k in range(n)]
    clip = piv.DeviceArray.from_numpy(np.stack(frames))
    fs = piv.PreparedFrame.from_device_batch(an, clip)
    g = an.output_grid(h, w)
    dx, dy, pr = (piv.DeviceArray((n, g.ny, g.nx), np.float32) for _ in range(3))
    fl = piv.DeviceArray((n, g.ny, g.nx), np.uint8)
    an.analyze_batch_device(fs[:-1], fs[1:], dx, dy, pr, fl)
    bdx = dx.numpy()
    lib = L.load()
    rc = lib.piv_analyze_batch_noalloc(an._h, C.c_void_p(clip.device_ptr),
                                       C.c_void_p(clip.device_ptr + h * w), L.PIXEL_U8, n, h, w,
                                       w, h * w, L.ptr(dx), L.ptr(dy), L.ptr(pr), L.ptr(fl), None,
                                       None)
    assert rc == 0, lib.piv_last_error()
    an.synchronize()
    ndx = dx.numpy()
    srv = piv.PivServer(hbm_budget=1 << 30)
    try:
        sid = srv.add_stream(cfg, h, w, np.uint8, piv.PreprocessConfig.raw())
        ids = [srv.submit(sid, frames[k], frames[k + 1]) for k in range(n)]
        got = {}
        while len(got) < n:
            for r in srv.poll(16, wait_ms=2000):
                got[r.request_id] = r
    finally:
        srv.close()
    for k in range(n):
        ref = singles[k].field
        assert np.array_equal(bits(bdx[k]), bits(ref.dx)), k
        assert np.array_equal(bits(ndx[k]), bits(ref.dx)), k
        assert np.array_equal(bits(got[ids[k]].field.dx), bits(ref.dx)), k
        assert np.array_equal(got[ids[k]].field.flags, ref.flags), k


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_dft_gemm_with_deformation_against_fp64_oracle(piv, oracle, vortex, precision):
    bench, pairs, field = vortex
    a, b = pairs[1]
    cfg = piv.PivConfig(windows=[64, 32, 16], overlap=0.5, final_iterations=1,
                        deformation="bilinear")
    an = piv.PivAnalyzer(cfg, piv.PreprocessConfig.raw())
    an.set_precision(precision)
    r = an.analyze(an.prepare(a), an.prepare(b)).field
    ocfg = oracle.PivConfig(windows=[64, 32, 16], overlap=0.5, final_iterations=1)
    ref, _ = oracle.analyze(oracle.prepare(a), oracle.prepare(b), ocfg, precision=64,
                            deformation="bilinear")
    both = ((r.flags & 23) == 0) & ((ref.flags & 23) == 0)
    assert both.mean() > 0.9
    d2 = (r.dx - ref.dx)[both] ** 2 + (r.dy - ref.dy)[both] ** 2
    rms, mx = float(np.sqrt(d2.mean())), float(np.sqrt(d2.max()))
    rt = truth_rms(bench, r, r.grid, field)
    print(f"{precision} + bilinear deformation: rms {rms:.5f} max {mx:.4f} px vs FP64 oracle; "
          f"rms {rt:.4f} px vs truth")
    assert rms <= 0.02

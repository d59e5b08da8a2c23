"""Serving loop (continuous batching) on the GPU: every served result equals the single-pair
analysis of the same pair and configuration, bitwise (pairs are independent in the reference,
multipass.rs:136-150), whatever batch and streams it ran with; admission follows the HBM
budget; repeated launch structures replay captured graphs with the same results."""
import numpy as np
import pytest

import piv_amd.synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def piv():
    import piv_amd

    if piv_amd.device_count() < 1:
        pytest.skip("no GPU")
    return piv_amd


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _streams(piv):
    """Four streams modelled on the mixed serving workload, at test sizes."""
    raw = piv.PreprocessConfig.raw()
    bg = (np.full((192, 256), 3.0, np.float32) / np.float32(65535)).astype(np.float32)
    return [
        (piv.PivConfig(windows=[32], final_iterations=1), raw, 256, 256, np.uint8, None, 0),
        (piv.PivConfig(windows=[64, 32, 16], final_iterations=1), raw, 320, 384, np.uint8, None, 0),
        (piv.PivConfig(windows=[48, 24], final_iterations=1), raw, 200, 320, np.uint16, None, 12),
        (piv.PivConfig(windows=[64, 32], final_iterations=2), piv.PreprocessConfig(), 192, 256,
         np.uint16, bg, 12),
    ]


def _pair(h, w, dtype, bits_, seed):
    field = S.lamb_oseen_shear(h, w, vmax=4.0)
    if dtype is np.uint8:
        return S.particle_frames(h, w, seed, field=field)
    return S.particle_frames(h, w, seed, field=field, bits=bits_, container=np.uint16, amp_scale=16.0)


def _alone(piv, cfg, pre, a, b, bg):
    an = piv.PivAnalyzer(cfg, pre)
    return an.analyze(an.prepare(a, bg), an.prepare(b, bg))


def _same(res, ref):
    f = res.field
    assert np.array_equal(bits(f.dx), bits(ref.field.dx))
    assert np.array_equal(bits(f.dy), bits(ref.field.dy))
    assert np.array_equal(bits(f.peak_ratio), bits(ref.field.peak_ratio))
    assert np.array_equal(f.flags, ref.field.flags)
    assert res.summary == ref.summary


def _run_mixed(piv, srv, specs, sids, rounds, seed0, paused):
    jobs = []
    for rnd in range(rounds):
        for si, (cfg, pre, h, w, dt, bg, b) in enumerate(specs):
            jobs.append((si,) + _pair(h, w, dt, b, seed0 + 10 * rnd + si))
    if paused:
        srv.pause(True)
    pairs = {}
    for si, a_, b_ in jobs:
        pairs[srv.submit(sids[si], a_, b_)] = (si, a_, b_)
    if paused:
        srv.pause(False)
    got = {}
    while len(got) < len(pairs):
        res = srv.poll(64, wait_ms=10000)
        assert res, "no result within 10 s"
        for r in res:
            assert r.error is None, r.error
            got[r.request_id] = r
    for rid, (si, a_, b_) in pairs.items():
        cfg, pre, h, w, dt, bg, _ = specs[si]
        r = got[rid]
        assert r.stream_id == sids[si]
        assert r.latency_ms >= 0
        _same(r, _alone(piv, cfg, pre, a_, b_, bg))
    return got


def test_mixed_streams_equal_single_pair_analysis(piv):
    """A burst of 20 requests from four streams, held and released together, runs as two
    mixed batches (16 + 4); then a free-running burst on two lanes.  Every result equals the
    pair analysed alone."""
    srv = piv.PivServer(max_batch_pairs=16, max_inflight=2)
    specs = _streams(piv)
    sids = [srv.add_stream(cfg, h, w, dt, pre, bg) for cfg, pre, h, w, dt, bg, _ in specs]
    got = _run_mixed(piv, srv, specs, sids, 5, 1000, paused=True)
    assert sorted(r.batch_pairs for r in got.values())[-1] == 16
    got2 = _run_mixed(piv, srv, specs, sids, 5, 2000, paused=False)
    st = srv.stats()
    assert st["submitted"] == st["completed"] == len(got) + len(got2)
    assert st["rejected"] == 0
    srv.close()


def test_repeated_structure_replays_graph(piv):
    """Same stream, one request at a time: every batch has the same launch structure, so the
    first batch runs eagerly, the second is captured and later ones replay that graph; eager
    single-pair work runs between the replays and every replay matches it."""
    srv = piv.PivServer(max_batch_pairs=4, max_inflight=1)
    cfg = piv.PivConfig(windows=[64, 32, 16], final_iterations=1)
    sid = srv.add_stream(cfg, 256, 320, np.uint8, piv.PreprocessConfig.raw())
    for k in range(5):
        a, b = S.particle_frames(256, 320, 70 + k, field=S.uniform_field(1.0 + 0.3 * k, -0.4))
        srv.submit(sid, a, b)
        res = srv.poll(4, wait_ms=5000)
        assert len(res) == 1 and res[0].error is None
        _same(res[0], _alone(piv, cfg, piv.PreprocessConfig.raw(), a, b, None))
    assert srv.stats()["batches"] == 5
    srv.close()


def test_admission_by_hbm_budget(piv):
    cfg = piv.PivConfig(windows=[32], final_iterations=1)
    raw = piv.PreprocessConfig.raw()
    a, b = S.particle_frames(256, 256, 5, field=S.uniform_field(1.5, 0.5))
    tiny = piv.PivServer(hbm_budget=64 << 10)
    sid = tiny.add_stream(cfg, 256, 256, np.uint8, raw)
    with pytest.raises(piv.AdmissionError):
        tiny.submit(sid, a, b)
    assert tiny.stats()["rejected"] == 1
    tiny.close()
    # room for exactly one request slot: the second submission waits for the first release
    probe = piv.PivServer()
    psid = probe.add_stream(cfg, 256, 256, np.uint8, raw)
    probe.submit(psid, a, b)
    probe.poll(1, wait_ms=5000)
    one = probe.stats()["hbm_peak"]
    probe.close()
    srv = piv.PivServer(hbm_budget=one)
    sid = srv.add_stream(cfg, 256, 256, np.uint8, raw)
    srv.submit(sid, a, b)
    with pytest.raises(piv.AdmissionError):
        srv.submit(sid, a, b)
    res = srv.poll(1, wait_ms=5000)  # poll releases the result and its slot
    assert len(res) == 1 and res[0].error is None
    srv.submit(sid, a, b)
    assert len(srv.poll(1, wait_ms=5000)) == 1
    assert srv.stats()["hbm_peak"] <= one
    srv.close()


def test_device_submission_equals_host_submission(piv):
    """Frames already in HBM (read in place) give the host-staged result bitwise."""
    srv = piv.PivServer(max_batch_pairs=8)
    cfg = piv.PivConfig(windows=[48, 24], final_iterations=1)
    raw = piv.PreprocessConfig.raw()
    sid = srv.add_stream(cfg, 200, 320, np.uint16, raw)
    a, b = _pair(200, 320, np.uint16, 12, 4242)
    da, db = piv.DeviceArray.from_numpy(a), piv.DeviceArray.from_numpy(b)
    srv.pause(True)
    r_host = srv.submit(sid, a, b)
    r_dev = srv.submit_device(sid, da, db)
    srv.pause(False)
    got = {}
    while len(got) < 2:
        for r in srv.poll(4, wait_ms=10000):
            assert r.error is None
            got[r.request_id] = r
    _same(got[r_dev], _alone(piv, cfg, raw, a, b, None))
    _same(got[r_host], _alone(piv, cfg, raw, a, b, None))
    srv.close()


def test_device_batch_submission(piv):
    """Several device pairs submitted in one call: every result equals the pair analysed
    alone; under an HBM budget for one request slot only the first is admitted, and the rest
    go in after its result is released."""
    cfg = piv.PivConfig(windows=[48, 24], final_iterations=1)
    raw = piv.PreprocessConfig.raw()
    pairs = [_pair(200, 320, np.uint16, 12, 5000 + i) for i in range(5)]
    dev = [(piv.DeviceArray.from_numpy(a), piv.DeviceArray.from_numpy(b)) for a, b in pairs]
    srv = piv.PivServer(max_batch_pairs=8)
    sid = srv.add_stream(cfg, 200, 320, np.uint16, raw)
    ids = srv.submit_device_many([(sid, da, db) for da, db in dev])
    assert len(ids) == 5
    got = {}
    while len(got) < 5:
        res = srv.poll(8, wait_ms=10000)
        assert res, "no result within 10 s"
        for r in res:
            assert r.error is None, r.error
            got[r.request_id] = r
    for rid, (a, b) in zip(ids, pairs):
        _same(got[rid], _alone(piv, cfg, raw, a, b, None))
    srv.close()
    # one request slot's worth of HBM: the first request is admitted, the next refused
    probe = piv.PivServer()
    psid = probe.add_stream(cfg, 200, 320, np.uint16, raw)
    probe.submit_device(psid, *dev[0])
    probe.poll(1, wait_ms=5000)
    one = probe.stats()["hbm_peak"]
    probe.close()
    srv = piv.PivServer(hbm_budget=one)
    sid = srv.add_stream(cfg, 200, 320, np.uint16, raw)
    first = srv.submit_device_many([(sid, da, db) for da, db in dev[:3]])
    assert len(first) == 1
    res = srv.poll(1, wait_ms=10000)
    assert len(res) == 1 and res[0].error is None and res[0].request_id == first[0]
    _same(res[0], _alone(piv, cfg, raw, *pairs[0], None))
    rest = srv.submit_device_many([(sid, da, db) for da, db in dev[1:2]])
    assert len(rest) == 1
    assert len(srv.poll(1, wait_ms=10000)) == 1
    srv.close()


def test_concurrent_host_submission(piv):
    """Four client threads submit host frames to the same streams at once: the staging copies
    of different requests run concurrently, each into its own request slot, and every result
    equals the pair analysed alone."""
    import threading

    srv = piv.PivServer(max_batch_pairs=8, max_inflight=2)
    specs = _streams(piv)
    sids = [srv.add_stream(cfg, h, w, dt, pre, bg) for cfg, pre, h, w, dt, bg, _ in specs]
    jobs = []
    for rnd in range(6):
        for si, (cfg, pre, h, w, dt, bg, b) in enumerate(specs):
            jobs.append((si,) + _pair(h, w, dt, b, 3000 + 10 * rnd + si))
    pairs, lock, errors = {}, threading.Lock(), []

    def client(part):
        try:
            for si, a_, b_ in part:
                rid = srv.submit(sids[si], a_, b_)
                with lock:
                    pairs[rid] = (si, a_, b_)
        except Exception as e:  # surfaced below
            errors.append(e)

    ths = [threading.Thread(target=client, args=(jobs[j::4],)) for j in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errors, errors
    got = {}
    while len(got) < len(jobs):
        res = srv.poll(64, wait_ms=10000)
        assert res, "no result within 10 s"
        for r in res:
            assert r.error is None, r.error
            got[r.request_id] = r
    assert sorted(got) == sorted(pairs)
    for rid, (si, a_, b_) in pairs.items():
        cfg, pre, h, w, dt, bg, _ = specs[si]
        _same(got[rid], _alone(piv, cfg, pre, a_, b_, bg))
    srv.close()


def test_free_slots_move_between_streams_only_when_idle(piv):
    """Over the HBM budget, a stream's cached free slots go to another stream only while the
    stream holds more free slots than requests in flight; otherwise the submission is refused
    (and succeeds once results are released), so a saturated mix does not reallocate slots
    back and forth."""
    cfg = piv.PivConfig(windows=[32], final_iterations=1)
    raw = piv.PreprocessConfig.raw()
    a, b = S.particle_frames(256, 256, 5, field=S.uniform_field(1.5, 0.5))
    probe = piv.PivServer()
    psid = probe.add_stream(cfg, 256, 256, np.uint8, raw)
    probe.submit(psid, a, b)
    probe.poll(1, wait_ms=5000)
    one = probe.stats()["hbm_peak"]
    probe.close()

    def drain(srv, n):
        got = []
        while len(got) < n:
            res = srv.poll(8, wait_ms=5000)
            assert res, "no result within 5 s"
            assert all(r.error is None for r in res)
            got += res
        return got

    srv = piv.PivServer(hbm_budget=3 * one)
    sa = srv.add_stream(cfg, 256, 256, np.uint8, raw)
    sb = srv.add_stream(cfg, 256, 256, np.uint8, raw)
    srv.pause(True)
    for _ in range(3):
        srv.submit(sa, a, b)
    with pytest.raises(piv.AdmissionError):  # stream a holds the budget, all in flight
        srv.submit(sb, a, b)
    srv.pause(False)
    drain(srv, 3)  # stream a: three free slots, none in flight
    srv.submit(sb, a, b)  # takes one of them
    srv.pause(True)
    srv.submit(sa, a, b)  # stream a: one free slot, one request in flight
    with pytest.raises(piv.AdmissionError):
        srv.submit(sb, a, b)
    srv.pause(False)
    drain(srv, 2)
    srv.submit(sb, a, b)
    srv.submit(sb, a, b)  # stream a is idle again: its free slot moves to stream b
    drain(srv, 2)
    st = srv.stats()
    assert st["rejected"] == 2
    assert st["hbm_peak"] <= 3 * one
    srv.close()


def test_unreleased_results_do_not_pin_free_slots(piv):
    """A client that polls results through the C ABI and keeps some of them unreleased: those
    results hold their slots, but they are not in flight, so the stream's spare cached slots
    still go to another stream over the budget."""
    import ctypes as C

    from piv_amd import _lib as L

    lib = L.load()
    cfg = piv.PivConfig(windows=[32], final_iterations=1)
    raw = piv.PreprocessConfig.raw()
    a, b = S.particle_frames(256, 256, 6, field=S.uniform_field(1.0, 0.5))
    probe = piv.PivServer()
    psid = probe.add_stream(cfg, 256, 256, np.uint8, raw)
    probe.submit(psid, a, b)
    probe.poll(1, wait_ms=5000)
    one = probe.stats()["hbm_peak"]
    probe.close()

    srv = piv.PivServer(hbm_budget=4 * one)
    sa = srv.add_stream(cfg, 256, 256, np.uint8, raw)
    sb = srv.add_stream(cfg, 256, 256, np.uint8, raw)
    srv.pause(True)
    ids = [srv.submit(sa, a, b) for _ in range(4)]  # the whole budget, in stream a's slots
    srv.pause(False)
    arr = (L.piv_result * 8)()
    got = []
    while len(got) < 4:
        cnt = C.c_uint32()
        assert lib.piv_poll(srv._h, arr, 8, C.byref(cnt), 5000) == 0
        assert cnt.value > 0, "no result within 5 s"
        got += [arr[i].request_id for i in range(cnt.value)]
    assert sorted(got) == sorted(ids)
    for rid in ids[:2]:  # two slots back to stream a's cache; two results kept
        assert lib.piv_result_release(srv._h, rid) == 0
    rid = srv.submit(sb, a, b)  # admitted from stream a's spare cached slots
    res = srv.poll(1, wait_ms=5000)
    assert res and res[0].request_id == rid and res[0].error is None
    for r in ids[2:]:
        assert lib.piv_result_release(srv._h, r) == 0
    st = srv.stats()
    assert st["rejected"] == 0
    assert st["hbm_peak"] <= 4 * one
    srv.close()


def test_large_host_pairs_staged_with_helpers(piv):
    """Host pairs of 3 MiB and more are staged by the submitting thread and the server's
    staging helpers together, in row chunks of about 1 MiB (the last one partial here), from
    two client threads at once: every result equals the pair analysed alone."""
    import threading

    h, w = 1100, 1536  # 3.2 MiB per u16 frame
    cfg = piv.PivConfig(windows=[64, 32], final_iterations=1)
    raw = piv.PreprocessConfig.raw()
    srv = piv.PivServer(max_batch_pairs=4)
    sid = srv.add_stream(cfg, h, w, np.uint16, raw)
    pairs = [_pair(h, w, np.uint16, 12, 500 + k) for k in range(4)]
    ids, lock = {}, threading.Lock()

    def client(ks):
        for k in ks:
            rid = srv.submit(sid, *pairs[k])
            with lock:
                ids[rid] = k

    ths = [threading.Thread(target=client, args=(ks,)) for ks in ([0, 2], [1, 3])]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    got = {}
    while len(got) < len(pairs):
        res = srv.poll(8, wait_ms=10000)
        assert res, "no result within 10 s"
        for r in res:
            assert r.error is None, r.error
            got[r.request_id] = r
    srv.close()
    for rid, k in ids.items():
        _same(got[rid], _alone(piv, cfg, raw, *pairs[k], None))

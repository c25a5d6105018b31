"""Continuous-batching server: ring arena (CPU) and served == alone (GPU)."""
import collections
import random

import numpy as np
import pytest


def test_ring_never_overlaps_live_data():
    from astro_transit_gpu_amd.serving import _Ring
    rnd = random.Random(0)
    for trial in range(200):
        cap = rnd.randint(10, 200)
        r = _Ring(cap)
        live = collections.deque()
        occ = [False] * cap
        for step in range(400):
            if live and rnd.random() < 0.45:
                s0, n = live.popleft()
                r.free_oldest()
                for i in range(s0, s0 + n):
                    occ[i] = False
            else:
                n = rnd.randint(0, cap // 3)
                st = r.plan(n)
                if st is not None:
                    assert st + n <= cap
                    assert not any(occ[st:st + n])
                    r.commit(st, n)
                    live.append((st, n))
                    for i in range(st, st + n):
                        occ[i] = True


def _requests(seed, k):
    from astro_transit_gpu_amd import synth
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        t, y, dy = synth.ragged_curve(rng, 400, 5000, 5.0, 27.0)
        nd = int(rng.integers(1, 7))
        durs = np.linspace(rng.uniform(0.02, 0.05), rng.uniform(0.1, 0.22), nd)
        per = synth.capped_grid(t, durs, 40, 300)
        out.append((t, y.astype(np.float32), dy.astype(np.float32), per, durs))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["parity", "fast"])
def test_served_in_mix_equals_alone(mode):
    """Each request's result is bit-identical whether it is served in a
    continuous mix (small arena: wrap-around, admission waits, several
    batches in flight) or computed alone."""
    from astro_transit_gpu_amd import Context, DeviceBatch
    from astro_transit_gpu_amd.serving import BLSServer
    from gpu_util import assert_parity
    from oracle import oracle as O
    reqs = _requests(31, 24)
    srv = BLSServer(objective="likelihood", mode=mode,
                    capacity_points=40000, capacity_periods=2500,
                    capacity_durations=200, max_batch=5)
    handles = []
    for i, r in enumerate(reqs):
        handles.append(srv.submit(*r))
        if i % 3 == 2:
            srv.poll()
    srv.drain()
    ctx = Context(0)
    for i, (h, r) in enumerate(zip(handles, reqs)):
        assert h.status == 0
        b = DeviceBatch.from_host([r[0]], [r[1]], [r[2]], r[3], r[4],
                                  device=ctx.device)
        alone = ctx.run(b, objective="likelihood", mode=mode).to_host()
        for k in h.result:
            assert np.array_equal(h.result[k], alone[k]), (i, k)
        assert h.best_index == int(alone["best_index"][0])
        if mode == "parity" and i < 6:
            ref = O.power(r[0], r[1], r[2], r[3], r[4], "likelihood")
            assert_parity(h.result, ref, "request %d" % i)


@pytest.mark.gpu
def test_served_large_rounds_equal_alone():
    """Rounds of admitted requests large enough for the threaded staging
    (tens of MB of points in one pinned buffer, copied in pieces by worker
    threads): every request's result equals its alone run bit for bit."""
    from astro_transit_gpu_amd import Context, DeviceBatch, synth
    from astro_transit_gpu_amd.serving import BLSServer
    rng = np.random.default_rng(17)
    reqs = []
    for _ in range(12):
        t, y, dy = synth.ragged_curve(rng, 120000, 200000, 20.0, 60.0)
        durs = np.array([0.05, 0.1, 0.2])
        per = synth.capped_grid(t, durs, 40, 200)
        reqs.append((t, y.astype(np.float32), dy.astype(np.float32), per,
                     durs))
    srv = BLSServer(objective="likelihood", max_batch=4,
                    capacity_points=4000000, capacity_periods=10000,
                    capacity_durations=200)
    handles = [srv.submit(*r) for r in reqs]  # one round of ~40 MB
    srv.drain()
    srv.close()
    ctx = Context(0)
    for i, (h, r) in enumerate(zip(handles, reqs)):
        assert h.status == 0
        b = DeviceBatch.from_host([r[0]], [r[1]], [r[2]], r[3], r[4],
                                  device=ctx.device)
        alone = ctx.run(b, objective="likelihood").to_host()
        for k in h.result:
            assert np.array_equal(h.result[k], alone[k]), (i, k)
        assert h.best_index == int(alone["best_index"][0])


@pytest.mark.gpu
def test_served_mixed_weighting_equals_alone():
    """Requests with and without per-point dy, interleaved: batches split
    where the weighting changes and the dy ring holds only the requests
    that have one; every result equals its alone run bit for bit."""
    from astro_transit_gpu_amd import Context, DeviceBatch
    from astro_transit_gpu_amd.serving import BLSServer
    reqs = _requests(12, 41)
    reqs = [(t, y, None if i % 3 == 1 else dy, per, durs)
            for i, (t, y, dy, per, durs) in enumerate(reqs)]
    srv = BLSServer(objective="likelihood", capacity_points=200000,
                    capacity_periods=20000, capacity_durations=200,
                    max_batch=4)
    handles = [srv.submit(*r) for r in reqs]
    srv.drain()
    ctx = Context(0)
    for i, (h, r) in enumerate(zip(handles, reqs)):
        assert h.status == 0
        b = DeviceBatch.from_host([r[0]], [r[1]],
                                  None if r[2] is None else [r[2]], r[3],
                                  r[4], device=ctx.device)
        alone = ctx.run(b, objective="likelihood").to_host()
        for k in h.result:
            assert np.array_equal(h.result[k], alone[k]), (i, k)
        assert h.best_index == int(alone["best_index"][0])

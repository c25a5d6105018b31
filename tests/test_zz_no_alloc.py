"""atg_bls_run_noalloc (the entry point that allocates nothing) against
atg_bls_run: bit-identical results on two calls in a row, the second with
a smaller batch, through the C ABI; a batch beyond the static scratch is
refused before anything is enqueued."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FIELDS = ("power", "depth", "depth_err", "duration", "transit_time",
          "depth_snr", "log_likelihood", "best_index", "status")


def _batch(n_lc, n_points, n_periods, seed, device):
    from astro_transit_gpu_amd import DeviceBatch, synth
    rng = np.random.default_rng(seed)
    ts, ys, dys = [], [], []
    for _ in range(n_lc):
        t, y, dy, _ = synth.config2_curve(rng, n_points)
        ts.append(t)
        ys.append(y)
        dys.append(dy)
    # one curve with per-point dy (unpacked bins)
    dys[-1] = dys[-1] * rng.uniform(0.9, 1.1, dys[-1].size).astype(np.float32)
    periods = np.geomspace(0.5, 13.5, n_periods)
    return DeviceBatch.from_host(ts, ys, dys, periods, synth.DURATIONS_5,
                                 device=device)


def test_noalloc_equals_run_twice_shrinking():
    from astro_transit_gpu_amd import Context, _abi
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    ctx = Context(0)
    for call, (n_lc, n_pts, n_per, seed) in enumerate(
            ((24, 15000, 2000, 31), (5, 4000, 700, 32))):
        b = _batch(n_lc, n_pts, n_per, seed, ctx.device)
        need = ctx.workspace_bytes(b)
        print("call %d: %d curves x %d points x %d periods, workspace %d "
              "bytes" % (call, n_lc, n_pts, n_per, need), flush=True)
        assert need <= _abi.ATG_NOALLOC_SCRATCH_BYTES
        for mode in ("parity", "fast"):
            ref = ctx.run(b, objective="likelihood", mode=mode).to_host()
            got = ctx.run_noalloc(b, objective="likelihood",
                                  mode=mode).to_host()
            for k in FIELDS:
                assert np.array_equal(got[k], ref[k], equal_nan=True), (
                    call, mode, k)
            assert (got["status"] == 0).all()


def test_noalloc_refuses_batch_beyond_scratch():
    """A batch whose workspace exceeds the scratch is refused by the host
    check (no kernel is launched)."""
    from astro_transit_gpu_amd import Context, DeviceBatch, _abi
    ctx = Context(0)
    n = (_abi.ATG_NOALLOC_SCRATCH_BYTES // 48) + 1024
    t = np.arange(n, dtype=np.float64) * 1e-3
    y = np.ones(n, dtype=np.float32)
    b = DeviceBatch.from_host([t], [y], None, np.array([2.0, 3.0]),
                              np.array([0.1]), device=ctx.device)
    assert ctx.workspace_bytes(b) > _abi.ATG_NOALLOC_SCRATCH_BYTES
    with pytest.raises(_abi.MemoryBudgetError, match="static scratch"):
        ctx.run_noalloc(b)

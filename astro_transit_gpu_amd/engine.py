"""Device-side driver of libatg_bls.so over torch-ROCm buffers.

torch is used only to own device memory and streams; every computation on
the path runs in the HIP kernels of csrc/atg_bls.hip.

* :class:`DeviceBatch`  ragged CSR batch of light curves resident in HBM.
* :class:`Context`      one library context (device + stream + workspace
  budget + tiling) and :meth:`Context.run`.
"""
import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _abi

FIELDS = ("power", "depth", "depth_err", "duration", "transit_time",
          "depth_snr", "log_likelihood")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev(device):
    d = torch.device(device if device is not None else "cuda")
    if d.type != "cuda":
        raise ValueError("device must be a GPU device, got %s" % d)
    if d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


def release_host_cache():
    """Returns the pinned host blocks torch keeps cached (freed staging
    buffers) to the system."""
    torch._C._host_emptyCache()


def pinned_empty(n, dtype):
    """A pinned host tensor of n elements.  Freed pinned buffers stay in
    torch's host cache; when a new one does not fit beside them, the cache
    is released and the allocation tried once more."""
    try:
        return torch.empty(n, dtype=dtype, pin_memory=True)
    except RuntimeError as e:
        if "out of memory" not in str(e):
            raise
        release_host_cache()
        return torch.empty(n, dtype=dtype, pin_memory=True)


@dataclass
class DeviceBatch:
    """CSR batch: curve i owns points lc_offsets[i]:lc_offsets[i+1], periods
    period_offsets[i]:... and durations dur_offsets[i]:...  All tensors live
    on one GPU; offsets are int64, t/periods/durations FP64, y/dy FP32 or
    FP64 (dy may be None: unit weights)."""
    t: torch.Tensor
    y: torch.Tensor
    dy: Optional[torch.Tensor]
    lc_offsets: torch.Tensor
    periods: torch.Tensor
    period_offsets: torch.Tensor
    durations: torch.Tensor
    dur_offsets: torch.Tensor
    host_period_offsets: np.ndarray = field(repr=False)

    @property
    def n_lc(self):
        return int(self.lc_offsets.numel() - 1)

    @property
    def n_periods(self):
        return int(self.periods.numel())

    @staticmethod
    def from_host(ts: Sequence[np.ndarray], ys: Sequence[np.ndarray],
                  dys: Optional[Sequence[Optional[np.ndarray]]],
                  periods, durations, device=None, y_dtype=None):
        """Pack host arrays into a device batch.  ``periods`` and
        ``durations`` are either one array shared by every curve or one array
        per curve.  ``dys`` is None (unit weights for all) or one array per
        curve.  y/dy keep FP32 if every input is FP32, else FP64."""
        dev = _dev(device)
        n = len(ts)
        if n == 0:
            raise ValueError("empty batch")

        def per_curve(a):
            if isinstance(a, (list, tuple)) and len(a) > 0 and \
                    np.ndim(a[0]) >= 1:
                if len(a) != n:
                    raise ValueError("need one grid per light curve")
                return [np.atleast_1d(np.asarray(x, dtype=np.float64))
                        for x in a]
            a = np.atleast_1d(np.asarray(a, dtype=np.float64))
            return [a] * n

        pers = per_curve(periods)
        durs = per_curve(durations)
        if y_dtype is None:
            all32 = all(np.asarray(y).dtype == np.float32 for y in ys) and (
                dys is None or all(np.asarray(d).dtype == np.float32
                                   for d in dys))
            y_dtype = np.float32 if all32 else np.float64
        lc_off = np.zeros(n + 1, dtype=np.int64)
        lc_off[1:] = np.cumsum([np.size(t) for t in ts])
        p_off = np.zeros(n + 1, dtype=np.int64)
        p_off[1:] = np.cumsum([p.size for p in pers])
        d_off = np.zeros(n + 1, dtype=np.int64)
        d_off[1:] = np.cumsum([d.size for d in durs])

        def cat(arrs, dt):
            if len(arrs) == 0:
                return np.zeros(0, dt)
            return np.ascontiguousarray(np.concatenate(
                [np.asarray(a, dtype=dt).ravel() for a in arrs]))

        # one pinned staging buffer and one copy to the device; every array
        # starts at a 256-byte offset and is a typed view of the device copy
        arrays = [cat(ts, np.float64), cat(ys, y_dtype),
                  None if dys is None else cat(dys, y_dtype), lc_off,
                  cat(pers, np.float64), p_off, cat(durs, np.float64), d_off]
        offs, total = [], 0
        for a in arrays:
            offs.append(total)
            if a is not None:
                total += (a.nbytes + 255) // 256 * 256
        host = pinned_empty(max(total, 256), torch.uint8)
        hv = host.numpy()
        for a, o in zip(arrays, offs):
            if a is not None and a.nbytes:
                hv[o:o + a.nbytes] = a.view(np.uint8)
        buf = host.to(dev, non_blocking=True)

        def piece(i):
            a = arrays[i]
            if a is None:
                return None
            tdt = torch.from_numpy(a[:0]).dtype
            return buf[offs[i]:offs[i] + a.nbytes].view(tdt)

        b = DeviceBatch(
            t=piece(0), y=piece(1), dy=piece(2), lc_offsets=piece(3),
            periods=piece(4), period_offsets=piece(5), durations=piece(6),
            dur_offsets=piece(7), host_period_offsets=p_off,
        )
        torch.cuda.current_stream(dev).synchronize()
        return b


@dataclass
class DeviceResult:
    """Per-period fields (FP64, aligned with the batch periods), per-curve
    best period index and status."""
    power: torch.Tensor
    depth: torch.Tensor
    depth_err: torch.Tensor
    duration: torch.Tensor
    transit_time: torch.Tensor
    depth_snr: torch.Tensor
    log_likelihood: torch.Tensor
    best_index: torch.Tensor
    status: torch.Tensor
    period_offsets: np.ndarray

    @staticmethod
    def empty(n_periods, n_lc, period_offsets, device):
        f = [torch.empty(n_periods, dtype=torch.float64, device=device)
             for _ in FIELDS]
        return DeviceResult(*f,
                            best_index=torch.empty(n_lc, dtype=torch.int64,
                                                   device=device),
                            status=torch.empty(n_lc, dtype=torch.int32,
                                               device=device),
                            period_offsets=period_offsets)

    def to_host(self):
        # the seven FP64 fields in one copy
        h = torch.stack([getattr(self, k) for k in FIELDS]).cpu().numpy()
        out = {k: h[i] for i, k in enumerate(FIELDS)}
        out["best_index"] = self.best_index.cpu().numpy()
        out["status"] = self.status.cpu().numpy()
        return out

    def curve(self, i, host=None):
        """Fields of curve i as numpy arrays (from ``host`` if given)."""
        h = host if host is not None else self.to_host()
        a, b = self.period_offsets[i], self.period_offsets[i + 1]
        return {k: h[k][a:b] for k in FIELDS}


BEST_FIELDS = FIELDS + ("period",)


@dataclass
class DeviceBest:
    """Per-curve winner of a best-only run: the 7 fields of the best period,
    the period itself, its index within the curve's grid, the status."""
    power: torch.Tensor
    depth: torch.Tensor
    depth_err: torch.Tensor
    duration: torch.Tensor
    transit_time: torch.Tensor
    depth_snr: torch.Tensor
    log_likelihood: torch.Tensor
    period: torch.Tensor
    best_index: torch.Tensor
    status: torch.Tensor

    @staticmethod
    def empty(n_lc, device):
        f = [torch.empty(n_lc, dtype=torch.float64, device=device)
             for _ in BEST_FIELDS]
        return DeviceBest(*f,
                          best_index=torch.empty(n_lc, dtype=torch.int64,
                                                 device=device),
                          status=torch.empty(n_lc, dtype=torch.int32,
                                             device=device))

    def to_host(self):
        out = {k: getattr(self, k).cpu().numpy() for k in BEST_FIELDS}
        out["best_index"] = self.best_index.cpu().numpy()
        out["status"] = self.status.cpu().numpy()
        return out


class Context:
    """A libatg_bls context bound to one GPU and one stream."""

    def __init__(self, device=None, stream=None, hbm_budget_bytes=0,
                 tuning=None):
        self.device = _dev(device)
        self._lib = _abi.lib()
        self.stream = stream if stream is not None else \
            torch.cuda.current_stream(self.device)
        h = ctypes.c_void_p()
        rc = self._lib.atg_ctx_create(self.device.index,
                                      ctypes.c_void_p(self.stream.cuda_stream),
                                      int(hbm_budget_bytes), ctypes.byref(h))
        if rc != _abi.ATG_OK:
            raise _abi.AtgError(rc, "context creation failed on %s (gfx950 "
                                "required)" % self.device)
        self._h = h
        if tuning:
            self.set_tuning(**tuning)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.atg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != _abi.ATG_OK:
            msg = self._lib.atg_last_error(self._h).decode()
            if rc == _abi.ATG_ERR_BUDGET:
                raise _abi.MemoryBudgetError(rc, msg)
            raise _abi.AtgError(rc, msg)

    def set_stream(self, stream):
        self.stream = stream
        self._check(self._lib.atg_ctx_set_stream(
            self._h, ctypes.c_void_p(stream.cuda_stream)))

    def set_tuning(self, periods_per_item=0, lds_bins=0, wg_per_cu=0):
        t = _abi.Tuning(int(periods_per_item), int(lds_bins), int(wg_per_cu),
                        0)
        self._check(self._lib.atg_ctx_set_tuning(self._h, ctypes.byref(t)))

    def get_tuning(self):
        t = _abi.Tuning()
        self._check(self._lib.atg_ctx_get_tuning(self._h, ctypes.byref(t)))
        return dict(periods_per_item=t.periods_per_item, lds_bins=t.lds_bins,
                    wg_per_cu=t.wg_per_cu)

    def timing(self, enable=True, reset=True):
        """Average main-kernel duration (ms) over the runs since the last
        reset, from HIP events recorded on this context's stream."""
        ms = ctypes.c_double(0.0)
        n = self._lib.atg_timing_main_kernel(self._h, int(enable), int(reset),
                                             ctypes.byref(ms))
        return n, ms.value

    def _structs(self, batch, out, objective, mode, oversample):
        b = self._batch_struct(batch, objective, mode, oversample)
        o = _abi.Out(*[_ptr(getattr(out, k)) for k in FIELDS],
                     best_index=_ptr(out.best_index), status=_ptr(out.status))
        return b, o

    def _batch_struct(self, batch, objective, mode, oversample):
        if objective not in _abi.OBJECTIVES:
            raise ValueError("objective must be 'likelihood' or 'snr'")
        if mode not in _abi.MODES:
            raise ValueError("mode must be 'parity' or 'fast'")
        for name in ("t", "y", "lc_offsets", "periods", "period_offsets",
                     "durations", "dur_offsets"):
            x = getattr(batch, name)
            if x.device != self.device or not x.is_contiguous():
                raise ValueError("%s must be contiguous on %s" % (name,
                                                                  self.device))
        dt = {torch.float32: _abi.ATG_F32, torch.float64: _abi.ATG_F64}
        if batch.y.dtype not in dt or (batch.dy is not None and
                                       batch.dy.dtype not in dt):
            raise ValueError("y and dy must be float32 or float64")
        if batch.t.dtype != torch.float64:
            raise ValueError("t must be float64")
        b = _abi.Batch(
            n_lc=batch.n_lc, n_points=int(batch.t.numel()),
            n_periods=batch.n_periods,
            n_durations=int(batch.durations.numel()),
            t=_ptr(batch.t), y=_ptr(batch.y), dy=_ptr(batch.dy),
            y_dtype=dt[batch.y.dtype],
            dy_dtype=dt[batch.dy.dtype] if batch.dy is not None else 0,
            lc_offsets=_ptr(batch.lc_offsets), periods=_ptr(batch.periods),
            period_offsets=_ptr(batch.period_offsets),
            durations=_ptr(batch.durations),
            dur_offsets=_ptr(batch.dur_offsets),
            oversample=int(oversample),
            objective=_abi.OBJECTIVES[objective], mode=_abi.MODES[mode],
            reserved=0)
        return b

    def run(self, batch: DeviceBatch, objective="likelihood", mode="parity",
            oversample=10, out: Optional[DeviceResult] = None, sync=True):
        """Run the periodogram of ``batch``; returns a DeviceResult (device
        tensors).  With sync=False the work is only enqueued."""
        if out is None:
            out = DeviceResult.empty(batch.n_periods, batch.n_lc,
                                     batch.host_period_offsets, self.device)
        b, o = self._structs(batch, out, objective, mode, oversample)
        fn = self._lib.atg_bls_run_sync if sync else self._lib.atg_bls_run
        self._check(fn(self._h, ctypes.byref(b), ctypes.byref(o)))
        return out

    def run_noalloc(self, batch: DeviceBatch, objective="likelihood",
                    mode="parity", oversample=10,
                    out: Optional[DeviceResult] = None, sync=True):
        """``run`` through atg_bls_run_noalloc: the library allocates no
        device memory (its workspace is a static array of
        ``_abi.ATG_NOALLOC_SCRATCH_BYTES`` per GPU, shared by every call on
        that GPU, so overlapping calls race).  Results equal ``run``'s bit
        for bit; a batch whose workspace exceeds the scratch raises
        MemoryBudgetError.  Pass ``out`` to allocate nothing on this side
        either."""
        if out is None:
            out = DeviceResult.empty(batch.n_periods, batch.n_lc,
                                     batch.host_period_offsets, self.device)
        b, o = self._structs(batch, out, objective, mode, oversample)
        self._check(self._lib.atg_bls_run_noalloc(self._h, ctypes.byref(b),
                                                  ctypes.byref(o)))
        if sync:
            self.stream.synchronize()
        return out

    def run_best(self, batch: DeviceBatch, objective="likelihood",
                 mode="parity", oversample=10, out: Optional[DeviceBest] = None,
                 sync=True):
        """Best-only run (atg_bls_run_best): the winner of every curve,
        gathered on the device; the per-period arrays stay in the library's
        workspace.  Every field equals the full run's at its best_index."""
        if out is None:
            out = DeviceBest.empty(batch.n_lc, self.device)
        b = self._batch_struct(batch, objective, mode, oversample)
        o = _abi.Best(*[_ptr(getattr(out, k)) for k in BEST_FIELDS],
                      best_index=_ptr(out.best_index), status=_ptr(out.status))
        self._check(self._lib.atg_bls_run_best(self._h, ctypes.byref(b),
                                               ctypes.byref(o)))
        if sync:
            self.stream.synchronize()
        return out

    def capture(self, batch: DeviceBatch, out: DeviceResult,
                objective="likelihood", mode="parity", oversample=10):
        """Capture one run of ``batch`` into ``out`` as a HIP graph
        (torch.cuda.CUDAGraph).  The workspace is sized by an eager run first;
        the captured launch sequence holds no memset or allocation.  The
        graph holds the workspace's address, so the workspace is pinned
        until :meth:`CapturedRun.release`: a later call of this context that
        would need more raises MemoryBudgetError instead of moving it.
        Replays and eager calls of this context share the workspace and must
        not overlap (use another Context for concurrent eager work)."""
        self.run(batch, objective, mode, oversample, out=out, sync=True)
        g = torch.cuda.CUDAGraph()
        prev = self.stream
        cap = torch.cuda.Stream(self.device)
        try:
            with torch.cuda.graph(g, stream=cap):
                self.set_stream(torch.cuda.current_stream(self.device))
                self.run(batch, objective, mode, oversample, out=out,
                         sync=False)
        finally:
            self.set_stream(prev)
        self._lib.atg_ctx_pin_workspace(self._h, 1)
        return CapturedRun(self, g)

    def top_k(self, batch: DeviceBatch, result: DeviceResult, k=5,
              distance=10, sync=True):
        """Top-k peaks of every curve's periodogram, on the device
        (atg_bls_top_k; the rule of the reference's get_top_k_candidates,
        search/gpu_bls.py:146-174).  Returns (index, count): int64 (n_lc, k)
        period indices relative to each curve's grid, -1 past count[c]; int32
        (n_lc,) candidates per curve (0 for a curve whose status is not OK).
        """
        idx = torch.empty((batch.n_lc, int(k)), dtype=torch.int64,
                          device=self.device)
        cnt = torch.empty(batch.n_lc, dtype=torch.int32, device=self.device)
        b, o = self._structs(batch, result, "likelihood", "parity", 10)
        self._check(self._lib.atg_bls_top_k(
            self._h, ctypes.byref(b), ctypes.byref(o), int(k), int(distance),
            _ptr(idx), _ptr(cnt)))
        if sync:
            self.stream.synchronize()
        return idx, cnt

    def candidates(self, batch: DeviceBatch, result: DeviceResult, k=5,
                   distance=10):
        """Per curve, the top-k candidates as the reference's dicts
        (period, t0, duration, depth, power; gpu_bls.py:164-173), gathered
        on the device so that only k values per field leave it."""
        idx, cnt = self.top_k(batch, result, k, distance, sync=False)
        off = torch.as_tensor(batch.host_period_offsets[:-1],
                              device=self.device).unsqueeze(1)
        g = (idx.clamp(min=0) + off).reshape(-1)
        cols = [batch.periods, result.transit_time, result.duration,
                result.depth, result.power]
        vals = torch.stack([c[g] for c in cols], 1).reshape(
            batch.n_lc, int(k), 5).cpu().numpy()
        cnt = cnt.cpu().numpy()
        keys = ("period", "t0", "duration", "depth", "power")
        return [[dict(zip(keys, map(float, vals[c, r])))
                 for r in range(int(cnt[c]))] for c in range(batch.n_lc)]

    def compute_stats(self, batch: DeviceBatch, curve, period, duration,
                      transit_time, max_transits=1024):
        """Vetting statistics of candidates (atg_bls_stats; astropy
        BoxLeastSquares.compute_stats per candidate).  ``curve``, ``period``,
        ``duration``, ``transit_time``: one value per candidate.  Returns a
        list of dicts with astropy's keys, plus ``status``."""
        cl = np.atleast_1d(np.asarray(curve, dtype=np.int32))
        n = cl.size
        if n == 0:
            return []
        if cl.min() < 0 or cl.max() >= batch.n_lc:
            raise ValueError("candidate curve index out of range")
        vals = [np.array(np.broadcast_to(np.asarray(v, dtype=np.float64),
                                         (n,)))
                for v in (period, duration, transit_time)]
        if not (np.all(vals[0] > 0) and np.all(vals[1] > 0)):
            raise ValueError("periods and durations must be positive")
        if np.any(vals[1] >= vals[0]):
            raise ValueError("The maximum transit duration must be shorter "
                             "than the minimum period")
        dev = self.device
        up = [torch.as_tensor(np.ascontiguousarray(v), device=dev)
              for v in [cl] + vals]
        cap = int(max_transits)
        st = torch.empty((n, _abi.ATG_STATS_FIELDS), dtype=torch.float64,
                         device=dev)
        pc = torch.zeros((n, cap), dtype=torch.int32, device=dev)
        pl = torch.zeros((n, cap), dtype=torch.float64, device=dev)
        status = torch.empty(n, dtype=torch.int32, device=dev)
        b, _ = self._structs(batch, DeviceResult.empty(
            1, 1, batch.host_period_offsets, dev), "likelihood", "parity",
            10)
        self._check(self._lib.atg_bls_stats(
            self._h, ctypes.byref(b), n, *[_ptr(u) for u in up], cap,
            _ptr(st), _ptr(pc), _ptr(pl), _ptr(status)))
        self.stream.synchronize()
        st, pc, pl, status = (st.cpu().numpy(), pc.cpu().numpy(),
                              pl.cpu().numpy(), status.cpu().numpy())
        out = []
        for i in range(n):
            f = st[i]
            n_tr = int(f[15]) if status[i] != 1 else 0
            P, T0 = vals[0][i], vals[2][i]
            full = status[i] == 0
            out.append(dict(
                transit_times=P * np.arange(int(f[14]), int(f[14]) + n_tr)
                + T0,
                per_transit_count=pc[i, :n_tr].astype(np.int64) if full
                else None,
                per_transit_log_likelihood=pl[i, :n_tr].copy() if full
                else None,
                depth=(f[0], f[1]), depth_phased=(f[2], f[3]),
                depth_half=(f[4], f[5]), depth_odd=(f[6], f[7]),
                depth_even=(f[8], f[9]), harmonic_amplitude=float(f[10]),
                harmonic_delta_log_likelihood=float(f[11]),
                model_levels=(float(f[12]), float(f[13])),
                status=int(status[i])))
        return out

    def model(self, t, period, duration, transit_time, y_in=None,
              y_out=None):
        """Transit mask (and, given the two levels, the box model) at the
        times ``t`` (atg_bls_model).  Returns (mask, y_model or None) as
        numpy arrays."""
        dev = self.device
        tt = torch.as_tensor(np.ascontiguousarray(t, dtype=np.float64),
                             device=dev)
        n = int(tt.numel())
        mask = torch.empty(n, dtype=torch.uint8, device=dev)
        ym = None if y_in is None else torch.empty(n, dtype=torch.float64,
                                                    device=dev)
        self._check(self._lib.atg_bls_model(
            self._h, n, _ptr(tt), float(period), float(duration),
            float(transit_time), float(y_in or 0.0), float(y_out or 0.0),
            _ptr(mask), _ptr(ym)))
        self.stream.synchronize()
        return (mask.cpu().numpy().astype(bool),
                None if ym is None else ym.cpu().numpy())

    def workspace_bytes(self, batch: DeviceBatch):
        b, _ = self._structs(batch, DeviceResult.empty(
            1, 1, batch.host_period_offsets, self.device), "likelihood",
            "parity", 10)
        n = ctypes.c_uint64(0)
        self._check(self._lib.atg_bls_workspace_size(self._h, ctypes.byref(b),
                                                     ctypes.byref(n)))
        return int(n.value)

    def workspace_bytes_best(self, batch: DeviceBatch):
        """Workspace of a best-only run of ``batch``."""
        b = self._batch_struct(batch, "likelihood", "parity", 10)
        n = ctypes.c_uint64(0)
        self._check(self._lib.atg_bls_workspace_size_best(
            self._h, ctypes.byref(b), ctypes.byref(n)))
        return int(n.value)


class CapturedRun:
    """A captured launch sequence (``Context.capture``) and the pin it holds
    on its context's workspace."""

    def __init__(self, ctx: Context, graph):
        self.ctx = ctx
        self.graph = graph

    def replay(self):
        if self.graph is None:
            raise RuntimeError("captured run was released")
        self.graph.replay()

    def release(self):
        """Drop the graph and unpin the workspace (idempotent)."""
        if self.graph is not None:
            self.graph = None
            if getattr(self.ctx, "_h", None):
                self.ctx._lib.atg_ctx_pin_workspace(self.ctx._h, 0)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


_default_ctx = {}


def default_context(device=None) -> Context:
    """One lazily created context per device on torch's current stream."""
    d = _dev(device)
    c = _default_ctx.get(d.index)
    if c is None:
        c = Context(d)
        _default_ctx[d.index] = c
    s = torch.cuda.current_stream(d)
    if s.cuda_stream != c.stream.cuda_stream:
        c.set_stream(s)
    return c

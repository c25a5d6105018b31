"""Continuous-batching BLS server on one GPU.

Requests (one light curve each, with its own period grid and durations)
arrive at any time.  The server

* admits a request when its bytes fit the HBM budget: its arrays and grids
  are copied once into a device-resident ring arena (arrival order), on a
  copy stream, as soon as it is admitted;
* forms batches from admitted requests: because the arena is filled in
  arrival order, a batch is a contiguous slice of every ring, i.e. a ready
  CSR batch (no gather);
* keeps up to ``n_slots`` batches in flight, each on its own stream and
  library context, so the next batch's work overlaps the current one on the
  device while new requests keep arriving;
* returns each request's per-period fields and best period, and frees its
  arena space when its batch completes.

The reference has no serving path; its per-curve loop (src/astrotransit_gpu/
cli.py:465-466, scripts/sector_marathon_v3.py:24-28) is the "plain" form
this replaces.
"""
import collections
import concurrent.futures
import os
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from .engine import FIELDS, Context, DeviceBatch, DeviceResult, pinned_empty

def _al(n):
    # requests sit back to back: a batch's CSR ranges must be contiguous
    return n


# Host copies into the pinned staging buffer are split into pieces of about
# this many bytes and run on worker threads (numpy releases the GIL while it
# copies plain arrays); a round of admitted requests is tens of MB.
_STAGE_PIECE = 4 << 20
_STAGE_THREADS = max(1, min(8, (os.cpu_count() or 1) // 2))


@dataclass
class Request:
    t: np.ndarray
    y: np.ndarray
    dy: Optional[np.ndarray]
    periods: np.ndarray
    durations: np.ndarray
    arrival: float = 0.0
    done: float = 0.0
    result: Optional[dict] = None
    best_index: int = -1
    status: int = -1
    # arena placement
    pt: int = -1
    pp: int = -1
    pd: int = -1
    admitted: bool = False

    @property
    def latency(self):
        return self.done - self.arrival


def _fill(piece):
    out, arrs = piece
    if len(arrs) == 1:
        np.copyto(out, arrs[0], casting="no")
    else:
        np.concatenate(arrs, out=out, casting="no")


class _Ring:
    """Offsets of a ring buffer filled and freed in FIFO order.  Live data
    is [tail, head) when not wrapped, [tail, cap) + [0, head) when
    wrapped."""

    def __init__(self, cap):
        self.cap = cap
        self.live = collections.deque()   # (start, length), oldest first
        self.head = 0
        self.wrapped = False

    def plan(self, n):
        """Start offset for n elements, or None if they do not fit now."""
        if not self.live:
            return 0 if n <= self.cap else None
        tail = self.live[0][0]
        if not self.wrapped:
            if self.head + n <= self.cap:
                return self.head
            return 0 if n <= tail else None
        return self.head if self.head + n <= tail else None

    def commit(self, start, n):
        if not self.live:
            self.wrapped = False
        elif start != self.head:  # went back to offset 0
            self.wrapped = True
        self.live.append((start, n))
        self.head = start + n

    def free_oldest(self):
        start, _ = self.live.popleft()
        if not self.live:
            self.head = 0
            self.wrapped = False
        elif self.wrapped and self.live[0][0] < start:
            self.wrapped = False  # the tail moved into the low region


class BLSServer:
    """Continuous-batching BLS over a device-resident arena.

    ``capacity_points`` / ``capacity_periods`` / ``capacity_durations`` size
    the rings; ``hbm_budget_bytes`` (if given) derives them from a byte
    budget for requests shaped like ``typical_n`` points and
    ``typical_p`` periods.
    """

    def __init__(self, device=None, objective="likelihood", mode="parity",
                 oversample=10, hbm_budget_bytes=None, capacity_points=None,
                 capacity_periods=None, capacity_durations=None,
                 typical_n=15000, typical_p=5000, max_batch=1024, n_slots=2,
                 y_dtype=np.float32, tuning=None, keep_results=True):
        self.device = torch.device(device if device is not None else "cuda")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.objective, self.mode, self.oversample = objective, mode, oversample
        # keep_results=False: a request keeps its best index and status
        # only, so that a long stream does not hold every periodogram
        self.keep_results = keep_results
        self.max_batch = max_batch
        self.y_dtype = np.dtype(y_dtype)
        ytd = torch.float32 if self.y_dtype == np.float32 else torch.float64
        ysz = self.y_dtype.itemsize
        if hbm_budget_bytes is not None:
            # bytes per typical request: inputs + engine workspace (two
            # 24 B/point layouts) + grid + 7 FP64 outputs
            per_req = typical_n * (8 + 2 * ysz + 48) + typical_p * (8 + 56)
            n_req = max(1, int(hbm_budget_bytes // per_req))
            capacity_points = capacity_points or n_req * typical_n
            capacity_periods = capacity_periods or n_req * typical_p
            capacity_durations = capacity_durations or n_req * 64
        self.cap_t = int(capacity_points)
        self.cap_p = int(capacity_periods)
        self.cap_d = int(capacity_durations or 64 * 1024)
        dev = self.device
        self.a_t = torch.empty(self.cap_t, dtype=torch.float64, device=dev)
        self.a_y = torch.empty(self.cap_t, dtype=ytd, device=dev)
        self.a_dy = torch.empty(self.cap_t, dtype=ytd, device=dev)
        self.a_p = torch.empty(self.cap_p, dtype=torch.float64, device=dev)
        self.a_d = torch.empty(self.cap_d, dtype=torch.float64, device=dev)
        self.a_out = [torch.empty(self.cap_p, dtype=torch.float64, device=dev)
                      for _ in FIELDS]
        self.r_t, self.r_p, self.r_d = (_Ring(self.cap_t), _Ring(self.cap_p),
                                        _Ring(self.cap_d))
        self.copy_stream = torch.cuda.Stream(device=dev)
        self.slots = []
        for _ in range(n_slots):
            st = torch.cuda.Stream(device=dev)
            self.slots.append(dict(stream=st, ctx=Context(dev, stream=st,
                                                          tuning=tuning),
                                   busy=None))
        self._pool = None                     # host staging threads
        self.pending = collections.deque()    # submitted, not admitted
        self.admitted = collections.deque()   # admitted, not launched
        self.inflight = collections.deque()   # launched batches (FIFO)
        self.completed = 0

    # ------------------------------------------------------------ admission
    def submit(self, t, y, dy, periods, durations, arrival=None):
        r = Request(np.ascontiguousarray(t, dtype=np.float64),
                    np.ascontiguousarray(y, dtype=self.y_dtype),
                    None if dy is None else
                    np.ascontiguousarray(dy, dtype=self.y_dtype),
                    np.ascontiguousarray(np.atleast_1d(periods),
                                         dtype=np.float64),
                    np.ascontiguousarray(np.atleast_1d(durations),
                                         dtype=np.float64),
                    arrival=time.perf_counter() if arrival is None
                    else arrival)
        if (_al(r.t.size) > self.cap_t or _al(r.periods.size) > self.cap_p
                or _al(r.durations.size) > self.cap_d):
            raise MemoryError("request larger than the server's HBM arena")
        self.pending.append(r)
        return r

    def _admit(self):
        """Admit pending requests in arrival order while they fit the rings,
        then stage the whole round: per ring, the new requests' arrays are
        packed into one pinned buffer and copied with one H2D per contiguous
        run (a wrap splits a run)."""
        round_ = []
        while self.pending:
            r = self.pending[0]
            nt, np_, nd = (_al(r.t.size), _al(r.periods.size),
                           _al(r.durations.size))
            st, sp, sd = self.r_t.plan(nt), self.r_p.plan(np_), \
                self.r_d.plan(nd)
            if st is None or sp is None or sd is None:
                break  # over budget: wait for batches to complete
            self.r_t.commit(st, nt)
            self.r_p.commit(sp, np_)
            self.r_d.commit(sd, nd)
            r.pt, r.pp, r.pd = st, sp, sd
            r.admitted = True
            self.pending.popleft()
            round_.append(r)
        if not round_:
            return 0
        with torch.cuda.stream(self.copy_stream):
            self._stage(round_)
        self.admitted.extend(round_)
        return len(round_)

    def _stage(self, round_):
        """Copy a round's arrays to their rings: every contiguous run of a
        ring (a wrap splits a run) is packed into one pinned host buffer for
        the whole round, by worker threads, then copied with one H2D copy."""
        runs = []  # (ring tensor, start, arrays, numpy dtype)
        specs = [(self.a_t, "pt", lambda r: r.t, np.float64, round_),
                 (self.a_y, "pt", lambda r: r.y, self.y_dtype, round_),
                 (self.a_dy, "pt", lambda r: r.dy, self.y_dtype,
                  [r for r in round_ if r.dy is not None]),
                 (self.a_p, "pp", lambda r: r.periods, np.float64, round_),
                 (self.a_d, "pd", lambda r: r.durations, np.float64, round_)]
        for dst, attr, get, dtype, reqs in specs:
            i = 0
            while i < len(reqs):
                j = i + 1
                end = getattr(reqs[i], attr) + get(reqs[i]).size
                while j < len(reqs) and getattr(reqs[j], attr) == end:
                    end += get(reqs[j]).size
                    j += 1
                arrs = [get(r) for r in reqs[i:j]]
                if end > getattr(reqs[i], attr):
                    runs.append((dst, getattr(reqs[i], attr), arrs,
                                 np.dtype(dtype)))
                i = j
        if not runs:
            return
        offs, total = [], 0
        for _, start, arrs, dt in runs:
            offs.append(total)
            total += (sum(a.size for a in arrs) * dt.itemsize + 63) & ~63
        pin = pinned_empty(total, torch.uint8)
        host = pin.numpy()
        pieces = []  # (output view, arrays): contiguous sub-ranges of a run
        for (_, _, arrs, dt), off in zip(runs, offs):
            n = sum(a.size for a in arrs)
            out = host[off:off + n * dt.itemsize].view(dt)
            o, k = 0, 0
            while k < len(arrs):
                k0, o0, nb = k, o, 0
                while k < len(arrs) and (nb == 0 or nb < _STAGE_PIECE):
                    nb += arrs[k].size * dt.itemsize
                    o += arrs[k].size
                    k += 1
                pieces.append((out[o0:o], arrs[k0:k]))
        if total >= 2 * _STAGE_PIECE and _STAGE_THREADS > 1:
            if self._pool is None:
                self._pool = concurrent.futures.ThreadPoolExecutor(
                    max_workers=_STAGE_THREADS)
            list(self._pool.map(_fill, pieces))
        else:
            for pc in pieces:
                _fill(pc)
        tdt = {np.dtype(np.float64): torch.float64,
               np.dtype(np.float32): torch.float32}
        for (dst, start, arrs, dt), off in zip(runs, offs):
            n = sum(a.size for a in arrs)
            src = pin[off:off + n * dt.itemsize].view(tdt[dt])
            dst[start:start + n].copy_(src, non_blocking=True)

    # ------------------------------------------------------------ batching
    def _take_batch(self):
        """Longest prefix of admitted requests that is contiguous in every
        ring and shares the weighting (dy given or not)."""
        if not self.admitted:
            return []
        batch = [self.admitted.popleft()]
        has_dy = batch[0].dy is not None
        while self.admitted and len(batch) < self.max_batch:
            r, last = self.admitted[0], batch[-1]
            if (r.dy is not None) != has_dy:
                break
            if (r.pt != last.pt + _al(last.t.size) or
                    r.pp != last.pp + _al(last.periods.size) or
                    r.pd != last.pd + _al(last.durations.size)):
                break
            batch.append(self.admitted.popleft())
        return batch

    def _launch(self, slot, reqs):
        first = reqs[0]
        t0, p0, d0 = first.pt, first.pp, first.pd
        # CSR offsets relative to the slice; padding between requests is
        # skipped by the offsets (each request's range starts at its slot)
        lc = np.array([r.pt - t0 for r in reqs] +
                      [reqs[-1].pt - t0 + reqs[-1].t.size], dtype=np.int64)
        po = np.array([r.pp - p0 for r in reqs] +
                      [reqs[-1].pp - p0 + reqs[-1].periods.size],
                      dtype=np.int64)
        do = np.array([r.pd - d0 for r in reqs] +
                      [reqs[-1].pd - d0 + reqs[-1].durations.size],
                      dtype=np.int64)
        nt, np_, nd = int(lc[-1]), int(po[-1]), int(do[-1])
        st = slot["stream"]
        ev = torch.cuda.Event()
        ev.record(self.copy_stream)
        st.wait_event(ev)
        with torch.cuda.stream(st):
            off = torch.from_numpy(np.concatenate([lc, po, do])).pin_memory()
            off = off.to(self.device, non_blocking=True)
            k = len(reqs)
            batch = DeviceBatch(
                t=self.a_t[t0:t0 + nt], y=self.a_y[t0:t0 + nt],
                dy=self.a_dy[t0:t0 + nt] if first.dy is not None else None,
                lc_offsets=off[:k + 1], periods=self.a_p[p0:p0 + np_],
                period_offsets=off[k + 1:2 * k + 2],
                durations=self.a_d[d0:d0 + nd],
                dur_offsets=off[2 * k + 2:], host_period_offsets=po)
            out = DeviceResult(*[a[p0:p0 + np_] for a in self.a_out],
                               best_index=torch.empty(k, dtype=torch.int64,
                                                      device=self.device),
                               status=torch.empty(k, dtype=torch.int32,
                                                  device=self.device),
                               period_offsets=po)
            slot["ctx"].run(batch, objective=self.objective, mode=self.mode,
                            oversample=self.oversample, out=out, sync=False)
            # the seven result arrays in one pinned buffer, one row each
            hbuf = torch.empty((len(FIELDS), np_), dtype=torch.float64,
                               pin_memory=True)
            host = {k_: hbuf[i_] for i_, k_ in enumerate(FIELDS)}
            for k_ in FIELDS:
                host[k_].copy_(getattr(out, k_), non_blocking=True)
            bi = torch.empty(k, dtype=torch.int64, pin_memory=True)
            bi.copy_(out.best_index, non_blocking=True)
            stt = torch.empty(k, dtype=torch.int32, pin_memory=True)
            stt.copy_(out.status, non_blocking=True)
            done = torch.cuda.Event()
            done.record(st)
        slot["busy"] = dict(reqs=reqs, event=done, host=host, bi=bi, st=stt,
                            po=po, keep=(off, out))
        self.inflight.append(slot)

    def _finish(self, slot):
        b = slot["busy"]
        b["event"].synchronize()
        now = time.perf_counter()
        po = b["po"]
        h = {k: v.numpy() for k, v in b["host"].items()}
        for i, r in enumerate(b["reqs"]):
            # views of the batch's pinned result buffers (kept alive by them)
            if self.keep_results:
                r.result = {k: h[k][po[i]:po[i + 1]] for k in FIELDS}
            r.best_index = int(b["bi"][i])
            r.status = int(b["st"][i])
            r.done = now
            self.r_t.free_oldest()
            self.r_p.free_oldest()
            self.r_d.free_oldest()
        self.completed += len(b["reqs"])
        slot["busy"] = None

    # ------------------------------------------------------------- driving
    def poll(self, block=False):
        """One step of the serving loop: admit, launch on free slots, and
        retire finished batches.  Returns the number of requests retired."""
        self._admit()
        for slot in self.slots:
            if slot["busy"] is None and self.admitted:
                reqs = self._take_batch()
                if reqs:
                    self._launch(slot, reqs)
        retired = 0
        while self.inflight:
            slot = self.inflight[0]
            if not (block or slot["busy"]["event"].query()):
                break
            n = len(slot["busy"]["reqs"])
            self._finish(slot)
            self.inflight.popleft()
            retired += n
            block = False
            self._admit()
        return retired

    def drain(self):
        while self.pending or self.admitted or self.inflight:
            if not self.poll(block=False):
                if self.inflight:
                    self.poll(block=True)

    def close(self):
        """Stop the host staging threads (the server can still be used;
        they start again on the next large round)."""
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None

    def __del__(self):
        pool = getattr(self, "_pool", None)
        if pool is not None:
            pool.shutdown(wait=False)

    def run_alone(self, req_args):
        """Serve one request by itself (for equality checks)."""
        r = self.submit(*req_args)
        self.drain()
        return r


def serve_benchmark(d, seconds=8.0, batched_rate=4000.0, load=0.7,
                    objective="likelihood", seed=5, max_batch=256):
    """Poisson arrivals of the 15k-point curves of ``d`` (reused cyclically)
    at ``load`` x ``batched_rate`` requests/s for ``seconds``; reports the
    served rate and latency percentiles."""
    rng = np.random.default_rng(seed)
    rate = load * batched_rate
    n_req = int(rate * seconds)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    srv = BLSServer(objective=objective, hbm_budget_bytes=8e9,
                    max_batch=max_batch)
    n_lc = len(d["t"])
    reqs = []
    start = time.perf_counter()
    i = 0
    while i < n_req or srv.pending or srv.admitted or srv.inflight:
        now = time.perf_counter() - start
        while i < n_req and arrivals[i] <= now:
            j = i % n_lc
            reqs.append(srv.submit(d["t"][j], d["y"][j], d["dy"][j],
                                   d["periods"], d["durations"],
                                   arrival=start + arrivals[i]))
            i += 1
        if not srv.poll():
            time.sleep(0.0005)
    el = time.perf_counter() - start
    lat = np.array([r.latency for r in reqs]) * 1e3
    return {"value": round(len(reqs) / el, 2), "unit": "LC/s",
            "offered_rate": round(rate, 1), "requests": len(reqs),
            "seconds": round(el, 2), "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "arrivals": "Poisson", "max_batch": max_batch,
            "statuses_ok": int(sum(r.status == 0 for r in reqs))}

"""Sector-cache ingest: the reference's consolidated NPZ to a CSR batch.

The reference's SectorCache (src/astrotransit_gpu/data/sector_cache.py:
66-210) stores a sector either

* "vectorized": one common time axis ``time`` (FP32) and ``flux`` /
  ``flux_err`` matrices of shape (targets, 1312) padded with
  ``flux_err = 1.0`` (constants.py:13-14); padded points are those with
  ``flux_err >= 0.9`` (sector_cache.py:196-202), or
* "flat": concatenated ``time`` / ``flux`` / ``flux_err`` with ``offsets``.

``load_sector_cache`` reads it without unpickling anything and returns the
targets as a ragged host batch with the padding removed (the GPU path never
sees padded points), ready for :meth:`DeviceBatch.from_host`.
"""
import numpy as np

FLUX_ERR_PAD_THRESHOLD = 0.9


def load_sector_cache(path, pad_threshold=FLUX_ERR_PAD_THRESHOLD):
    """Returns dict(tic_ids, t, y, dy) with one array per target."""
    with np.load(path, allow_pickle=False) as d:
        files = set(d.files)
        time = np.asarray(d["time"])
        flux = np.asarray(d["flux"])
        err = np.asarray(d["flux_err"])
        tic = np.asarray(d["tic_ids"]) if "tic_ids" in files else \
            np.arange(flux.shape[0] if flux.ndim == 2 else 0)
        vectorized = bool(d["is_vectorized"]) if "is_vectorized" in files \
            else flux.ndim == 2
        offsets = np.asarray(d["offsets"]) if "offsets" in files else None
    ts, ys, dys = [], [], []
    if vectorized:
        t64 = time.astype(np.float64)
        for i in range(flux.shape[0]):
            keep = err[i] < pad_threshold
            ts.append(t64[keep])
            ys.append(flux[i][keep])
            dys.append(err[i][keep])
    else:
        if offsets is None:
            raise ValueError("flat sector cache without offsets")
        ends = list(offsets[1:]) + [len(flux)] if len(offsets) == len(tic) \
            else list(offsets[1:])
        for a, b in zip(offsets[:len(ends)], ends):
            keep = err[a:b] < pad_threshold
            ts.append(time[a:b].astype(np.float64)[keep])
            ys.append(flux[a:b][keep])
            dys.append(err[a:b][keep])
    return dict(tic_ids=tic, t=ts, y=ys, dy=dys)


def write_sector_cache(path, time, flux, flux_err, tic_ids, length=1312):
    """Writes host curves in the reference's vectorized layout (padding with
    flux_err = 1.0); used to build fixtures of that format."""
    n = len(flux)
    F = np.zeros((n, length), np.float32)
    E = np.zeros((n, length), np.float32)
    for i in range(n):
        m = min(length, len(flux[i]))
        F[i, :m] = flux[i][:m]
        E[i, :m] = flux_err[i][:m]
        E[i, m:] = 1.0
    np.savez(path, time=np.asarray(time, np.float32)[:length], flux=F,
             flux_err=E, tic_ids=np.asarray(tic_ids, np.int64),
             is_vectorized=True)

"""Sector-cache ingest: the reference's consolidated NPZ to a CSR batch.

The reference's SectorCache (src/astrotransit_gpu/data/sector_cache.py:
66-210) stores a sector either

* "vectorized": one common time axis ``time`` (FP32) and ``flux`` /
  ``flux_err`` matrices of shape (targets, 1312) padded with
  ``flux_err = 1.0`` (constants.py:13-14); padded points are those with
  ``flux_err >= 0.9`` (sector_cache.py:193-200), or
* "flat": concatenated ``time`` / ``flux`` / ``flux_err`` with ``offsets``
  (one start per target, optionally a final end); every point of a target
  is returned (sector_cache.py:201-210: no mask).

``load_sector_cache`` reads it without unpickling anything and returns the
targets as a ragged host batch (FP64 times, exact widenings of the stored
FP32), the vectorized padding removed, ready for
:meth:`DeviceBatch.from_host`.  ``target_data`` is the reference's
``SectorCache.get_target_data`` on that result.
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
        # target i spans offsets[i] .. offsets[i+1] (the last one to the
        # end when no final offset is stored), as get_target_data
        for i in range(len(tic)):
            a = int(offsets[i])
            b = int(offsets[i + 1]) if i + 1 < len(offsets) else len(flux)
            ts.append(time[a:b].astype(np.float64))
            ys.append(flux[a:b])
            dys.append(err[a:b])
    return dict(tic_ids=tic, t=ts, y=ys, dy=dys)


def target_data(loaded, tic_id):
    """time / flux / flux_err of the first target with this TIC id, or None
    (the reference's SectorCache.get_target_data,
    data/sector_cache.py:174-210, on load_sector_cache's result)."""
    idx = np.nonzero(np.asarray(loaded["tic_ids"]) == int(tic_id))[0]
    if idx.size == 0:
        return None
    i = int(idx[0])
    return {"time": loaded["t"][i], "flux": loaded["y"][i],
            "flux_err": loaded["dy"][i]}


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

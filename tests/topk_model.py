"""Reference model of the device top-k (test helper).

``reference_top_k`` is the reference's get_top_k_candidates index rule
(search/gpu_bls.py:146-174: scipy ``find_peaks(power, distance)``, argsort
fallback, strongest first) with numpy's stable sort, i.e. equal values rank
the later index first.  ``rounds_top_k`` restates the kernel's round-based
thinning in plain Python; the tests check that both agree with scipy.
"""
import numpy as np
from scipy.signal import find_peaks


def reference_top_k(x, k=5, distance=10):
    peaks, _ = find_peaks(x, distance=distance)
    if len(peaks) == 0:
        peaks = np.argsort(x, kind="stable")[-k:][::-1]
    return peaks[np.argsort(x[peaks], kind="stable")[::-1][:k]]


def local_maxima(x):
    n = len(x)
    out = []
    for i in range(1, n - 1):
        if x[i - 1] < x[i]:
            a = i + 1
            while a < n - 1 and x[a] == x[i]:
                a += 1
            if x[a] < x[i]:
                out.append((i + a - 1) // 2)
    return np.array(out, dtype=np.int64)


def rounds_top_k(x, k=5, distance=10):
    pos = local_maxima(x)
    if len(pos) == 0:
        chosen = sorted(range(len(x)), key=lambda i: (-x[i], -i))[:k]
        return np.array(sorted(chosen, key=lambda i: (-x[i], i)), dtype=int)

    def above(a, b):
        return x[pos[a]] > x[pos[b]] or (x[pos[a]] == x[pos[b]] and a > b)

    m = len(pos)
    st = np.zeros(m, int)
    while (st == 0).any():
        snap = st.copy()
        for j in range(m):
            if snap[j]:
                continue
            near = [q for q in range(m) if q != j and
                    abs(pos[q] - pos[j]) < distance]
            if all(snap[q] != 1 and not (snap[q] == 0 and above(q, j))
                   for q in near):
                st[j] = 1
        snap = st.copy()
        for j in range(m):
            if snap[j] == 0 and any(snap[q] == 1 for q in range(m)
                                    if abs(pos[q] - pos[j]) < distance):
                st[j] = 2
    kept = pos[st == 1]
    return np.array(sorted(kept, key=lambda i: (-x[i], -i))[:k], dtype=int)


def sample_arrays(rng, n_arrays):
    """Power-like arrays: plateaus and ties, random walks, noise, -inf
    stretches, too short for a peak, monotone (no peak), all equal."""
    out = []
    for a in range(n_arrays):
        n = int(rng.integers(3, 600))
        kind = a % 7
        if kind == 0:
            x = np.round(rng.standard_normal(n) * 3)
        elif kind == 1:
            x = np.cumsum(rng.standard_normal(n))
        elif kind == 2:
            x = rng.standard_normal(n)
        elif kind == 3:
            x = rng.standard_normal(n)
            x[rng.random(n) < 0.3] = -np.inf
        elif kind == 4:
            x = rng.standard_normal(int(rng.integers(1, 3)))
        elif kind == 5:
            x = np.sort(rng.standard_normal(n))
        else:
            x = np.full(n, 2.5)
        out.append(x.astype(np.float64))
    return out

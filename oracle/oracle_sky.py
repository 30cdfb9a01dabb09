"""CPU oracle of the sky-map binning (test infrastructure only: tests/ and smoke() import it).

Restates the binning part of ProcessRawSkymap (package/src/PyBlastAfterglowMag/skymap_process.py)
with numpy, in the reference's order of operations:

* ``get_skymap``     ProcessRawSkymap._get_skymap (:191-328), remove_zeros=False: shells whose
                     intensities sum to zero are dropped; each shell's arrays split at len/2 into
                     the principal-jet entry and the counter-jet entry; all principal entries
                     first, then all counter entries.
* ``centroid``       _compute_position_of_the_flux_centroid (:98-105): np.average over every point.
* ``data_extent``    _get_data_extend (:107-189).
* ``edges``          np.linspace(min * extend_grid, max * extend_grid, num=n) (:480-481).
* ``hist_image``     _total_hist_skymaps (:371-383) for the principal and the counter entries,
                     summed as _combine_images_from_shells does (:402-463).

``linspace_two_roundings`` and ``hist_brute`` are the closed definitions the tests hold the numpy
routines to (i * step + start with separate roundings; half-open bins with a closed last bin).
"""
import numpy as np


class SkymapError(ValueError):
    """The reference raises (ValueError / ZeroDivisionError) for these images."""


def get_skymap(shells):
    """shells: list of (xrs, yrs, intensity) of one raw_skymap file, 2 ncells entries each.
    Returns (xs, ys, datas), principal entries then counter entries."""
    kept = [tuple(np.asarray(a, dtype=np.float64) for a in sh) for sh in shells if np.sum(sh[2]) != 0]
    for x, _, _ in kept:
        if len(x) % 2:
            raise SkymapError(f"len(xrs) is expected to be even, got {len(x)}")
    pj = [tuple(a[:len(a) // 2] for a in sh) for sh in kept]
    cj = [tuple(a[len(a) // 2:] for a in sh) for sh in kept]
    ent = pj + cj
    return [e[0] for e in ent], [e[1] for e in ent], [e[2] for e in ent]


def centroid(xs, ys, datas):
    w = np.concatenate(datas)
    if np.sum(w) == 0:
        raise SkymapError("weights sum to zero")
    return (float(np.average(np.concatenate(xs), weights=w)), float(np.average(np.concatenate(ys), weights=w)))


def data_extent(xs, ys, datas):
    """(xmin, xmax, ymin, ymax) over the entries with >= 2 points and a non-zero intensity sum."""
    xmin, xmax, ymin, ymax = [], [], [], []
    for x, y, d in zip(xs, ys, datas):
        x, y, d = np.asarray(x), np.asarray(y), np.asarray(d)
        if np.sum(d) == 0 or len(x) < 2 or len(y) < 2:
            continue
        if np.min(x) == np.max(x) or np.min(y) == np.max(y):
            raise SkymapError("flat entry: interpolation would fail")
        xmin.append(x.min()); xmax.append(x.max()); ymin.append(y.min()); ymax.append(y.max())
    if not xmin:
        raise SkymapError("no entry with data")
    return float(np.min(xmin)), float(np.max(xmax)), float(np.min(ymin)), float(np.max(ymax))


def edges(lo, hi, extend_grid, n):
    return np.linspace(lo * extend_grid, hi * extend_grid, num=n)


def linspace_two_roundings(start, stop, n):
    """Element i = fl(fl(i * step) + start), step = fl(fl(stop - start) / (n - 1)), last = stop;
    a zero step takes fl(fl(fl(i / (n - 1)) * delta) + start)."""
    div = float(n - 1)
    delta = stop - start
    step = delta / div
    out = np.empty(n)
    for i in range(n):
        y = (i / div) * delta if step == 0 else float(i) * step
        out[i] = y + start
    out[-1] = stop
    return out


def hist_entry(x, y, w, ex, ey):
    h, _ = np.histogramdd((np.asarray(x), np.asarray(y)), bins=(ex, ey), weights=np.asarray(w))
    return h


def count_entry(x, y, ex, ey):
    h, _ = np.histogramdd((np.asarray(x), np.asarray(y)), bins=(ex, ey))
    return h


def hist_image(xs, ys, datas, ex, ey, weighted=True):
    """Principal-jet entries are the first half of the lists, counter-jet entries the second."""
    n = len(xs) // 2
    if len(xs) % 2:
        raise SkymapError("expected 2 * nshells entries")
    shape = (len(ex) - 1, len(ey) - 1)
    parts = []
    for sl in (slice(0, n), slice(n, 2 * n)):
        acc = np.zeros(shape)
        for x, y, d in zip(xs[sl], ys[sl], datas[sl]):
            if len(d) > 0:
                acc += hist_entry(x, y, d, ex, ey) if weighted else count_entry(x, y, ex, ey)
        parts.append(acc)
    return parts[0] + parts[1]


def hist_brute(x, y, w, ex, ey):
    """Histogram by definition: bin i holds ex[i] <= v < ex[i+1], the last bin also v == ex[-1]."""
    def find(e, v):
        for i in range(len(e) - 1):
            if e[i] <= v < e[i + 1] or (i == len(e) - 2 and v == e[-1]):
                return i
        return -1
    h = np.zeros((len(ex) - 1, len(ey) - 1))
    for xi, yi, wi in zip(x, y, w):
        i, j = find(ex, xi), find(ey, yi)
        if i >= 0 and j >= 0:
            h[i, j] += wi
    return h

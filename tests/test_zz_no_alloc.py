"""GPU test of the allocation-free sky-map binning entry (pbagpu_skymap_hist_noalloc) against the
allocating one (pbagpu_skymap_hist), through the C ABI.

Two calls of each in a row, the second on a smaller batch (fewer images, entries and points, a
coarser grid), so the second no-alloc call runs on scratch that still holds the first call's
tables.  Bin counts, edges, extents, statuses and centroids come from the same kernels in the same
order and must agree bit for bit; the weighted bins are atomic sums whose order differs from run
to run: relative 1e-12.  A call whose tables do not fit the fixed scratch is rejected."""
import ctypes as C

import numpy as np
import pytest

from pyblastafterglowmag_amd import gpu

pytestmark = pytest.mark.gpu
REL = 1e-12


def _batch(rng, n_images, n_entries, n_pts):
    images = []
    for i in range(n_images):
        ents = []
        for k in range(n_entries):
            n = int(rng.integers(n_pts // 2, n_pts + 1)) if k else 0   # first entry empty
            s = 1.0 + i + 0.3 * k
            ents.append((rng.normal(size=n) * s, rng.normal(size=n) * 0.5 * s, rng.lognormal(size=n)))
        ents.append((rng.normal(size=9000), rng.normal(size=9000), rng.lognormal(size=9000)))  # > one work item
        images.append(ents)
    return images


def _same(a, b):
    assert np.array_equal(a["status"], b["status"])
    assert np.array_equal(a["edges_x"], b["edges_x"]) and np.array_equal(a["edges_y"], b["edges_y"])
    assert np.array_equal(a["moments"], b["moments"])
    assert np.array_equal(a["count"], b["count"])
    h0, h1 = a["hist"], b["hist"]
    assert np.all(np.abs(h0 - h1) <= REL * np.maximum(np.abs(h0), np.abs(h1)))
    assert (a["status"] == 0).all() and a["count"].sum() > 0


def test_noalloc_matches_allocating_entry_on_two_calls():
    rng = np.random.default_rng(20261204)
    big = _batch(rng, 5, 7, 6000)
    small = _batch(rng, 2, 3, 1500)
    print("[noalloc] call 1: 5 images, 256 x 128 bins; call 2: 2 images, 63 x 31 bins, extend 2", flush=True)
    r1 = gpu.skymap_hist(big, 257, 129, 1.0, want_count=True, noalloc=True)
    r2 = gpu.skymap_hist(small, 64, 32, 2.0, want_count=True, noalloc=True)
    _same(r1, gpu.skymap_hist(big, 257, 129, 1.0, want_count=True))
    _same(r2, gpu.skymap_hist(small, 64, 32, 2.0, want_count=True))
    # given edges after the other calls (the scratch holds the larger batch's tables underneath)
    ex, ey = np.linspace(-3, 3, 41), np.linspace(-2, 2, 21)
    r3 = gpu.skymap_hist(small, 41, 21, edges=(ex, ey), want_count=True, noalloc=True)
    _same(r3, gpu.skymap_hist(small, 41, 21, edges=(ex, ey), want_count=True))


def test_noalloc_rejects_tables_beyond_its_scratch():
    import torch
    L = gpu.lib()
    n_segs = 60000   # 60000 empty entries are 60000 work items: about 5.3 MB of tables
    seg = (C.c_int * 2)(0, n_segs)
    off = (C.c_long * (n_segs + 1))(*([0] * (n_segs + 1)))
    d = torch.zeros(64, dtype=torch.float64, device="cuda")
    p = C.c_void_p(d.data_ptr())
    rc = L.pbagpu_skymap_hist_noalloc(1, seg, off, None, None, None, 4, 4, 1.0, 0, p, p, p, None, p, p,
                                      C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -1 and b"scratch" in L.pbagpu_last_error()

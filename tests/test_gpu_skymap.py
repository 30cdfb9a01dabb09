"""GPU parity of the sky-map binning (pbagpu_skymap_hist) against the oracle (numpy's histogramdd,
linspace and average in the reference's order), on the reference's own raw sky maps and on ragged
synthetic batches with every error status.

Tolerances: bin indices, point counts, extents and edges are bit-exact; the weighted bins and the
centroid are sums whose order differs (atomic adds, a tree reduction): relative 1e-12."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle_sky as osky  # noqa: E402

from pyblastafterglowmag_amd import gpu, h5io  # noqa: E402
from pyblastafterglowmag_amd.skymap_process import ProcessRawSkymap, entries_of  # noqa: E402

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(ROOT, "tests", "golden")
REL = 1e-12
K = {n: i for i, n in enumerate(gpu.SKY_MOMENTS)}
CONF = {"nx": 64, "ny": 32, "extend_grid": 2, "fwhm_fac": 0.5, "lat_dist_method": "integ",
        "intp_filter": {"type": None, "sigma": 2, "mode": "reflect"},
        "hist_filter": {"type": None, "sigma": 2, "mode": "reflect"}}


def golden_shells(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    return [[(z[f"sky{i}_xrs"], z[f"sky{i}_yrs"], z[f"sky{i}_intensity"])] for i in range(int(z["sky_n"]))]


def close(a, b, rel=REL):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= rel * np.maximum(np.abs(a), np.abs(b)))


def check_image(res, i, ents, ex=None, ey=None, nx=None, ny=None, eg=None):
    """Compare image i of a pbagpu_skymap_hist result with the oracle."""
    xs, ys, ds = [e[0] for e in ents], [e[1] for e in ents], [e[2] for e in ents]
    m = res["moments"][i]
    assert res["status"][i] == 0, (i, res["status"][i])
    if ex is None:
        ext = osky.data_extent(xs, ys, ds)
        assert (m[K["xmin"]], m[K["xmax"]], m[K["ymin"]], m[K["ymax"]]) == ext
        ex, ey = osky.edges(ext[0], ext[1], eg, nx), osky.edges(ext[2], ext[3], eg, ny)
    assert np.array_equal(res["edges_x"][i], ex) and np.array_equal(res["edges_y"][i], ey)
    xc, yc = osky.centroid(xs, ys, ds)
    sx = max(np.max(np.abs(np.concatenate(xs))), 1e-300)
    sy = max(np.max(np.abs(np.concatenate(ys))), 1e-300)
    assert abs(m[K["xc"]] - xc) <= 1e-13 * sx and abs(m[K["yc"]] - yc) <= 1e-13 * sy
    assert m[K["npts"]] == sum(len(d) for d in ds)
    cnt = sum(osky.count_entry(x, y, ex, ey) for x, y in zip(xs, ys))
    hist = sum(osky.hist_entry(x, y, d, ex, ey) for x, y, d in zip(xs, ys, ds))
    if res["count"] is not None:
        assert np.array_equal(res["count"][i].astype(np.float64), cnt)
    assert close(res["hist"][i], hist)
    return hist


@pytest.mark.parametrize("name", ["tophat_pw_small", "gauss_pw_small", "gauss64_pw_small"])
@pytest.mark.parametrize("nx,ny,eg", [(257, 257, 1.0), (64, 32, 2.0)])
def test_reference_raw_maps(name, nx, ny, eg):
    """All raw maps of one run in one call: 256 x 256 bins (the large piece-wise workload's maps)
    and the reference package's default grid (nx=64, ny=32, extend_grid=2)."""
    images = [entries_of(sh) for sh in golden_shells(name)]
    print(f"[sky] {name} nx={nx} ny={ny} images={len(images)}", flush=True)
    res = gpu.skymap_hist(images, nx, ny, eg, want_count=True)
    for i, ents in enumerate(images):
        h = check_image(res, i, ents, nx=nx, ny=ny, eg=eg)
        # principal + counter jet in the reference's order, as _combine_images_from_shells sums them
        xs, ys, ds = osky.get_skymap(golden_shells(name)[i])
        assert close(h, osky.hist_image(xs, ys, ds, res["edges_x"][i], res["edges_y"][i]))


def _ragged_batch(rng):
    def pts(n, s=1.0):
        return (rng.normal(size=n) * s, rng.normal(size=n) * s * 0.5, rng.lognormal(size=n))
    good0 = [pts(500), (np.zeros(0),) * 3, pts(1), pts(700, 3.0), (np.array([0.1, 0.2]), np.array([0.0, 1.0]), np.zeros(2))]
    good1 = [pts(4096, 0.2), pts(4096, 0.3)]
    zero_w = [(np.array([0.0, 1.0]), np.array([0.0, 1.0]), np.zeros(2))]
    flat = [pts(50), (np.array([1.0, 1.0, 1.0]), np.array([0.0, 1.0, 2.0]), np.ones(3))]
    nonfin = [pts(50), (np.array([np.nan, 1.0]), np.array([0.0, 1.0]), np.ones(2))]
    no_data = [pts(1), pts(1)]
    return [good0, zero_w, good1, flat, nonfin, no_data], [0, 2, 0, 3, 4, 1]


def test_ragged_batch_and_statuses():
    rng = np.random.default_rng(11)
    images, want = _ragged_batch(rng)
    print(f"[sky] ragged batch of {len(images)} images", flush=True)
    res = gpu.skymap_hist(images, 101, 77, 1.5, want_count=True)
    assert list(res["status"]) == want
    for i, ents in enumerate(images):
        if want[i] == 0:
            check_image(res, i, ents, nx=101, ny=77, eg=1.5)
        else:
            assert not np.any(res["hist"][i]) and not np.any(res["count"][i])


def test_given_edges():
    """Edges passed in (process_multiples' common grids): points on edges, on the last edge and
    outside; a flat entry is no error there; decreasing edges are."""
    ex = np.arange(-4.0, 5.0)          # 9 edges, 8 bins
    ey = np.arange(-2.0, 3.0) * 0.5    # 5 edges, 4 bins
    rng = np.random.default_rng(5)
    x = np.concatenate([ex, ex, rng.uniform(-6, 6, 300)])
    y = np.concatenate([ey[rng.integers(0, 5, 9)], np.full(9, ey[-1]), rng.uniform(-2, 2, 300)])
    w = rng.lognormal(size=len(x))
    img_a = [(x, y, w)]
    img_b = [(np.full(4, 0.5), np.array([-1.0, 0.0, 0.5, 1.0]), np.ones(4)), (x[:40], y[:40], w[:40])]
    img_c = [(x, y, w)]
    EX = np.stack([ex, ex, ex[::-1]])
    EY = np.stack([ey, ey, ey])
    print("[sky] given edges", flush=True)
    res = gpu.skymap_hist([img_a, img_b, img_c], len(ex), len(ey), 1.0, edges=(EX, EY), want_count=True)
    assert list(res["status"]) == [0, 0, 5]
    check_image(res, 0, img_a, ex, ey)
    check_image(res, 1, img_b, ex, ey)
    assert res["count"][0].sum() == np.sum((x >= ex[0]) & (x <= ex[-1]) & (y >= ey[0]) & (y <= ey[-1]))


def test_largest_grid_and_many_images():
    """nx + ny at the limit (4096 edges) with 300k points, then 16 maps of 8192 points at 256 x 256."""
    rng = np.random.default_rng(9)
    n = 300_000
    big = [(rng.normal(size=n), rng.normal(size=n) * 2, rng.lognormal(size=n))]
    print("[sky] 2048 x 2048 edges, 300k points", flush=True)
    res = gpu.skymap_hist([big], 2048, 2048, 1.0, want_count=True)
    check_image(res, 0, big, nx=2048, ny=2048, eg=1.0)
    many = [[(rng.normal(size=4096), rng.normal(size=4096), rng.lognormal(size=4096)),
             (rng.normal(size=4096), rng.normal(size=4096), rng.lognormal(size=4096))] for _ in range(16)]
    print("[sky] 16 x 8192 points, 256 x 256 bins", flush=True)
    res = gpu.skymap_hist(many, 257, 257, 1.0)
    for i, ents in enumerate(many):
        check_image(res, i, ents, nx=257, ny=257, eg=1.0)
    with pytest.raises(RuntimeError):
        gpu.skymap_hist([big], 2048, 2049, 1.0)


def _write_raw(dirpath, name):
    """The reference's raw maps of one run as raw_skymap_<i>.h5 (image i = ifreq * 2 + it)."""
    times, freqs = [1e5, 1e7], [1e9, 1e18]
    os.makedirs(dirpath, exist_ok=True)
    for i, sh in enumerate(golden_shells(name)):
        with h5io.Writer(os.path.join(dirpath, f"raw_skymap_{i}.h5")) as w:
            for v, a in zip(("xrs", "yrs", "intensity"), sh[0]):
                w.dataset("shell=0", v, a)
            for k, v in (("nshells", 1.0), ("time", times[i % 2]), ("freq", freqs[i // 2]), ("d_l", 3.0e26)):
                w.attr("/", k, v)
    return times, freqs


def test_process_singles_and_multiples(tmp_path):
    """skymap.h5 from the host mirror: GPU histogram images and centroids against the oracle, the
    layout of the reference's output, and common grids over two runs."""
    print("[sky] ProcessRawSkymap", flush=True)
    da, db = str(tmp_path / "a"), str(tmp_path / "b")
    times, freqs = _write_raw(da, "tophat_pw_small")
    _write_raw(db, "gauss_pw_small")
    prep = ProcessRawSkymap(conf=CONF)
    out = str(tmp_path / "skymap.h5")
    prep.process_singles(infpaths=os.path.join(da, "raw_skymap_*.h5"), outfpath=out)
    assert np.array_equal(h5io.read(out, "times"), [times[i % 2] for i in range(4)])
    shells = golden_shells("tophat_pw_small")
    for i in range(4):
        g = "time={:.4e} freq={:.4e}".format(times[i % 2], freqs[i // 2])
        xs, ys, ds = osky.get_skymap(shells[i])
        ext = osky.data_extent(xs, ys, ds)
        ex, ey = osky.edges(ext[0], ext[1], 2, 64), osky.edges(ext[2], ext[3], 2, 32)
        assert np.array_equal(h5io.read(out, f"{g}/grid_x"), 0.5 * (ex[1:] + ex[:-1]))
        assert close(h5io.read(out, f"{g}/image_hist").reshape(63, 31), osky.hist_image(xs, ys, ds, ex, ey))
        xc, yc = osky.centroid(xs, ys, ds)
        assert abs(h5io.attr(out, "xc", g) - xc) <= 1e-13 * np.max(np.abs(np.concatenate(xs)))
        assert h5io.attr(out, "nshells", g) == 1.0 and h5io.attr(out, "d_l", g) == 3.0e26
        x1, x2 = h5io.attr(out, "x1", g), h5io.attr(out, "x2", g)
        assert ex[0] <= x1 <= x2 <= ex[-1]
    oa, ob = str(tmp_path / "ma.h5"), str(tmp_path / "mb.h5")
    prep.process_multiples(infpaths=(os.path.join(da, "raw_skymap_*.h5"), os.path.join(db, "raw_skymap_*.h5")),
                           outfpaths=(oa, ob))
    sa, sb = golden_shells("tophat_pw_small"), golden_shells("gauss_pw_small")
    for i in range(4):
        g = "time={:.4e} freq={:.4e}".format(times[i % 2], freqs[i // 2])
        ea = osky.data_extent(*osky.get_skymap(sa[i]))
        eb = osky.data_extent(*osky.get_skymap(sb[i]))
        ex = np.linspace(min(ea[0], eb[0]) * 2, max(ea[1], eb[1]) * 2, num=64)
        ey = np.linspace(min(ea[2], eb[2]) * 2, max(ea[3], eb[3]) * 2, num=32)
        for path, sh in ((oa, sa[i]), (ob, sb[i])):
            assert np.array_equal(h5io.read(path, f"{g}/grid_y"), 0.5 * (ey[1:] + ey[:-1]))
            assert close(h5io.read(path, f"{g}/image_hist").reshape(63, 31), osky.hist_image(*osky.get_skymap(sh), ex, ey))

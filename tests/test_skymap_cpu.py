"""Sky-map binning on the CPU: the oracle held to closed definitions, the reference's raw sky maps
split into entries, and the host-side parts of the ProcessRawSkymap mirror (FWHM, lateral
distributions, interpolation) against closed forms."""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle_sky as osky  # noqa: E402

from pyblastafterglowmag_amd import skymap_process as sp  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def golden_shells(name):
    """Raw sky maps the reference wrote (raw_skymap_<i>.h5, one shell each)."""
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    return [[(z[f"sky{i}_xrs"], z[f"sky{i}_yrs"], z[f"sky{i}_intensity"])] for i in range(int(z["sky_n"]))]


def test_linspace_restatement_is_numpy_bit_for_bit():
    """The edge formula the binning kernel evaluates (two roundings, last edge = stop) is np.linspace's;
    the algebraically equal start + i * delta / div is not, so the check has teeth."""
    rng = np.random.default_rng(7)
    differs = 0
    for _ in range(300):
        a, b = np.sort(rng.normal(size=2) * 10 ** rng.uniform(-3, 3, size=2))
        n = int(rng.integers(2, 300))
        ref = np.linspace(a, b, n)
        assert np.array_equal(osky.linspace_two_roundings(a, b, n), ref)
        wrong = np.array([a + i * (b - a) / (n - 1) for i in range(n)])
        wrong[-1] = b
        differs += not np.array_equal(wrong, ref)
    assert differs > 0
    tiny = 5e-324   # zero step: the denormal branch
    assert np.array_equal(osky.linspace_two_roundings(0.0, tiny, 5), np.linspace(0.0, tiny, 5))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_histogram_matches_definition(seed):
    """np.histogramdd (the routine the reference calls) against half-open bins with a closed last
    bin, summed in input order: bit-exact, including points on edges, on the last edge and outside."""
    rng = np.random.default_rng(seed)
    ex = np.linspace(-2.0, 3.0, 11)
    ey = np.linspace(-1.0, 1.0, 7)
    x = np.concatenate([rng.uniform(-3, 4, 400), ex, [ex[-1]] * 3, [ex[0] - 1e-12]])
    y = np.concatenate([rng.uniform(-1.5, 1.5, 400), ey[rng.integers(0, len(ey), len(ex))], [ey[-1]] * 3, [0.0]])
    w = rng.lognormal(size=len(x))
    assert np.array_equal(osky.hist_entry(x, y, w, ex, ey), osky.hist_brute(x, y, w, ex, ey))
    ones = np.ones_like(w)
    assert np.array_equal(osky.count_entry(x, y, ex, ey), osky.hist_brute(x, y, ones, ex, ey))


def test_centroid_and_extent_closed_forms():
    phi = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    xs, ys, ds = [3.0 + np.cos(phi)], [-1.0 + 2 * np.sin(phi)], [np.ones(64)]
    xc, yc = osky.centroid(xs, ys, ds)
    assert abs(xc - 3.0) < 1e-14 and abs(yc + 1.0) < 1e-14
    # a single weighted point moves the centroid to the weighted mean
    xc, yc = osky.centroid([np.array([0.0, 1.0])], [np.array([0.0, 4.0])], [np.array([1.0, 3.0])])
    assert (xc, yc) == (0.75, 3.0)
    # empty and one-point entries do not count towards the extent
    xs = [np.array([0.0, 2.0]), np.array([9.0]), np.array([-5.0, 7.0])]
    ys = [np.array([0.0, 1.0]), np.array([9.0]), np.array([-1.0, 3.0])]
    ds = [np.array([1.0, 1.0]), np.array([1.0]), np.array([0.0, 0.0])]
    assert osky.data_extent(xs, ys, ds) == (0.0, 2.0, 0.0, 1.0)
    with pytest.raises(osky.SkymapError):
        osky.data_extent([np.array([1.0, 1.0])], [np.array([0.0, 1.0])], [np.array([1.0, 1.0])])
    with pytest.raises(osky.SkymapError):
        osky.data_extent([np.array([1.0])], [np.array([0.0])], [np.array([1.0])])
    with pytest.raises(osky.SkymapError):
        osky.centroid([np.array([1.0, 2.0])], [np.array([1.0, 2.0])], [np.zeros(2)])


@pytest.mark.parametrize("name", ["tophat_pw_small", "gauss_pw_small"])
def test_entries_of_reference_raw_maps(name):
    """The package's split of the reference's raw maps into principal / counter entries."""
    for shells in golden_shells(name):
        got = sp.entries_of(shells)
        xs, ys, ds = osky.get_skymap(shells)
        assert len(got) == len(xs) == 2
        for (x, y, w), xo, yo, do in zip(got, xs, ys, ds):
            assert np.array_equal(x, xo) and np.array_equal(y, yo) and np.array_equal(w, do)
        # the two jets' entries together hold every point of the shell
        assert sum(len(e[0]) for e in got) == len(shells[0][0])


def test_fwhm_of_gaussian():
    """Lateral FWHM of a Gaussian: the outermost grid points inside +-sigma sqrt(2 ln 2)."""
    sigma = 1.7
    grid = np.linspace(-10, 10, 2001)
    dist = np.exp(-0.5 * ((grid - 0.4) / sigma) ** 2)
    x1, x2 = sp.skymap_fwhm(grid, dist, 0.4, 0.5)
    half = sigma * math.sqrt(2 * math.log(2))
    h = grid[1] - grid[0]
    assert abs(x1 - (0.4 - half)) <= h and abs(x2 - (0.4 + half)) <= h


def test_lateral_distribution_of_separable_image():
    """integ: trapezoid over the other axis; for f(x) g(y) that is f(x) times the integral of g."""
    gx = np.linspace(-1, 1, 41)
    gy = np.linspace(0, 2, 31)
    f = 1 + gx ** 2
    g = np.exp(-gy)
    img = np.outer(f, g)   # [x][y]
    ig = np.sum(0.5 * (g[1:] + g[:-1]) * np.diff(gy))
    assert np.allclose(sp.lateral_distribution(gx, gy, img, "y", "integ"), f * ig, rtol=1e-13)
    fi = np.sum(0.5 * (f[1:] + f[:-1]) * np.diff(gx))
    assert np.allclose(sp.lateral_distribution(gx, gy, img, "x", "integ"), g * fi, rtol=1e-13)
    assert np.allclose(sp.lateral_distribution(gx, gy, img, "y", "max"), f * g.max())


def test_interp_image_reproduces_linear_field():
    """Linear scattered-data interpolation is exact for a linear intensity inside the hull and 0 outside."""
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1, 1, size=(2, 300))
    lin = lambda x, y: 2.0 + 0.5 * x - 0.25 * y  # noqa: E731
    ent = [(pts[0], pts[1], lin(pts[0], pts[1])), (np.zeros(0), np.zeros(0), np.zeros(0))]
    ex, ey = np.linspace(-1.2, 1.2, 25), np.linspace(-1.2, 1.2, 17)
    img = sp.interp_image(ent, ex, ey)
    cx, cy = 0.5 * (ex[1:] + ex[:-1]), 0.5 * (ey[1:] + ey[:-1])
    X, Y = np.meshgrid(cx, cy, indexing="ij")
    inside = img != 0
    assert inside.sum() > 100
    assert np.allclose(img[inside], lin(X, Y)[inside], rtol=1e-12)
    assert not inside[0, 0] and not inside[-1, -1]


def _oracle_skymap_hist(images, nx, ny, extend_grid=1.0, edges=None, want_count=False, device=0):
    """The oracle in place of gpu.skymap_hist, with the same result layout (host-logic tests only)."""
    n = len(images)
    out = dict(hist=np.zeros((n, nx - 1, ny - 1)), count=None, edges_x=np.zeros((n, nx)), edges_y=np.zeros((n, ny)),
               moments=np.zeros((n, 8)), status=np.zeros(n, dtype=np.int32))
    for i, ents in enumerate(images):
        xs, ys, ds = [e[0] for e in ents], [e[1] for e in ents], [e[2] for e in ents]
        ext = osky.data_extent(xs, ys, ds)
        if edges is None:
            ex, ey = osky.edges(ext[0], ext[1], extend_grid, nx), osky.edges(ext[2], ext[3], extend_grid, ny)
        else:
            ex, ey = np.asarray(edges[0])[i], np.asarray(edges[1])[i]
        out["edges_x"][i], out["edges_y"][i] = ex, ey
        out["hist"][i] = sum(osky.hist_entry(x, y, d, ex, ey) for x, y, d in zip(xs, ys, ds))
        xc, yc = osky.centroid(xs, ys, ds)
        out["moments"][i] = [xc, yc, ext[0], ext[1], ext[2], ext[3], np.concatenate(ds).sum(),
                             sum(len(d) for d in ds)]
    return out


def _write_raw_maps(dirpath, name):
    from pyblastafterglowmag_amd import h5io
    os.makedirs(dirpath, exist_ok=True)
    for i, sh in enumerate(golden_shells(name)):
        with h5io.Writer(os.path.join(dirpath, f"raw_skymap_{i}.h5")) as w:
            for v, a in zip(("xrs", "yrs", "intensity"), sh[0]):
                w.dataset("shell=0", v, a)
            for k, v in (("nshells", 1.0), ("time", [1e5, 1e7][i % 2]), ("freq", [1e9, 1e18][i // 2]), ("eps_e", 0.1)):
                w.attr("/", k, v)


def test_process_raw_skymaps_host_logic(tmp_path, monkeypatch):
    """ProcessRawSkymap's file handling, group names, attributes and common grids, with the oracle
    standing in for the GPU binning call: files sorted by index, one "time=.. freq=.." group each,
    the raw attributes copied, nshells = kept shells, and process_multiples' edges spanning both runs."""
    from pyblastafterglowmag_amd import gpu, h5io
    monkeypatch.setattr(gpu, "skymap_hist", _oracle_skymap_hist)
    da, db = str(tmp_path / "a"), str(tmp_path / "b")
    _write_raw_maps(da, "tophat_pw_small")
    _write_raw_maps(db, "gauss_pw_small")
    prep = sp.ProcessRawSkymap(conf=sp.DEFAULT_CONF)
    out = str(tmp_path / "skymap.h5")
    prep.process_singles(infpaths=os.path.join(da, "raw_skymap_*.h5"), outfpath=out)
    assert np.array_equal(h5io.read(out, "times"), [1e5, 1e7, 1e5, 1e7])
    assert np.array_equal(h5io.read(out, "freqs"), [1e9, 1e9, 1e18, 1e18])
    groups = sorted(h5io.members(out, "/"))
    assert groups == sorted(["freqs", "times"] + ["time={:.4e} freq={:.4e}".format(t, f)
                                                  for f in (1e9, 1e18) for t in (1e5, 1e7)])
    g = "time=1.0000e+07 freq=1.0000e+18"
    assert sorted(h5io.members(out, g)) == ["dist_x", "dist_y", "grid_x", "grid_y", "image_hist", "image_intp"]
    assert h5io.attr(out, "nshells", g) == 1.0 and h5io.attr(out, "eps_e", g) == 0.1
    for k in ("xc", "yc", "x1", "x2", "y1", "y2", "time", "freq"):
        h5io.attr(out, k, g)
    sh = golden_shells("tophat_pw_small")[3]
    xs, ys, ds = osky.get_skymap(sh)
    ext = osky.data_extent(xs, ys, ds)
    ex, ey = osky.edges(ext[0], ext[1], 2, 64), osky.edges(ext[2], ext[3], 2, 32)
    assert np.array_equal(h5io.read(out, f"{g}/image_hist").reshape(63, 31), osky.hist_image(xs, ys, ds, ex, ey))
    assert h5io.read(out, f"{g}/image_intp").size == 63 * 31
    x1, x2 = h5io.attr(out, "x1", g), h5io.attr(out, "x2", g)
    assert ex[0] <= x1 <= x2 <= ex[-1]
    # common grids over two runs
    oa, ob = str(tmp_path / "ma.h5"), str(tmp_path / "mb.h5")
    prep.process_multiples(infpaths=(os.path.join(da, "raw_skymap_*.h5"), os.path.join(db, "raw_skymap_*.h5")),
                           outfpaths=(oa, ob))
    ea = osky.data_extent(*osky.get_skymap(golden_shells("tophat_pw_small")[3]))
    eb = osky.data_extent(*osky.get_skymap(golden_shells("gauss_pw_small")[3]))
    gx = np.linspace(min(ea[0], eb[0]) * 2, max(ea[1], eb[1]) * 2, num=64)
    for path in (oa, ob):
        assert np.array_equal(h5io.read(path, f"{g}/grid_x"), 0.5 * (gx[1:] + gx[:-1]))
    with pytest.raises(IndexError):
        prep.process_multiples(infpaths=(da,), outfpaths=(oa, ob))
    with pytest.raises(FileNotFoundError):
        prep.process_singles(infpaths=os.path.join(str(tmp_path), "none_*.h5"), outfpath=out)


def test_entries_of_several_shells():
    """Several shells (_get_skymap, :191-328): a shell whose intensities sum to zero is dropped, the
    principal-jet halves of the kept shells come first, then their counter-jet halves."""
    rng = np.random.default_rng(4)
    sh = [tuple(rng.normal(size=6) for _ in range(3)) for _ in range(3)]
    sh[1] = (sh[1][0], sh[1][1], np.zeros(6))
    got = sp.entries_of(sh)
    xs, ys, ds = osky.get_skymap(sh)
    assert len(got) == 4 and [len(e[0]) for e in got] == [3, 3, 3, 3]
    assert np.array_equal(got[0][0], sh[0][0][:3]) and np.array_equal(got[1][0], sh[2][0][:3])
    assert np.array_equal(got[2][0], sh[0][0][3:]) and np.array_equal(got[3][2], sh[2][2][3:])
    for (x, y, w), xo, yo, do in zip(got, xs, ys, ds):
        assert np.array_equal(x, xo) and np.array_equal(y, yo) and np.array_equal(w, do)
    with pytest.raises(ValueError):
        sp.entries_of([(np.ones(5), np.ones(5), np.ones(5))])

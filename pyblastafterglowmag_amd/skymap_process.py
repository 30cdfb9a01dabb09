"""Post-processing of raw sky maps into ``skymap.h5`` with the histogram images binned on the GPU.

Mirrors ProcessRawSkymap (package/src/PyBlastAfterglowMag/skymap_process.py:14-673): the same
``conf`` keys (nx, ny, extend_grid, fwhm_fac, lat_dist_method, intp_filter, hist_filter), the same
``process_singles`` / ``process_multiples`` entry points and the same output layout (one group
``"time=%.4e freq=%.4e"`` per raw file with image_hist, image_intp, grid_x, grid_y, dist_x, dist_y
and the attributes xc, yc, x1, x2, y1, y2 plus the raw file's attributes; root datasets times and
freqs).

The intensity-weighted histogram images, the flux centroids, the data extents and the histogram
edges of every raw file come from one pbagpu_skymap_hist call (there is no CPU path for them: the
call raises without the GPU library).  The scattered-data interpolation onto the bin centres
(scipy.interpolate.griddata, a Delaunay triangulation per entry), the optional smoothing filters,
the lateral distributions and the FWHM search stay host-side with the same scipy / numpy calls
as the reference (:35-96, :330-369, :385-400).
"""
import glob
import os
import re

import numpy as np

from . import gpu, h5io


# the reference package's default configuration (interface.py:1001-1003)
DEFAULT_CONF = {"nx": 64, "ny": 32, "extend_grid": 2, "fwhm_fac": 0.5, "lat_dist_method": "integ",
                "intp_filter": {"type": None, "sigma": 2, "mode": "reflect"},
                "hist_filter": {"type": None, "sigma": 2, "mode": "reflect"}}


def _sorted_files(pattern):
    files = glob.glob(pattern)
    if not files:
        raise FileNotFoundError(f" Not raw skympas found. Searching: {pattern}")
    return sorted(files, key=lambda key: int(re.findall(r"\d+", key)[-2]))


def _read_raw(path):
    """Shell arrays and attributes of one raw_skymap file (saveRawImage, src/image.h:250-300)."""
    nshells = int(h5io.attr(path, "nshells"))
    shells = []
    for ish in range(nshells):
        g = f"shell={ish}"
        shells.append(tuple(h5io.read(path, f"{g}/{v}") for v in ("xrs", "yrs", "intensity")))
    attrs = {k: h5io.attr(path, k) for k in h5io.attr_names(path, "/")}
    return shells, attrs


def entries_of(shells):
    """ProcessRawSkymap._get_skymap (:191-328, remove_zeros=False): drop shells whose intensities
    sum to zero, split each shell at len/2 into principal- and counter-jet entries, principal
    entries first."""
    kept = [sh for sh in shells if np.sum(sh[2]) != 0]
    for sh in kept:
        if len(sh[0]) % 2:
            raise ValueError(f"len(xrs) is expected to be even (2*ncells). Got={len(sh[0])}")
    pj = [tuple(a[:len(a) // 2] for a in sh) for sh in kept]
    cj = [tuple(a[len(a) // 2:] for a in sh) for sh in kept]
    return pj + cj


def apply_filter(image, conf):
    """ProcessRawSkymap._apply_filter (:35-52)."""
    from scipy import ndimage
    kind = conf["type"]
    if kind is None:
        return image
    if kind == "uniform":
        return ndimage.uniform_filter(image, size=conf["size"])
    if kind == "gaussian":
        return ndimage.gaussian_filter(image, sigma=[conf["sigma"]] * image.ndim, mode=conf["mode"])
    raise KeyError(f"apply_filter type is not recognize: {kind} Available: None 'uniform' 'gaussian' ")


def skymap_fwhm(grid, dist, cc, fac=0.5):
    """ProcessRawSkymap._get_skymap_fwhm (:55-71): the outermost grid points where the lateral
    distribution first reaches fac times its value at the centroid."""
    if len(grid) != len(dist):
        raise ValueError("grid and distribution differ in length")
    val = np.interp(cc, grid, dist)
    x1 = grid[np.argmin(dist < val * fac)]
    x2 = grid[::-1][np.argmin(dist[::-1] < val * fac)]
    if x2 < x1:
        raise ValueError(f"x2={x2} < x1={x1}")
    return x1, x2


def lateral_distribution(grid_x, grid_y, image, collapse_axis="y", method="integ"):
    """ProcessRawSkymap._lateral_distributions (:330-369) on an image indexed [x][y]."""
    if collapse_axis == "x":
        cols = [image[:, i] for i in range(image.shape[1])]
        grid = grid_x
    elif collapse_axis == "y":
        cols = [image[i, :] for i in range(image.shape[0])]
        grid = grid_y
    else:
        raise KeyError(collapse_axis)
    if method == "integ":
        return np.array([np.trapezoid(c, grid, dx=np.diff(grid)[0]) for c in cols])
    if method == "ave":
        return np.array([c.mean() for c in cols])
    if method == "max":
        return np.array([c.max() for c in cols])
    raise KeyError(f"method {method} is not recognized")


def interp_image(entries, edges_x, edges_y, method="linear"):
    """_total_intep_skymap (:385-400) for principal and counter entries, summed: scipy griddata of
    every entry onto the bin centres, fill value 0; indexed [x][y]."""
    from scipy import interpolate
    cx = 0.5 * (edges_x[1:] + edges_x[:-1])
    cy = 0.5 * (edges_y[1:] + edges_y[:-1])
    X, Y = np.meshgrid(cx, cy)
    n = len(entries) // 2
    total = []
    for part in (entries[:n], entries[n:]):
        acc = np.zeros_like(X).T
        for (x, y, w) in part:
            if len(w) == 0:
                continue
            try:
                acc += interpolate.griddata(np.vstack((x, y)).T, w, np.array((X, Y)).T, method=method, fill_value=0.)
            except Exception:   # the reference returns zeros for a failed interpolation (:88-94)
                pass
        total.append(acc)
    return total[0] + total[1]


class ProcessRawSkymap:
    """ProcessRawSkymap with GPU-binned histogram images (see the module docstring)."""

    def __init__(self, conf, verbose=False, device=0):
        self.conf = conf
        self.verb = verbose
        self.device = device
        self.conf_intp_filter = conf["intp_filter"]
        self.conf_hist_filter = conf["hist_filter"]

    def _bin(self, entry_lists, edges_x=None, edges_y=None):
        nx, ny = int(self.conf["nx"]), int(self.conf["ny"])
        edges = None if edges_x is None else (np.asarray(edges_x), np.asarray(edges_y))
        res = gpu.skymap_hist(entry_lists, nx, ny, float(self.conf["extend_grid"]), edges=edges, device=self.device)
        for i, st in enumerate(res["status"]):
            if st != 0:
                raise ValueError(f"sky map {i}: {gpu.SKY_STATUS.get(int(st), st)}")
        return res

    def _group(self, w, key, entries, attrs, res, i):
        m = dict(zip(gpu.SKY_MOMENTS, res["moments"][i]))
        ex, ey = res["edges_x"][i], res["edges_y"][i]
        grid_x = 0.5 * (ex[1:] + ex[:-1])
        grid_y = 0.5 * (ey[1:] + ey[:-1])
        image_hist = apply_filter(res["hist"][i], self.conf_hist_filter)
        image_intp = apply_filter(interp_image(entries, ex, ey), self.conf_intp_filter)
        dist_x = lateral_distribution(grid_x, grid_y, image_intp, "y", self.conf["lat_dist_method"])
        dist_y = lateral_distribution(grid_x, grid_y, image_intp, "x", self.conf["lat_dist_method"])
        x1, x2 = skymap_fwhm(grid_x, dist_x, m["xc"], self.conf["fwhm_fac"])
        y1, y2 = skymap_fwhm(grid_y, dist_y, m["yc"], self.conf["fwhm_fac"])
        w.dataset2d(key, "image_hist", image_hist)
        w.dataset2d(key, "image_intp", image_intp)
        for name, v in (("grid_x", grid_x), ("grid_y", grid_y), ("dist_x", dist_x), ("dist_y", dist_y)):
            w.dataset(key, name, v)
        for name, v in (("xc", m["xc"]), ("yc", m["yc"]), ("x1", x1), ("x2", x2), ("y1", y1), ("y2", y2),
                        ("nshells", len(entries) / 2)):
            w.attr(key, name, v)
        for k, v in attrs.items():   # the raw file's attributes; nshells keeps the value above
            w.attr(key, k, v)

    def _process_files(self, in_files, outfpath="skymap.h5", edges_x=None, edges_y=None):
        """ProcessRawSkymap._process_files / _process_file (:465-559); one GPU call for all files."""
        if edges_x is not None and len(edges_x) != len(in_files):
            raise ValueError(f"len(edges_x)={len(edges_x)} != len(in_files)={len(in_files)}")
        raws = [_read_raw(f) for f in in_files]
        ents = [entries_of(sh) for sh, _ in raws]
        res = self._bin(ents, edges_x, edges_y)
        times, freqs = [], []
        with h5io.Writer(outfpath) as w:
            for i, (f, (_, attrs)) in enumerate(zip(in_files, raws)):
                time, freq = float(attrs["time"]), float(attrs["freq"])
                if self.verb:
                    print(f"Processing {f}")
                self._group(w, "time={:.4e} freq={:.4e}".format(time, freq), ents[i], attrs, res, i)
                times.append(time)
                freqs.append(freq)
            w.dataset("/", "times", np.array(times))
            w.dataset("/", "freqs", np.array(freqs))

    def process_singles(self, infpaths="raw_skymap_*.h5", outfpath="skymap.h5", remove_input=False,
                        edges_x=None, edges_y=None):
        """ProcessRawSkymap.process_singles (:561-591)."""
        files = _sorted_files(infpaths)
        self._process_files(files, outfpath, edges_x, edges_y)
        if remove_input:
            for f in files:
                os.remove(f)

    def process_multiples(self, infpaths=("raw_skymap_*.h5",), outfpaths=("skymap.h5",)):
        """ProcessRawSkymap.process_multiples (:641-673): maps of several simulations on common
        grids, whose edges span every simulation's extent at the same (time, freq)."""
        if len(infpaths) != len(outfpaths):
            raise IndexError(f"N(infpaths) must be equal to N(outfpaths). Given N(infpaths) ={len(infpaths)}  "
                             f"N(outfpaths)={len(outfpaths)}")
        all_files = [_sorted_files(p) for p in infpaths]
        if any(len(f) != len(all_files[0]) for f in all_files):
            raise ValueError("every simulation needs the same number of raw sky maps")
        if len(all_files) <= 1:
            raise ValueError("expected that there are more than 1 simulations for a givem skymap")
        # _get_combined_data_edges (:593-639): extents of every map from one GPU call
        n_maps = len(all_files[0])
        raws = [[_read_raw(f) for f in files] for files in all_files]
        for j in range(n_maps):
            t0, f0 = raws[0][j][1]["time"], raws[0][j][1]["freq"]
            for r in raws[1:]:
                if r[j][1]["time"] != t0 or r[j][1]["freq"] != f0:
                    raise ValueError(f"raw sky maps {j} differ in time or frequency across simulations")
        res = self._bin([entries_of(sh) for sim in raws for sh, _ in sim])
        mom = res["moments"].reshape(len(all_files), n_maps, -1)
        k = {n: i for i, n in enumerate(gpu.SKY_MOMENTS)}
        eg = self.conf["extend_grid"]
        edges_x = [np.linspace(np.min(mom[:, j, k["xmin"]]) * eg, np.max(mom[:, j, k["xmax"]]) * eg,
                               num=self.conf["nx"]) for j in range(n_maps)]
        edges_y = [np.linspace(np.min(mom[:, j, k["ymin"]]) * eg, np.max(mom[:, j, k["ymax"]]) * eg,
                               num=self.conf["ny"]) for j in range(n_maps)]
        for files, out in zip(all_files, outfpaths):
            self._process_files(files, out, edges_x, edges_y)

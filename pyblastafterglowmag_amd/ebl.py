"""Attenuation of the light curves by the extragalactic background light (EBL).

The reference multiplies every layer's flux density at (t_obs, nu_obs) by exp(-tau(nu_obs, z))
after the surface integral (EATS::evalFluxDensA, src/eats.h:849-853).  tau comes from a table of
optical depths on a (frequency, redshift) grid, read from the HDF5 file `ebl_tbl_fpath`
(datasets freq [nf], z [nz], tau [nf][nz]; EBL::load_h5_table, src/microphysics/common.h:20-66)
and interpolated bilinearly in (z, nu) (EBL::interpolate, common.h:67-82, with
Interp2d::InterpolateBilinear, src/utilitites/interpolators.h:724-808):
  - a frequency or redshift below the table gives tau = 0;
  - a frequency or redshift above it is clamped to the table's last value;
  - the cell is (iz, iz + 1) x (if, if + 1), iz the last index with z > z[iz] (0 if none), and
    the same for the frequency;
  - tau = 1 / ((z2 - z1)(f2 - f1)) (q11 (z2 - z)(f2 - f) + q21 (z - z1)(f2 - f)
          + q12 (z2 - z)(f - f1) + q22 (z - z1)(f - f1)),  q_ab = tau[f_b][z_a].
The factors are per observer pair and per run; the light curves are scaled once the observer
stage has run.
"""
import os

import numpy as np

from . import h5io


class EblTable:
    """Optical-depth table: freq [nf], z [nz] ascending, tau [nf][nz]."""

    def __init__(self, freq, z, tau):
        self.freq = np.ascontiguousarray(freq, dtype=np.float64)
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        tau = np.ascontiguousarray(tau, dtype=np.float64).reshape(len(self.freq), len(self.z))
        if len(self.freq) < 2 or len(self.z) < 2:
            raise ValueError("EBL table: at least 2 frequencies and 2 redshifts")
        if np.any(np.diff(self.freq) <= 0) or np.any(np.diff(self.z) <= 0):
            raise ValueError("EBL table: frequency and redshift grids must increase")
        self.tau_table = tau

    @classmethod
    def load(cls, path):
        """The reference's table file (datasets freq, z, tau)."""
        return cls(h5io.read(path, "freq"), h5io.read(path, "z"), h5io.read(path, "tau"))

    @staticmethod
    def _lower(grid, x):
        # last i with x > grid[i] (0 if none): the reference's linear scan, at most len - 2
        i = np.searchsorted(grid, x, side="left") - 1
        return np.clip(i, 0, len(grid) - 2)

    def tau(self, nu, z):
        """Optical depth at observer frequencies nu (array) for redshift z."""
        nu = np.atleast_1d(np.asarray(nu, dtype=np.float64)).copy()
        out = np.zeros_like(nu)
        if z < self.z[0]:
            return out
        zz = min(float(z), float(self.z[-1]))
        live = ~(nu < self.freq[0])
        f = np.where(nu > self.freq[-1], self.freq[-1], nu)[live]
        iz = int(self._lower(self.z, np.array([zz]))[0])
        jf = self._lower(self.freq, f)
        x1, x2 = self.z[iz], self.z[iz + 1]
        y1, y2 = self.freq[jf], self.freq[jf + 1]
        q11 = self.tau_table[jf, iz]
        q12 = self.tau_table[jf + 1, iz]
        q21 = self.tau_table[jf, iz + 1]
        q22 = self.tau_table[jf + 1, iz + 1]
        x2x1, y2y1 = x2 - x1, y2 - y1
        x2x, y2y, yy1, xx1 = x2 - zz, y2 - f, f - y1, zz - x1
        out[live] = 1.0 / (x2x1 * y2y1) * (q11 * x2x * y2y + q21 * xx1 * y2y + q12 * x2x * yy1 + q22 * xx1 * yy1)
        return out

    def attenuation(self, nu, z):
        """exp(-tau(nu, z)): the factor on the flux density at each observer frequency."""
        return np.exp(-self.tau(nu, z))


def resolve_table_path(path, workdir=None):
    """The table file the reference would open for `ebl_tbl_fpath` (None for "none"): the path as
    given (relative paths against the working directory of the process), else against the run's
    working directory, else data/EBL/Franceschini18/table.h5 under the PyBlastAfterglowMag folder
    the current directory lies in (EBL::load_h5_table, src/microphysics/common.h:20-33)."""
    if path is None or path == "none":
        return None
    if os.path.exists(path):
        return path
    if workdir is not None and os.path.exists(os.path.join(workdir, path)):
        return os.path.join(workdir, path)
    cwd = os.getcwd()
    key = "/PyBlastAfterglowMag/"
    if key in cwd + "/":
        root = (cwd + "/")[:(cwd + "/").index(key)] + key
        alt = os.path.join(root, "data", "EBL", "Franceschini18", "table.h5")
        if os.path.exists(alt):
            return alt
    raise FileNotFoundError(f"EBL table not found: {path}")

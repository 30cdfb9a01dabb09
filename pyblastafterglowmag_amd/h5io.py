"""ctypes wrapper over ``libpba_h5io.so`` (HDF5 file contract of the afterglow code)."""
import ctypes
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpba_h5io.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"HDF5 helper library missing: {_LIB_PATH} (run __graft_entry__.build())")
        _lib = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.h5io_write_id.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [dp] * 6 + [ctypes.c_double] * 2
        _lib.h5io_dataset_len.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        _lib.h5io_dataset_len.restype = ctypes.c_long
        _lib.h5io_read_dataset.argtypes = [ctypes.c_char_p, ctypes.c_char_p, dp, ctypes.c_long]
        _lib.h5io_dataset_dims.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long), ctypes.c_int]
        _lib.h5io_read_attr.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, dp]
        _lib.h5io_write_lc.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                       dp, dp, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), dp]
        _lib.h5io_create.argtypes = [ctypes.c_char_p]
        _lib.h5io_create.restype = ctypes.c_long
        _lib.h5io_add_dataset.argtypes = [ctypes.c_long, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_long, dp]
        _lib.h5io_add_dataset2d.argtypes = [ctypes.c_long, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, dp]
        _lib.h5io_add_attr.argtypes = [ctypes.c_long, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double]
        _lib.h5io_close.argtypes = [ctypes.c_long]
        _lib.h5io_attr_count.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        _lib.h5io_attr_name.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        _lib.h5io_link_count.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        _lib.h5io_link_name.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    return _lib


class Writer:
    """Incremental HDF5 writer: 1-D / 2-D float64 datasets and scalar float64 attributes in groups."""

    def __init__(self, path):
        self.f = _load().h5io_create(path.encode())
        if self.f < 0:
            raise IOError(f"cannot create {path}")
        self.path = path

    def dataset(self, group, name, data):
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel())
        if _load().h5io_add_dataset(self.f, group.encode(), name.encode(), a.size, _dp(a)) != 0:
            raise IOError(f"{self.path}: dataset {group}/{name}")

    def dataset2d(self, group, name, data):
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float64))
        if a.ndim != 2 or _load().h5io_add_dataset2d(self.f, group.encode(), name.encode(), a.shape[0], a.shape[1],
                                                      _dp(a)) != 0:
            raise IOError(f"{self.path}: dataset {group}/{name} {a.shape}")

    def attr(self, group, name, value):
        if _load().h5io_add_attr(self.f, group.encode(), name.encode(), float(value)) != 0:
            raise IOError(f"{self.path}: attribute {group}@{name}")

    def close(self):
        if self.f is not None:
            _load().h5io_close(self.f)
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def attr_names(path, obj="/"):
    lib = _load()
    n = lib.h5io_attr_count(path.encode(), obj.encode())
    if n < 0:
        raise KeyError(f"{obj} in {path}: rc={n}")
    out = []
    buf = ctypes.create_string_buffer(256)
    for i in range(n):
        if lib.h5io_attr_name(path.encode(), obj.encode(), i, buf, 256) == 0:
            out.append(buf.value.decode())
    return out


def members(path, group="/"):
    """Names of the datasets and groups in a group (name order)."""
    lib = _load()
    n = lib.h5io_link_count(path.encode(), group.encode())
    if n < 0:
        raise KeyError(f"{group} in {path}: rc={n}")
    out = []
    buf = ctypes.create_string_buffer(256)
    for i in range(n):
        if lib.h5io_link_name(path.encode(), group.encode(), i, buf, 256) == 0:
            out.append(buf.value.decode())
    return out


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def write_id(path, r, mom, theta, ctheta, ek, mass, theta_core, theta_wing, one_d=True):
    """Write an initial-data file: 1-D arrays with theta_core / theta_wing attributes
    (``use_1d_id = yes``, as the reference package writes it) or [nshells x nlayers] arrays."""
    arrs = [np.ascontiguousarray(np.atleast_2d(np.asarray(a, dtype=np.float64))) for a in (r, mom, theta, ctheta, ek, mass)]
    ns, nl = arrs[0].shape
    rc = _load().h5io_write_id(path.encode(), 1 if one_d else 0, ns, nl, *[_dp(a) for a in arrs],
                               float(theta_core), float(theta_wing))
    if rc != 0:
        raise IOError(f"h5io_write_id({path}) failed rc={rc}")


def shape(path, dset):
    """Extents of a dataset (a tuple, as numpy gives it)."""
    dims = (ctypes.c_long * 8)()
    rank = _load().h5io_dataset_dims(path.encode(), dset.encode(), dims, 8)
    if rank < 0:
        raise KeyError(f"{dset} not found in {path} (rc={rank})")
    return tuple(int(dims[i]) for i in range(rank))


def has(path, dset):
    return _load().h5io_dataset_len(path.encode(), dset.encode()) >= 0


def read(path, dset):
    lib = _load()
    n = lib.h5io_dataset_len(path.encode(), dset.encode())
    if n < 0:
        raise KeyError(f"{dset} not found in {path} (rc={n})")
    out = np.empty(n, dtype=np.float64)
    rc = lib.h5io_read_dataset(path.encode(), dset.encode(), _dp(out), n)
    if rc != 0:
        raise IOError(f"read {dset} from {path} failed rc={rc}")
    return out


def attr(path, name, obj="/"):
    v = np.zeros(1)
    rc = _load().h5io_read_attr(path.encode(), obj.encode(), name.encode(), _dp(v))
    if rc != 0:
        raise KeyError(f"attribute {name} on {obj} in {path}: rc={rc}")
    return float(v[0])


def write_lc(path, flux, times, freqs, attrs=None):
    """flux: [nlayers, npairs] (one shell)."""
    flux = np.ascontiguousarray(np.asarray(flux, dtype=np.float64))
    if flux.ndim == 1:
        flux = flux[None, :]
    nl, npairs = flux.shape
    times = np.ascontiguousarray(times, dtype=np.float64)
    freqs = np.ascontiguousarray(freqs, dtype=np.float64)
    attrs = attrs or {}
    keys = (ctypes.c_char_p * max(1, len(attrs)))(*[k.encode() for k in attrs])
    vals = np.ascontiguousarray([float(v) for v in attrs.values()] or [0.0])
    rc = _load().h5io_write_lc(path.encode(), 1, nl, npairs, _dp(flux), _dp(times), _dp(freqs),
                               len(attrs), keys, _dp(vals))
    if rc != 0:
        raise IOError(f"h5io_write_lc({path}) failed rc={rc}")

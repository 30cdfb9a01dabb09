"""The C ABI called from C++ with no Python in the process: tests/cabi/pba_gpu_run.cpp (built by
pyblastafterglowmag_amd/_build.py with hipcc, linked against libpbagpu.so) allocates with hipMalloc
and runs the three stages of one run, as the integration notes' run_on_gpu does inside the
reference program.  Its light curves against the reference program's (golden fixtures, 1e-6
relative) and against the same stages driven from Python through ctypes (bit for bit: the same
kernels on the same inputs)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import golden_run, gpu_pipeline, load_golden, rel_err
from pyblastafterglowmag_amd import abi

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pyblastafterglowmag_amd", "_lib",
                      "pba_gpu_run")


def _write_input(path, run, lc_times, lc_freqs):
    pw = run.method == "piece-wise"
    with open(path, "wb") as f:
        f.write(np.array([1 if pw else 0, run.n_layers, len(run.tgrid), len(lc_times), len(lc_freqs)],
                         dtype=np.int32).tobytes())
        f.write(bytes(run.params))
        for li in run.inits:
            f.write(bytes(li))
        for a in (run.tgrid, lc_times, lc_freqs):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        if pw:
            f.write(np.array([abi.cells_in_layer(li.ilayer) for li in run.inits], dtype=np.int32).tobytes())


@pytest.mark.parametrize("name", ["tophat_small", "tophat_pw_small"])
def test_cpp_caller_matches_reference_and_python(name):
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"C++ caller missing: {DRIVER} (run __graft_entry__.build())")
    from pyblastafterglowmag_amd import gpu
    g = load_golden(name)
    run = golden_run(g)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        ts = np.unique(run.tobs)
        fs = run.nuobs[:len(run.tobs) // len(ts)]
        _write_input(fin, run, ts, fs)
        print(f"[cabi] {name}: {run.n_layers} layers, {len(ts)} x {len(fs)} pairs", flush=True)
        p = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=600)
        print(p.stdout.strip(), p.stderr.strip(), flush=True)
        assert p.returncode == 0, p.stderr
        raw = open(fout, "rb").read()
    nbw, npair = np.frombuffer(raw[:8], dtype=np.int32)
    out = np.frombuffer(raw[8:], dtype=np.float64).reshape(nbw, npair)
    assert nbw == run.n_layers and npair == len(run.tobs)
    for k, il in enumerate(g["lc_layers"]):
        e = rel_err(out[il], g["lc"][k], 1e-30)
        print(f"  layer {il}: vs reference {e:.3e}", flush=True)
        assert e <= 1e-6, (name, il, e)
    _, pl = gpu_pipeline(name)   # the same run through the Python host path
    flux, _ = gpu.radiate(pl, precision=0)
    np.testing.assert_array_equal(out, flux[0])

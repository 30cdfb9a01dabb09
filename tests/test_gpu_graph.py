"""Electron stage with its per-step launches captured in one graph (library option
PBAGPU_OPT_ELE_GRAPH) against the same stage launched eagerly: the kernels and their arguments
are the same, so every output is bit-identical.  Three graph calls in a row, each a fresh capture
and launch, with eager library and torch work on the same device between them."""
import time

import numpy as np
import pytest

from conftest import gpu_pipeline

pytestmark = pytest.mark.gpu


def _outputs(pl):
    """The electron stage's outputs: the rows of the populations (the emissivity and absorption
    tables are allocated for two populations per blastwave and only those rows are written)."""
    npop = len(pl.pops)
    out = {"jall": pl.d_jall[:npop], "aall": pl.d_aall[:npop] if pl.d_aall is not None else None,
           "fall": pl.d_fall, "est": pl.d_est[:npop]}
    return {k: v.detach().cpu().numpy().copy() for k, v in out.items() if v is not None}


@pytest.mark.parametrize("precision", [0, 1])
def test_electron_graph_matches_eager(precision):
    import torch
    from pyblastafterglowmag_amd import gpu
    # the fixture's pipeline runs on a created stream, where the option captures
    run, pl = gpu_pipeline("gauss5_small")
    assert pl.sp is not None and pl.sp.value
    pl.precision = precision
    print(f"[gpu] electron stage graph vs eager, precision {precision}, {len(pl.pops)} populations, "
          f"{pl.nt} steps", flush=True)
    assert gpu.get_option(gpu.OPT_ELE_GRAPH) == 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pl.electrons(want_f=True)
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0
    ref = _outputs(pl)
    assert not ref["est"].any(), ref["est"]
    # the same tables on the null stream with the option on: launched directly, same results
    pl0 = gpu.Pipeline([run], precision=precision)
    pl0.use_tables(pl.d_table.cpu().numpy(), pl.res)
    try:
        gpu.set_option(gpu.OPT_ELE_GRAPH, 1)
        pl0.electrons(want_f=True)
    finally:
        gpu.set_option(gpu.OPT_ELE_GRAPH, 0)
    got0 = _outputs(pl0)
    for k, v in ref.items():
        assert np.array_equal(got0[k], v, equal_nan=True), (precision, "null stream", k)
    try:
        for rep in range(3):
            gpu.set_option(gpu.OPT_ELE_GRAPH, 1)
            assert gpu.get_option(gpu.OPT_ELE_GRAPH) == 1
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pl.electrons(want_f=True)
            torch.cuda.synchronize()
            t_graph = time.perf_counter() - t0
            got = _outputs(pl)
            for k, v in ref.items():
                assert np.array_equal(got[k], v, equal_nan=True), (precision, rep, k)
            print(f"[gpu]   call {rep}: bit-identical; {t_graph * 1e3:.1f} ms (eager {t_eager * 1e3:.1f} ms)",
                  flush=True)
            # eager work between the graph calls: the library's eager stage and a torch reduction
            gpu.set_option(gpu.OPT_ELE_GRAPH, 0)
            pl.electrons()
            s = torch.ones(1 << 20, dtype=torch.float64, device=pl.dev).sum().item()
            assert s == float(1 << 20)
            assert np.array_equal(pl.d_jall[:len(pl.pops)].cpu().numpy(), ref["jall"])
    finally:
        gpu.set_option(gpu.OPT_ELE_GRAPH, 0)

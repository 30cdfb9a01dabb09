"""Learned SSC emissivity on the GPU: the BF16 MFMA product, the network's forward and backward
passes against the CPU oracle (oracle/oracle_mlp.py, itself pinned by finite differences), the
batched plain SSC contraction against the oracle's, training on pairs from a real run, and the
electron stage with the surrogate and its sampled check.

Bounds: the product, relative to the scale sum_k |A_ik||B_kj|, 1e-5 (float32 accumulation in
another order); forward outputs and gradients, relative Frobenius norm 2e-3 / 2e-2 (a layer input
that rounds to the other bfloat16 neighbour moves the next layer by one bfloat16 ulp, 4e-3); the
plain contraction 1e-11."""
import math

import numpy as np
import pytest

from conftest import golden_run, gpu_pipeline, load_golden
import oracle_mlp as om
import pyoracle

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("M,N,K,ta,tb", [(37, 45, 70, 0, 0), (128, 96, 297, 0, 0), (64, 33, 100, 0, 1),
                                         (300, 64, 48, 1, 0), (16, 16, 32, 0, 0)])
def test_gemm_bf16_matches_numpy(M, N, K, ta, tb):
    import ctypes as C
    import torch
    from pyblastafterglowmag_amd import gpu
    rng = np.random.default_rng(M * 7 + N)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    dA = _t(A.T.copy() if ta else A)
    dB = _t(B.T.copy() if tb else B)
    dC = torch.empty((M, N), dtype=torch.float32, device="cuda")
    gpu._check(gpu.lib().pbagpu_gemm_bf16(M, N, K, gpu._ptr(dA), ta, gpu._ptr(dB), tb, gpu._ptr(dC), None), "gemm")
    ref = om.to_bf16(A).astype(np.float64) @ om.to_bf16(B).astype(np.float64)
    scale = np.abs(om.to_bf16(A)).astype(np.float64) @ np.abs(om.to_bf16(B)).astype(np.float64)
    err = np.max(np.abs(dC.cpu().numpy() - ref) / scale)
    print(f"[gpu] gemm {M}x{N}x{K} ta={ta} tb={tb}: {err:.2e}", flush=True)
    assert err <= 1e-5, err


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64)) / np.linalg.norm(b))


def test_mlp_forward_and_gradients_match_oracle():
    import torch
    from pyblastafterglowmag_amd import surrogate
    dims = [113, 64, 48, 40]
    net = surrogate.Net(dims, seed=4)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(33, dims[0])).astype(np.float32)
    t = rng.normal(size=(33, dims[-1])).astype(np.float32)
    ws = [w.cpu().numpy() for w in net.w]
    bs = [b.cpu().numpy() + 0.05 for b in net.b]
    for l in range(net.L):
        net.b[l].copy_(torch.from_numpy(bs[l]))
    y = net.forward(_t(x)).cpu().numpy()
    yo = om.forward(ws, bs, x)[-1]
    e_f = _rel(y, yo)
    loss = net.train_step(_t(x), _t(t), lr=0.0)   # gradients only: lr 0 leaves the weights
    lo, gwo, gbo = om.loss_and_grads(ws, bs, x, t)
    print(f"[gpu] mlp forward {e_f:.2e}, loss {float(loss.item()):.6e} vs {lo:.6e}", flush=True)
    assert e_f <= 2e-3
    assert float(loss.item()) == pytest.approx(lo, rel=2e-3)
    for l in range(net.L):
        eg, eb = _rel(net.gw[l].cpu().numpy(), gwo[l]), _rel(net.gb[l].cpu().numpy(), gbo[l])
        print(f"  layer {l}: dW {eg:.2e} db {eb:.2e}", flush=True)
        assert eg <= 2e-2 and eb <= 2e-2, (l, eg, eb)
        np.testing.assert_array_equal(net.w[l].cpu().numpy(), ws[l])


def test_adam_training_reduces_loss():
    from pyblastafterglowmag_amd import surrogate
    rng = np.random.default_rng(7)
    x = rng.normal(size=(2048, 16)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]) + x[:, 1] * x[:, 2], np.tanh(x[:, 3] - x[:, 4])], axis=1).astype(np.float32)
    net = surrogate.Net([16, 64, 64, 2], seed=1)
    dx, dt = _t(x), _t(t)
    l0 = float(net.train_step(dx, dt, 0.0).item())
    for _ in range(400):
        loss = net.train_step(dx, dt, 3e-3)
    l1 = float(loss.item())
    print(f"[gpu] adam: loss {l0:.3e} -> {l1:.3e}", flush=True)
    assert l1 < 0.05 * l0


def _cells(ngam, nfreq, n, seed):
    """Electron spectra (broken power laws) and photon weights (smooth humps) of n cells."""
    rng = np.random.default_rng(seed)
    step = math.exp(math.log(1e8) / ngam)
    g = np.array([step ** i for i in range(ngam)])
    f = np.zeros((n, ngam))
    for c in range(n):
        gm, p = 10 ** rng.uniform(0.5, 4), rng.uniform(2.05, 2.8)
        gM = gm * 10 ** rng.uniform(1, 3.5)
        m = (g >= gm) & (g <= gM)
        f[c, m] = 10 ** rng.uniform(-2, 4) * (g[m] / gm) ** -p
    k = np.arange(nfreq)
    w = 10 ** (rng.uniform(-3, 3, (n, 1)) - ((k[None, :] - rng.uniform(5, nfreq - 5, (n, 1))) / 8.0) ** 2)
    return f, w


def _shock(ngam, nfreq):
    from pyblastafterglowmag_amd import abi
    s = abi.ShockParams()
    s.ngam, s.nfreq, s.gam2, s.freq1, s.freq2, s.ssc = ngam, nfreq, 1e8, 1e5, 1e32, 1
    return s


def test_ssc_emissivity_matches_oracle():
    from pyblastafterglowmag_amd import surrogate
    sh = _shock(64, 48)
    f, w = _cells(64, 48, 70, seed=2)
    j = surrogate.plain_ssc(sh, _t(f), _t(w)).cpu().numpy()
    jo = pyoracle.ssc_emissivity(sh, f, w)
    m = np.abs(jo) > 1e-300
    err = float(np.max(np.abs(j[m] - jo[m]) / np.abs(jo[m])))
    print(f"[gpu] plain SSC contraction vs oracle {err:.2e}", flush=True)
    assert err <= 1e-11
    x, sc = surrogate.features(sh, _t(f), _t(w), _t(np.full(len(f), 0.3)))
    x = x.cpu().numpy()
    np.testing.assert_allclose(sc.cpu().numpy(), f.max(1) * w[:, :-1].max(1), rtol=1e-15)
    assert np.all(x[:, :-1] <= 1.0 + 1e-6) and np.all(x[:, :-1] >= 0.0)
    assert x[:, -1] == pytest.approx(math.log10(0.3) / 10, rel=1e-6)


def test_surrogate_on_a_run_and_fallback():
    """Train on the SSC cells of one FP64 run (gauss5_small: 5 layers, both shocks with SSC), then run
    the electron stage with the surrogate: its emissivity against the FP64 stage's, and, with every
    population checked every step at a tolerance no network meets, the plain fallback reproducing
    the FP64 stage bit for bit."""
    import torch
    from pyblastafterglowmag_amd import gpu, surrogate
    run, pl = gpu_pipeline("gauss5_small")
    pl.precision = 0
    pl.electrons(want_f=True)
    j64 = pl.d_jall[:len(pl.pops)].clone()
    f, w, B, sh = surrogate.cells_from_pipeline(pl)
    x, sc, ylog, jex = surrogate.targets(sh, f, w, B)
    n = x.shape[0]
    perm = torch.randperm(n, generator=torch.Generator().manual_seed(0)).cuda()
    tr, va = perm[: int(0.8 * n)], perm[int(0.8 * n):]
    sur, hist = surrogate.train(x[tr], ylog[tr], steps=1500, batch=2048)
    print(f"[gpu] {n} cells; loss {hist[0][1]:.3e} -> {hist[-1][1]:.3e}", flush=True)
    assert hist[-1][1] < 0.05 * hist[0][1]
    pred = sur.predict(x[va], sc[va])
    err = surrogate.relative_error(pred, jex[va]).cpu().numpy()
    print(f"  held-out cells: median max-rel error {np.median(err):.3e}, 90% {np.quantile(err, 0.9):.3e}", flush=True)
    assert np.median(err) < 0.2
    # the electron stage with the surrogate (1 % of the cells checked per step)
    st = pl.electrons_surrogate(sur)
    print(f"  stage: inferred {st.inferred} checked {st.checked} failed {st.failed} fallback {st.fallback} "
          f"max err {st.max_err:.3e}", flush=True)
    assert st.inferred > 0 and st.checked > 0
    assert not pl.status()[1].any()
    # fallback: every population checked at its first step with emission and failed -> plain
    sur_bad = surrogate.SscSurrogate(sur.net, sur.y_offset.cpu().numpy(), sur.y_scale, check_tol=1e-30,
                                     check_every=1)
    st2 = pl.electrons_surrogate(sur_bad)
    print(f"  fallback run: inferred {st2.inferred} failed {st2.failed} fallback {st2.fallback}", flush=True)
    # the only network values kept are those of checked steps whose plain emissivity is all zero
    # (no seed photons yet), where the prediction is zero too (the network's scale is 0 there)
    assert st2.failed > 0 and st2.inferred + st2.failed == st2.checked
    assert torch.equal(pl.d_jall[:len(pl.pops)], j64)

"""The surrogate network's CPU oracle (oracle/oracle_mlp.py) pinned by finite differences and by
the definition of bfloat16 rounding; the oracle's SSC contraction of given cells pinned by its
bilinearity and by the kernel values of the Compton kernel (orc_ssc_nava)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_mlp as om
import pyoracle


def _net(dims, seed=0):
    rng = np.random.default_rng(seed)
    ws = [rng.normal(0, 1 / math.sqrt(a), (a, b)) for a, b in zip(dims[:-1], dims[1:])]
    bs = [rng.normal(0, 0.1, b) for b in dims[1:]]
    return ws, bs


def test_gradients_match_finite_differences():
    ws, bs = _net([5, 7, 6, 3])
    rng = np.random.default_rng(1)
    x, t = rng.normal(size=(4, 5)), rng.normal(size=(4, 3))
    loss, gw, gb = om.loss_and_grads(ws, bs, x, t, bf16=False)
    eps = 1e-6
    for l in range(len(ws)):
        for idx in [(0, 0), (1, 2), (ws[l].shape[0] - 1, ws[l].shape[1] - 1)]:
            wp = [w.copy() for w in ws]; wm = [w.copy() for w in ws]
            wp[l][idx] += eps; wm[l][idx] -= eps
            num = (om.loss_and_grads(wp, bs, x, t, bf16=False)[0] - om.loss_and_grads(wm, bs, x, t, bf16=False)[0]) / (2 * eps)
            assert gw[l][idx] == pytest.approx(num, rel=1e-6, abs=1e-10), (l, idx)
        bp = [b.copy() for b in bs]; bm = [b.copy() for b in bs]
        bp[l][0] += eps; bm[l][0] -= eps
        num = (om.loss_and_grads(ws, bp, x, t, bf16=False)[0] - om.loss_and_grads(ws, bm, x, t, bf16=False)[0]) / (2 * eps)
        assert gb[l][0] == pytest.approx(num, rel=1e-6, abs=1e-10), l


def test_bf16_rounding():
    # bfloat16 keeps 8 significant bits: 1 + 2^-8 is a tie and rounds to the even neighbour 1;
    # 1 + 3 * 2^-9 rounds up to 1 + 2^-7; 1 + 2^-7 and 2^-130 (subnormal) are exact
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7, -2.5, 2 ** -130, 3.0e38], dtype=np.float32)
    got = om.to_bf16(x)
    assert list(got[:5]) == [1.0, 1.0, 1 + 2 ** -7, 1 + 2 ** -7, -2.5]
    assert got[5] == np.float32(2 ** -130)
    assert np.all((got.view(np.uint32) & 0xFFFF) == 0)


def _shock(ngam=12, nfreq=10):
    from pyblastafterglowmag_amd import abi
    s = abi.ShockParams()
    s.ngam, s.nfreq, s.gam2, s.freq1, s.freq2, s.ssc = ngam, nfreq, 1e8, 1e5, 1e32, 1
    return s


def test_oracle_ssc_contraction():
    sh = _shock()
    rng = np.random.default_rng(3)
    f = rng.random((3, sh.ngam)) * 10 ** rng.uniform(-3, 3, (3, sh.ngam))
    w = rng.random((3, sh.nfreq)) * 10 ** rng.uniform(-3, 3, (3, sh.nfreq))
    j = pyoracle.ssc_emissivity(sh, f, w)
    np.testing.assert_allclose(pyoracle.ssc_emissivity(sh, 2 * f, 3 * w), 6 * j, rtol=1e-13)
    # the contraction against the kernel's values: j_i = sum_j f_j h nu_i (3/4 sigma_T c / g_j^2) dg_j
    # sum_{k < nf-1} w_k K(nu_i, g_j, nu_k), with the grids of the electron stage
    L = pyoracle.lib()
    L.orc_ssc_nava.restype = C.c_double
    L.orc_ssc_nava.argtypes = [C.c_double] * 3
    step = math.exp(math.log(sh.gam2) / sh.ngam)
    g = np.array([step ** i for i in range(sh.ngam)])
    nu = np.empty(sh.nfreq)
    cur, ratio = 10 ** math.log10(sh.freq1), 10 ** ((math.log10(sh.freq2) - math.log10(sh.freq1)) / sh.nfreq)
    for i in range(sh.nfreq):
        nu[i] = cur; cur *= ratio
    K = np.array([[[L.orc_ssc_nava(nu[i], g[jj], nu[k]) for k in range(sh.nfreq)] for jj in range(sh.ngam)]
                  for i in range(sh.nfreq)])
    h, sT, c = 6.6260755e-27, 6.6524e-25, 2.99792458e10
    pre = 0.75 * sT * c / g[:-1] ** 2 * np.diff(g)
    ref = np.einsum("cj,j,i,ijk,ck->ci", f[:, :-1], pre, h * nu, K[:, :-1, :-1], w[:, :-1])
    np.testing.assert_allclose(j, ref, rtol=1e-12, atol=0)

"""Learned SSC emissivity: the multilayer perceptron, its training pairs and its training on the GPU.

The network stands in for the SSC contraction of the electron stage (computeSSCSpectrum,
src/microphysics/numeric_model.h:247-285); it has no counterpart in the reference.  Everything
numerical runs in the library (include/pbagpu.h): the BF16/MFMA forward and backward passes and
the Adam update (pbagpu_mlp_train_step), the input features (pbagpu_ssc_features) and the plain
contraction that produces the training targets (pbagpu_ssc_emissivity).  This module allocates
the weights, builds the training pairs from runs of the FP64 electron stage and drives the
training loop.

Training pairs.  A pair is one cell (population, step) of an FP64 run: its electron spectrum f
(the stage's f_all) and the Compton photon weights of the contraction at that step,
w_k = n_k / nu_k (nu_{k+1} - nu_k), where n is the photon density of the previous step.  That
density is proportional to the previous step's total emissivity divided by nu, so
w_k is proportional to j_{k, it-1} / nu_k^2 (nu_{k+1} - nu_k); the constant cancels in the
features, which take f / max f and w / max w.  The target is the plain contraction of (f, w),
divided by max f max w (exact, as the contraction is bilinear), in log10 with a floor 6 decades
below the cell's peak, standardised per frequency.
"""
import ctypes as C
import math

import numpy as np

from . import abi, gpu

FLOOR_DECADES = 6.0


def _torch():
    return gpu._torch()


class Net:
    """Weights (FP32 master copy), gradients and Adam moments of a pba_mlp on the device."""

    def __init__(self, dims, seed=0, device=0, train=True):
        torch = _torch()
        self.dims = list(dims)
        self.L = len(dims) - 1
        if self.L < 1 or self.L > abi.PBA_MLP_MAXL:
            raise ValueError("layer count out of range")
        dev = torch.device("cuda", device)
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.w, self.b = [], []
        for l in range(self.L):
            # scaled for unit-variance SiLU activations
            std = math.sqrt(2.0 / dims[l]) if l < self.L - 1 else math.sqrt(1.0 / dims[l])
            self.w.append((torch.randn(dims[l], dims[l + 1], generator=g) * std).to(dev).contiguous())
            self.b.append(torch.zeros(dims[l + 1], device=dev))
        self.train = train
        if train:
            self.gw = [torch.zeros_like(t) for t in self.w]
            self.gb = [torch.zeros_like(t) for t in self.b]
            self.mw = [torch.zeros_like(t) for t in self.w]
            self.vw = [torch.zeros_like(t) for t in self.w]
            self.mb = [torch.zeros_like(t) for t in self.b]
            self.vb = [torch.zeros_like(t) for t in self.b]
        self.steps = 0
        self._spec = None

    def spec(self):
        m = abi.Mlp()
        m.n_layers = self.L
        for l, d in enumerate(self.dims):
            m.dims[l] = d
        names = ("w", "b", "gw", "gb", "mw", "vw", "mb", "vb") if self.train else ("w", "b")
        for name in names:
            arr = getattr(m, name)
            for l, t in enumerate(getattr(self, name)):
                arr[l] = t.data_ptr()
        self._spec = m
        return m

    def forward(self, x):
        torch = _torch()
        x = x.contiguous().float()
        y = torch.empty((x.shape[0], self.dims[-1]), dtype=torch.float32, device=x.device)
        m = self.spec()
        gpu._check(gpu.lib().pbagpu_mlp_forward(C.byref(m), x.shape[0], gpu._ptr(x), gpu._ptr(y), None), "mlp_forward")
        return y

    def train_step(self, x, t, lr):
        torch = _torch()
        loss = torch.zeros(1, dtype=torch.float32, device=x.device)
        self.steps += 1
        m = self.spec()
        gpu._check(gpu.lib().pbagpu_mlp_train_step(C.byref(m), x.shape[0], gpu._ptr(x.contiguous()),
                                                   gpu._ptr(t.contiguous()), float(lr), self.steps, gpu._ptr(loss),
                                                   None), "mlp_train_step")
        return loss

    def state(self):
        return {f"w{l}": self.w[l].cpu().numpy() for l in range(self.L)} | \
               {f"b{l}": self.b[l].cpu().numpy() for l in range(self.L)}

    def load(self, state):
        torch = _torch()
        for l in range(self.L):
            self.w[l].copy_(torch.from_numpy(np.asarray(state[f"w{l}"], dtype=np.float32)))
            self.b[l].copy_(torch.from_numpy(np.asarray(state[f"b{l}"], dtype=np.float32)))


class SscSurrogate:
    """A trained network with its output standardisation and the check settings (pba_ssc_surrogate)."""

    def __init__(self, net, y_offset, y_scale, check_tol=0.05, check_every=100):
        torch = _torch()
        self.net = net
        self.y_offset = torch.as_tensor(np.asarray(y_offset, dtype=np.float32)).to(net.w[0].device).contiguous()
        self.y_scale = float(y_scale)
        self.check_tol = float(check_tol)
        self.check_every = int(check_every)

    def spec(self):
        s = abi.SscSurrogate()
        s.net = self.net.spec()
        s.y_offset = self.y_offset.data_ptr()
        s.y_scale = self.y_scale
        s.check_tol = self.check_tol
        s.check_every = self.check_every
        s.pad_ = 0
        return s

    def predict(self, x, scale):
        """j_ssc [n][nfreq] (torch, FP64) of cells with features x and scales max f max w."""
        y = self.net.forward(x).double()
        return scale[:, None] * 10.0 ** (self.y_offset.double()[None, :] + self.y_scale * y)


def features(shock, f, w, B):
    """Surrogate input features and scales of cells (torch FP64 device tensors f [n][ngam],
    w [n][nfreq], B [n]) by the library's routine."""
    torch = _torch()
    n = f.shape[0]
    x = torch.empty((n, shock.ngam + shock.nfreq + 1), dtype=torch.float32, device=f.device)
    sc = torch.empty(n, dtype=torch.float64, device=f.device)
    gpu._check(gpu.lib().pbagpu_ssc_features(C.byref(shock), n, gpu._ptr(f.contiguous()), gpu._ptr(w.contiguous()),
                                             gpu._ptr(B.contiguous()), gpu._ptr(x), gpu._ptr(sc), None),
               "ssc_features")
    return x, sc


def plain_ssc(shock, f, w):
    """The plain SSC contraction of cells (pbagpu_ssc_emissivity), torch FP64 [n][nfreq]."""
    torch = _torch()
    j = torch.empty((f.shape[0], shock.nfreq), dtype=torch.float64, device=f.device)
    gpu._check(gpu.lib().pbagpu_ssc_emissivity(C.byref(shock), f.shape[0], gpu._ptr(f.contiguous()),
                                               gpu._ptr(w.contiguous()), gpu._ptr(j), None), "ssc_emissivity")
    return j


def comoving_grid(shock):
    """Comoving frequency grid of a shock (MakeLogspaceVec, src/utilitites/utils.h:91-105)."""
    nu = np.empty(shock.nfreq)
    cur = 10.0 ** math.log10(shock.freq1)
    ratio = 10.0 ** ((math.log10(shock.freq2) - math.log10(shock.freq1)) / shock.nfreq)
    for i in range(shock.nfreq):
        nu[i] = cur
        cur *= ratio
    return nu


def cells_from_pipeline(pl, step_stride=1):
    """(f, w, B, shock) of every SSC population-step of a pipeline whose FP64 electron stage ran with
    want_f=True: f from f_all, w from the previous step's emissivity, B from the dynamics table."""
    torch = _torch()
    out = []
    nt = pl.nt
    for q, P in enumerate(pl.pops):
        sh = pl.params[P.run].fs if P.shock == 0 else pl.params[P.run].rs
        if not sh.ssc:
            continue
        out.append((q, P, sh))
    if not out:
        raise ValueError("no population of the batch has SSC")
    sh0 = out[0][2]
    nu = torch.from_numpy(comoving_grid(sh0)).to(pl.dev)
    wfac = torch.zeros_like(nu)
    wfac[:-1] = (nu[1:] - nu[:-1]) / (nu[:-1] * nu[:-1])
    fs, ws, bs = [], [], []
    its = torch.arange(1, nt, step_stride, device=pl.dev)
    for q, P, sh in out:
        f = pl.d_fall[q, its]                       # [n][ng]
        w = pl.d_jall[q, its - 1] * wfac[None, :]   # [n][nf]
        bvar = abi.PV["B"] if P.shock == 0 else abi.PV["B_rs"]
        B = pl.d_table[P.bw, bvar, its]
        keep = (f.max(dim=1).values > 0) & (w.max(dim=1).values > 0) & (B > 0)
        fs.append(f[keep]); ws.append(w[keep]); bs.append(B[keep])
    return torch.cat(fs), torch.cat(ws), torch.cat(bs), sh0


def targets(shock, f, w, B):
    """Features x, scales and log10 targets (floored) of cells."""
    torch = _torch()
    x, sc = features(shock, f, w, B)
    j = plain_ssc(shock, f, w)
    g = j / sc[:, None].clamp_min(1e-300)
    gmax = g.max(dim=1, keepdim=True).values.clamp_min(1e-300)
    ylog = torch.log10(torch.maximum(g, gmax * 10.0 ** -FLOOR_DECADES))
    keep = (sc > 0) & torch.isfinite(ylog).all(dim=1)
    return x[keep], sc[keep], ylog[keep], j[keep]


def train(x, ylog, hidden=(512, 512, 512, 512), steps=3000, batch=4096, lr=1e-3, lr_min=1e-5, seed=0,
          log_every=500, log=None):
    """Train a network on features x [n][d_in] and log10 targets ylog [n][nfreq] (torch, device).
    Returns (SscSurrogate, list of (step, loss))."""
    torch = _torch()
    y_off = ylog.mean(dim=0)
    y_sc = float((ylog - y_off).std())
    yt = ((ylog - y_off) / y_sc).float().contiguous()
    dims = [x.shape[1], *hidden, ylog.shape[1]]
    net = Net(dims, seed=seed, device=x.device.index or 0)
    g = torch.Generator(device="cpu").manual_seed(seed + 1)
    n = x.shape[0]
    hist = []
    for s in range(1, steps + 1):
        idx = torch.randint(0, n, (min(batch, n),), generator=g).to(x.device)
        lr_s = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * (s - 1) / max(steps - 1, 1)))
        loss = net.train_step(x[idx].contiguous(), yt[idx].contiguous(), lr_s)
        if s % log_every == 0 or s == 1 or s == steps:
            v = float(loss.item())
            hist.append((s, v))
            if log:
                log(f"[surrogate] step {s}/{steps} loss {v:.4e} lr {lr_s:.2e}")
    return SscSurrogate(net, y_off.cpu().numpy(), y_sc), hist


def relative_error(pred, exact, peak_frac=1e-6):
    """Per-cell max relative error over frequencies above peak_frac of the cell's peak (torch)."""
    peak = exact.abs().max(dim=1, keepdim=True).values
    m = exact.abs() > peak_frac * peak
    e = torch_where(m, (pred - exact).abs() / exact.abs().clamp_min(1e-300), 0.0)
    return e.max(dim=1).values


def torch_where(m, a, b):
    torch = _torch()
    return torch.where(m, a, torch.full_like(a, b))

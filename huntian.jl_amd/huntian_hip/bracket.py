"""The learned RTHL bracket: features, network, training (all on the GPU).

No counterpart in the reference.  The network (18 -> 128 -> 128 -> 128 -> 3, SiLU, BF16 on
the matrix cores with FP32 accumulation) reads 18 features of a peak's first 12 shells and
proposes an RTHL bracket and a collapse probability: log(1 + lo), lo the shell holding RTHL;
log(1 + hi), hi the last shell the exact search reads (the gather extent); the collapse
logit.  hnt_shell_analyse(order="learned") orders the peaks by hi so that a wavefront's
lanes finish their gathers together, then runs the exact search inside the bracket: the
gather until the exact stop rule holds, the walk from the crossing down to lo - 2 first; a
peak whose gather runs past hi + 2 or whose walk passes lo - 2 without collapsing fails
verification and is finished by the full scan.  The exact rules decide every result: results
never depend on the prediction.  hnt_results.bracket records each peak's verification.

Training targets come from the plain method itself: the shell holding RTHL (1 without a
collapse), nshells_used and RTHL > 0 of hnt_shell_analyse.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import abi

NI, H, NO = abi.MLP_INPUTS, abi.MLP_HIDDEN, abi.MLP_OUTPUTS
FEATURE_NAMES = tuple([f"dF{s}" for s in range(1, 13)] +
                      ["e_in", "p_in", "e_out", "p_out", "Fbar12", "Rf_over_alatt"])
SHAPES = (("W1", (NI, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)),
          ("b3", (H,)), ("W4", (H, NO)), ("b4", (NO,)))


def unflatten(p: np.ndarray) -> dict:
    """Flat parameter vector -> dict of arrays (row-major W[k][o])."""
    out, off = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = p[off:off + n].reshape(shp)
        off += n
    assert off == abi.MLP_NPARAM
    return out


def flatten(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float32).ravel() for name, _ in SHAPES])


def init_params(seed: int = 0) -> np.ndarray:
    """He-style uniform init, zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            lim = math.sqrt(6.0 / shp[0])
            d[name] = rng.uniform(-lim, lim, size=shp).astype(np.float32)
        else:
            d[name] = np.zeros(shp, np.float32)
    d["W4"] *= 0.1
    return flatten(d)


def rthl_shell(RTHL, tables):
    """The shell holding each RTHL (first shell s with rad[s] >= RTHL, 1-based), 1 where there
    is no collapse; RTHL a device float64 tensor in lattice units."""
    import torch

    rad = torch.sqrt(torch.as_tensor(np.asarray(tables.shell_r2, np.float64),
                                     device=RTHL.device))
    s = torch.searchsorted(rad, RTHL.contiguous(), right=False) + 1
    return torch.where(RTHL > 0, s, torch.ones_like(s))


def targets(result: dict, tables, npk: int | None = None):
    """(log1p RTHL shell, log1p gather extent, collapsed) from a plain-method result dict of
    device tensors -> device float32 [npk, 3]."""
    import torch

    n = npk if npk is not None else len(result["RTHL"])
    t = torch.empty((n, NO), dtype=torch.float32, device=result["RTHL"].device)
    t[:, 0] = torch.log1p(rthl_shell(result["RTHL"][:n], tables).float())
    t[:, 1] = torch.log1p(result["nshells_used"][:n].float())
    t[:, 2] = (result["RTHL"][:n] > 0).float()
    return t


class BracketNet:
    """Device-resident network (FP32 master weights, input normalisation, Adam state)."""

    def __init__(self, params: np.ndarray | None = None, norm: np.ndarray | None = None,
                 seed: int = 0, device="cuda"):
        import torch

        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        assert p.shape == (abi.MLP_NPARAM,)
        if norm is None:
            norm = np.concatenate([np.zeros(NI, np.float32), np.ones(NI, np.float32)])
        self.params = torch.tensor(p, device=device)
        self.norm = torch.tensor(np.asarray(norm, np.float32), device=device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self._s = abi.Mlp(abi.as_ptr(self.params.data_ptr(), C.c_float),
                          abi.as_ptr(self.norm.data_ptr(), C.c_float),
                          abi.as_ptr(self.m.data_ptr(), C.c_float),
                          abi.as_ptr(self.v.data_ptr(), C.c_float), 0)

    def struct(self) -> abi.Mlp:
        return self._s

    @property
    def step(self) -> int:
        return int(self._s.step)

    def set_norm_from(self, x):
        """Input normalisation from a feature sample (mean, std with a floor)."""
        import torch

        mu = x.mean(dim=0)
        sd = x.std(dim=0).clamp_min(1e-3)
        self.norm.copy_(torch.cat([mu, sd]).float())

    def host_params(self) -> np.ndarray:
        return self.params.cpu().numpy()

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(), norm=self.norm.cpu().numpy())

    @classmethod
    def load(cls, path, device="cuda"):
        with np.load(path, allow_pickle=False) as z:
            return cls(z["params"], z["norm"], device=device)

    # ---------------------------------------------------------------- library calls
    def forward(self, x):
        import torch

        from .api import context

        ctx = context()
        x = x.contiguous().float()
        out = torch.empty((x.shape[0], NO), dtype=torch.float32, device=x.device)
        abi.check(ctx.lib, ctx.lib.hnt_mlp_forward(ctx.handle, C.byref(self._s),
                                                   abi.as_ptr(x.data_ptr(), C.c_float),
                                                   int(x.shape[0]),
                                                   abi.as_ptr(out.data_ptr(), C.c_float)))
        return out

    def gradient(self, x, t):
        import torch

        from .api import context

        ctx = context()
        x = x.contiguous().float()
        t = t.contiguous().float()
        g = torch.empty(abi.MLP_NPARAM, dtype=torch.float32, device=x.device)
        loss = C.c_double()
        abi.check(ctx.lib, ctx.lib.hnt_mlp_gradient(ctx.handle, C.byref(self._s),
                                                    abi.as_ptr(x.data_ptr(), C.c_float),
                                                    abi.as_ptr(t.data_ptr(), C.c_float),
                                                    int(x.shape[0]),
                                                    abi.as_ptr(g.data_ptr(), C.c_float),
                                                    C.byref(loss)))
        return g, loss.value

    def train_step(self, x, t, lr: float, want_loss=False):
        from .api import context

        ctx = context()
        loss = C.c_double()
        abi.check(ctx.lib, ctx.lib.hnt_mlp_train_step(
            ctx.handle, C.byref(self._s), abi.as_ptr(x.data_ptr(), C.c_float),
            abi.as_ptr(t.data_ptr(), C.c_float), int(x.shape[0]), float(lr),
            C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def fit(self, x, t, *, epochs=20, batch=4096, lr=2e-3, seed=0, log=None):
        """Adam over shuffled minibatches; the last partial batch of an epoch is dropped.
        Returns the mean loss of each epoch."""
        import torch

        g = torch.Generator(device=x.device)
        g.manual_seed(seed)
        n = x.shape[0]
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g, device=x.device)
            tot, cnt = 0.0, 0
            for b0 in range(0, n - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                xb = x[idx].contiguous()
                tb = t[idx].contiguous()
                lr_ep = lr * 0.5 * (1.0 + math.cos(math.pi * ep / epochs))
                tot += self.train_step(xb, tb, lr_ep, want_loss=True)
                cnt += 1
            hist.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4f}")
        return hist


def features(fields: dict, pi, pj, pk, Rf, tables, alatt: float):
    """hnt_bracket_features: device float32 [npk, 18]."""
    import torch

    from .api import _dev1, _fields_struct, _ptr, context

    ctx = context()
    dev = torch.device("cuda", ctx.device)
    fs, keep = _fields_struct(fields)
    npk = int(len(pi))
    pi_ = _dev1(pi, torch.int32, dev)
    pj_ = _dev1(pj, torch.int32, dev)
    pk_ = _dev1(pk, torch.int32, dev)
    Rf_ = _dev1(Rf if isinstance(Rf, torch.Tensor) else np.broadcast_to(
        np.asarray(Rf, np.float64), (npk,)), torch.float64, dev)
    b = abi.Batch(npk, abi.as_ptr(_ptr(pi_), C.c_int32), abi.as_ptr(_ptr(pj_), C.c_int32),
                  abi.as_ptr(_ptr(pk_), C.c_int32), abi.as_ptr(_ptr(Rf_), C.c_double),
                  abi.as_ptr(0, C.c_int32), abi.as_ptr(0, C.c_double),
                  abi.as_ptr(0, C.c_double), 1.0, 1.0)
    out = torch.empty((max(npk, 1), NI), dtype=torch.float32, device=dev)
    ts = tables.struct()
    abi.check(ctx.lib, ctx.lib.hnt_bracket_features(ctx.handle, C.byref(fs), C.byref(b),
                                                    C.byref(ts), float(alatt),
                                                    abi.as_ptr(_ptr(out), C.c_float)))
    del keep
    return out[:npk]


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(__file__), "data", "bracket_mlp.npz")

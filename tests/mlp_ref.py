"""TEST INFRASTRUCTURE: CPU restatement of the bracket network (numpy), used to check the
GPU forward pass, gradients and features.  Not imported by the library.

forward/backward in FP64; with bf16=True the inputs, weights, activations and output
gradients are rounded to bfloat16 (round to nearest even) at the points where the GPU
stores them in bfloat16, so the remaining difference is FP32-versus-FP64 accumulation.
"""
import numpy as np

NI, H, NO = 18, 128, 3


def bf16_round(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), np.float32(np.nan), out).astype(np.float64)


def unflatten(p):
    shapes = (("W1", (NI, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)),
              ("b3", (H,)), ("W4", (H, NO)), ("b4", (NO,)))
    out, off = {}, 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[off:off + n], np.float64).reshape(shp)
        off += n
    return out


def flatten(d):
    return np.concatenate([d[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4",
                                                   "b4")])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, norm, x, bf16=False):
    """Returns (y [B, 3], cache)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    w = unflatten(p)
    x = np.asarray(x, np.float64)
    xn = (x - norm[:NI]) / norm[NI:]
    if bf16:  # the GPU normalises in FP32
        xn = ((np.asarray(x, np.float32) - np.asarray(norm[:NI], np.float32))
              / np.asarray(norm[NI:], np.float32)).astype(np.float64)
    a0 = r(xn)
    z1 = a0 @ r(w["W1"]) + w["b1"]
    a1 = r(_silu(z1))
    z2 = a1 @ r(w["W2"]) + w["b2"]
    a2 = r(_silu(z2))
    z3 = a2 @ r(w["W3"]) + w["b3"]
    a3 = r(_silu(z3))
    y = a3 @ r(w["W4"]) + w["b4"]
    return y, dict(w=w, a0=a0, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3, r=r)


def loss_and_grad(p, norm, x, t, bf16=False):
    """Mean over the batch of (y0-t0)^2 + (y1-t1)^2 + BCE-with-logits(y2, t2)."""
    y, c = forward(p, norm, x, bf16)
    r, w = c["r"], c["w"]
    t = np.asarray(t, np.float64)
    B = len(x)
    bce = np.maximum(y[:, 2], 0) - y[:, 2] * t[:, 2] + np.log1p(np.exp(-np.abs(y[:, 2])))
    loss = float(np.mean((y[:, 0] - t[:, 0]) ** 2 + (y[:, 1] - t[:, 1]) ** 2 + bce))
    dy = np.zeros_like(y)
    dy[:, 0] = 2 * (y[:, 0] - t[:, 0]) / B
    dy[:, 1] = 2 * (y[:, 1] - t[:, 1]) / B
    dy[:, 2] = (1.0 / (1.0 + np.exp(-y[:, 2])) - t[:, 2]) / B
    g = {"b4": dy.sum(0)}
    dyr = r(dy)
    g["W4"] = c["a3"].T @ dyr
    dz3 = (dyr @ r(w["W4"]).T) * _silu_grad(c["z3"])
    g["b3"] = dz3.sum(0)
    dz3r = r(dz3)
    g["W3"] = c["a2"].T @ dz3r
    dz2 = (dz3r @ r(w["W3"]).T) * _silu_grad(c["z2"])
    g["b2"] = dz2.sum(0)
    dz2r = r(dz2)
    g["W2"] = c["a1"].T @ dz2r
    dz1 = (dz2r @ r(w["W2"]).T) * _silu_grad(c["z1"])
    g["b1"] = dz1.sum(0)
    g["W1"] = c["a0"].T @ r(dz1)
    return loss, flatten(g)


def features_ref(f, pi, pj, pk, Rf, alatt, t):
    """Brute-force features from the fields dict and oracle shell tables (1-based peaks)."""
    n = f["delta"].shape[0]
    out = np.zeros((len(pi), NI))
    for q in range(len(pi)):
        T = np.zeros((3, 3))
        tot, ntot = 0.0, 0
        for s in range(1, 13):
            st, nc = t["shell_start"][s - 1], t["shell_count"][s - 1]
            F, nin = 0.0, 0
            for c in range(st, st + nc):
                d = (t["di"][c], t["dj"][c], t["dk"][c])
                i1, i2, i3 = pi[q] + d[0], pj[q] + d[1], pk[q] + d[2]
                if not (1 <= i1 <= n and 1 <= i2 <= n and 1 <= i3 <= n):
                    continue
                dv = float(f["delta"][i1 - 1, i2 - 1, i3 - 1])
                F += dv
                nin += 1
                T += dv * np.outer(d, d)
            out[q, s - 1] = F / nin if nin else 0.0
            tot += F
            ntot += nin
            if s == 6:
                out[q, 12:14] = _ep(T)
        out[q, 14:16] = _ep(T)
        out[q, 16] = tot / ntot if ntot else 0.0
        out[q, 17] = Rf[q] / alatt
    return out


def _ep(T):
    lam = np.linalg.eigvalsh(0.5 * (T + T.T))
    F = lam.sum()
    if F <= 0:
        return 0.0, 0.0
    return 0.5 * (lam[2] - lam[0]) / F, 0.5 * (lam[2] + lam[0] - 2 * lam[1]) / F

"""Synthetic inputs of the halo-finding path (the producer side, not the hot path).

Restates the reference's field synthesis closely enough to make realistic inputs:
  * Threefry-2x64-20 white noise (Salmon et al., SC'11; the Random123 algorithm that
    src/InitialConditions/RandomField.jl:45-80 calls with key (seed, 0), counter (pair, 0)),
    Box-Muller to two deviates per counter, cell idx (1-based) taking the first deviate when
    idx - 1 is even.
  * BBKS linear power spectrum P(k) ~ k^ns T(k/Gamma)^2 normalised to sigma8 with the
    top-hat window, stored in the reference's (2 pi)^3-divided convention
    (data/pk_websky.dat); convolution amp = sqrt(P(k) dk^3 n^3), DC zeroed
    (RandomField.jl:134-171).
  * 1LPT and 2LPT displacements (src/InitialConditions/LPT.jl:21-196), Nyquist zeroed.
  * Top-hat filter-bank smoothing, k = 0 left unchanged (src/HaloFinder/Filters.jl:72-99).
Oracle and GPU always consume the same synthesised arrays, so parity does not depend on
this module reproducing FFTW or Random123 bit for bit.

Arrays follow the reference (Julia) layout: shape (n1, n2, n3) indexed [i, j, k] with
column-major storage (numpy order="F"; torch tensors with strides (1, n1, n1*n2)).
"""
from __future__ import annotations

import math

import numpy as np

M64 = (1 << 64) - 1
_R = (16, 42, 12, 31, 16, 32, 24, 21)
_PARITY = 0x1BD11BDAA9FC1A22


# ---------------------------------------------------------------- Threefry-2x64-20

def threefry2x64_20(c0, c1, k0, k1):
    """Threefry-2x64 with 20 rounds on uint64 numpy arrays (or Python ints)."""
    c0 = np.asarray(c0, dtype=np.uint64)
    c1 = np.asarray(c1, dtype=np.uint64)
    k0 = np.uint64(k0) if np.ndim(k0) == 0 else np.asarray(k0, np.uint64)
    k1 = np.uint64(k1) if np.ndim(k1) == 0 else np.asarray(k1, np.uint64)
    ks2 = np.uint64(_PARITY) ^ k0 ^ k1
    ks = (k0, k1, ks2)
    with np.errstate(over="ignore"):
        x0 = c0 + ks[0]
        x1 = c1 + ks[1]
        for r in range(20):
            x0 = x0 + x1
            rot = np.uint64(_R[r % 8])
            x1 = (x1 << rot) | (x1 >> (np.uint64(64) - rot))
            x1 = x1 ^ x0
            if r % 4 == 3:
                s = (r + 1) // 4
                x0 = x0 + ks[s % 3]
                x1 = x1 + ks[(s + 1) % 3] + np.uint64(s)
    return x0, x1


def _u64_to_unit(u):
    # (x >>> 11) + 1, times 2^-53: a double in (0, 1]
    return ((u >> np.uint64(11)).astype(np.float64) + 1.0) * 1.1102230246251565e-16


def white_noise(n: int, seed: int, dtype=np.float32) -> np.ndarray:
    """Gaussian white noise of the Threefry stream, Fortran-ordered (n, n, n)."""
    n3 = n * n * n
    npair = (n3 + 1) // 2
    out = np.empty(n3, dtype=np.float64)
    step = 1 << 22
    for a in range(0, npair, step):
        b = min(npair, a + step)
        pair = np.arange(a, b, dtype=np.uint64)
        r0, r1 = threefry2x64_20(pair, np.zeros_like(pair), seed, 0)
        u1 = _u64_to_unit(r0)
        u2 = _u64_to_unit(r1)
        r = np.sqrt(-2.0 * np.log(u1))
        th = 2.0 * math.pi * u2
        g1 = r * np.cos(th)
        g2 = r * np.sin(th)
        lo = 2 * a
        hi = min(n3, 2 * b)
        seg = np.empty(2 * (b - a))
        seg[0::2] = g1
        seg[1::2] = g2
        out[lo:hi] = seg[: hi - lo]
    return out.astype(dtype).reshape((n, n, n), order="F")


def white_noise_torch(n: int, seed: int, device):
    """The same Threefry noise computed on the GPU with int64 arithmetic (wraps like uint64).
    Returns a torch float32 tensor of shape (n, n, n) with column-major strides."""
    import torch

    n3 = n * n * n
    npair = (n3 + 1) // 2

    def u(x):
        return torch.tensor(x - (1 << 64) if x >= (1 << 63) else x, dtype=torch.int64,
                            device=device)

    def rotl(x, r):
        # logical shift right emulated on int64
        return (x << r) | ((x >> (64 - r)) & ((1 << r) - 1))

    k0, k1 = seed & M64, 0
    ks = (u(k0), u(k1), u(_PARITY ^ k0 ^ k1))
    flat = torch.empty(2 * npair, dtype=torch.float32, device=device)
    step = 1 << 25
    for a in range(0, npair, step):
        b = min(npair, a + step)
        x0 = torch.arange(a, b, dtype=torch.int64, device=device) + ks[0]
        x1 = torch.zeros(b - a, dtype=torch.int64, device=device) + ks[1]
        for r in range(20):
            x0 = x0 + x1
            x1 = rotl(x1, _R[r % 8])
            x1 = x1 ^ x0
            if r % 4 == 3:
                s = (r + 1) // 4
                x0 = x0 + ks[s % 3]
                x1 = x1 + ks[(s + 1) % 3] + s

        def unit(x):
            hi = (x >> 11) & ((1 << 53) - 1)
            return (hi.to(torch.float64) + 1.0) * 1.1102230246251565e-16

        u1 = unit(x0)
        u2 = unit(x1)
        rr = torch.sqrt(-2.0 * torch.log(u1))
        th = 2.0 * math.pi * u2
        seg = torch.stack([rr * torch.cos(th), rr * torch.sin(th)], dim=1).reshape(-1)
        flat[2 * a: 2 * b] = seg.to(torch.float32)
    return flat[:n3].reshape(n, n, n).permute(2, 1, 0)


# ---------------------------------------------------------------- power spectrum

def bbks_transfer(k, gamma=0.21):
    q = np.maximum(np.asarray(k, np.float64) / gamma, 1e-30)
    return (np.log(1.0 + 2.34 * q) / (2.34 * q)) * (
        1.0 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25


def tophat_w(x):
    x = np.asarray(x, np.float64)
    out = np.ones_like(x)
    nz = x != 0
    xs = x[nz]
    out[nz] = 3.0 * (np.sin(xs) - xs * np.cos(xs)) / xs ** 3
    return out


class PowerSpectrum:
    """P(k) [(Mpc/h)^3] ~ k^ns T_BBKS^2, normalised to sigma8 at R = 8 Mpc/h with
    sigma^2(R) = 1/(2 pi^2) int k^2 P W^2 dk (src/MassFunction/MassFunction.jl:28-38).
    file(k) returns the (2 pi)^3-divided value the convolution uses."""

    def __init__(self, ns=0.965, gamma=0.21, sigma8=0.81):
        self.ns, self.gamma = ns, gamma
        lk = np.linspace(np.log(1e-5), np.log(1e3), 200001)
        k = np.exp(lk)
        integrand = k ** 3 * self._shape(k) * tophat_w(8.0 * k) ** 2
        s2 = np.trapezoid(integrand, lk) / (2.0 * np.pi ** 2)
        self.amp = sigma8 ** 2 / s2

    def _shape(self, k):
        return np.asarray(k, np.float64) ** self.ns * bbks_transfer(k, self.gamma) ** 2

    def phys(self, k):
        return self.amp * self._shape(k)

    def file(self, k):
        return self.phys(k) / (2.0 * np.pi) ** 3


# ---------------------------------------------------------------- fields (numpy)

def _kgrid(n, boxsize):
    dk = 2.0 * np.pi / boxsize
    kx = np.fft.rfftfreq(n, 1.0 / n) * dk  # along i (the real axis)
    ky = np.fft.fftfreq(n, 1.0 / n) * dk
    kz = np.fft.fftfreq(n, 1.0 / n) * dk
    # C-order view b[k, j, i] of the Fortran array a[i, j, k]
    return kz[:, None, None], ky[None, :, None], kx[None, None, :], dk


def gaussian_field(n, boxsize, seed, pk: PowerSpectrum | None = None):
    """delta on an n^3 periodic grid; returns (delta Fortran array, delta_k C-order view)."""
    pk = pk or PowerSpectrum()
    noise = white_noise(n, seed, np.float32)
    b = np.ascontiguousarray(noise.T)  # [k, j, i]
    fk = np.fft.rfftn(b.astype(np.float64))
    KZ, KY, KX, dk = _kgrid(n, boxsize)
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k = np.sqrt(k2)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.sqrt(pk.file(np.where(k2 > 0, k, 1.0)) * dk ** 3 * float(n) ** 3)
    amp = np.where(k2 > 0, amp, 0.0)
    fk = fk * amp
    d = np.fft.irfftn(fk, s=b.shape, axes=(0, 1, 2)).astype(np.float32)
    dk_ = np.fft.rfftn(d.astype(np.float64))
    return np.asfortranarray(d.T), dk_


def lpt_fields(delta_k, n, boxsize, second_order=True):
    """psi_1 (and psi_2) from delta_k (C-order rfft of delta[k, j, i])."""
    KZ, KY, KX, _ = _kgrid(n, boxsize)
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    nyq = n // 2
    kidx_x = np.arange(n // 2 + 1)[None, None, :]
    kidx_y = np.arange(n)[None, :, None]
    kidx_z = np.arange(n)[:, None, None]
    zero_nyq = (kidx_x == nyq) | (kidx_y == nyq) | (kidx_z == nyq)
    inv = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1.0), 0.0)
    out = {}
    for name, kc in (("etax", KX), ("etay", KY), ("etaz", KZ)):
        pk_ = 1j * kc * inv * delta_k
        pk_ = np.where(zero_nyq | (k2 == 0), 0.0, pk_)
        out[name] = np.asfortranarray(np.fft.irfftn(pk_, s=(n, n, n), axes=(0, 1, 2)).astype(np.float32).T)
    if second_order:
        comps = {}
        for a, ka in (("x", KX), ("y", KY), ("z", KZ)):
            for b_, kb in (("x", KX), ("y", KY), ("z", KZ)):
                if a + b_ in ("xx", "yy", "zz", "xy", "xz", "yz"):
                    comps[a + b_] = np.fft.irfftn(-ka * kb * inv * delta_k, s=(n, n, n), axes=(0, 1, 2))
        src = (comps["xx"] * comps["yy"] - comps["xy"] ** 2 + comps["xx"] * comps["zz"]
               - comps["xz"] ** 2 + comps["yy"] * comps["zz"] - comps["yz"] ** 2)
        sk = np.fft.rfftn(src)
        for name, kc in (("eta2x", KX), ("eta2y", KY), ("eta2z", KZ)):
            pk_ = -1j * kc * inv * sk
            pk_ = np.where(zero_nyq | (k2 == 0), 0.0, pk_)
            out[name] = np.asfortranarray(np.fft.irfftn(pk_, s=(n, n, n), axes=(0, 1, 2)).astype(np.float32).T)
    return out


def laplacian_field(delta_k, n, boxsize):
    KZ, KY, KX, _ = _kgrid(n, boxsize)
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    return np.asfortranarray(np.fft.irfftn(-k2 * delta_k, s=(n, n, n), axes=(0, 1, 2)).astype(np.float32).T)


def smooth_tophat(delta_k, n, boxsize, R):
    """Top-hat smoothing at radius R; k = 0 untouched (Filters.jl:72-99)."""
    KZ, KY, KX, _ = _kgrid(n, boxsize)
    k = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    w = tophat_w(k * R).astype(np.float32).astype(np.float64)
    w = np.where(k == 0, 1.0, w)
    return np.asfortranarray(np.fft.irfftn(delta_k * w, s=(n, n, n), axes=(0, 1, 2)).astype(np.float32).T)


def filter_radii(rmin, rmax, nfilt):
    """Log-spaced radii, largest first (the order the drivers process them in)."""
    return list(np.geomspace(rmin, rmax, nfilt)[::-1])


def ir2min_of(Rf, alatt):
    """ir2min = min(floor((1.75 Rf/alatt)^2), floor((40/alatt - 1)^2))
    (src/MultiResolution.jl:575-576)."""
    return min(int(math.floor((1.75 * Rf / alatt) ** 2)), int(math.floor((40.0 / alatt - 1) ** 2)))


# ---------------------------------------------------------------- fields (torch, on the GPU)

def _kgrid_torch(n, boxsize, device):
    import torch

    dk = 2.0 * np.pi / boxsize
    kx = torch.arange(n // 2 + 1, device=device, dtype=torch.float32) * dk
    f = torch.fft.fftfreq(n, d=1.0 / n, device=device, dtype=torch.float32) * dk
    return kx[None, None, :], f[None, :, None], f[:, None, None]


def _power_torch(pk: PowerSpectrum, k):
    import torch

    q = torch.clamp(k / pk.gamma, min=1e-30)
    T = torch.log(1.0 + 2.34 * q) / (2.34 * q) * (
        1.0 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25
    return pk.amp * k ** pk.ns * T * T / (2.0 * np.pi) ** 3


def fields_torch(n, boxsize, seed, device, second_order=True, pk: PowerSpectrum | None = None):
    """delta, psi_1 (and psi_2) on the GPU in float32 / complex64, same recipe as the numpy
    path.  Returns (dict of column-major (n,n,n) tensors, delta_k (C-order rfft))."""
    import torch

    pk = pk or PowerSpectrum()
    noise = white_noise_torch(n, seed, device)          # [i,j,k] view of C-order [k,j,i]
    fk = torch.fft.rfftn(noise.permute(2, 1, 0))
    del noise
    KX, KY, KZ = _kgrid_torch(n, boxsize, device)
    dk = 2.0 * np.pi / boxsize
    for z0 in range(0, n, 64):  # slab-wise to bound temporaries
        sl = slice(z0, min(n, z0 + 64))
        k2 = KX ** 2 + KY ** 2 + KZ[sl] ** 2
        k = torch.sqrt(k2)
        amp = torch.sqrt(_power_torch(pk, torch.where(k2 > 0, k, torch.ones_like(k)))
                         * dk ** 3 * float(n) ** 3)
        fk[sl] *= torch.where(k2 > 0, amp, torch.zeros_like(amp))
    delta = torch.fft.irfftn(fk, s=(n, n, n))
    del fk
    out = {"delta": delta.permute(2, 1, 0)}
    dkk = torch.fft.rfftn(delta)
    nyq = n // 2
    ix = torch.arange(n // 2 + 1, device=device)[None, None, :]
    iy = torch.arange(n, device=device)[None, :, None]
    iz = torch.arange(n, device=device)[:, None, None]

    def kfilter(fn):
        res = torch.empty_like(dkk)
        for z0 in range(0, n, 64):
            sl = slice(z0, min(n, z0 + 64))
            k2 = KX ** 2 + KY ** 2 + KZ[sl] ** 2
            inv = torch.where(k2 > 0, 1.0 / torch.where(k2 > 0, k2, torch.ones_like(k2)),
                              torch.zeros_like(k2))
            res[sl] = fn(sl, inv)
        return res

    def nyq_zero(sl, v):
        z = (ix == nyq) | (iy == nyq) | (iz[sl] == nyq)
        return torch.where(z, torch.zeros_like(v), v)

    for name, kc in (("etax", 0), ("etay", 1), ("etaz", 2)):
        comp = (KX, KY, KZ)[kc]
        pkk = kfilter(lambda sl, inv: nyq_zero(sl, 1j * (comp if kc < 2 else comp[sl]) * inv * dkk[sl]))
        out[name] = torch.fft.irfftn(pkk, s=(n, n, n)).permute(2, 1, 0)
        del pkk
    if second_order:
        def phi(a, b):
            ka, kb = (KX, KY, KZ)[a], (KX, KY, KZ)[b]
            pkk = kfilter(lambda sl, inv: -(ka if a < 2 else ka[sl]) * (kb if b < 2 else kb[sl])
                          * inv * dkk[sl])
            return torch.fft.irfftn(pkk, s=(n, n, n))

        p11, p22, p33 = phi(0, 0), phi(1, 1), phi(2, 2)
        src = p11 * p22 + p11 * p33 + p22 * p33
        del p11, p22, p33
        for a, b in ((0, 1), (0, 2), (1, 2)):
            p = phi(a, b)
            src -= p * p
            del p
        sk = torch.fft.rfftn(src)
        del src
        for name, kc in (("eta2x", 0), ("eta2y", 1), ("eta2z", 2)):
            comp = (KX, KY, KZ)[kc]
            pkk = kfilter(lambda sl, inv: nyq_zero(sl, -1j * (comp if kc < 2 else comp[sl]) * inv * sk[sl]))
            out[name] = torch.fft.irfftn(pkk, s=(n, n, n)).permute(2, 1, 0)
            del pkk
        del sk
    return out, dkk


def smooth_tophat_torch(delta_k, n, boxsize, R):
    """Top-hat smoothing on the GPU (k = 0 untouched), column-major float32 result."""
    import torch

    KX, KY, KZ = _kgrid_torch(n, boxsize, delta_k.device)
    out = torch.empty_like(delta_k)
    for z0 in range(0, n, 64):
        sl = slice(z0, min(n, z0 + 64))
        k = torch.sqrt(KX ** 2 + KY ** 2 + KZ[sl] ** 2)
        x = k * R
        xs = torch.where(x > 0, x, torch.ones_like(x))
        w = 3.0 * (torch.sin(xs) - xs * torch.cos(xs)) / xs ** 3
        w = torch.where(x > 0, w, torch.ones_like(w))
        out[sl] = delta_k[sl] * w
    return torch.fft.irfftn(out, s=(n, n, n)).permute(2, 1, 0)


# ---------------------------------------------------------------- benchmark workloads

# name: n, box (Mpc/h), Threefry seed, filter bank (rmin, rmax, count; log-spaced, largest
# first), buffer cells, hunt radius in cells.  Peak threshold 1.5 for every filter, collapse
# at z = 0 with HomelTab_websky.dat, alatt = box / n = 2 Mpc/h throughout.
WORKLOADS = {
    "tile1024": dict(n=1024, box=2048.0, seed=2024, rmin=1.5, rmax=34.0, nfilt=24, nbuff=64,
                     nhunt=30),
    "tile512": dict(n=512, box=1024.0, seed=7, rmin=1.5, rmax=20.0, nfilt=16, nbuff=32,
                    nhunt=20),
    "tile128": dict(n=128, box=256.0, seed=12345, rmin=1.5, rmax=20.0, nfilt=16, nbuff=16,
                    nhunt=15),
    # 64 tiles, sides drawn from {96, 128, 192, 256}, buffers 16..48, seeds 100..163
    "batched64": dict(ntiles=64, sides=(96, 128, 192, 256), nbuff=(16, 48), seeds=(100, 163),
                      rmin=1.5, rmax=20.0, nfilt=16, nhunt=15, alatt=2.0),
}


def workload_tile(name: str, device, smoothed: bool = True, seed: int | None = None):
    """Inputs of one tile workload, synthesised on the GPU: the fields (delta, psi_1, psi_2)
    and, with smoothed, the filter bank's smoothed deltas (largest radius first)."""
    import torch

    w = dict(WORKLOADS[name])
    if seed is not None:
        w["seed"] = seed
    n, box = w["n"], w["box"]
    fields, dk = fields_torch(n, box, w["seed"], device)
    radii = filter_radii(w["rmin"], w["rmax"], w["nfilt"])
    sm = [smooth_tophat_torch(dk, n, box, R) for R in radii] if smoothed else []
    del dk
    torch.cuda.synchronize(device)
    w.update(fields=fields, smoothed=sm, radii=radii, alatt=box / n)
    return w


def batched_specs(name: str = "batched64"):
    """(side, nbuff, seed) of every tile of a batched workload: sides and buffers drawn with
    numpy's default generator seeded by the first tile seed, seeds consecutive."""
    w = WORKLOADS[name]
    rng = np.random.default_rng(w["seeds"][0])
    nt = w["ntiles"]
    sides = rng.choice(w["sides"], size=nt)
    nbuffs = rng.integers(w["nbuff"][0], w["nbuff"][1] + 1, size=nt)
    return [(int(sides[q]), int(nbuffs[q]), w["seeds"][0] + q) for q in range(nt)]

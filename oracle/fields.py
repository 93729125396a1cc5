"""Oracle for the tile field synthesis -- TEST INFRASTRUCTURE ONLY (the library never imports
it; only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may).

numpy restatement, in FP64 except where noted, of the GPU methods that build a tile's fields:

  isolated_convolve   isolated_convolve_gpu(_multi) (ext/CUDAExt.jl:2475-2622): zero-pad to
                      2 n_input, rfft, transfer (_transfer_kernel!, :2367-2432) with the
                      tabulated sqrt P(k) (_interp_pk_table, :2351-2364, Float32 as there) or,
                      with table=None, P(k) evaluated directly as the CPU _isolated_convolve
                      (src/MultiResolution.jl:237-270); Hermitian DC / Nyquist planes
                      (_hermitize_planes_kernel!, :2439-2473); irfft; crop.
  periodic_kernel     _periodic_kernel! (ext/CUDAExt.jl:2629-2677)
  compute_2lpt        compute_2lpt_gpu (ext/CUDAExt.jl:2724-2791)
  compute_laplacian   compute_laplacian_gpu (ext/CUDAExt.jl:2793-2825)
  isolated_poisson_psi, isolated_short_pot   (ext/CUDAExt.jl:2679-2722)
  interpolate_to_tile _tricubic_kernel! / _interpolate_to_tile (ext/CUDAExt.jl:3083-3140,
                      src/MultiResolution.jl:91-185), in Float32 with the reference's
                      operation order (bit for bit).

Arrays are indexed [i, j, k] (x first, the reference's column-major order); transforms run
over axes (2, 1, 0) so that the halved axis is i, as in the reference's rfft.  Pinned by
closed forms, explicit DFT sums and identities in tests/test_fields_cpu.py."""
from __future__ import annotations

import numpy as np

AX = (2, 1, 0)


def rfft3(a):
    return np.fft.rfftn(np.asarray(a, np.float64), axes=AX)


def irfft3(A, n):
    return np.fft.irfftn(A, s=(n, n, n), axes=AX)


def kgrid(nk, n, dk, nyquist_negative):
    """Float64 kx (>= 0), ky, kz of a half spectrum [ix, iy, iz]; the index n/2 of ky, kz is
    -n/2 when nyquist_negative (fftfreq, _transfer_kernel!) and +n/2 otherwise
    (_periodic_kernel!)."""
    idx = np.arange(n)
    half = n // 2
    if nyquist_negative:
        signed = np.where(idx < half, idx, idx - n)
    else:
        signed = np.where(idx <= half, idx, idx - n)
    kx = (np.arange(nk) * dk)[:, None, None]
    ky = (signed * dk)[None, :, None]
    kz = (signed * dk)[None, None, :]
    return kx, ky, kz


def interp_pk_table(tab, log_k, log_k_min, d_log_k):
    """_interp_pk_table in Float32 (ext/CUDAExt.jl:2351-2364)."""
    f32 = np.float32
    tab = np.asarray(tab, f32)
    fpos = (log_k.astype(f32) - f32(log_k_min)) / f32(d_log_k)
    i0 = np.floor(fpos).astype(np.int64) + 1
    nt = len(tab)
    ic = np.clip(i0, 1, nt - 1)
    frac = fpos - (i0 - 1).astype(f32)
    mid = tab[ic - 1] + frac * (tab[ic] - tab[ic - 1])
    return np.where(i0 < 1, tab[0], np.where(i0 >= nt, tab[nt - 1], mid))


def transfer(kid, d1, d2, kx, ky, kz, k2):
    """The transfer of kernel_fn_id (complex FP64)."""
    pick = lambda d: kx if d == 1 else (ky if d == 2 else kz)
    with np.errstate(divide="ignore", invalid="ignore"):
        if kid == 0:
            return np.ones_like(k2)
        if kid == 1:
            return 1j * pick(d1) / k2
        if kid == 2:
            return -1j * pick(d1) / k2
        if kid == 3:
            return -pick(d1) * pick(d2) / k2
        if kid == 4:
            return k2
        if kid == 5:
            return -1.0 / k2
    raise ValueError(kid)


def hermitize_planes(A):
    """Average each conjugate pair (iy, iz) ~ (-iy, -iz) of the kx = 0 and kx = Nyquist
    planes, and make the self-conjugate modes real."""
    n2 = A.shape[1]
    out = A.copy()
    idx = np.arange(n2)
    conj_idx = (-idx) % n2
    for ix in (0, A.shape[0] - 1):
        P = A[ix]
        Pc = np.conj(P[np.ix_(conj_idx, conj_idx)])
        out[ix] = 0.5 * (P + Pc)
    return out


def isolated_convolve(noise, pk, boxsize_local, n, kernel=(0, 0, 0), nshell=0, table=None):
    """One output of the isolated convolution.  table = (values, log_k_min, d_log_k) gives the
    GPU method's tabulated sqrt(P dk^3 n2^3); table=None evaluates sqrt(pk(k) dk^3 n2^3)."""
    noise = np.asarray(noise, np.float64)
    n_input = noise.shape[0]
    n2 = 2 * n_input
    dx = boxsize_local / n
    dk = 2 * np.pi / (n2 * dx)
    padded = np.zeros((n2, n2, n2))
    padded[:n_input, :n_input, :n_input] = noise
    A = rfft3(padded)
    kx, ky, kz = kgrid(A.shape[0], n2, dk, nyquist_negative=True)
    k2 = kx * kx + ky * ky + kz * kz
    with np.errstate(divide="ignore", invalid="ignore"):
        if table is None:
            amp = np.sqrt(np.asarray(pk(np.sqrt(np.where(k2 > 0, k2, 1.0))), np.float64)
                          * dk ** 3 * float(n2) ** 3)
        else:
            lk = (np.log(np.where(k2 > 0, k2, 1.0)) * 0.5).astype(np.float32)
            amp = interp_pk_table(table[0], lk, table[1], table[2]).astype(np.float64)
        T = amp * transfer(*kernel, kx, ky, kz, k2)
    A = np.where(k2 == 0.0, 0.0, A * np.where(k2 == 0.0, 0.0, T))
    A = hermitize_planes(A)
    r = irfft3(A, n2)
    s = nshell
    return r[s:s + n, s:s + n, s:s + n]


def periodic_kernel(A, n, dk, kid, d1=0, d2=0):
    """_periodic_kernel!: k = 0 zeroed; the Nyquist planes zeroed except for phi_ij (id 3);
    id 2 -i k_d1/k^2, 3 -k_d1 k_d2/k^2, 4 k^2."""
    kx, ky, kz = kgrid(A.shape[0], n, dk, nyquist_negative=False)
    k2 = kx * kx + ky * ky + kz * kz
    nk, nyq = n // 2 + 1, n // 2
    ix = np.arange(nk)[:, None, None]
    iy = np.arange(n)[None, :, None]
    iz = np.arange(n)[None, None, :]
    zero = k2 == 0.0
    if kid != 3:
        zero = zero | (ix == nk - 1) | (iy == nyq) | (iz == nyq)
    T = transfer(kid, d1, d2, kx, ky, kz, k2)
    return np.where(zero, 0.0, A * np.where(zero, 0.0, T))


def compute_2lpt(delta, n, boxsize_local):
    """(psi2_x, psi2_y, psi2_z) on the periodic tile (compute_2lpt_gpu)."""
    delta = np.asarray(delta, np.float64)
    dk = 2 * np.pi / boxsize_local
    D = rfft3(delta)
    src2 = delta ** 2 * 0.5
    for di, dj, coef in ((1, 1, 0.5), (2, 2, 0.5), (3, 3, 0.5), (1, 2, 1.0), (1, 3, 1.0),
                         (2, 3, 1.0)):
        phi = irfft3(periodic_kernel(D, n, dk, 3, di, dj), n)
        src2 -= phi ** 2 * coef
    S = rfft3(src2)
    return tuple(irfft3(periodic_kernel(S, n, dk, 2, d), n) for d in (1, 2, 3))


def compute_laplacian(delta, n, boxsize_local):
    dk = 2 * np.pi / boxsize_local
    return irfft3(periodic_kernel(rfft3(delta), n, dk, 4), n)


def isolated_poisson_psi(dS, boxsize_local):
    dS = np.asarray(dS, np.float64)
    n = dS.shape[0]
    n2 = 2 * n
    padded = np.zeros((n2, n2, n2))
    padded[:n, :n, :n] = dS
    F = rfft3(padded)
    dk = 2 * np.pi / (2 * boxsize_local)
    return tuple(-irfft3(periodic_kernel(F, n2, dk, 2, d), n2)[:n, :n, :n] for d in (1, 2, 3))


def isolated_short_pot(delta, boxsize_local, rs):
    delta = np.asarray(delta, np.float64)
    n = delta.shape[0]
    n2 = 2 * n
    padded = np.zeros((n2, n2, n2))
    padded[:n, :n, :n] = delta
    F = rfft3(padded)
    dk = 2 * np.pi / (2 * boxsize_local)
    kx, ky, kz = kgrid(F.shape[0], n2, dk, nyquist_negative=False)
    k2 = kx * kx + ky * ky + kz * kz
    rs2 = rs * rs
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = np.where(k2 == 0.0, -rs2, -(1.0 - np.exp(-k2 * rs2)) / k2)
    return irfft3(F * fac, n2)[:n, :n, :n]


def _catmull_rom(t):
    f32 = np.float32
    t2 = t * t
    t3 = t2 * t
    return (f32(-0.5) * t3 + t2 - f32(0.5) * t,
            f32(1.5) * t3 - f32(2.5) * t2 + f32(1.0),
            f32(-1.5) * t3 + f32(2.0) * t2 + f32(0.5) * t,
            f32(0.5) * t3 - f32(0.5) * t2)


def interpolate_to_tile(coarse, it, jt, kt, nsub, nmesh, N, M):
    """Tri-cubic Catmull-Rom interpolation, Float32 in the reference's order: the tile's
    (nmesh,)*3 Float32 array indexed [i, j, k]."""
    coarse = np.asarray(coarse, np.float32)
    block = N // M
    nbuff = (nmesh - nsub) // 2
    starts = [(t - 1) * nsub + 1 - nbuff for t in (it, jt, kt)]
    axes = []
    for s in starts:
        l = np.arange(1, nmesh + 1)
        g = (s + l - 2) % N + 1                          # mod1(s + l - 1, N)
        f = (g.astype(np.float64) - 0.5) / block + 0.5
        i = np.floor(f).astype(np.int64)
        w = _catmull_rom((f - i).astype(np.float32))
        idx = [(i + q - 2) % M for q in range(4)]       # mod1(i - 1 + q, M) - 1
        axes.append((w, idx))
    (wx, xs), (wy, ys), (wz, zs) = axes
    val = np.zeros((nmesh, nmesh, nmesh), np.float32)
    for c in range(4):
        for b in range(4):
            for a in range(4):
                w = ((wx[a][:, None, None] * wy[b][None, :, None]) * wz[c][None, None, :])
                val = val + w * coarse[np.ix_(xs[a], ys[b], zs[c])]
    return val

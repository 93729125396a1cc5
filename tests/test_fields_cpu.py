"""Tile field synthesis: the oracle (oracle/fields.py) pinned by closed forms (white noise
through a unit transfer, plane waves through the periodic 2LPT and Laplacian), explicit DFT
sums on tiny grids (conventions, normalisation, the Hermitian DC / Nyquist planes), the trace
identity sum_i phi_ii = -delta, Catmull-Rom's partition of unity and node reproduction; and
the host sqrt P(k) table (huntian_hip.api.pk_table) against direct evaluation."""
import itertools

import numpy as np
import pytest

from oracle import fields as F


def unit_pk(n, n_input, boxsize):
    """P(k) that makes sqrt(P dk^3 n2^3) = 1 at every k."""
    n2 = 2 * n_input
    dk = 2 * np.pi / (n2 * boxsize / n)
    return lambda k: np.full_like(np.asarray(k, np.float64), 1.0 / (dk ** 3 * n2 ** 3))


def unit_table(n=16):
    return np.ones(n, np.float32), np.float32(-5.0), np.float32(0.5)


def test_unit_transfer_subtracts_the_padded_mean():
    """With sqrt P = 1 and kernel 0 the transfer is the identity except k = 0, which is set to
    0: the output is the noise less its mean over the padded box, cropped."""
    rng = np.random.default_rng(1)
    for n, nshell in ((6, 0), (4, 2)):
        n_input = n + 2 * nshell
        noise = rng.normal(size=(n_input,) * 3)
        want = (noise - noise.sum() / (2 * n_input) ** 3)[nshell:nshell + n, nshell:nshell + n,
                                                           nshell:nshell + n]
        for table in (None, unit_table()):
            got = F.isolated_convolve(noise, unit_pk(n, n_input, 50.0), 50.0, n, (0, 0, 0),
                                      nshell, table=table)
            np.testing.assert_allclose(got, want, atol=1e-12)


def brute_isolated(noise, n, boxsize, kernel, nshell=0):
    """Explicit DFT sums: X(k) = sum_x pad(x) e^{-ik.x}; Y = X T on the half spectrum
    (kx >= 0; ky, kz of index n2/2 negative), the kx = 0 and kx = n2/2 planes averaged with
    their conjugate partners, the other half filled by Y(-k) conjugated; out = (1/n2^3) Re
    sum_k Y(k) e^{ik.x}.  Unit sqrt P."""
    n_input = noise.shape[0]
    n2 = 2 * n_input
    dk = 2 * np.pi / (n2 * boxsize / n)
    pad = np.zeros((n2,) * 3)
    pad[:n_input, :n_input, :n_input] = noise
    m = np.arange(n2)
    E = np.exp(-2j * np.pi * np.outer(m, m) / n2)          # E[k, x]
    X = np.einsum("ax,by,cz,xyz->abc", E, E, E, pad)       # full DFT [kx, ky, kz]
    signed = np.where(m < n2 // 2, m, m - n2)
    Y = np.zeros_like(X)
    half = n2 // 2
    for a, b, c in itertools.product(range(half + 1), range(n2), range(n2)):
        kx, ky, kz = a * dk, signed[b] * dk, signed[c] * dk
        k2 = kx * kx + ky * ky + kz * kz
        if k2 == 0:
            continue
        Y[a, b, c] = X[a, b, c] * F.transfer(*kernel, kx, ky, kz, k2)
    for a in (0, half):
        P = Y[a].copy()
        Y[a] = 0.5 * (P + np.conj(P[np.ix_((-m) % n2, (-m) % n2)]))
    for a in range(half + 1, n2):
        Y[a] = np.conj(Y[(-a) % n2][np.ix_((-m) % n2, (-m) % n2)])
    Einv = np.conj(E)
    r = np.einsum("ax,by,cz,abc->xyz", Einv, Einv, Einv, Y).real / n2 ** 3
    s = nshell
    return r[s:s + n, s:s + n, s:s + n]


@pytest.mark.parametrize("kernel", [(0, 0, 0), (1, 1, 0), (1, 2, 0), (2, 3, 0), (3, 1, 2),
                                    (3, 3, 3), (4, 0, 0), (5, 0, 0)])
def test_isolated_convolve_matches_explicit_dft_sums(kernel):
    rng = np.random.default_rng(5)
    n, nshell = 3, 0
    noise = rng.normal(size=(n,) * 3)
    got = F.isolated_convolve(noise, unit_pk(n, n, 30.0), 30.0, n, kernel, nshell)
    np.testing.assert_allclose(got, brute_isolated(noise, n, 30.0, kernel), atol=1e-12)
    n, nshell = 2, 1
    noise = rng.normal(size=(4,) * 3)
    got = F.isolated_convolve(noise, unit_pk(n, 4, 30.0), 30.0, n, kernel, nshell)
    np.testing.assert_allclose(got, brute_isolated(noise, n, 30.0, kernel, nshell), atol=1e-12)


def test_phi_trace_and_lpt_sign_identities():
    rng = np.random.default_rng(9)
    n = 8
    noise = rng.normal(size=(n,) * 3).astype(np.float32)
    pk = lambda k: 100.0 / (1.0 + (k * 50.0) ** 4)
    d = F.isolated_convolve(noise, pk, 100.0, n, (0, 0, 0))
    tr = sum(F.isolated_convolve(noise, pk, 100.0, n, (3, q, q)) for q in (1, 2, 3))
    np.testing.assert_allclose(tr, -d, atol=1e-12 * np.abs(d).max())
    for q in (1, 2, 3):
        np.testing.assert_array_equal(F.isolated_convolve(noise, pk, 100.0, n, (2, q, 0)),
                                      -F.isolated_convolve(noise, pk, 100.0, n, (1, q, 0)))
    # phi_ij is symmetric in (i, j)
    np.testing.assert_allclose(F.isolated_convolve(noise, pk, 100.0, n, (3, 1, 2)),
                               F.isolated_convolve(noise, pk, 100.0, n, (3, 2, 1)), atol=1e-15)


def test_signs_from_the_physics():
    """A positive Gaussian bump: the 1LPT displacement (kernel 1, -grad of the potential phi
    with lap phi = delta) points at the bump, phi (kernel 5, -1/k^2) is lowest at it, and
    kernel 4 (+k^2, minus the Laplacian) is positive at its top; kernel 2 is the opposite
    of kernel 1."""
    n = 16
    g = np.arange(n) - 7.5
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    bump = np.exp(-(X * X + Y * Y + Z * Z) / 4.0)
    pk = unit_pk(n, n, 32.0)
    psi = F.isolated_convolve(bump, pk, 32.0, n, (1, 1, 0))
    assert psi[10, 7, 7] < 0 < psi[5, 7, 7]
    assert np.all(np.sign(psi[9:13, 7, 7]) == -1) and np.all(np.sign(psi[3:7, 7, 7]) == 1)
    phi = F.isolated_convolve(bump, pk, 32.0, n, (5, 0, 0))
    assert phi[7, 7, 7] < phi[0, 0, 0] and phi[7, 7, 7] < 0
    lap = F.isolated_convolve(bump, pk, 32.0, n, (4, 0, 0))
    assert lap[7, 7, 7] > 0


def test_pk_table_against_direct_evaluation():
    """The GPU method's table (8192 points in ln k, Float32 linear interpolation) against
    sqrt(P dk^3 n2^3) evaluated directly, over the modes of an n = 32 tile; and the whole
    convolution with either, to the reference's own GPU-vs-CPU tolerance."""
    from huntian_hip import api

    pk = lambda k: 100.0 / (1.0 + (k * 50.0) ** 4)
    n, box = 32, 100.0
    vals, lkmin, dlk = api.pk_table(pk, box, n, n)
    assert vals.dtype == np.float32 and len(vals) == 8192
    n2 = 2 * n
    dk = 2 * np.pi / (n2 * box / n)
    k = dk * np.arange(1, n2) * np.sqrt(3) / 2
    lk = (np.log(k * k) * 0.5).astype(np.float32)
    got = F.interp_pk_table(vals, lk, lkmin, dlk)
    want = np.sqrt(pk(k) * dk ** 3 * n2 ** 3)
    np.testing.assert_allclose(got, want, rtol=2e-5)
    rng = np.random.default_rng(42)
    noise = rng.normal(size=(n,) * 3).astype(np.float32)
    a = F.isolated_convolve(noise, pk, box, n, (1, 2, 0), table=(vals, lkmin, dlk))
    b = F.isolated_convolve(noise, pk, box, n, (1, 2, 0), table=None)
    assert np.linalg.norm(a - b) <= 5e-6 * np.linalg.norm(b)


def test_periodic_2lpt_and_laplacian_closed_forms():
    """delta = A cos(a x) + B cos(b y): phi_xx = -A cos(ax), phi_yy = -B cos(by), phi_xy = 0,
    src2 = AB cos(ax) cos(by), psi2_x = AB a/(a^2+b^2) sin(ax) cos(by) (and x <-> y);
    a single plane wave has no 2LPT term; the Laplacian kernel k^2 gives +k^2 delta."""
    n, L = 16, 80.0
    x = np.arange(n) * L / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    a, b = 2 * np.pi * 2 / L, 2 * np.pi * 3 / L
    A, B = 0.7, -1.3
    delta = A * np.cos(a * X) + B * np.cos(b * Y)
    px, py, pz = F.compute_2lpt(delta, n, L)
    np.testing.assert_allclose(px, A * B * a / (a * a + b * b) * np.sin(a * X) * np.cos(b * Y),
                               atol=1e-12)
    np.testing.assert_allclose(py, A * B * b / (a * a + b * b) * np.cos(a * X) * np.sin(b * Y),
                               atol=1e-12)
    np.testing.assert_allclose(pz, 0.0, atol=1e-12)
    for comp in F.compute_2lpt(A * np.cos(a * X + b * Z), n, L):
        np.testing.assert_allclose(comp, 0.0, atol=1e-12)
    lap = F.compute_laplacian(delta, n, L)
    np.testing.assert_allclose(lap, A * a * a * np.cos(a * X) + B * b * b * np.cos(b * Y),
                               atol=1e-10)


def brute_padded(f, boxsize, mult, nyquist_zero):
    """Explicit sums for the 2n-padded Poisson / short-range solves: mult(kx, ky, kz, k2) on
    every mode (ky, kz of index n2/2 positive), Nyquist planes zeroed when asked."""
    n = f.shape[0]
    n2 = 2 * n
    dk = 2 * np.pi / (2 * boxsize)
    pad = np.zeros((n2,) * 3)
    pad[:n, :n, :n] = f
    m = np.arange(n2)
    E = np.exp(-2j * np.pi * np.outer(m, m) / n2)
    X = np.einsum("ax,by,cz,xyz->abc", E, E, E, pad)
    signed = np.where(m <= n2 // 2, m, m - n2)
    Y = np.zeros_like(X)
    for a, b, c in itertools.product(range(n2 // 2 + 1), range(n2), range(n2)):
        if nyquist_zero and (a == n2 // 2 or b == n2 // 2 or c == n2 // 2):
            continue
        kx, ky, kz = a * dk, signed[b] * dk, signed[c] * dk
        Y[a, b, c] = X[a, b, c] * mult(kx, ky, kz, kx * kx + ky * ky + kz * kz)
    for a in range(n2 // 2 + 1, n2):
        Y[a] = np.conj(Y[(-a) % n2][np.ix_((-m) % n2, (-m) % n2)])
    Einv = np.conj(E)
    return (np.einsum("ax,by,cz,abc->xyz", Einv, Einv, Einv, Y).real / n2 ** 3)[:n, :n, :n]


def test_isolated_poisson_and_short_potential_match_explicit_sums():
    rng = np.random.default_rng(3)
    f = rng.normal(size=(3, 3, 3))
    psi = F.isolated_poisson_psi(f, 12.0)
    for d in range(3):
        want = brute_padded(f, 12.0, lambda kx, ky, kz, k2, d=d:
                            0.0 if k2 == 0 else 1j * (kx, ky, kz)[d] / k2, True)
        np.testing.assert_allclose(psi[d], want, atol=1e-12)
    rs = 1.7
    want = brute_padded(f, 12.0, lambda kx, ky, kz, k2: -rs * rs if k2 == 0 else
                        -(1 - np.exp(-k2 * rs * rs)) / k2, False)
    np.testing.assert_allclose(F.isolated_short_pot(f, 12.0, rs), want, atol=1e-12)


def test_tricubic_partition_of_unity_and_nodes():
    """Catmull-Rom weights sum to 1 (a constant coarse field comes back within Float32
    rounding); with an odd block every third fine cell sits on a coarse centre, where the
    weights are (0, 1, 0, 0) and the coarse value comes back exactly."""
    M, N = 6, 18                      # block 3
    nsub, nmesh = 6, 10               # nbuff 2
    c = np.full((M,) * 3, 2.5, np.float32)
    t = F.interpolate_to_tile(c, 1, 2, 3, nsub, nmesh, N, M)
    assert t.dtype == np.float32 and t.shape == (nmesh,) * 3
    np.testing.assert_allclose(t, 2.5, rtol=4e-7)
    rng = np.random.default_rng(0)
    c = rng.normal(size=(M,) * 3).astype(np.float32)
    for tile in ((1, 1, 1), (3, 2, 1)):
        t = F.interpolate_to_tile(c, *tile, nsub, nmesh, N, M)
        starts = [(q - 1) * nsub + 1 - 2 for q in tile]
        for li in range(nmesh):
            gi = [(s + li - 1) % N + 1 for s in starts]        # global cell, 1-based
            if all((g - 2) % 3 == 0 for g in gi):               # g = 3 m - 1: centre of m
                ci = [((g + 1) // 3 - 1) % M for g in gi]
                assert t[li, li, li] == c[ci[0], ci[1], ci[2]]
    # a smooth periodic field is reproduced to the scheme's accuracy
    M, N = 16, 64
    x = (np.arange(M) + 0.5) / M
    cx = np.sin(2 * np.pi * x).astype(np.float32)
    c = np.broadcast_to(cx[:, None, None], (M, M, M)).copy()
    t = F.interpolate_to_tile(c, 1, 1, 1, 48, 64, N, M)
    g = (np.arange(64) + 1 - 8 - 1) % N + 1
    exact = np.sin(2 * np.pi * ((g - 0.5) / N))
    np.testing.assert_allclose(t[:, 0, 0], exact, atol=3e-3)

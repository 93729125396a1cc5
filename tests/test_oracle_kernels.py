"""Pins the oracle's small kernels (shells, hRinteg, atab4, kernel_strain, eigenvalues,
collapse-table lookup, fsc_of_z) against the reference's known answers
(test/test_radialshell.jl, test/test_ellipsoidalcollapse.jl, test/test_shell_gpu.jl),
independent library routines (numpy LAPACK, scipy) and closed forms."""
import math

import numpy as np
import pytest

from huntian_hip.homeltab import make_ctab


# ---------------------------------------------------------------- shells

def test_precompute_shells_counts(oracle):
    # test/test_radialshell.jl:49-63
    t1 = oracle.shell_tables(1)
    assert t1["ncells"] == 7 and t1["shell_r2"][0] == 0
    t2 = oracle.shell_tables(2)
    assert t2["ncells"] == 33
    assert list(t2["shell_count"]) == [1, 6, 12, 8, 6]
    # r^2 = 1 shell holds exactly the 6 face neighbours (test/test_shell_gpu.jl:80-82)
    assert t2["shell_count"][1] == 6


@pytest.mark.parametrize("rmax", [3, 8, 15])
def test_shells_brute_force(oracle, rmax):
    t = oracle.shell_tables(rmax)
    # generation order dk, dj, di then a stable sort by r^2: numpy's stable argsort
    d = np.arange(-rmax, rmax + 1)
    dk, dj, di = np.meshgrid(d, d, d, indexing="ij")
    r2 = (di ** 2 + dj ** 2 + dk ** 2).ravel()
    keep = r2 <= rmax * rmax
    di, dj, dk, r2 = di.ravel()[keep], dj.ravel()[keep], dk.ravel()[keep], r2[keep]
    o = np.argsort(r2, kind="stable")
    assert np.array_equal(t["di"], di[o]) and np.array_equal(t["dj"], dj[o])
    assert np.array_equal(t["dk"], dk[o])
    u, c = np.unique(r2, return_counts=True)
    assert np.array_equal(t["shell_r2"], u) and np.array_equal(t["shell_count"], c)
    assert np.array_equal(t["shell_start"], np.concatenate([[0], np.cumsum(c)[:-1]]))


def test_shell_counts_at_baseline_radii(oracle):
    # distinct-r^2 shells and cells at the hunt radii of the benchmark configurations
    for rmax, S, Cc in [(15, 190, 14147), (20, 336, 33401), (30, 753, 113081)]:
        t = oracle.shell_tables(rmax)
        assert (t["nshells"], t["ncells"]) == (S, Cc)


# ---------------------------------------------------------------- hRinteg / atab4

def test_hRinteg_known_answers(oracle):
    # test/test_radialshell.jl:3-30
    assert oracle.hRinteg(0.0, 0.0) == pytest.approx(-31.0 / 35.0, abs=1e-12)
    assert oracle.hRinteg(0.0, 1.5) == pytest.approx(1.4 * 1.5 - 31.0 / 35.0, abs=1e-12)
    assert oracle.hRinteg(2.0, 0.0) == 0.0
    assert oracle.hRinteg(3.0, 1.0) == 0.0
    assert oracle.hRinteg(1.0, 0.0) == pytest.approx(-3.2 / 7.0, abs=1e-12)
    assert oracle.hRinteg(0.5, 2.0) == oracle.hRinteg(-0.5, 2.0)
    assert oracle.hRinteg(0.5, 1.0) != oracle.hRinteg(0.5, 2.0)


def test_hRinteg_is_the_kernel_integral(oracle):
    # hRinteg(u0, u12) = int_{|u0|}^{2} 4 pi W(x) x (u12 - x^2) dx with the M4 kernel
    # W(x) = (1/pi)(1 - 1.5x^2 + 0.75x^3) on [0,1], (1/4pi)(2-x)^3 on [1,2]
    from scipy.integrate import quad

    def W(x):
        return (1 - 1.5 * x * x + 0.75 * x ** 3) / math.pi if x < 1 else (2 - x) ** 3 / (4 * math.pi)

    for u0 in (0.0, 0.3, 0.9, 1.0, 1.4, 1.95):
        for u12 in (0.0, 0.7, 3.1):
            ref = sum(quad(lambda x: 4 * math.pi * W(x) * x * (u12 - x * x), a, b,
                           epsabs=1e-13, epsrel=1e-13)[0]
                      for a, b in [(u0, max(u0, 1.0)), (max(u0, 1.0), 2.0)] if b > a)
            assert oracle.hRinteg(u0, u12) == pytest.approx(ref, abs=1e-11)


def test_atab4(oracle):
    # test/test_radialshell.jl:32-47
    akk = oracle.atab4()
    ca = 3.0 / (2.0 * math.pi)
    assert len(akk) == 1101
    assert akk[0] == pytest.approx(ca * 2.0 / 3.0, abs=1e-14)
    assert akk[400] == 0.0 and akk[499] == 0.0 and akk[1099] == 0.0 and akk[1100] == 0.0
    # continuity at r = 1 and the closed form W(sqrt(u)) = (1/pi)(1 - 1.5u + 0.75 u^1.5)
    assert ca * (2 / 3 - 1 + 0.5) == pytest.approx((2 - 1) ** 3 / (4 * math.pi), abs=1e-14)
    u = np.arange(1100) * 0.01
    W = np.where(u < 1, (1 - 1.5 * u + 0.75 * u ** 1.5) / math.pi,
                 np.where(u <= 4, np.clip(2 - np.sqrt(u), 0, None) ** 3 / (4 * math.pi), 0))
    np.testing.assert_allclose(akk[:1100], W, rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- kernel_strain

def _profiles(ns, seed):
    rng = np.random.default_rng(seed)
    rad = np.sqrt(np.arange(ns, dtype=np.float64))  # 0, 1, sqrt2, ...
    return rad, rng.standard_normal((ns, 3)), rng.standard_normal((ns, 3)), \
        rng.standard_normal((ns, 3, 3))


def test_kernel_strain_single_shell_closed_form(oracle):
    # window of one shell at m1 = mp: u0 = 0, u12 = 2 r^2, hRinteg(0,u12) = 1.4 u12 - 31/35
    rad, G, Gf, SR = _profiles(20, 1)
    mp = 9
    wR, aR = 0.37, 2.5
    E, g, gf = oracle.kernel_strain(rad, G, Gf, SR, mp, mp, mp, wR, aR)
    r = rad[mp - 1]
    u12 = 2 * r * r
    wt = (1.4 * u12 - 31 / 35) / u12 ** 2 * wR / r
    np.testing.assert_allclose(g, wt * G[mp - 1], rtol=1e-13)
    np.testing.assert_allclose(gf, wt * Gf[mp - 1], rtol=1e-13)
    np.testing.assert_allclose(E, -0.5 * wt * (SR[mp - 1] + SR[mp - 1].T), rtol=1e-13)


def test_kernel_strain_centre_branch_closed_form(oracle):
    # mp == 1 with one shell at r^2 = 1: con = 100 exactly, aww = atab4 at u = 1
    rad, G, Gf, SR = _profiles(8, 2)
    E, g, gf = oracle.kernel_strain(rad, G, Gf, SR, 1, 2, 2, 0.3, 1.7)
    aww = (1 - 1.5 + 0.75) / math.pi
    wt = 0.5 * aww / 1.0 * 1.7
    np.testing.assert_allclose(g, wt * G[1], rtol=1e-12)
    np.testing.assert_allclose(E, -0.5 * wt * (SR[1] + SR[1].T), rtol=1e-12)


def test_kernel_strain_linear_and_symmetric(oracle):
    rad, G, Gf, SR = _profiles(30, 3)
    _, G2, Gf2, SR2 = _profiles(30, 4)
    for (mp, lo, hi) in [(5, 3, 7), (10, 7, 13), (15, 10, 18), (3, 2, 5), (1, 2, 8)]:
        a = oracle.kernel_strain(rad, G, Gf, SR, mp, lo, hi, 1.0, 2.5)
        b = oracle.kernel_strain(rad, G2, Gf2, SR2, mp, lo, hi, 1.0, 2.5)
        c = oracle.kernel_strain(rad, G + G2, Gf + Gf2, SR + SR2, mp, lo, hi, 1.0, 2.5)
        for x, y, z in zip(a, b, c):
            np.testing.assert_allclose(x + y, z, rtol=1e-10, atol=1e-12)
        assert np.array_equal(a[0], a[0].T)
        # a shell with |r - r_mp| >= 2 contributes nothing
        far = np.abs(rad - rad[mp - 1]) >= 2
        if mp > 1 and far[lo - 1:hi].any():
            SRz = SR.copy()
            SRz[~far] = 0
            Gz = G.copy()
            Gz[~far] = 0
            e, gg, _ = oracle.kernel_strain(rad, Gz, Gf, SRz, mp, lo, hi, 1.0, 2.5)
            assert np.all(e == 0) and np.all(gg == 0)


# ---------------------------------------------------------------- eigenvalues

def _reference_matrices():
    # test/test_shell_gpu.jl:357-420
    mats = [np.eye(3), np.diag([2.0, 3.0, 5.0]),
            np.array([[1, 0.5, 0.25], [0.5, 2, 0.75], [0.25, 0.75, 3]]),
            np.array([[5.0, 4, 1], [4, 5, 1], [1, 1, 2]]), np.diag([1.0, 1.0, 2.0]),
            0.01 * np.eye(3), np.array([[1.0, 1e-4, 0], [1e-4, 1.0, 0], [0, 0, 3.0]])]
    rng = np.random.default_rng(91)
    for _ in range(550):
        A = rng.standard_normal((3, 3))
        mats.append((A + A.T) / 2)
    for _ in range(50):  # scaled and near-degenerate cases
        A = rng.standard_normal((3, 3))
        Q, _ = np.linalg.qr(A)
        lam = np.sort(rng.standard_normal(3)) * 10.0 ** rng.integers(-6, 6)
        lam[1] = lam[0] * (1 + 1e-9 * rng.random())
        mats.append(Q @ np.diag(lam) @ Q.T)
    return mats


def test_eig3_vs_lapack(oracle):
    for A in _reference_matrices():
        got = oracle.eig3(A)
        ref = np.linalg.eigvalsh(A)
        scale = np.abs(ref).max() + 1e-300
        np.testing.assert_allclose(got, ref, rtol=0, atol=4e-15 * scale)
        assert np.all(np.diff(got) >= 0)


def test_eig3_known_answers(oracle):
    # test/test_radialshell.jl:65-74
    np.testing.assert_allclose(oracle.eig3(np.eye(3)), [1, 1, 1], atol=1e-12)
    np.testing.assert_allclose(oracle.eig3(np.diag([1.0, 2, 3])), [1, 2, 3], atol=1e-12)
    # upper triangle only (Symmetric(mat) semantics)
    A = np.array([[1.0, 2, 3], [99, 4, 5], [99, 99, 6]])
    np.testing.assert_allclose(oracle.eig3(A), np.linalg.eigvalsh(np.triu(A) + np.triu(A, 1).T),
                               atol=1e-13)


# ---------------------------------------------------------------- collapse table

def test_interp3_linear_table_exact(oracle):
    # test/test_ellipsoidalcollapse.jl:115-137
    i, j, k = np.meshgrid(np.arange(11), np.arange(11), np.arange(11), indexing="ij")
    t = (0.1 * i + 0.1 * j + 0.1 * k).astype(np.float32)
    ct = make_ctab(t, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    assert oracle.interp3(ct, 0.0, 0.0, 0.0) == pytest.approx(0.0, abs=1e-5)
    assert oracle.interp3(ct, 0.5, 0.0, 0.0) == pytest.approx(0.5, abs=1e-5)
    assert oracle.interp3(ct, 0.0, 0.5, 0.0) == pytest.approx(0.5, abs=1e-5)
    assert oracle.interp3(ct, 0.25, 0.25, 0.25) == pytest.approx(0.75, abs=1e-4)
    assert oracle.interp3(ct, -0.1, 0.0, 0.0) == -1.0
    assert oracle.interp3(ct, 1.1, 0.0, 0.0) == -1.0
    assert oracle.interp3(ct, 0.5, -0.2, 0.5) == -1.0
    assert oracle.interp3(ct, 1.0, 1.0, 1.0) == pytest.approx(3.0, abs=1e-6)


def test_interp3_vs_scipy(oracle):
    from scipy.interpolate import RegularGridInterpolator

    xs, ys, zs = np.linspace(-2, 2, 21), np.linspace(0, 3, 17), np.linspace(1, 5, 13)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    t = (np.sin(X) * Y + Z * 0.1).astype(np.float32)
    ct = make_ctab(t, -2.0, 2.0, 0.0, 3.0, 1.0, 5.0)
    rgi = RegularGridInterpolator((xs, ys, zs), t.astype(np.float64))
    rng = np.random.default_rng(999)
    pts = np.stack([rng.uniform(-2, 2, 300), rng.uniform(0, 3, 300), rng.uniform(1, 5, 300)], 1)
    got = np.array([oracle.interp3(ct, *p) for p in pts])
    np.testing.assert_allclose(got, rgi(pts), rtol=1e-12, atol=1e-12)


def test_fsc_of_z(oracle, websky):
    # fsc_of_z bisects the table at e = p/e = 0; the crossing value for HomelTab_websky.dat
    fc = oracle.fsc_of_z(websky, 0.0)
    assert fc == pytest.approx(1.716402, abs=2e-6)
    x = math.log10(fc)
    step = (websky["x2"] - websky["x1"]) * 2.0 ** -49
    assert oracle.interp3(websky, x, 0.0, 0.0) <= 1.0 < oracle.interp3(websky, x + step, 0, 0)
    # a z = 1 collapse needs a larger linear overdensity
    assert oracle.fsc_of_z(websky, 1.0) > fc

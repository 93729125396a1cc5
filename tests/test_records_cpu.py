"""Halo records: the host cosmology tables (huntian_hip.cosmology), the oracle's restatement
of the record loop (Dlinear_ab, chi_to_z, MultiResolution.jl:1159-1219) and the .pksc
layout, pinned by closed forms (Einstein-de Sitter growth and distance, the exact
Lambda-CDM growth integral, Lagrange interpolation at table nodes), the reference's own
test values (test/test_cosmology.jl) and hand-computed records."""
import math
import os

import numpy as np
import pytest

from huntian_hip import catalog, cosmology


def test_growth_table_is_exact_for_einstein_de_sitter():
    """Omega_m = 1 (no baryons): D(a) = a, so D/a = 1 and dlnD/dlna = 1 at every node."""
    at, Dt, ft = cosmology.dlinear_tables(cosmology.CosmologyParams(1.0, 0.0, 0.0, 0.7))
    assert at[0] >= 1 / 1001 and at[-1] <= 2.0 and np.all(np.diff(at) > 0)
    assert len(at) == 904  # a = 1e-7 .. 2 in 2000 log steps, kept for a >= 1/1001
    np.testing.assert_allclose(Dt, 1.0, rtol=1e-12)
    np.testing.assert_allclose(ft, 1.0, rtol=1e-12)


def _growth_integral(a, Om, OL):
    """Linear growth of flat matter + Lambda, D(a) ~ H(a) int_0^a da' / (a' H(a'))^3."""
    from scipy.integrate import quad

    E = lambda x: math.sqrt(Om / x ** 3 + OL)
    val, _ = quad(lambda x: 1.0 / (x * E(x)) ** 3, 0.0, a, epsabs=0.0, epsrel=1e-13, limit=200)
    return E(a) * val


def test_growth_table_matches_the_lcdm_growth_integral(oracle):
    Om, OL = 0.3138, 0.6862
    growth = cosmology.dlinear_tables(cosmology.CosmologyParams(Om, 0.0, OL, 0.6735))
    norm = _growth_integral(1.0, Om, OL)
    for a in (0.01, 0.1, 0.25, 0.5, 0.8, 1.0, 1.3, 1.9):
        exact = _growth_integral(a, Om, OL) / norm
        assert abs(oracle.growth_D(growth, a) - exact) < 2e-6 * exact, a


def test_growth_D_reference_bounds_and_table_nodes(oracle):
    """test/test_cosmology.jl:26-27 bounds D(z=0.5), D(z=1) for Planck 2018 parameters;
    at a table node the 4-point Lagrange weights are exactly 0 and 1."""
    c = cosmology.CosmologyParams(0.2645 + 0.0493, 0.0493, 1.0 - 0.2645 - 0.0493, 0.6735)
    growth = cosmology.dlinear_tables(c)
    assert 0.75 < oracle.growth_D(growth, 1 / 1.5) < 0.82
    assert 0.58 < oracle.growth_D(growth, 1 / 2.0) < 0.65
    assert abs(oracle.growth_D(growth, 1.0) - 1.0) < 1e-6
    at, Dt = growth[0], growth[1]
    for k in (5, 100, 777, len(at) - 3):
        assert oracle.growth_D(growth, at[k]) == Dt[k] * at[k], k
    # outside the table D/a is held at its end values
    assert oracle.growth_D(growth, 3.0) == Dt[-1] * 3.0


def test_comoving_distance_and_its_inverse(oracle):
    """Einstein-de Sitter chi(z) = 2 c/H0 (1 - 1/sqrt(1+z)); the reference's own checks of
    chi(1) and of the chi -> z table (test/test_cosmology.jl:33-37, 66-73)."""
    eds = cosmology.CosmologyParams(1.0, 0.0, 0.0, 0.7)
    for z in (0.1, 1.0, 3.0):
        exact = 2 * 2998.0 * (1 - 1 / math.sqrt(1 + z))
        assert abs(cosmology.chi(z, eds) - exact) < 1e-9 * exact
    c = cosmology.CosmologyParams(0.2645 + 0.0493, 0.0493, 1.0 - 0.2645 - 0.0493, 0.6735)
    assert 2250 < cosmology.chi(1.0, c) < 2350
    chi2z = cosmology.build_chi_to_z(c, z_max=6.0, npts=2000)
    for z in (0.0, 0.1, 0.5, 1.0, 2.0, 4.0, 5.9):
        assert abs(oracle.chi_to_z(chi2z, cosmology.chi(z, c)) - z) < 0.005
    zs, chis, _ = chi2z
    for q in (1, 500, 1999):
        assert oracle.chi_to_z(chi2z, chis[q]) == zs[q]
    assert oracle.chi_to_z(chi2z, -1.0) == 0.0 and oracle.chi_to_z(chi2z, 1e9) == 6.0


def _results(npk, rng):
    r = {k: rng.normal(size=npk) for k in ("RTHL", "Fbarx", "e_v", "p_v", "Srb", "d2F",
                                            "zvir_half")}
    r.update({k: rng.normal(size=(npk, 3)) for k in ("eigs", "Sbar", "Sbar2", "gradpk",
                                                     "gradpkf", "gradpkrf")})
    m = rng.normal(size=(npk, 3, 3))
    r["strain"] = m + np.swapaxes(m, 1, 2)  # symmetric, six distinct entries
    r["nshells_used"] = np.zeros(npk, np.int32)
    r["crossing_shell"] = np.zeros(npk, np.int32)
    return r


def test_records_by_hand(oracle):
    """D/a = 1 (so D(a) = a), ZZon = 2: D = 1/2; Omega_m = 1, Omega_L = 0: the 2LPT factor
    is -3/7 D^2.  Peaks with RTHL <= 0 give no record; fields in Catalog.jl order."""
    rng = np.random.default_rng(3)
    npk = 5
    r = _results(npk, rng)
    r["RTHL"] = np.array([2.5, -1.0, 1.25, 0.0, 3.0])
    pi, pj, pk = np.array([3, 4, 5, 6, 7]), np.array([1, 2, 3, 4, 10]), np.array([9, 8, 7, 6, 5])
    Rf = np.array([4.0, 3.0, 2.0, 1.5, 8.0])
    FcRf = rng.normal(size=npk).astype(np.float32)
    d2 = rng.normal(size=npk).astype(np.float32)
    growth = (np.linspace(0.001, 2.0, 200), np.ones(200))
    kw = dict(alatt=2.0, n=10, xbx=100.0, ybx=-50.0, zbx=7.0, ZZon=2.0, ilpt=2, Om=1.0, OL=0.0)
    rec, nh = oracle.pack_records(pi, pj, pk, Rf, r, FcRf, d2Rf=d2, growth=growth, ioutshear=1,
                                  **kw)
    keep = [0, 2, 4]
    assert nh == 3 and rec.shape == (3, 33)
    f32 = np.float32
    D = f32(0.5)
    coef = f32(-3.0 / 7.0 * 1.0 ** (-1.0 / 143) * 0.25)
    for row, q in zip(rec, keep):
        assert row[0] == f32(100.0 + 2.0 * (pi[q] - 5.5))
        assert row[1] == f32(-50.0 + 2.0 * (pj[q] - 5.5))
        assert row[2] == f32(7.0 + 2.0 * (pk[q] - 5.5))
        np.testing.assert_array_equal(row[3:6], r["Sbar"][q].astype(f32) * D)
        assert row[6] == f32(float(f32(r["RTHL"][q])) * 2.0)
        np.testing.assert_array_equal(row[7:10], r["Sbar2"][q].astype(f32) * coef)
        assert row[10] == f32(r["Fbarx"][q])
        assert row[11] == f32(r["e_v"][q]) and row[12] == f32(r["p_v"][q])
        s = r["strain"][q]
        np.testing.assert_array_equal(row[13:19], np.array(
            [s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]], f32))
        assert row[19] == f32(r["d2F"][q]) and row[20] == f32(r["zvir_half"][q])
        np.testing.assert_array_equal(row[21:24], r["gradpk"][q].astype(f32))
        np.testing.assert_array_equal(row[24:27], r["gradpkf"][q].astype(f32))
        assert row[27] == f32(Rf[q]) and row[28] == FcRf[q] and row[29] == d2[q]
        np.testing.assert_array_equal(row[30:33], r["gradpkrf"][q].astype(f32))
    basic, nb = oracle.pack_records(pi, pj, pk, Rf, r, FcRf, growth=growth, ioutshear=0,
                                    **dict(kw, ilpt=1))
    assert nb == 3 and basic.shape == (3, 11)
    np.testing.assert_array_equal(basic[:, :7], rec[:, :7])
    assert np.all(basic[:, 7:10] == 0) and np.array_equal(basic[:, 10], rec[:, 10])
    cut, nc = oracle.pack_records(pi, pj, pk, Rf, r, FcRf, growth=growth, capacity=2,
                                  ioutshear=0, **kw)
    assert nc == 3 and cut.shape == (2, 11) and np.array_equal(cut, rec[:2, :11])


def test_lightcone_records(oracle):
    """ievol = 1: each peak's z from its distance to the observer through the chi -> z
    table; peaks beyond z_max give no record, the others use D(1/(1+z))."""
    c = cosmology.CosmologyParams(0.31, 0.049, 0.69, 0.68)
    growth = cosmology.dlinear_tables(c)
    chi2z = cosmology.build_chi_to_z(c, z_max=1.0, npts=400)
    rng = np.random.default_rng(5)
    npk = 40
    r = _results(npk, rng)
    r["RTHL"] = np.abs(r["RTHL"]) + 0.1
    pi = rng.integers(1, 65, npk)
    pj = rng.integers(1, 65, npk)
    pk = rng.integers(1, 65, npk)
    kw = dict(alatt=8.0, n=64, xbx=300.0, ybx=0.0, zbx=0.0, ievol=1, obs=(0.0, 0.0, 0.0),
              z_max=0.15, ilpt=2, Om=c.Om, OL=c.OL)
    rec, nh = oracle.pack_records(pi, pj, pk, 2.0, r, np.zeros(npk, np.float32), growth=growth,
                                  chi2z=chi2z, ioutshear=0, **kw)
    x = 300.0 + 8.0 * (pi - 32.5)
    y = 8.0 * (pj - 32.5)
    z = 8.0 * (pk - 32.5)
    dist = np.sqrt(x * x + y * y + z * z)
    zpk = np.array([oracle.chi_to_z(chi2z, d) for d in dist])
    kept = np.nonzero(zpk <= 0.15)[0]
    assert 0 < nh == len(kept) < npk
    for row, q in zip(rec, kept):
        D = np.float32(oracle.growth_D(growth, 1.0 / (1.0 + zpk[q])))
        np.testing.assert_array_equal(row[3:6], r["Sbar"][q].astype(np.float32) * D)


def test_pksc_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(1)
    for nf in (11, 33):
        rec = rng.normal(size=(7, nf)).astype(np.float32)
        path = os.path.join(tmp_path, f"h{nf}.pksc")
        catalog.write_pksc(path, rec, 60.0, 0.5)
        raw = open(path, "rb").read()
        assert len(raw) == 12 + 7 * nf * 4
        assert np.frombuffer(raw[:4], "<i4")[0] == 7
        assert np.array_equal(np.frombuffer(raw[4:12], "<f4"), np.array([60.0, 0.5], np.float32))
        back, rmax, zout = catalog.read_pksc(path)
        assert back.shape == (7, nf) and np.array_equal(back, rec)
        assert (rmax, zout) == (60.0, 0.5)
        tab = catalog.as_table(back)
        assert list(tab)[6] == "RTHL" and np.array_equal(tab["overdensity"], rec[:, 10])
    with pytest.raises(ValueError):
        catalog.write_pksc(os.path.join(tmp_path, "bad.pksc"), np.zeros((2, 12)), 1.0, 0.0)

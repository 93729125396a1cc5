"""finalize_eulerian: the oracle's restatement of Merger.finalize_eulerian
(src/Merger/Merger.jl:107-166) checked against independent physics rather than its own
formulas, as test/test_finalize_am.jl:10-57 does: linear theory at z = 0 (v = 100 f psi, with
the growth rate f ~ Omega_m^0.55, Linder 2005), the lightcone velocity factor a H(a) f(a) at
z = 1, bookkeeping of the consumed 2LPT fields, and the extra ExtHaloRecord fields.  The
growth rate itself is pinned against the closed-form Lambda-CDM growth integral
D ~ E(a) int_0^a da' / (a' E(a'))^3."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from huntian_hip import cosmology

COSMO = cosmology.CosmologyParams(0.31, 0.049, 0.69, 0.68)
F0 = 0.31 ** 0.55


@pytest.fixture(scope="module")
def tabs():
    return dict(growth=cosmology.dlinear_tables(COSMO),
                chi2z=cosmology.build_chi_to_z(COSMO, z_max=10.0))


def rec(q, d1, R, d2, od, extra=()):
    return np.array([[*q, *d1, R, *d2, od, *extra]], np.float32)


def fin(oracle, tabs, r, obs=(0.0, 0.0, 0.0), **kw):
    return oracle.finalize_eulerian(r, tabs["growth"], obs, Om=COSMO.Om, OL=COSMO.OL,
                                    chi2z=tabs["chi2z"], **kw)


Q, D1, D2 = (10.0, -20.0, 30.0), (1.0, -2.0, 0.5), (0.1, 0.05, -0.2)


def test_linear_theory_at_z0(oracle, tabs):
    out = fin(oracle, tabs, rec(Q, D1, 3.0, D2, 0.7), ievol=0, z_out=0.0)[0]
    for c in range(3):
        assert out[c] == pytest.approx(Q[c] + D1[c] + D2[c], rel=1e-6)
    for c, tol in zip(range(3), (0.01, 0.01, 0.02)):
        assert out[3 + c] == pytest.approx(100 * F0 * (D1[c] + 2 * D2[c]), rel=tol)


def test_bookkeeping(oracle, tabs):
    h = rec(Q, D1, 3.0, D2, 0.7)
    out = fin(oracle, tabs, h, ievol=0)[0]
    assert tuple(out[7:10]) == (0.0, 0.0, 0.0)            # 2LPT consumed
    assert out[6] == h[0, 6] and out[10] == h[0, 10]       # RTHL, overdensity kept
    z = fin(oracle, tabs, rec((1, 2, 3), (0, 0, 0), 1.0, (0, 0, 0), 0.0), ievol=0)[0]
    assert tuple(z[:6]) == (1.0, 2.0, 3.0, 0.0, 0.0, 0.0)
    assert fin(oracle, tabs, np.zeros((0, 11), np.float32)).shape == (0, 11)


def test_lightcone_redshift_from_eulerian_distance(oracle, tabs):
    obs = (-500.0, 100.0, 50.0)
    r1 = cosmology.chi(1.0, COSMO)
    hz = rec((obs[0] + r1, obs[1], obs[2]), (1, 0, 0), 3.0, (0, 0, 0), 0.0)
    o = fin(oracle, tabs, hz, obs, ievol=1)[0]
    a = 0.5
    E = math.sqrt(0.31 / a ** 3 + 0.69)
    om_a = 0.31 / a ** 3 / E ** 2
    assert o[3] == pytest.approx(a * 100 * E * om_a ** 0.55, rel=0.01)
    o0 = fin(oracle, tabs, hz, obs, ievol=0, z_out=0.0)[0]
    assert o[3] / o0[3] == pytest.approx(a * E * om_a ** 0.55 / F0, rel=0.01)


def test_lightcone_uses_the_eulerian_not_the_lagrangian_distance(oracle, tabs):
    """A halo displaced from chi(z = 0.5) to chi(z = 1) takes z = 1's velocity factor."""
    r05, r1 = cosmology.chi(0.5, COSMO), cosmology.chi(1.0, COSMO)
    d = r1 - r05
    o = fin(oracle, tabs, rec((r05, 0, 0), (d, 0, 0), 3.0, (0, 0, 0), 0.0), ievol=1)[0]
    a = 0.5
    E = math.sqrt(0.31 / a ** 3 + 0.69)
    f = oracle.growth_f(tabs["growth"], a)
    assert o[0] == pytest.approx(r1, rel=1e-6)
    assert o[3] == pytest.approx(a * 100 * E * f * d, rel=1e-5)
    # the observer itself: a = 1
    at0 = fin(oracle, tabs, rec((0, 0, 0), (0, 0, 0), 3.0, (0, 0, 0), 0.0), ievol=1)[0]
    assert tuple(at0[:6]) == (0.0,) * 6


def test_ext_records_keep_their_extra_fields(oracle, tabs):
    extra = np.arange(1, 23, dtype=np.float32)
    e = rec(Q, D1, 3.0, D2, 0.7, extra)
    eo = fin(oracle, tabs, e, ievol=0)[0]
    out = fin(oracle, tabs, rec(Q, D1, 3.0, D2, 0.7), ievol=0)[0]
    assert np.array_equal(eo[:11], out)
    assert np.array_equal(eo[11:], extra)


def test_growth_rate_against_the_lcdm_growth_integral(oracle):
    """Without baryons the tabulated growth is the exact Lambda-CDM solution, whose rate is
    dlnD/dlna = -3/2 Om a^-3 / E^2 + 1 / (a^2 E^3 I(a)), I(a) = int_0^a da' / (a' E)^3.
    The interpolated rate matches it between the table's nodes (a wrong node, weight or
    column would not)."""
    c = cosmology.CosmologyParams(0.31, 0.0, 0.69, 0.68)
    g = cosmology.dlinear_tables(c)

    def E(a):
        return math.sqrt(c.Om / a ** 3 + c.OL)

    for a in (0.0123, 0.1, 0.33333, 0.5, 0.77777, 1.0, 1.5):
        I, _ = quad(lambda x: 1.0 / (x * E(x)) ** 3, 0.0, a, epsabs=0.0, epsrel=1e-12,
                    limit=200)
        f = -1.5 * c.Om / a ** 3 / E(a) ** 2 + 1.0 / (a * a * E(a) ** 3 * I)
        assert oracle.growth_f(g, a) == pytest.approx(f, rel=2e-6), a
    # at a node the interpolation returns the tabulated value
    i = 600
    assert oracle.growth_f(g, g[0][i]) == pytest.approx(g[2][i], rel=1e-14)
    # beyond the table: its end values
    assert oracle.growth_f(g, 5.0) == g[2][-1]
    assert oracle.growth_f(g, 1e-5) == g[2][0]


def test_einstein_de_sitter_rate_is_one(oracle):
    c = cosmology.CosmologyParams(1.0, 0.0, 0.0, 0.68)
    g = cosmology.dlinear_tables(c)
    for a in (0.01, 0.25, 1.0):
        assert oracle.growth_f(g, a) == pytest.approx(1.0, abs=1e-9)
    o = oracle.finalize_eulerian(rec(Q, D1, 3.0, D2, 0.7), g, Om=1.0, OL=0.0, ievol=0,
                                 z_out=3.0)[0]
    # a 100 E f = 100 a a^-3/2 = 100 / sqrt(a), a = 1/4
    assert o[3] == pytest.approx(200.0 * (D1[0] + 2 * D2[0]), rel=1e-6)


def test_julia_negative_cube_is_compensated(oracle):
    """E(a) = sqrt(Om a^-3 + OL) (Merger.jl:139) with a^-3 evaluated as Julia's literal power
    does: a compensated square-and-multiply whose result is the correctly rounded a^-3 except
    in rare near-ties, where (1/a)^3 is off by an ulp or two in a large share of cases.
    Pinned by exact rational arithmetic (Python's float(Fraction) rounds correctly)."""
    from fractions import Fraction

    L = oracle.lib()
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(0.05, 1.0, 3000), rng.uniform(1.0, 40.0, 500)])
    exact_hits = naive_hits = 0
    for x in xs:
        exact = float(Fraction(float(x)) ** -3)
        got = L.hro_pow_m3(float(x))
        r = 1.0 / float(x)
        naive = r * r * r
        assert abs(got - exact) <= np.spacing(exact), (x, got, exact)
        exact_hits += got == exact
        naive_hits += naive == exact
    assert exact_hits >= 0.995 * len(xs), exact_hits
    assert naive_hits <= 0.9 * len(xs), naive_hits  # the test tells the two apart
    for x, want in ((1.0, 1.0), (0.5, 8.0), (2.0, 0.125), (0.25, 64.0)):
        assert L.hro_pow_m3(x) == want
    assert L.hro_pow_m3(float("inf")) == 0.0
    assert L.hro_pow_m3(0.0) == float("inf")

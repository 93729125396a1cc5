"""Pins the oracle's radial-shell collapse: the Fortran golden vector of
test/test_radialshell_fortran.jl, the synthetic collapse / no-collapse cases of
test/test_radialshell.jl:85-139, a numpy brute-force shell gather, the documented
differences between the reference's variants, early-exit invariance and symmetry."""
import numpy as np
import pytest

from fieldgen import FORTRAN_GOLDEN, bumps_field, fortran_golden_peak, nontrivial_ctab
from huntian_hip.homeltab import make_ctab

VARIANTS = ("A", "B", "C")


@pytest.mark.parametrize("variant", VARIANTS)
def test_fortran_golden_vector(oracle, fortran_table, variant):
    g = FORTRAN_GOLDEN
    t = oracle.shell_tables(8)
    r = oracle.analyse_peaks(fortran_golden_peak(), [71], [71], [71], 4.0, 4, t, fortran_table,
                             alatt=1.0, ZZon=1.0, fcrit=g["fcrit"], variant=variant)
    rel = lambda a, b: abs(a - b) / abs(b)  # noqa: E731
    assert r["RTHL"][0] > 0 and rel(r["RTHL"][0], g["RTHL"]) < 0.01
    assert rel(r["Srb"][0], g["Srb"]) < 0.01
    assert rel(r["Fbarx"][0], g["Fbarx"]) < 0.01
    np.testing.assert_allclose(r["Sbar"][0], g["Sbar"], atol=0.01)
    np.testing.assert_allclose(r["Sbar2"][0], g["Sbar2"], atol=0.01)
    for a in range(3):
        for b in range(3):
            v = g["strain"][a, b]
            if abs(v) > 1e-10:
                assert rel(r["strain"][0, a, b], v) < 0.01
            else:
                assert abs(r["strain"][0, a, b]) < 0.01
    np.testing.assert_allclose(r["gradpk"][0], g["gradpk"], atol=0.01)
    np.testing.assert_allclose(r["gradpkrf"][0], g["gradpkrf"], atol=0.01)
    assert rel(r["e_v"][0], g["e_v"]) < 0.01
    assert abs(r["p_v"][0] - g["p_v"]) < 0.01
    if variant != "B":
        assert abs(r["zvir_half"][0] - g["zvir_half"]) < 0.01
    # the FP64 restatement is much closer to the Fortran numbers than the 1 % of the test
    assert rel(r["RTHL"][0], g["RTHL"]) < 1e-7 and rel(r["Srb"][0], g["Srb"]) < 1e-6
    assert rel(r["Fbarx"][0], g["Fbarx"]) < 1e-6 and rel(r["e_v"][0], g["e_v"]) < 1e-6


def _small_grid_peak(fill, table_value):
    n = 20
    f = {k: np.zeros((n, n, n), np.float32, order="F") for k in ("delta", "etax", "etay", "etaz")}
    if fill is None:
        f["delta"][9, 9, 9] = 10
        f["delta"][8:11, 8:11, 8:11] = 5
        f["delta"][9, 9, 9] = 5
    else:
        f["delta"][:] = fill
    ct = make_ctab(np.full((5, 3, 3), table_value, np.float32), np.log10(1.5), np.log10(8.0),
                   0.0, 0.5, -1 + 1e-4, 1 - 1e-4)
    return f, ct


def test_synthetic_collapse_and_no_collapse(oracle):
    # test/test_radialshell.jl:85-139 (fcrit from fsc_of_z of the flat table)
    t = oracle.shell_tables(8)
    f, ct = _small_grid_peak(None, 1.0)
    fc = oracle.fsc_of_z(ct, 0.0 - 1.0)
    r = oracle.analyse_peaks(f, [10], [10], [10], 4.0, 4, t, ct, alatt=1.0, ZZon=0.0,
                             fcrit=fc, variant="A")
    assert r["RTHL"][0] > 0 and r["Srb"][0] > 0
    f, ct = _small_grid_peak(0.5, 100.0)
    fc = oracle.fsc_of_z(ct, 0.0 - 1.0)
    assert fc == pytest.approx(1.5, rel=1e-6)
    r = oracle.analyse_peaks(f, [10], [10], [10], 4.0, 4, t, ct, alatt=1.0, ZZon=0.0,
                             fcrit=fc, variant="A")
    assert r["RTHL"][0] == -1.0 and r["zvir_half"][0] == -1.0


def gather_numpy(f, pi, pj, pk, t):
    """Brute-force per-shell sums in the reference arithmetic (Float32 products summed in
    Float64), vectorised over cells with numpy: test/test_shell_gpu.jl:241-300."""
    n = f["delta"].shape[0]
    ns = t["nshells"]
    out = {k: np.zeros((len(pi), ns) + s) for k, s in
           (("F", ()), ("S", (3,)), ("S2", (3,)), ("G", (3,)), ("Gf", (3,)), ("SR", (3, 3)),
            ("lapd", ()))}
    out["n"] = np.zeros((len(pi), ns), np.int64)
    d = np.stack([t["di"], t["dj"], t["dk"]], 1)
    sid = np.repeat(np.arange(ns), t["shell_count"])
    for p, c in enumerate(zip(pi, pj, pk)):
        iv = np.asarray(c)[None, :] + d  # 1-based
        ok = np.all((iv >= 1) & (iv <= n), 1)
        a = iv[ok] - 1
        s = sid[ok]
        dd = f["delta"][a[:, 0], a[:, 1], a[:, 2]]
        dT = f["delta"][a[:, 2], a[:, 0], a[:, 1]]
        e = np.stack([f[k][a[:, 0], a[:, 1], a[:, 2]] for k in ("etax", "etay", "etaz")], 1)
        ix = d[ok].astype(np.float32)
        # sequential FP64 accumulation per shell in cell order
        for q in range(ns):
            w = s == q
            out["n"][p, q] = w.sum()
            out["F"][p, q] = np.cumsum(dd[w].astype(np.float64))[-1] if w.any() else 0.0
            for L in range(3):
                out["S"][p, q, L] = np.sum(e[w, L].astype(np.float64)) if w.any() else 0.0
                out["G"][p, q, L] = np.sum((dd[w] * ix[w, L]).astype(np.float64))
                out["Gf"][p, q, L] = np.sum((dT[w] * ix[w, L]).astype(np.float64))
                for K in range(3):
                    out["SR"][p, q, L, K] = np.sum((e[w, L] * ix[w, K]).astype(np.float64))
    return out


def test_gather_vs_numpy(oracle):
    f = bumps_field(24, 5, nbumps=4, margin=4)
    t = oracle.shell_tables(6)
    pi, pj, pk = [12, 2, 23, 5], [12, 3, 23, 20], [12, 22, 1, 6]  # interior and boundary
    got = oracle.shell_gather_full(f, pi, pj, pk, t)
    ref = gather_numpy(f, pi, pj, pk, t)
    assert np.array_equal(got["n"], ref["n"])
    for k in ("F", "S", "G", "Gf", "SR"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-12, atol=1e-10, err_msg=k)


def _peaks(f, nbuff, fcrit):
    from oracle import oracle as O

    m = np.zeros(f["delta"].shape, np.int8, order="F")
    return O.find_peaks(f["delta"], m, nbuff, fcrit)


@pytest.fixture(scope="module")
def bumps():
    return bumps_field(40, 31415, nbumps=12, margin=8)


def _run(oracle, f, peaks, t, ct, **kw):
    i, j, k, _ = peaks
    return oracle.analyse_peaks(f, i, j, k, 4.0, 4, t, ct, alatt=1.0, ZZon=0.0, fcrit=1.0, **kw)


def test_variants_agree_where_the_reference_says(oracle, bumps):
    ct = nontrivial_ctab(make_ctab)
    t = oracle.shell_tables(8)
    pk = _peaks(bumps, 1, 0.8)
    assert len(pk[0]) > 5
    A = _run(oracle, bumps, pk, t, ct, variant="A")
    B = _run(oracle, bumps, pk, t, ct, variant="B")
    C = _run(oracle, bumps, pk, t, ct, variant="C")
    assert (A["RTHL"] > 0).sum() >= 3
    same = ("RTHL", "Fbarx", "e_v", "p_v", "Srb", "Sbar", "Sbar2", "gradpk", "gradpkf",
            "strain", "eigs")
    for k in same:  # A and B differ only in mrf, d2F and zvir_half
        assert np.array_equal(A[k], B[k]) and np.array_equal(B[k], C[k]), k
    assert np.array_equal(A["zvir_half"], C["zvir_half"])  # C = B + the formation redshift
    assert np.all(B["zvir_half"][B["RTHL"] > 0] == 0.0)
    for k in ("gradpkrf", "d2F"):
        assert np.array_equal(B[k], C[k])


def test_early_exit_is_exact(oracle, bumps):
    ct = nontrivial_ctab(make_ctab)
    t = oracle.shell_tables(10)
    pk = _peaks(bumps, 1, 0.8)
    for v in VARIANTS:
        full = _run(oracle, bumps, pk, t, ct, variant=v, early_exit=False)
        ee = _run(oracle, bumps, pk, t, ct, variant=v, early_exit=True)
        for k in full:
            if k not in ("nshells_used", "crossing_shell"):
                assert np.array_equal(full[k], ee[k]), (v, k)
        assert np.array_equal(full["nshells_used"], ee["nshells_used"])
        assert np.array_equal(full["crossing_shell"], ee["crossing_shell"])
        assert (full["nshells_used"] < t["nshells"] - 1).any()


def test_mask_marks_core_cells_within_RTHL(oracle, bumps):
    ct = nontrivial_ctab(make_ctab)
    t = oracle.shell_tables(8)
    pk = _peaks(bumps, 1, 0.8)
    i, j, k, _ = pk
    nbuff = 3
    mask = np.zeros(bumps["delta"].shape, np.int8, order="F")
    r = oracle.analyse_peaks(bumps, i, j, k, 4.0, 4, t, ct, alatt=1.0, ZZon=0.0, fcrit=1.0,
                             variant="C", mask=mask, nbuff=nbuff)
    n = bumps["delta"].shape[0]
    X, Y, Z = np.meshgrid(*[np.arange(1, n + 1)] * 3, indexing="ij")
    core = (X > nbuff) & (X <= n - nbuff) & (Y > nbuff) & (Y <= n - nbuff) & (Z > nbuff) & (Z <= n - nbuff)
    exp = np.zeros_like(core)
    for q in range(len(i)):
        if r["RTHL"][q] > 0:
            d2 = (X - i[q]) ** 2 + (Y - j[q]) ** 2 + (Z - k[q]) ** 2
            exp |= np.sqrt(d2) <= r["RTHL"][q]
    assert np.array_equal(mask.astype(bool), exp & core)


def test_spherical_symmetry(oracle):
    # delta depending on r only and radial displacements: isotropic strain, zero gradients
    n = 40
    X, Y, Z = np.meshgrid(*[np.arange(n) - 20] * 3, indexing="ij")
    r2 = (X ** 2 + Y ** 2 + Z ** 2).astype(np.float64)
    g = np.exp(-r2 / 30.0)
    f = {"delta": np.asfortranarray((4.0 * g).astype(np.float32))}
    for name, c in (("etax", X), ("etay", Y), ("etaz", Z)):
        f[name] = np.asfortranarray((0.05 * c * g).astype(np.float32))
    ct = make_ctab(np.full((5, 3, 3), 1.5, np.float32), -2, 2, 0, 0.5, -1 + 1e-4, 1 - 1e-4)
    t = oracle.shell_tables(12)
    r = oracle.analyse_peaks(f, [21], [21], [21], 4.0, 4, t, ct, alatt=1.0, ZZon=1.0,
                             fcrit=1.0, variant="A")
    assert r["RTHL"][0] > 0
    assert abs(r["e_v"][0]) < 1e-12 and abs(r["p_v"][0]) < 1e-12
    np.testing.assert_allclose(r["gradpk"][0], 0, atol=1e-12)
    np.testing.assert_allclose(r["gradpkf"][0], 0, atol=1e-12)
    s = r["strain"][0]
    assert abs(s[0, 1]) < 1e-12 and abs(s[0, 0] - s[2, 2]) < 1e-12
    # strain trace equals Fbarx after the rescale, up to the last-shell rescaling
    assert np.trace(s) == pytest.approx(r["Fbarx"][0], rel=1e-12)


def test_early_exit_is_exact_on_a_random_field(oracle, websky):
    """Early exit against the full hunt radius on a Gaussian random field (2 Mpc/h cells,
    every candidate of a 6-filter bank): shells the early exit does not gather are NaN in
    the oracle, so any read of one would show up in the results."""
    from huntian_hip import synth

    n, box, nbuff, nhunt = 48, 96.0, 13, 12
    delta, dk = synth.gaussian_field(n, box, 777)
    f = {"delta": delta}
    f.update(synth.lpt_fields(dk, n, box))
    f["lapd"] = synth.laplacian_field(dk, n, box)
    radii = synth.filter_radii(1.5, 12.0, 6)
    sm = [synth.smooth_tophat(dk, n, box, R) for R in radii]
    pk, tot = oracle.peak_find_multi(sm, [1.5] * len(radii), nbuff,
                                     np.zeros((n, n, n), np.int8, order="F"))
    assert tot > 100
    Rf = np.array([radii[q] for q in pk["filt"]])
    ir2 = np.array([synth.ir2min_of(r, 2.0) for r in Rf], np.int32)
    t = oracle.shell_tables(nhunt)
    fcrit = oracle.fsc_of_z(websky, 0.0)
    for v in VARIANTS:
        kw = dict(alatt=2.0, ZZon=1.0, fcrit=fcrit, nbuff=nbuff, variant=v, nthreads=4)
        full = oracle.analyse_peaks(f, pk["i"], pk["j"], pk["k"], Rf, ir2, t, websky,
                                    early_exit=False, **kw)
        ee = oracle.analyse_peaks(f, pk["i"], pk["j"], pk["k"], Rf, ir2, t, websky,
                                  early_exit=True, **kw)
        assert (full["RTHL"] > 0).sum() > 20
        for k in full:
            assert np.array_equal(full[k], ee[k], equal_nan=True), (v, k)
        assert (full["nshells_used"] < t["nshells"] - 1).mean() > 0.5

"""GPU hot path vs the FP64 oracle on the same seeded inputs.

Bars: peak lists (indices, filter, value) and masks equal; collapse flags and halo counts
equal; float results within 1e-10 relative of the oracle.  The required tolerance is
1e-5 for RTHL / overdensity and 1e-4 for strain, e_v, p_v; the two sides run the same FP64
operations in the same order and differ only where the device and host log10 / pow differ
by an ulp, which reaches ~1e-12 after the RTHL interpolation and the eigen-solve."""
import numpy as np
import pytest

from fieldgen import FORTRAN_GOLDEN, bumps_field, fortran_golden_peak, nontrivial_ctab
from huntian_hip import api, synth
from huntian_hip.homeltab import make_ctab

pytestmark = pytest.mark.gpu

FLOAT_KEYS = ("RTHL", "Fbarx", "e_v", "p_v", "Srb", "d2F", "zvir_half", "eigs", "Sbar",
              "Sbar2", "gradpk", "gradpkf", "gradpkrf", "strain")
RTOL = 1e-10


@pytest.fixture(params=[1, 16, 64], ids=["lane", "gang16", "wave"])
def gang(request):
    """Run a test with one lane per peak (the reference's summation order), with sixteen
    lanes per peak (the same sums in lane order) and with one wavefront per peak (cells
    staged in LDS, each sum in the reference's order by its own lane); the bar is the same."""
    ctx = api.context()
    ctx.set_tuning(gang=request.param)
    yield request.param
    ctx.set_tuning(gang=1)


def to_host(r):
    out = {k: v.cpu().numpy() for k, v in r.items() if v is not None}
    out["strain"] = np.swapaxes(out["strain"], 1, 2)  # [p, K, L] storage -> [p, L, K]
    return out


def compare(gpu, ref, label=""):
    n = len(ref["RTHL"])
    assert np.array_equal(gpu["RTHL"] > 0, ref["RTHL"] > 0), label
    worst = {}
    for k in FLOAT_KEYS:
        a = gpu[k][:n].reshape(n, -1)
        b = ref[k].reshape(n, -1)
        # NaN where the reference itself makes one (gradpkrf when the filter-scale shell is
        # the centre: kernel_strain divides by rad[1] = 0, RadialShell.jl:259-260)
        assert np.array_equal(np.isnan(a), np.isnan(b)), (label, k, "NaN pattern")
        a = np.where(np.isnan(a), 0.0, a)
        b = np.where(np.isnan(b), 0.0, b)
        scale = np.maximum(np.abs(b), 1e-300)
        # entries that are tiny against the rest of their vector are compared absolutely
        vscale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-30)
        err = np.abs(a - b) / np.maximum(scale, 1e-9 * vscale)
        worst[k] = float(err.max()) if err.size else 0.0
        assert worst[k] <= RTOL, (label, k, worst[k])
    exact = sum(np.array_equal(gpu[k][:n], ref[k], equal_nan=True) for k in FLOAT_KEYS)
    print(f"{label}: {n} peaks, {exact}/{len(FLOAT_KEYS)} fields bit-identical, worst rel "
          f"{max(worst.values()):.2e} ({max(worst, key=worst.get)})")
    return worst


def oracle_run(oracle, f, pi, pj, pk, Rf, ir2min, t, ct, **kw):
    return oracle.analyse_peaks(f, pi, pj, pk, Rf, ir2min, t, ct, nthreads=8, **kw)


# ---------------------------------------------------------------- peak finding

def _mask(n):
    return np.zeros((n, n, n), np.int8, order="F")


@pytest.mark.parametrize("n,nbuff,seed,nan", [(24, 2, 1, False), (37, 3, 2, False),
                                               (64, 16, 3, False), (70, 1, 4, False),
                                               (100, 2, 5, False), (45, 4, 6, True)])
def test_peak_find_matches_oracle(oracle, n, nbuff, seed, nan):
    """Random (rough) fields: many maxima, ties on a plateau, a 3% mask; inner cubes that are
    not multiples of the kernel's 62 x 8 x 32 cells per wavefront, several of them along
    every axis at n=100; with nan, 0.5% NaN cells, whose comparisons are all false (a NaN
    neighbour never blocks a peak, a NaN cell passes every test)."""
    import torch

    rng = np.random.default_rng(seed)
    fields = [np.asfortranarray((rng.standard_normal((n, n, n)) * s).astype(np.float32))
              for s in (1.0, 1.4, 0.7, 1.1)]
    fields[2][nbuff + 2:nbuff + 5, nbuff + 2:nbuff + 5, nbuff + 2:nbuff + 5] = 5.0  # plateau
    if nan:
        for a in fields:
            a[rng.random((n, n, n)) < 0.005] = np.nan
    fc = [0.6, 1.0, 0.3, 1.7]
    mask0 = _mask(n)
    mask0[rng.random((n, n, n)) < 0.03] = 1
    mref = mask0.copy(order="F")
    ref, tot = oracle.peak_find_multi(fields, fc, nbuff, mref)
    dev = [api.to_device_field(a) for a in fields]
    md = api.to_device_field(mask0, dtype=torch.int8)
    got = api.peak_find(dev, fc, nbuff, md, capacity=tot + 10)
    assert got["total"] == tot > 0
    for k in ("i", "j", "k", "filt"):
        assert np.array_equal(got[k].cpu().numpy(), ref[k]), k
    assert np.array_equal(got["delta"].cpu().numpy(), ref["delta"], equal_nan=True)
    if nan:
        assert np.isnan(ref["delta"]).sum() > 10
    assert np.array_equal(md.cpu().numpy(), mref)


def test_peak_find_capacity(oracle):
    import torch

    n, nbuff = 32, 2
    rng = np.random.default_rng(11)
    fields = [np.asfortranarray(rng.standard_normal((n, n, n)).astype(np.float32))]
    mref = _mask(n)
    ref, tot = oracle.peak_find_multi(fields, [0.0], nbuff, mref, cap=50)
    md = api.to_device_field(_mask(n), dtype=torch.int8)
    got = api.peak_find([api.to_device_field(fields[0])], [0.0], nbuff, md, capacity=50)
    assert got["total"] == tot > 50 and len(got["i"]) == 50
    for k in ("i", "j", "k"):
        assert np.array_equal(got[k].cpu().numpy(), ref[k])
    assert np.array_equal(md.cpu().numpy(), mref)  # only the kept peaks are masked


def test_peak_find_rejects_bad_arguments():
    import torch

    from huntian_hip.abi import HuntianError

    f = api.to_device_field(np.zeros((8, 8, 8), np.float32))
    md = api.to_device_field(_mask(8), dtype=torch.int8)
    with pytest.raises(HuntianError):
        api.peak_find([f], [1.0], 0, md)  # nbuff must be >= 1


# ---------------------------------------------------------------- shell analysis

@pytest.fixture(scope="module")
def bumps():
    return bumps_field(56, 31415, nbumps=40, margin=10)


def _bump_peaks(oracle, f, nbuff=1, fcrit=0.12):
    m = _mask(f["delta"].shape[0])
    return oracle.find_peaks(f["delta"], m, nbuff, fcrit)


@pytest.mark.parametrize("variant", ["A", "B", "C"])
@pytest.mark.parametrize("early", [True, False])
def test_shell_analyse_bumps(oracle, bumps, variant, early, gang):
    import torch

    ct = nontrivial_ctab(make_ctab)
    rmax = 10
    t = oracle.shell_tables(rmax)
    stab = api.build_shell_tables(rmax)
    pi, pj, pk, _ = _bump_peaks(oracle, bumps)
    assert len(pi) > 10
    rng = np.random.default_rng(5)
    Rf = rng.choice([1.5, 2.2, 3.0, 4.0, np.sqrt(2.5)], size=len(pi))  # incl. an mrf tie
    ir2 = np.array([4 if r < 3 else 9 for r in Rf], np.int32)
    mref = _mask(56)
    ref = oracle_run(oracle, bumps, pi, pj, pk, Rf, ir2, t, ct, alatt=1.0, ZZon=0.0, fcrit=1.0,
                     variant=variant, early_exit=early, mask=mref, nbuff=2)
    md = api.to_device_field(_mask(56), dtype=torch.int8)
    r = api.shell_analyse(bumps, pi, pj, pk, Rf, ir2, stab, ct, alatt=1.0, ZZon=0.0, fcrit=1.0,
                          variant=variant, early_exit=early, mask=md, nbuff=2)
    g = to_host(r)
    compare(g, ref, f"{variant} early={early}")
    assert np.array_equal(g["nshells_used"], ref["nshells_used"])
    assert np.array_equal(g["crossing_shell"], ref["crossing_shell"])
    assert np.array_equal(md.cpu().numpy(), mref)
    assert (ref["RTHL"] > 0).sum() >= 3


def test_shell_analyse_boundary_and_optional_fields(oracle, bumps, gang):
    """Peaks whose shells leave the grid (bounds-checked path), no psi_2, no lapd,
    per-peak ZZon / fcrit."""
    ct = nontrivial_ctab(make_ctab)
    f = dict(bumps)
    f["eta2x"] = f["eta2y"] = f["eta2z"] = None
    f["lapd"] = None
    rmax = 9
    t = oracle.shell_tables(rmax)
    stab = api.build_shell_tables(rmax)
    pi, pj, pk, _ = _bump_peaks(oracle, bumps)
    pi = np.concatenate([pi, [1, 56, 3, 55]]).astype(np.int32)
    pj = np.concatenate([pj, [1, 20, 54, 2]]).astype(np.int32)
    pk = np.concatenate([pk, [24, 56, 2, 30]]).astype(np.int32)
    rng = np.random.default_rng(8)
    zz = rng.uniform(-0.5, 0.3, len(pi))
    fc = rng.uniform(0.7, 1.3, len(pi))
    for variant in ("A", "C"):
        ref = oracle_run(oracle, f, pi, pj, pk, 3.0, 9, t, ct, alatt=1.0, ZZon=0.0, fcrit=1.0,
                         variant=variant, early_exit=True, ZZon_pp=zz, fcrit_pp=fc)
        r = api.shell_analyse(f, pi, pj, pk, 3.0, 9, stab, ct, alatt=1.0, ZZon=0.0, fcrit=1.0,
                              variant=variant, early_exit=True, ZZon_pp=zz, fcrit_pp=fc)
        g = to_host(r)
        compare(g, ref, variant)
        assert np.all(g["Sbar2"] == 0) and np.all(g["d2F"] == 0)


def test_fortran_golden_vector_on_gpu(oracle, fortran_table, gang):
    gold = FORTRAN_GOLDEN
    stab = api.build_shell_tables(8)
    for variant in ("A", "B", "C"):
        r = to_host(api.shell_analyse(fortran_golden_peak(), [71], [71], [71], 4.0, 4, stab,
                                      fortran_table, alatt=1.0, ZZon=1.0, fcrit=gold["fcrit"],
                                      variant=variant))
        ref = oracle.analyse_peaks(fortran_golden_peak(), [71], [71], [71], 4.0, 4,
                                   oracle.shell_tables(8), fortran_table, alatt=1.0, ZZon=1.0,
                                   fcrit=gold["fcrit"], variant=variant)
        compare(r, ref, variant)
        assert abs(r["RTHL"][0] - gold["RTHL"]) / gold["RTHL"] < 1e-7
        assert abs(r["Srb"][0] - gold["Srb"]) / gold["Srb"] < 1e-6
        np.testing.assert_allclose(np.diag(r["strain"][0]), np.diag(gold["strain"]), rtol=1e-6)


def test_shell_analyse_rejects_bad_input(bumps):
    from huntian_hip.abi import HuntianError

    ct = nontrivial_ctab(make_ctab)
    stab = api.build_shell_tables(6)
    with pytest.raises(HuntianError):
        api.shell_analyse(bumps, [0], [5], [5], 2.0, 4, stab, ct, alatt=1.0, ZZon=0.0, fcrit=1.0)
    with pytest.raises(HuntianError):
        api.shell_analyse(bumps, [5], [5], [5], 2.0, 4, stab, ct, alatt=1.0, ZZon=-1.0,
                          fcrit=1.0)
    r = api.shell_analyse(bumps, [], [], [], 2.0, 4, stab, ct, alatt=1.0, ZZon=0.0, fcrit=1.0)
    assert len(r["RTHL"]) == 0


# ---------------------------------------------------------------- realistic end to end

@pytest.fixture(scope="module")
def cosmo64():
    """A 64^3 Gaussian random field (2 Mpc/h cells, BBKS, sigma8 0.81, seed 12345) with
    1LPT / 2LPT displacements and a 10-filter top-hat bank, 1.5..20 Mpc/h."""
    n, box = 64, 128.0
    delta, dk = synth.gaussian_field(n, box, 12345)
    f = {"delta": delta}
    f.update(synth.lpt_fields(dk, n, box))
    f["lapd"] = synth.laplacian_field(dk, n, box)
    radii = synth.filter_radii(1.5, 20.0, 10)
    sm = [synth.smooth_tophat(dk, n, box, R) for R in radii]
    return f, radii, sm


def test_end_to_end_realistic_field(oracle, websky, cosmo64, gang):
    """Peak finding over the filter bank, then every candidate's collapse, GPU vs oracle."""
    import torch

    f, radii, sm = cosmo64
    n, alatt, nbuff, nhunt = 64, 2.0, 11, 10
    fc = [1.5] * len(radii)
    mref = _mask(n)
    ref_pk, tot = oracle.peak_find_multi(sm, fc, nbuff, mref)
    md = api.to_device_field(_mask(n), dtype=torch.int8)
    got_pk = api.peak_find([api.to_device_field(a) for a in sm], fc, nbuff, md, capacity=tot + 1)
    assert tot > 30
    for k in ("i", "j", "k", "filt", "delta"):
        assert np.array_equal(got_pk[k].cpu().numpy(), ref_pk[k])
    assert np.array_equal(md.cpu().numpy(), mref)

    Rf = np.array([radii[q] for q in ref_pk["filt"]])
    ir2 = np.array([api.ir2min_of(r, alatt) for r in Rf], np.int32)
    fcrit = api.fsc_of_z(websky, 0.0)
    t = oracle.shell_tables(nhunt)
    stab = api.build_shell_tables(nhunt)
    for variant in ("A", "C"):
        mref2 = _mask(n)
        ref = oracle_run(oracle, f, ref_pk["i"], ref_pk["j"], ref_pk["k"], Rf, ir2, t, websky,
                         alatt=alatt, ZZon=1.0, fcrit=fcrit, variant=variant, mask=mref2,
                         nbuff=nbuff)
        md2 = api.to_device_field(_mask(n), dtype=torch.int8)
        r = api.shell_analyse(f, got_pk["i"], got_pk["j"], got_pk["k"], Rf, ir2, stab, websky,
                              alatt=alatt, ZZon=1.0, fcrit=fcrit, variant=variant, mask=md2,
                              nbuff=nbuff)
        g = to_host(r)
        compare(g, ref, variant)
        assert (g["RTHL"] > 0).sum() == (ref["RTHL"] > 0).sum() > 0
        assert np.array_equal(md2.cpu().numpy(), mref2)


def test_sbar_read_below_the_crossing(oracle, websky, cosmo64):
    """Sbar and Sbar2 are the running sums at m0 - 1, the shell below where the walk
    collapses (RadialShell.jl:738-754), and the walk can collapse anywhere below the crossing
    shell, also past ir2min.  The shell kernel keeps those sums only for the shells up to the
    crossing: the 64^3 field holds halos whose m0 - 1 lies past ir2min and two or more shells
    below the crossing, and their Sbar / Sbar2 equal the oracle's, in both schedules."""
    f, radii, sm = cosmo64
    n, alatt, nbuff, nhunt = 64, 2.0, 11, 10
    pk, _ = oracle.peak_find_multi(sm, [1.5] * len(radii), nbuff, _mask(n))
    Rf = np.array([radii[q] for q in pk["filt"]])
    ir2 = np.array([api.ir2min_of(r, alatt) for r in Rf], np.int32)
    fcrit = api.fsc_of_z(websky, 0.0)
    t = oracle.shell_tables(nhunt)
    kw = dict(alatt=alatt, ZZon=1.0, fcrit=fcrit, variant="C", nbuff=nbuff)
    ref = oracle_run(oracle, f, pk["i"], pk["j"], pk["k"], Rf, ir2, t, websky, **kw)
    r2 = t["shell_r2"]
    rad = np.concatenate([[0.0], np.sqrt(r2.astype(np.float64))])  # rad[s], s from 1
    halo = ref["RTHL"] > 0
    below = np.searchsorted(rad[1:], ref["RTHL"])  # m0 - 1, m0 the first s with rad >= RTHL
    sel = halo & (r2[np.maximum(below - 1, 0)] > ir2) & (below <= ref["crossing_shell"] - 2)
    assert sel.sum() >= 20, sel.sum()
    stab = api.build_shell_tables(nhunt)
    for order in ("crossing", "plain"):
        g = to_host(api.shell_analyse(f, pk["i"], pk["j"], pk["k"], Rf, ir2, stab, websky,
                                      order=order, **kw))
        compare(g, ref, f"Sbar below the crossing, {order}")
        for k in ("Sbar", "Sbar2"):
            a, b = g[k][: len(sel)][sel], ref[k][sel]
            assert np.all(np.abs(a - b) <= RTOL * np.maximum(np.abs(b), 1e-30)), (order, k)


def test_multirf_mirror(oracle, bumps):
    """analyse_peaks_gpu_cuda_multirf returns the reference NamedTuple fields per batch."""
    ct_np = nontrivial_ctab(make_ctab)
    stab = api.build_shell_tables(8)
    pi, pj, pk, _ = _bump_peaks(oracle, bumps)
    h = len(pi) // 2
    batches = [(pi[:h], pj[:h], pk[:h], 3.0, 9), (pi[h:], pj[h:], pk[h:], 2.0, 4)]
    mask = _mask(56)
    res, m = api.analyse_peaks_gpu_cuda_multirf(
        bumps["delta"], bumps["etax"], bumps["etay"], bumps["etaz"], bumps["eta2x"],
        bumps["eta2y"], bumps["eta2z"], batches, stab, ct_np["table"], -2.0, 2.0, 0.0, 1.0,
        0.0, 2.0, 1.0, 0.0, fcrit_override=1.0, lapd=bumps["lapd"], mask=mask, nbuff=2)
    assert len(res) == 2 and m.shape == (56, 56, 56)
    t = oracle.shell_tables(8)
    for (bi, bj, bk, Rf, ir2), d in zip(batches, res):
        ref = oracle_run(oracle, bumps, bi, bj, bk, Rf, ir2, t, ct_np, alatt=1.0, ZZon=0.0,
                         fcrit=1.0, variant="C")
        assert d["strain_final"].shape == (3, 3, len(bi)) and d["Sbar"].shape == (3, len(bi))
        np.testing.assert_allclose(d["RTHL"], ref["RTHL"].astype(np.float32), rtol=1e-6)
        np.testing.assert_allclose(d["strain_final"], np.transpose(ref["strain"], (1, 2, 0)),
                                   rtol=1e-5, atol=1e-6)


def test_fp32_kahan_sums_within_tolerance(oracle, websky, cosmo64):
    """Per-shell sums in FP32 with Kahan compensation against the FP64 oracle: same halo
    set, RTHL and overdensity within 1e-5 relative, strain, e_v, p_v within 1e-4."""
    f, radii, sm = cosmo64
    n, alatt, nbuff, nhunt = 64, 2.0, 11, 10
    ref_pk, tot = oracle.peak_find_multi(sm, [1.5] * len(radii), nbuff, _mask(n))
    Rf = np.array([radii[q] for q in ref_pk["filt"]])
    ir2 = np.array([api.ir2min_of(r, alatt) for r in Rf], np.int32)
    fcrit = api.fsc_of_z(websky, 0.0)
    t = oracle.shell_tables(nhunt)
    ref = oracle_run(oracle, f, ref_pk["i"], ref_pk["j"], ref_pk["k"], Rf, ir2, t, websky,
                     alatt=alatt, ZZon=1.0, fcrit=fcrit, variant="C", nbuff=nbuff)
    r = api.shell_analyse(f, ref_pk["i"], ref_pk["j"], ref_pk["k"], Rf, ir2,
                          api.build_shell_tables(nhunt), websky, alatt=alatt, ZZon=1.0,
                          fcrit=fcrit, variant="C", nbuff=nbuff, precision="fp32_kahan")
    g = to_host(r)
    assert np.array_equal(g["RTHL"] > 0, ref["RTHL"] > 0)
    halo = ref["RTHL"] > 0
    worst = {}
    for k, tol in (("RTHL", 1e-5), ("Fbarx", 1e-5), ("strain", 1e-4), ("e_v", 1e-4),
                   ("p_v", 1e-4)):
        a = g[k][halo].reshape(halo.sum(), -1)
        b = ref[k][halo].reshape(halo.sum(), -1)
        vs = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-30)
        floor = 1e-3 * vs if k in ("RTHL", "Fbarx", "strain") else 1e-3  # e_v, p_v near 0
        err = np.abs(a - b) / np.maximum(np.abs(b), floor)
        worst[k] = float(err.max())
        assert worst[k] <= tol, (k, worst[k])
    print(f"fp32_kahan vs FP64 oracle over {halo.sum()} halos: worst relative {worst}")


def test_boundary_peaks_with_lapd(oracle, bumps, gang):
    """d2F of peaks whose shells leave the grid: variant A stops adding the Laplacian at the
    first out-of-bounds cell (RadialShell.jl:784-799), B and C skip such cells."""
    ct = nontrivial_ctab(make_ctab)
    assert bumps.get("lapd") is not None
    rmax = 9
    t = oracle.shell_tables(rmax)
    stab = api.build_shell_tables(rmax)
    pi, pj, pk, _ = _bump_peaks(oracle, bumps)
    pi = np.concatenate([pi, [1, 56, 3, 55, 5, 52]]).astype(np.int32)
    pj = np.concatenate([pj, [1, 20, 54, 2, 28, 30]]).astype(np.int32)
    pk = np.concatenate([pk, [24, 56, 2, 30, 6, 50]]).astype(np.int32)
    for variant in ("A", "B", "C"):
        ref = oracle_run(oracle, bumps, pi, pj, pk, 3.0, 9, t, ct, alatt=1.0, ZZon=0.0,
                         fcrit=1.0, variant=variant, early_exit=True)
        r = api.shell_analyse(bumps, pi, pj, pk, 3.0, 9, stab, ct, alatt=1.0, ZZon=0.0,
                              fcrit=1.0, variant=variant, early_exit=True)
        g = to_host(r)
        compare(g, ref, f"{variant} boundary+lapd")
        assert np.any(g["d2F"] != 0)


@pytest.mark.parametrize("order", ["crossing", "plain"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_kahan"])
def test_wave_gather_bit_identical_to_lane_gather(websky, cosmo64, order, precision):
    """The wavefront gather runs every sum in the reference's cell order, like the one-lane
    gather, so every output is bit-identical -- with the crossing pass's stop shells bounding
    its chunks and without them (plain order), in FP64 and in FP32-Kahan."""
    import torch

    f, radii, sm = cosmo64
    n, alatt, nbuff, nhunt = 64, 2.0, 11, 10
    md = api.to_device_field(_mask(n), dtype=torch.int8)
    pk = api.peak_find([api.to_device_field(a) for a in sm], [1.5] * len(radii), nbuff, md,
                       capacity=n ** 3)
    Rf = np.array([radii[q] for q in pk["filt"].cpu().numpy()])
    ir2 = np.array([api.ir2min_of(r, alatt) for r in Rf], np.int32)
    stab = api.build_shell_tables(nhunt)
    kw = dict(alatt=alatt, ZZon=1.0, fcrit=api.fsc_of_z(websky, 0.0), variant="C",
              nbuff=nbuff, order=order, precision=precision)
    args = (f, pk["i"], pk["j"], pk["k"], Rf, ir2, stab, websky)
    ctx = api.context()
    try:
        ctx.set_tuning(gang=1)
        one = api.shell_analyse(*args, **kw)
        ctx.set_tuning(gang=64)
        wave = api.shell_analyse(*args, **kw)
    finally:
        ctx.set_tuning(gang=1)
    assert len(pk["i"]) > 30
    for k in one:
        assert torch.equal(torch.nan_to_num(one[k], nan=7.0), torch.nan_to_num(wave[k], nan=7.0)), k

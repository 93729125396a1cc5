"""The oracle's rmax2rs > 0 shell limit (variants B and C): ShellAnalysisGPU.jl:90-99 (the
leading shells with r^2 <= (rmax2rs Rf / alatt)^2, no collapse with fewer than 2),
ShellAnalysisGPU.jl:112-117 (variant B's pre-mask skip), ext/CUDAExt.jl:2176-2185 and
2235-2236 (the batched path: same limit, every shell up to it gathered).

Pins: a limit that cuts the table after shell s gives what the table built only up to
shell s + 1 gives with no limit (the analysis uses shells 1 .. NS - 1 of any table); a limit
beyond every shell changes nothing; below two shells no peak collapses; variant B skips
exactly the peaks whose cell value exceeds floor(Rf / alatt rmax2rs)."""
import numpy as np
import pytest

from fieldgen import bumps_field, nontrivial_ctab
from huntian_hip.homeltab import make_ctab

RESULT_KEYS = ("RTHL", "Fbarx", "e_v", "p_v", "Srb", "d2F", "zvir_half", "eigs", "Sbar",
               "Sbar2", "gradpk", "gradpkf", "gradpkrf", "strain", "nshells_used",
               "crossing_shell")


@pytest.fixture(scope="module")
def bumps():
    return bumps_field(40, 31415, nbumps=12, margin=8)


@pytest.fixture(scope="module")
def peaks(oracle, bumps):
    m = np.zeros(bumps["delta"].shape, np.int8, order="F")
    i, j, k, _ = oracle.find_peaks(bumps["delta"], m, 1, 0.8)
    assert len(i) > 5
    return i, j, k


def _run(oracle, f, pk, t, ct, Rf=4.0, **kw):
    i, j, k = pk
    kw.setdefault("early_exit", False)
    return oracle.analyse_peaks(f, i, j, k, Rf, 4, t, ct, alatt=1.0, ZZon=0.0, fcrit=1.0, **kw)


def _same(a, b, label):
    for key in RESULT_KEYS:
        assert np.array_equal(a[key], b[key], equal_nan=True), (label, key)


def test_small_table_is_a_prefix_of_a_large_one(oracle):
    small, big = oracle.shell_tables(6), oracle.shell_tables(10)
    ns = small["nshells"]
    for key in ("shell_start", "shell_count", "shell_r2"):
        assert np.array_equal(small[key], big[key][:ns]), key
    nc = small["ncells"]
    for key in ("di", "dj", "dk"):
        assert np.array_equal(small[key], big[key][:nc]), key


@pytest.mark.parametrize("variant", ["B", "C"])
@pytest.mark.parametrize("Rf", [4.0, 2.5])
def test_limit_equals_the_truncated_table(oracle, bumps, peaks, variant, Rf):
    ct = nontrivial_ctab(make_ctab)
    small, big = oracle.shell_tables(6), oracle.shell_tables(10)
    ns = small["nshells"]
    # the limit keeps shells 1 .. ns - 1 of the big table: r2[ns - 2] <= lim < r2[ns - 1]
    lim = float(big["shell_r2"][ns - 2]) + 0.5
    assert lim < big["shell_r2"][ns - 1]
    rmax2rs = np.sqrt(lim) / Rf
    got = _run(oracle, bumps, peaks, big, ct, Rf=Rf, variant=variant, rmax2rs=rmax2rs)
    want = _run(oracle, bumps, peaks, small, ct, Rf=Rf, variant=variant)
    _same(got, want, (variant, Rf))
    assert (got["RTHL"] > 0).sum() >= 3
    # and the limit did change something against the big table without one
    free = _run(oracle, bumps, peaks, big, ct, Rf=Rf, variant=variant)
    assert not np.array_equal(free["nshells_used"], got["nshells_used"])


@pytest.mark.parametrize("variant", ["B", "C"])
def test_limit_beyond_every_shell_changes_nothing(oracle, bumps, peaks, variant):
    ct = nontrivial_ctab(make_ctab)
    t = oracle.shell_tables(8)
    free = _run(oracle, bumps, peaks, t, ct, variant=variant)
    lim = _run(oracle, bumps, peaks, t, ct, variant=variant, rmax2rs=1e6)
    _same(lim, free, variant)
    # early exit is ignored under a limit: all shells gathered, as the reference GPU path
    ee = _run(oracle, bumps, peaks, t, ct, variant=variant, rmax2rs=1e6, early_exit=True)
    _same(ee, free, variant)


@pytest.mark.parametrize("rmax2rs", [0.1, 0.24])
def test_fewer_than_two_shells_never_collapse(oracle, bumps, peaks, rmax2rs):
    # Rf = 4: (rmax2rs Rf)^2 < 1 keeps the centre shell only
    ct = nontrivial_ctab(make_ctab)
    t = oracle.shell_tables(8)
    r = _run(oracle, bumps, peaks, t, ct, variant="C", rmax2rs=rmax2rs)
    assert np.all(r["RTHL"] == -1.0) and np.all(r["zvir_half"] == -1.0)
    assert np.all(r["nshells_used"] == 0) and np.all(r["strain"] == 0.0)
    # r^2 = 1 (rmax2rs = 0.25) already gives two shells
    r = _run(oracle, bumps, peaks, t, ct, variant="C", rmax2rs=0.25)
    assert np.all(r["nshells_used"] == 2)


def test_variant_b_skips_peaks_the_mask_already_holds(oracle, bumps, peaks):
    """The skip threshold floor(Rf / alatt rmax2rs): with Rf = 4 and rmax2rs = 0.6 it is 2,
    so a cell value of 3 skips its peak and a value of 2 does not."""
    ct = nontrivial_ctab(make_ctab)
    t = oracle.shell_tables(8)
    i, j, k = peaks
    n = bumps["delta"].shape[0]
    mask = np.zeros((n, n, n), np.int8, order="F")
    free = _run(oracle, bumps, peaks, t, ct, variant="B", rmax2rs=0.6,
                mask=mask.copy(order="F"))
    hal = np.nonzero(free["RTHL"] > 0)[0]
    assert len(hal) >= 3
    q3, q2 = hal[0], hal[1]
    mask[i[q3] - 1, j[q3] - 1, k[q3] - 1] = 3
    mask[i[q2] - 1, j[q2] - 1, k[q2] - 1] = 2
    mB = mask.copy(order="F")
    B = _run(oracle, bumps, peaks, t, ct, variant="B", rmax2rs=0.6, mask=mB)
    assert B["RTHL"][q3] == -1.0 and B["nshells_used"][q3] == 0
    assert B["RTHL"][q2] == free["RTHL"][q2]
    rest = np.ones(len(i), bool)
    rest[q3] = False
    for key in RESULT_KEYS:
        assert np.array_equal(B[key][rest], free[key][rest], equal_nan=True), key
    # variant C (the batched reference path) has no such skip
    C = _run(oracle, bumps, peaks, t, ct, variant="C", rmax2rs=0.6,
             mask=mask.copy(order="F"))
    assert C["RTHL"][q3] == free["RTHL"][q3]


def test_rejected_settings(oracle, bumps, peaks):
    ct = nontrivial_ctab(make_ctab)
    t = oracle.shell_tables(6)
    with pytest.raises(Exception):
        _run(oracle, bumps, peaks, t, ct, variant="A", rmax2rs=1.0)
    with pytest.raises(Exception):
        _run(oracle, bumps, peaks, t, ct, variant="C", rmax2rs=-1.0)

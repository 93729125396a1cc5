"""rmax2rs > 0 through the C-ABI against the oracle: each peak's shells limited to
r^2 <= (rmax2rs Rf / alatt)^2 (ShellAnalysisGPU.jl:90-99, ext/CUDAExt.jl:2176-2185), every
one of them gathered (ext/CUDAExt.jl:2235-2236), and variant B's skip of peaks whose cell
the mask holds beyond floor(Rf / alatt rmax2rs) (ShellAnalysisGPU.jl:112-117).

Bars as the other shell-analysis parity tests: collapse flags, shell counts and masks
equal, float results within 1e-10 relative of the FP64 oracle."""
import numpy as np
import pytest

from checks import compare, to_host
from fieldgen import bumps_field, nontrivial_ctab
from huntian_hip import api
from huntian_hip.homeltab import make_ctab

pytestmark = pytest.mark.gpu

N = 56


@pytest.fixture(scope="module")
def bumps():
    return bumps_field(N, 31415, nbumps=40, margin=10)


@pytest.fixture(scope="module")
def case(oracle, bumps):
    m = np.zeros((N, N, N), np.int8, order="F")
    pi, pj, pk, _ = oracle.find_peaks(bumps["delta"], m, 1, 0.12)
    assert len(pi) > 10
    rng = np.random.default_rng(5)
    Rf = rng.choice([1.5, 2.2, 3.0, 4.0, np.sqrt(2.5)], size=len(pi))
    ir2 = np.array([4 if r < 3 else 9 for r in Rf], np.int32)
    # a caller's mask holding values above 1 at some peak cells (variant B skips a peak whose
    # value exceeds floor(Rf / alatt rmax2rs)), 1 at others
    mask0 = np.zeros((N, N, N), np.int8, order="F")
    for q in range(0, len(pi), 3):
        mask0[pi[q] - 1, pj[q] - 1, pk[q] - 1] = 1 + q % 5
    return pi, pj, pk, Rf, ir2, mask0


def _both(oracle, bumps, case, variant, rmax2rs, rmax=10):
    import torch

    pi, pj, pk, Rf, ir2, mask0 = case
    ct = nontrivial_ctab(make_ctab)
    mref = mask0.copy(order="F")
    ref = oracle.analyse_peaks(bumps, pi, pj, pk, Rf, ir2, oracle.shell_tables(rmax), ct,
                               alatt=1.0, ZZon=0.0, fcrit=1.0, variant=variant,
                               early_exit=False, mask=mref, nbuff=2, rmax2rs=rmax2rs,
                               nthreads=8)
    md = api.to_device_field(mask0, dtype=torch.int8)
    r = api.shell_analyse(bumps, pi, pj, pk, Rf, ir2, api.build_shell_tables(rmax), ct,
                          alatt=1.0, ZZon=0.0, fcrit=1.0, variant=variant, early_exit=True,
                          mask=md, nbuff=2, rmax2rs=rmax2rs)
    return to_host(r), ref, md.cpu().numpy(), mref


@pytest.mark.parametrize("variant", ["B", "C"])
@pytest.mark.parametrize("rmax2rs", [0.24, 0.6, 1.3, 2.5, 1e6])
def test_limit_matches_oracle(oracle, bumps, case, variant, rmax2rs):
    print(f"variant {variant} rmax2rs {rmax2rs}", flush=True)
    g, ref, mg, mref = _both(oracle, bumps, case, variant, rmax2rs)
    compare(g, ref, f"{variant} rmax2rs={rmax2rs}")
    assert np.array_equal(g["nshells_used"], ref["nshells_used"])
    assert np.array_equal(g["crossing_shell"], ref["crossing_shell"])
    assert np.array_equal(mg, mref)
    if rmax2rs == 0.24:  # at most the centre shell for Rf <= 4: nothing collapses
        assert np.all(ref["RTHL"] == -1.0) and np.all(ref["nshells_used"] == 0)
    if rmax2rs in (1.3, 2.5):
        assert (ref["RTHL"] > 0).sum() >= 3


def test_skip_and_limit_with_every_gang_setting(oracle, bumps, case):
    """The ganged and wavefront kernels analyse every shell: with a limit the library takes
    the lane-per-peak kernel whatever the tuning says."""
    ctx = api.context()
    try:
        for gang in (16, 64):
            ctx.set_tuning(gang=gang)
            g, ref, mg, mref = _both(oracle, bumps, case, "B", 0.6)
            compare(g, ref, f"B gang={gang}")
            assert np.array_equal(mg, mref)
    finally:
        ctx.set_tuning(gang=1)


def test_plan_replays_take_the_mask_as_they_find_it(oracle, bumps, case):
    """A plan snapshots variant B's skip flags in its graph: a replay after the mask is reset
    in place equals the first run; a replay over the mask the first run left (halo cells
    marked 1) equals an eager call on that mask."""
    import torch

    pi, pj, pk, Rf, ir2, mask0 = case
    ct = nontrivial_ctab(make_ctab)
    stab = api.build_shell_tables(10)
    md = api.to_device_field(mask0, dtype=torch.int8)
    tile = dict(fields=bumps, pi=pi, pj=pj, pk=pk, Rf=Rf, ir2min=ir2, mask=md)
    kw = dict(alatt=1.0, ZZon=0.0, fcrit=1.0, nbuff=2, variant="B", rmax2rs=0.6)
    plan = api.ShellPlan([tile], stab, ct, **kw)
    first = {k: v.clone() for k, v in plan.results[0].items()}
    m_after = md.clone()
    g, ref, _, mref = _both(oracle, bumps, case, "B", 0.6)
    compare(to_host(first), ref, "plan first run")
    assert np.array_equal(m_after.cpu().numpy(), mref)
    md.copy_(api.to_device_field(mask0, dtype=torch.int8))
    for r in plan.results:
        r["RTHL"].fill_(-5.0)
    plan.launch()
    torch.cuda.synchronize()
    for k in first:
        assert torch.equal(torch.nan_to_num(plan.results[0][k], nan=7.0),
                           torch.nan_to_num(first[k], nan=7.0)), k
    # replay over the marked mask: the same as an eager call on it
    md.copy_(m_after)
    plan.launch()
    torch.cuda.synchronize()
    m2 = m_after.clone()
    eager = api.shell_analyse(bumps, pi, pj, pk, Rf, ir2, stab, ct, early_exit=True, mask=m2,
                              **kw)
    for k in first:
        assert torch.equal(torch.nan_to_num(plan.results[0][k], nan=7.0),
                           torch.nan_to_num(eager[k], nan=7.0)), k
    plan.close()


def test_variant_a_and_negative_limits_are_rejected(bumps, case):
    from huntian_hip.abi import HuntianError

    pi, pj, pk, Rf, ir2, _ = case
    ct = nontrivial_ctab(make_ctab)
    stab = api.build_shell_tables(6)
    for variant, rmax2rs in (("A", 1.0), ("C", -1.0), ("C", float("nan"))):
        with pytest.raises(HuntianError):
            api.shell_analyse(bumps, pi, pj, pk, Rf, ir2, stab, ct, alatt=1.0, ZZon=0.0,
                              fcrit=1.0, variant=variant, rmax2rs=rmax2rs)

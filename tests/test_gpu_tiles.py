"""hnt_shell_analyse_tiles: tiles of uneven size and candidate count in one persistent launch
give exactly the results (and masks) of one hnt_shell_analyse call per tile, which the
parity tests tie to the oracle; one tile is also checked against the oracle directly."""
import numpy as np
import pytest

from huntian_hip import api, synth

pytestmark = pytest.mark.gpu


def _tile(n, seed, nbuff, oracle):
    box = 2.0 * n
    delta, dk = synth.gaussian_field(n, box, seed)
    f = {"delta": delta}
    f.update(synth.lpt_fields(dk, n, box))
    f["lapd"] = synth.laplacian_field(dk, n, box)
    radii = synth.filter_radii(1.5, 12.0, 5)
    sm = [synth.smooth_tophat(dk, n, box, R) for R in radii]
    pk, _ = oracle.peak_find_multi(sm, [1.5] * len(radii), nbuff,
                                   np.zeros((n, n, n), np.int8, order="F"))
    Rf = np.array([radii[q] for q in pk["filt"]])
    ir2 = np.array([api.ir2min_of(r, 2.0) for r in Rf], np.int32)
    return dict(fields=f, pi=pk["i"], pj=pk["j"], pk=pk["k"], Rf=Rf, ir2min=ir2)


@pytest.mark.parametrize("order", ["crossing", "plain"])
def test_tiles_equal_one_call_per_tile(oracle, websky, order):
    import torch

    nbuff, nhunt = 11, 10
    specs = [(32, 101), (48, 102), (24, 103), (40, 104)]
    tiles = [_tile(n, s, nbuff, oracle) for n, s in specs]
    # an empty tile in the middle
    e = tiles[2]
    tiles.insert(2, dict(fields=e["fields"], pi=e["pi"][:0], pj=e["pj"][:0], pk=e["pk"][:0],
                         Rf=e["Rf"][:0], ir2min=e["ir2min"][:0]))
    stab = api.build_shell_tables(nhunt)
    fcrit = api.fsc_of_z(websky, 0.0)
    kw = dict(alatt=2.0, ZZon=1.0, fcrit=fcrit, nbuff=nbuff, order=order)
    for t in tiles:
        n = t["fields"]["delta"].shape[0]
        t["mask"] = api.to_device_field(np.zeros((n, n, n), np.int8), dtype=torch.int8)
    got = api.shell_analyse_tiles(tiles, stab, websky, **kw)
    assert sum(len(t["pi"]) for t in tiles) > 200
    for t, g in zip(tiles, got):
        n = t["fields"]["delta"].shape[0]
        m1 = api.to_device_field(np.zeros((n, n, n), np.int8), dtype=torch.int8)
        one = api.shell_analyse(t["fields"], t["pi"], t["pj"], t["pk"], t["Rf"], t["ir2min"],
                                stab, websky, mask=m1, **kw)
        for k in one:
            assert torch.equal(torch.nan_to_num(one[k], nan=123.0),
                               torch.nan_to_num(g[k], nan=123.0)), k
        assert torch.equal(m1, t["mask"])
    # first tile against the oracle directly
    t = tiles[0]
    ref = oracle.analyse_peaks(t["fields"], t["pi"], t["pj"], t["pk"], t["Rf"], t["ir2min"],
                               oracle.shell_tables(nhunt), websky, alatt=2.0, ZZon=1.0,
                               fcrit=fcrit, variant="C", nbuff=nbuff, early_exit=True)
    g0 = {k: v.cpu().numpy() for k, v in got[0].items()}
    assert np.array_equal(g0["RTHL"] > 0, ref["RTHL"] > 0)
    assert np.array_equal(g0["nshells_used"], ref["nshells_used"])
    np.testing.assert_allclose(g0["RTHL"], ref["RTHL"], rtol=1e-10)


def test_plan_replays_match_eager_calls(oracle, websky):
    """A plan (hipGraph of the repacks and the persistent kernel) replayed several times,
    with other GPU work in between, reproduces the eager results; after the fields change in
    place, a replay equals a fresh eager call on the changed fields."""
    import torch

    nbuff, nhunt = 11, 10
    tiles = [_tile(n, s, nbuff, oracle) for n, s in [(32, 201), (40, 202), (24, 203)]]
    stab = api.build_shell_tables(nhunt)
    fcrit = api.fsc_of_z(websky, 0.0)
    kw = dict(alatt=2.0, ZZon=1.0, fcrit=fcrit, nbuff=nbuff)
    plan = api.ShellPlan(tiles, stab, websky, **kw)
    eager = api.shell_analyse_tiles(tiles, stab, websky, **kw)

    def same(a, b):
        for k in a:
            assert torch.equal(torch.nan_to_num(a[k], nan=7.0), torch.nan_to_num(b[k], nan=7.0)), k

    for a, b in zip(plan.results, eager):
        same(a, b)
    for rep in range(3):
        for r in plan.results:
            r["RTHL"].fill_(-5.0)
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for a, b in zip(plan.results, eager):
            same(a, b)
    # change the fields in place (the plan's own device copies) and in a fresh set of inputs
    changed = []
    for t, (kf, kb) in zip(tiles, zip(plan.inputs[0::2], plan.inputs[1::2])):
        kf["delta"].mul_(1.07)
        kf["etax"].mul_(0.9)
        f2 = {k: v.clone() for k, v in kf.items()}
        changed.append(dict(t, fields=f2))
    plan.launch()
    torch.cuda.synchronize()
    fresh = api.shell_analyse_tiles(changed, stab, websky, **kw)
    for a, b in zip(plan.results, fresh):
        same(a, b)
    assert not torch.equal(plan.results[0]["RTHL"], eager[0]["RTHL"])
    plan.close()


def test_plan_keeps_its_tables_after_other_calls(oracle, websky):
    """A plan reads its own copies of the shell, collapse and SPH kernel tables: eager calls
    on the same context with another hunt radius and another collapse table (which replace
    the context's uploads, reallocating them) leave later replays exact."""
    import torch
    from fieldgen import nontrivial_ctab
    from huntian_hip.homeltab import make_ctab

    nbuff, nhunt = 11, 10
    tiles = [_tile(n, s, nbuff, oracle) for n, s in [(32, 211), (40, 212)]]
    stab = api.build_shell_tables(nhunt)
    fcrit = api.fsc_of_z(websky, 0.0)
    kw = dict(alatt=2.0, ZZon=1.0, fcrit=fcrit, nbuff=nbuff)
    plan = api.ShellPlan(tiles, stab, websky, **kw)
    ref = [{k: v.clone() for k, v in r.items()} for r in plan.results]
    # other calls in between: larger and smaller shell tables, another collapse table
    other_ct = nontrivial_ctab(make_ctab)
    for rmax, ct in ((14, other_ct), (6, websky), (9, other_ct)):
        api.shell_analyse_tiles(tiles, api.build_shell_tables(rmax), ct, **kw)
    for r in plan.results:
        for v in r.values():
            v.fill_(-3)
    plan.launch()
    torch.cuda.synchronize()
    for a, b in zip(plan.results, ref):
        for k in a:
            assert torch.equal(torch.nan_to_num(a[k], nan=7.0), torch.nan_to_num(b[k], nan=7.0)), k
    plan.close()

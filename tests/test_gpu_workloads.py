"""The benchmark workloads end to end on the GPU, against the FP64 oracle on the same inputs.

Each workload is synthesised on the GPU exactly as bench.py makes it (huntian_hip.synth):
Threefry noise, BBKS P(k) with sigma8 0.81, 1LPT + 2LPT displacements, top-hat filter bank,
threshold 1.5, collapse at z = 0 with HomelTab_websky.dat, alatt 2 Mpc/h.  The fields are
then copied to the host and the oracle runs on the very same arrays.

Bars (the required ones are the peak list and halo count bit-exact, RTHL and overdensity
within 1e-5, strain / e_v / p_v within 1e-4 relative): the peak list (i, j, k, filter,
value) and the mask equal the oracle's; the halo set equals the oracle's; every float field
within 1e-8 relative for the FP64 sums (they differ from the oracle only by device vs host
log10 / pow ulps, which stay below 1e-10 on all but a few of millions of peaks).  For the
compensated-FP32 option the errors against FP64 are printed and the share of halos beyond
the required bars is held below 1e-5.

  tile128   128^3, box 256, seed 12345, 16 filters 1.5-20, nbuff 16, hunt radius 15:
            every candidate, variants A (full scan, the reference CPU path) and C (early exit)
  tile512   512^3, box 1024, seed 7, 16 filters, nbuff 32, hunt radius 20: the full peak
            list and every candidate (variant C, early exit)
  tile1024  1024^3, box 2048, seed 2024, 24 filters 1.5-34, nbuff 64, hunt radius 30: the
            full peak list (filter by filter, one smoothed field on the host at a time) and a
            uniform sample of 250 000 candidates; the compensated-FP32 sums on every candidate
  batched64 64 tiles of sides 96-256, buffers 16-48, seeds 100-163: one hipGraph plan over all
            tiles, a per-tile sample of its candidates against the oracle
The learned schedule on its evaluation tile (seed 99) is in test_gpu_bracket.py.
"""
import os

import numpy as np
import pytest
import torch

from checks import BARS, bar_errors, compare, rel_errors, to_host
from huntian_hip import api, synth
from huntian_hip.homeltab import WEBSKY_TABLE, read_homeltab

pytestmark = pytest.mark.gpu

NTHREADS = max(1, min(16, len(os.sched_getaffinity(0))))
RTOL = 1e-8


@pytest.fixture(scope="module")
def websky_ct():
    return read_homeltab(WEBSKY_TABLE)


def _release():
    api.context().release_scratch()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _host_fields(fields):
    """Column-major (n,n,n) device tensors -> Fortran-ordered numpy arrays (same bytes)."""
    out = {}
    for k, v in fields.items():
        a = v.permute(2, 1, 0).contiguous().cpu().numpy()  # C order [k, j, i]
        out[k] = a.T  # Fortran-ordered [i, j, k] view, no copy
        assert out[k].flags.f_contiguous
    return out


def _peaks_gpu(w, mask):
    n, nbuff = w["n"], w["nbuff"]
    mask.zero_()
    pf = api.peak_find(w["smoothed"], [1.5] * len(w["radii"]), nbuff, mask,
                       capacity=(n - 2 * nbuff) ** 3)
    return pf


def _peak_inputs(w, pf, dev):
    Rf_tab = torch.tensor(w["radii"], dtype=torch.float64, device=dev)
    ir2_tab = torch.tensor([api.ir2min_of(R, w["alatt"]) for R in w["radii"]], dtype=torch.int32,
                           device=dev)
    f = pf["filt"].long()
    return Rf_tab[f], ir2_tab[f]


def _check_peaks_against_oracle(oracle, w, pf, streamed=False):
    """The GPU peak list and mask against the oracle's filter loop on the same smoothed
    fields (copied to the host all at once, or one field at a time with the mask carried)."""
    n, nbuff = w["n"], w["nbuff"]
    fc = [1.5] * len(w["radii"])
    mask_h = np.zeros((n, n, n), np.int8, order="F")
    cap = (n - 2 * nbuff) ** 3
    if not streamed:
        sm = [_host_fields({"s": s})["s"] for s in w["smoothed"]]
        ref, tot = oracle.peak_find_multi(sm, fc, nbuff, mask_h, cap=cap, nthreads=NTHREADS)
        del sm
    else:
        parts, tot = [], 0
        for q, s in enumerate(w["smoothed"]):
            sh = _host_fields({"s": s})["s"]
            r, t = oracle.peak_find_multi([sh], [fc[q]], nbuff, mask_h, cap=cap,
                                          nthreads=NTHREADS)
            r["filt"][:] = q
            parts.append(r)
            tot += t
            del sh
            print(f"  oracle peak scan, filter {q}: {t} peaks", flush=True)
        ref = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    got = {k: v.cpu().numpy() for k, v in pf.items() if k in ref}
    print(f"peak list: {tot} peaks (oracle), {len(got['i'])} (GPU)", flush=True)
    assert len(got["i"]) == tot == len(ref["i"])
    for k in ("i", "j", "k", "filt", "delta"):
        assert np.array_equal(got[k], ref[k]), k
    return mask_h


def _assert_reduced_precision(worst, flips, nhalo):
    """Compensated FP32 shell sums against FP64: their rounding (~1e-7 per shell sum) is
    amplified where the kernel-weighted strain nearly cancels or two walk steps give nearly
    the same zvir, so a few halos in 10^6 exceed the bars (RTHL, overdensity 1e-5; strain,
    e_v, p_v 1e-4; the worst values are printed); at most one halo flag in 10^5 flips."""
    lim = max(3, nhalo // 100_000)
    assert flips <= lim, flips
    for k in BARS:
        assert worst[k + "_over_bar"] <= lim, (k, worst[k + "_over_bar"], worst[k])


def test_tile128_every_candidate_variants_A_B_and_C(oracle, websky_ct):
    dev = torch.device("cuda", 0)
    w = synth.workload_tile("tile128", dev)
    n, nbuff, nhunt, alatt = w["n"], w["nbuff"], w["nhunt"], w["alatt"]
    mask = api.fortran_empty((n, n, n), torch.int8, dev)
    pf = _peaks_gpu(w, mask)
    mask_ref = _check_peaks_against_oracle(oracle, w, pf)
    assert np.array_equal(mask.cpu().numpy(), mask_ref)
    npk = len(pf["i"])
    assert npk > 5000
    Rf, ir2 = _peak_inputs(w, pf, dev)
    hf = _host_fields(w["fields"])
    fcrit = api.fsc_of_z(websky_ct, 0.0)
    pi, pj, pk = (pf[k].cpu().numpy() for k in ("i", "j", "k"))
    Rf_h, ir2_h = Rf.cpu().numpy(), ir2.cpu().numpy()
    t = oracle.shell_tables(nhunt)
    stab = api.build_shell_tables(nhunt)
    kw = dict(alatt=alatt, ZZon=1.0, fcrit=fcrit, nbuff=nbuff)
    for variant, early in (("A", False), ("B", True), ("C", True)):
        mask.zero_()
        r = api.shell_analyse(w["fields"], pf["i"], pf["j"], pf["k"], Rf, ir2, stab, websky_ct,
                              variant=variant, early_exit=early, mask=mask, **kw)
        mref = np.zeros((n, n, n), np.int8, order="F")
        ref = oracle.analyse_peaks(hf, pi, pj, pk, Rf_h, ir2_h, t, websky_ct, variant=variant,
                                   early_exit=early, mask=mref, nthreads=NTHREADS, **kw)
        g = to_host(r)
        compare(g, ref, f"tile128 variant {variant} early_exit={early}", rtol=RTOL)
        assert np.array_equal(mask.cpu().numpy(), mref), variant
        if early:
            assert np.array_equal(g["nshells_used"], ref["nshells_used"])
            assert np.array_equal(g["crossing_shell"], ref["crossing_shell"])
    # compensated FP32 shell sums against the FP64 oracle (variant C): the required bars
    r = api.shell_analyse(w["fields"], pf["i"], pf["j"], pf["k"], Rf, ir2, stab, websky_ct,
                          variant="C", precision="fp32_kahan", **kw)
    g = to_host(r)
    flips = int(((g["RTHL"] > 0) != (ref["RTHL"] > 0)).sum())
    worst = bar_errors(g, ref)
    print(f"tile128 fp32_kahan vs FP64 oracle: halo flag flips {flips}, worst "
          f"{worst}; every field, plain relative: "
          f"{ {k: f'{v:.1e}' for k, v in rel_errors(g, ref).items()} }", flush=True)
    _assert_reduced_precision(worst, flips, int((ref["RTHL"] > 0).sum()))
    del w, r
    _release()


@pytest.mark.timeout(900)
def test_tile512_full_peak_list_and_every_candidate(oracle, websky_ct):
    dev = torch.device("cuda", 0)
    w = synth.workload_tile("tile512", dev)
    n, nbuff, nhunt, alatt = w["n"], w["nbuff"], w["nhunt"], w["alatt"]
    mask = api.fortran_empty((n, n, n), torch.int8, dev)
    pf = _peaks_gpu(w, mask)
    mask_ref = _check_peaks_against_oracle(oracle, w, pf)
    assert np.array_equal(mask.cpu().numpy(), mask_ref)
    w["smoothed"].clear()
    torch.cuda.empty_cache()
    Rf, ir2 = _peak_inputs(w, pf, dev)
    fcrit = api.fsc_of_z(websky_ct, 0.0)
    kw = dict(alatt=alatt, ZZon=1.0, fcrit=fcrit, nbuff=nbuff)
    mask.zero_()
    r = api.shell_analyse(w["fields"], pf["i"], pf["j"], pf["k"], Rf, ir2,
                          api.build_shell_tables(nhunt), websky_ct, variant="C",
                          early_exit=True, mask=mask, **kw)
    g = to_host(r)
    hf = _host_fields(w["fields"])
    mref = np.zeros((n, n, n), np.int8, order="F")
    ref = oracle.analyse_peaks(hf, pf["i"].cpu().numpy(), pf["j"].cpu().numpy(),
                               pf["k"].cpu().numpy(), Rf.cpu().numpy(), ir2.cpu().numpy(),
                               oracle.shell_tables(nhunt), websky_ct, variant="C",
                               early_exit=True, mask=mref, nthreads=NTHREADS, **kw)
    compare(g, ref, "tile512 variant C", rtol=RTOL)
    assert np.array_equal(g["nshells_used"], ref["nshells_used"])
    assert np.array_equal(mask.cpu().numpy(), mref)
    del w, r, hf
    _release()


@pytest.mark.timeout(1200)
def test_tile1024_full_peak_list_and_sampled_candidates(oracle, websky_ct):
    dev = torch.device("cuda", 0)
    w = synth.workload_tile("tile1024", dev)
    n, nbuff, nhunt, alatt = w["n"], w["nbuff"], w["nhunt"], w["alatt"]
    print("tile1024 inputs synthesised", flush=True)
    mask = api.fortran_empty((n, n, n), torch.int8, dev)
    pf = _peaks_gpu(w, mask)
    mask_ref = _check_peaks_against_oracle(oracle, w, pf, streamed=True)
    assert np.array_equal(mask.cpu().numpy(), mask_ref)
    del mask_ref
    w["smoothed"].clear()
    torch.cuda.empty_cache()
    npk = len(pf["i"])
    Rf, ir2 = _peak_inputs(w, pf, dev)
    fcrit = api.fsc_of_z(websky_ct, 0.0)
    kw = dict(alatt=alatt, ZZon=1.0, fcrit=fcrit, nbuff=nbuff)
    stab = api.build_shell_tables(nhunt)
    mask.zero_()
    r = api.shell_analyse(w["fields"], pf["i"], pf["j"], pf["k"], Rf, ir2, stab, websky_ct,
                          variant="C", early_exit=True, mask=mask, **kw)
    sel = np.sort(np.random.default_rng(0).choice(npk, size=min(npk, 250_000), replace=False))
    g = to_host(r, sel)
    hf = _host_fields(w["fields"])
    print(f"tile1024: {npk} candidates analysed on the GPU, fields on the host", flush=True)
    ref = oracle.analyse_peaks(hf, pf["i"].cpu().numpy()[sel], pf["j"].cpu().numpy()[sel],
                               pf["k"].cpu().numpy()[sel], Rf.cpu().numpy()[sel],
                               ir2.cpu().numpy()[sel], oracle.shell_tables(nhunt), websky_ct,
                               variant="C", early_exit=True, nthreads=NTHREADS, **kw)
    del hf
    compare(g, ref, f"tile1024 variant C, {len(sel)} of {npk} candidates", rtol=RTOL)
    assert np.array_equal(g["nshells_used"], ref["nshells_used"])
    # compensated FP32 sums on every candidate against the FP64 sums (equal to the oracle)
    full = to_host(r)
    nh = int((full["RTHL"] > 0).sum())
    r32 = api.shell_analyse(w["fields"], pf["i"], pf["j"], pf["k"], Rf, ir2, stab, websky_ct,
                            variant="C", early_exit=True, precision="fp32_kahan", **kw)
    k32 = to_host(r32)
    flips = int(((k32["RTHL"] > 0) != (full["RTHL"] > 0)).sum())
    worst = bar_errors(k32, full)
    print(f"tile1024 fp32_kahan vs FP64, {npk} candidates, {nh} halos: halo flag flips "
          f"{flips}, halo count {int((k32['RTHL'] > 0).sum())}, worst {worst}", flush=True)
    _assert_reduced_precision(worst, flips, nh)
    del w, r, r32, full, k32
    _release()


@pytest.mark.timeout(900)
def test_batched64_plan_against_oracle_per_tile(oracle, websky_ct):
    dev = torch.device("cuda", 0)
    W = synth.WORKLOADS["batched64"]
    specs = synth.batched_specs("batched64")
    radii = synth.filter_radii(W["rmin"], W["rmax"], W["nfilt"])
    alatt, nhunt = W["alatt"], W["nhunt"]
    fcrit = api.fsc_of_z(websky_ct, 0.0)
    Rf_tab = torch.tensor(radii, dtype=torch.float64, device=dev)
    ir2_tab = torch.tensor([api.ir2min_of(R, alatt) for R in radii], dtype=torch.int32,
                           device=dev)
    tiles = []
    for n, nb, seed in specs:
        box = alatt * n
        fields, dk = synth.fields_torch(n, box, seed, dev)
        sm = [synth.smooth_tophat_torch(dk, n, box, R) for R in radii]
        del dk
        mask = api.fortran_empty((n, n, n), torch.int8, dev)
        mask.zero_()
        pf = api.peak_find(sm, [1.5] * len(radii), nb, mask, capacity=(n - 2 * nb) ** 3)
        del sm
        f = pf["filt"].long()
        tiles.append(dict(fields=fields, pi=pf["i"], pj=pf["j"], pk=pf["k"], Rf=Rf_tab[f],
                          ir2min=ir2_tab[f], nbuff=nb))
        if len(tiles) % 16 == 0:
            print(f"  batched64: {len(tiles)} tiles synthesised", flush=True)
    counts = [len(t["pi"]) for t in tiles]
    print(f"batched64: {sum(counts)} candidates, per tile {min(counts)}..{max(counts)}",
          flush=True)
    assert max(counts) > 20 * max(1, min(counts))
    stab = api.build_shell_tables(nhunt)
    nbuff_all = min(nb for _, nb, _ in specs)
    kw = dict(alatt=alatt, ZZon=1.0, fcrit=fcrit, variant="C")
    plan = api.ShellPlan(tiles, stab, websky_ct, nbuff=nbuff_all, **kw)
    for res in plan.results:
        res["RTHL"].fill_(-9.0)
    plan.launch()
    torch.cuda.synchronize()
    t_or = oracle.shell_tables(nhunt)
    rng = np.random.default_rng(1)
    checked = 0
    for q, (t, res) in enumerate(zip(tiles, plan.results)):
        npk = len(t["pi"])
        if npk == 0:
            continue
        sel = np.sort(rng.choice(npk, size=min(npk, 1500), replace=False))
        hf = _host_fields(t["fields"])
        ref = oracle.analyse_peaks(hf, t["pi"].cpu().numpy()[sel], t["pj"].cpu().numpy()[sel],
                                   t["pk"].cpu().numpy()[sel], t["Rf"].cpu().numpy()[sel],
                                   t["ir2min"].cpu().numpy()[sel], t_or, websky_ct,
                                   early_exit=True, nbuff=nbuff_all, nthreads=NTHREADS, **kw)
        g = to_host(res, sel)
        assert np.array_equal(g["RTHL"] > 0, ref["RTHL"] > 0), q
        worst = rel_errors(g, ref)
        assert max(worst.values()) <= RTOL, (q, worst)
        assert np.array_equal(g["nshells_used"], ref["nshells_used"]), q
        checked += len(sel)
        if q % 8 == 7:
            print(f"  batched64: {q + 1} tiles checked", flush=True)
    print(f"batched64 plan replay: {checked} sampled candidates of 64 tiles equal the oracle",
          flush=True)
    plan.close()
    del tiles, plan
    _release()

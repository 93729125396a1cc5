"""Benchmark of the halo-finding hot path on one MI355X.

Workload (default, the headline configuration): a 1024^3 Float32 tile on a 2048 Mpc/h box
(alatt 2 Mpc/h), Threefry seed 2024, BBKS P(k) with sigma8 0.81, 1LPT + 2LPT displacements,
24 top-hat filters log-spaced 1.5..34 Mpc/h processed largest first, peak threshold 1.5,
64-cell buffer, collapse at z = 0 with HomelTab_websky.dat, shells to 60 Mpc/h (hunt radius
30 cells, 753 distinct-r^2 shells).  Inputs are synthesised on the GPU before timing and
stay resident in HBM.

One step = one pass of the hot path: mask reset, multi-filter peak finding over the 24
smoothed fields (hnt_peak_find), mask reset, radial-shell collapse of every candidate
(hnt_shell_analyse: validation, record repack, peak sort, gather + radial logic).
value = candidates / step time.  The CPU baseline is the oracle's FP64 restatement of the
reference CPU path (RadialShell.analyse_peak, no early exit) on a bounded random sample of
the same candidates, OpenMP over peaks on the host cores.
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "huntian.jl_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
METRIC = "peaks resolved/s at 1024^3 tile: learned-bracket vs plain shell scan, same GPU"


def _configs():
    from huntian_hip import synth

    return synth.WORKLOADS


CONFIGS = _configs()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="tile1024", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle baseline sample (0 = skip)")
    ap.add_argument("--variant", default="C", choices=["A", "B", "C"])
    ap.add_argument("--order", default=None, choices=["crossing", "plain", "learned"],
                    help="wavefront schedule of the shell analysis (never changes a result); "
                         "default: the autotuner's choice")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (library defaults, schedule 'crossing')")
    ap.add_argument("--gang", type=int, default=None, choices=[1, 16],
                    help="lanes per peak in the shell gather (library default if omitted)")
    ap.add_argument("--brick-log2", type=int, default=None,
                    help="peak-order brick size 2^v cells (library default if omitted)")
    ap.add_argument("--no-ablation", action="store_true",
                    help="skip timing the shell analysis under the other schedules")
    ap.add_argument("--no-filter-bank", action="store_true",
                    help="skip timing tile peak finding from the raw tile (smoothing included)")
    ap.add_argument("--no-field-synthesis", action="store_true",
                    help="skip timing the tile field synthesis")
    ap.add_argument("--no-subworkloads", action="store_true",
                    help="skip the 512^3 tile, the 64 batched tiles and the learned-bracket "
                         "tile reported beside the headline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a read-only reduction of a 4 GiB array and a plain 4 GiB "
                         "device copy (3 times each), the reference kernels for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE and TCC_EA0_RDREQ_* counts")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): the library has "
              "no multi-GPU path; the per-tile work would shard by tiles, one process per GPU",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from huntian_hip import api, synth
    from huntian_hip.homeltab import WEBSKY_TABLE, read_homeltab

    if args.copy_calibration:
        a = torch.ones(1 << 30, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        tot = 0.0
        for _ in range(3):
            tot += float(a.sum())  # read-only: every byte of a once
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_bytes_read_per_pass": a.numel() * 4,
                          "sum": tot}), flush=True)
        return

    cfg = CONFIGS[args.config]
    if args.config == "batched64":
        return run_batched(args, cfg)
    n, box, nbuff, nhunt = cfg["n"], cfg["box"], cfg["nbuff"], cfg["nhunt"]
    alatt = box / n
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(0)

    # ---------------- inputs (untimed)
    t0 = time.time()
    fields, dk = synth.fields_torch(n, box, cfg["seed"], dev)
    radii = synth.filter_radii(cfg["rmin"], cfg["rmax"], cfg["nfilt"])
    smoothed = [synth.smooth_tophat_torch(dk, n, box, R) for R in radii]
    del dk
    torch.cuda.synchronize()
    log(f"synthesised {n}^3 fields + {len(radii)} smoothed fields in {time.time() - t0:.1f} s; "
        f"GPU memory {torch.cuda.memory_allocated() / 2**30:.1f} GiB")
    ct = read_homeltab(WEBSKY_TABLE)
    fcrit = api.fsc_of_z(ct, 0.0)
    fc_peak = [1.5] * len(radii)
    stab = api.build_shell_tables(nhunt)
    Rf_tab = torch.tensor(radii, dtype=torch.float64, device=dev)
    ir2_tab = torch.tensor([api.ir2min_of(R, alatt) for R in radii], dtype=torch.int32,
                           device=dev)
    mask = api.fortran_empty((n, n, n), torch.int8, dev)
    cap = 48_000_000
    ctx = api.context(0)
    knobs = {k: v for k, v in (("gang", args.gang), ("brick_log2", args.brick_log2))
             if v is not None}
    if knobs:
        ctx.set_tuning(**knobs)
    from huntian_hip import bracket

    net = None
    if os.path.exists(bracket.DEFAULT_WEIGHTS):
        net = bracket.BracketNet.load(bracket.DEFAULT_WEIGHTS, device=dev)
    elif args.order == "learned":
        raise SystemExit(f"--order learned needs trained weights at {bracket.DEFAULT_WEIGHTS}")

    results = None
    tuning = None
    if not args.no_autotune and args.order is None and not knobs:
        from huntian_hip import autotune

        mask.zero_()
        pf0 = api.peak_find(smoothed, fc_peak, nbuff, mask, capacity=cap)
        n0 = len(pf0["i"])
        res0 = api.alloc_results(n0, dev)
        Rf0 = Rf_tab[pf0["filt"].long()]
        ir0 = ir2_tab[pf0["filt"].long()]

        def run(order):
            mask.zero_()
            api.shell_analyse(fields, pf0["i"], pf0["j"], pf0["k"], Rf0, ir0, stab, ct,
                              alatt=alatt, ZZon=1.0, fcrit=fcrit, nbuff=nbuff,
                              variant=args.variant, early_exit=True, mask=mask, results=res0,
                              order=order, bracket=net)

        t1 = time.time()
        tuning = autotune.tune(run, ctx, has_bracket=net is not None, log=log)
        log(f"autotune: best {tuning['best']} after {time.time() - t1:.1f} s")
        del res0, pf0
        args.order = tuning["best"]["order"]
    if args.order is None:
        args.order = "crossing"

    def step():
        nonlocal results
        mask.zero_()
        pf = api.peak_find(smoothed, fc_peak, nbuff, mask, capacity=cap)
        npk = len(pf["i"])
        if results is None or len(results["RTHL"]) < npk:
            results = api.alloc_results(int(npk * 1.25) + 1024, dev)
        Rf = Rf_tab[pf["filt"].long()]
        ir2 = ir2_tab[pf["filt"].long()]
        mask.zero_()  # the reference tile driver clears the mask between peak finding and shells
        r = api.shell_analyse(fields, pf["i"], pf["j"], pf["k"], Rf, ir2, stab, ct, alatt=alatt,
                              ZZon=1.0, fcrit=fcrit, nbuff=nbuff, variant=args.variant,
                              early_exit=True, mask=mask, results=results, order=args.order,
                              bracket=net)
        return pf, r, npk

    for w in range(args.warmup):
        t1 = time.time()
        pf, r, npk = step()
        torch.cuda.synchronize()
        log(f"warmup {w}: {npk} candidates, {time.time() - t1:.2f} s")

    ctx.set_profiling(True)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for s in range(args.steps):
        pf, r, npk = step()
    ev1.record()
    torch.cuda.synchronize()
    total_ms = ev0.elapsed_time(ev1)
    prof = ctx.profile()
    ctx.set_profiling(False)
    ms_step = total_ms / args.steps
    log(f"timed {args.steps} steps: {ms_step:.2f} ms/step, {npk} candidates; stages {prof}")

    # ---------------- schedule ablation: shell analysis alone under each schedule, every
    # result field compared with the step's (the bench's default schedule)
    ablation = {}
    if not args.no_ablation:
        ablation = schedule_ablation(args, ctx, fields, pf, r, npk, Rf_tab, ir2_tab, stab, ct,
                                     alatt, fcrit, nbuff, mask, net, dev)

    # ---------------- halo records of the step's results (hnt_pack_records: 33-float
    # ExtHaloRecord with the 2LPT terms, z = 0), timed once after a warm call
    from huntian_hip import cosmology

    cos = cosmology.CosmologyParams(0.31, 0.049, 0.69, 0.68)
    growth = cosmology.dlinear_tables(cos)
    Rf_pk = Rf_tab[pf["filt"].long()]
    rkw = dict(alatt=alatt, n=n, ZZon=1.0, ilpt=2, Om=cos.Om, OL=cos.OL, ioutshear=1)
    recs = api.pack_records(pf["i"], pf["j"], pf["k"], Rf_pk, r, pf["delta"], growth=growth,
                            **rkw)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    recs = api.pack_records(pf["i"], pf["j"], pf["k"], Rf_pk, r, pf["delta"], growth=growth,
                            **rkw)
    e1.record()
    torch.cuda.synchronize()
    records = dict(ms=round(e0.elapsed_time(e1), 3), halos=int(recs.shape[0]),
                   floats_per_record=int(recs.shape[1]),
                   note="hnt_pack_records over the step's results: flags, scan, scatter")
    log(f"records: {records}")

    # ---------------- catalog merge of those records (hnt_merge_catalog): Lagrangian
    # exclusion, then with volume reduction as production runs use it; warm call, then timed
    merge = {}
    for reduce in (False, True):
        api.merge_catalog(recs, volume_reduction=reduce)
        torch.cuda.synchronize()
        e0.record()
        merged = api.merge_catalog(recs, volume_reduction=reduce)
        e1.record()
        torch.cuda.synchronize()
        merge["reduction" if reduce else "exclusion"] = dict(
            ms=round(e0.elapsed_time(e1), 3), halos_in=int(recs.shape[0]),
            halos_out=int(merged.shape[0]))
    log(f"merge: {merge}")
    del recs

    # ---------------- Eulerian positions and velocities of the merged catalog
    # (hnt_finalize_eulerian, lightcone mode: each halo's z from its distance to an observer
    # 2000 Mpc/h away), in place; warm call, then timed
    fkw = dict(Om=cos.Om, OL=cos.OL, ievol=1, chi2z=cosmology.build_chi_to_z(cos, z_max=10.0))
    obs = (0.0, 0.0, -2000.0)
    fin_in = merged.clone()
    api.finalize_eulerian(fin_in, growth, obs, out=fin_in, **fkw)
    torch.cuda.synchronize()
    e0.record()
    api.finalize_eulerian(merged, growth, obs, out=merged, **fkw)
    e1.record()
    torch.cuda.synchronize()
    finalize = dict(ms=round(e0.elapsed_time(e1), 3), halos=int(merged.shape[0]),
                    note="hnt_finalize_eulerian in place, lightcone mode")
    log(f"finalize: {finalize}")
    del merged, fin_in

    # ---------------- tile peak finding from the raw tile: filter-bank smoothing (FFT
    # rfft once, window + irfft per filter) fused with the peak test (hnt_peak_find_tile);
    # reported beside the step, whose peak test reads the pre-smoothed fields
    tile_pf = None
    if not args.no_filter_bank:
        smoothed.clear()  # the shell inputs no longer need them
        torch.cuda.empty_cache()
        delta = fields["delta"]
        mask.zero_()
        api.peak_find_tile(delta, radii, fc_peak, nbuff, mask, alatt=alatt, capacity=cap)
        torch.cuda.synchronize()
        ctx.set_profiling(True)
        mask.zero_()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        ft = api.peak_find_tile(delta, radii, fc_peak, nbuff, mask, alatt=alatt, capacity=cap)
        e1.record()
        torch.cuda.synchronize()
        fprof = ctx.profile()
        ctx.set_profiling(False)
        tile_pf = dict(ms=round(e0.elapsed_time(e1), 3), candidates=len(ft["i"]),
                       stage_ms={k: round(v[0], 3) for k, v in fprof.items() if v[1]},
                       filters=len(radii),
                       note="one rfft, then per filter a window table and an irfft of the "
                            "spectrum times the table to the inner cube and its halo (library "
                            "FFT, Float32) and the peak test; candidates differ from the "
                            "step's (torch-smoothed input) only by FFT rounding")
        del ft
        log(f"tile peak finding: {tile_pf}")

    # ---------------- results summary and algorithmic bytes
    RTHL = r["RTHL"][:npk].cpu().numpy()
    used = r["nshells_used"][:npk].cpu().numpy().astype(np.int64)
    cross = r["crossing_shell"][:npk].cpu().numpy().astype(np.int64)
    nhalo = int((RTHL > 0).sum())
    cum_cells = np.concatenate([[0], np.cumsum(stab.shell_count)])  # cells in shells 1..s
    cells = cum_cells[used]
    # algorithmic bytes follow the reference's gather extent (ext/CUDAExt.jl:896-917): every
    # shell up to and including the first with rad >= rad[crossing] + 3, all shells if there
    # is no crossing; 32 B per cell (delta, delta^T, psi1 xyz, psi2 xyz; no lapd here)
    rad = np.sqrt(np.asarray(stab.shell_r2, np.float64))  # rad[s-1] of 1-based shell s
    ns = len(rad)
    ref_stop = np.where(cross > 0,
                        np.minimum(np.searchsorted(rad, rad[np.maximum(cross, 1) - 1] + 3.0,
                                                   side="left") + 1, ns),
                        ns)
    ref_cells = cum_cells[ref_stop]
    rec_bytes = 32
    sa_bytes = float(ref_cells.sum()) * rec_bytes
    m = n - 2 * nbuff
    pf_bytes = float(len(radii)) * m ** 3 * 4 + 2.0 * m ** 3  # fields + mask read + claims
    stage_ms = {k: v[0] / max(v[1], 1) for k, v in prof.items()}
    t_pf = stage_ms["pf_mark"] + stage_ms["pf_compact"]
    # shell analysis = validation, repack, first sort, crossing pass and work-ordered sort,
    # gather + radial logic
    t_sa = (stage_ms["sa_prep"] + stage_ms.get("sa_cross", 0.0) + stage_ms["sa_kernel"] +
            stage_ms.get("sa_mask", 0.0))
    kern = {
        "shell_analyse_kernel": dict(ms=stage_ms["sa_kernel"], bytes=sa_bytes),
        "pf_mark": dict(ms=stage_ms["pf_mark"], bytes=float(len(radii)) * m ** 3 * 4 + m ** 3),
    }
    for k, v in kern.items():
        v["GBs"] = v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0.0
    dom = max(kern, key=lambda k: kern[k]["ms"])
    # HBM bytes per launch of the dominant kernel from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs (scaled by the copy-kernel calibration), when they cover this workload
    traffic, traffic_note = None, None
    tf = os.path.join(REPO, "profiles", "hbm_traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            ent = json.load(fh).get(args.config, {}).get(dom)
        if ent:
            traffic = float(ent["traffic_bytes"])
            traffic_note = (f"{ent['source']}: {traffic / 1e9:.1f} GB per launch in "
                            f"{ent['kernel_ms']} ms ({traffic / ent['kernel_ms'] / 1e9:.2f} TB/s "
                            f"measured)")
    roof = dict(bound="hbm", achieved=round(kern[dom]["GBs"], 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(kern[dom]["GBs"] / HBM_PEAK_GBS, 4), traffic=traffic,
                traffic_note=traffic_note, kernel=dom, kernel_ms=round(kern[dom]["ms"], 3),
                algorithmic_bytes=kern[dom]["bytes"])

    # ---------------- CPU baseline: oracle, variant A (the reference CPU path), sample; then
    # the oracle in the GPU's variant on the same sample, checked against the GPU's results
    cpu, oracle_check = None, None
    if args.cpu_seconds > 0:
        cpu, oracle_check = cpu_baseline_and_check(args, fields, pf, r, npk, radii, alatt,
                                                   nhunt, ct, fcrit, nbuff)

    # ---------------- tile field synthesis of one tile of this size (the step before the
    # path), after the step's buffers are freed
    synth_t = None
    if not args.no_field_synthesis:
        fields.clear()
        smoothed.clear()
        del mask, r, pf
        results = None
        ctx.release_scratch()
        torch.cuda.empty_cache()
        synth_t = time_field_synthesis(n, box, nbuff, dev)
        log(f"field synthesis: {synth_t}")

    # ---------------- the other workloads, each on its own after the headline's buffers are
    # released: the 512^3 tile, the 64 uneven tiles (per-tile calls, one launch, hipGraph
    # plan), and the learned bracket on its evaluation tile (128^3, seed 99)
    subs = None
    if not args.no_subworkloads and args.config == "tile1024":
        fields.clear()
        smoothed.clear()
        mask = r = pf = results = None
        ctx.release_scratch()
        torch.cuda.empty_cache()
        subs = {"tile512": tile_result(args, "tile512"),
                "batched64": batched_result(args, "batched64"),
                "tile128_seed99_learned_bracket": tile_result(args, "tile128", seed=99,
                                                              net=net, ablate=True)}

    out = {
        "metric": METRIC,
        "value": round(npk / (ms_step * 1e-3), 1),
        "unit": "peaks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (sums of f32 inputs and f32 products, as the reference CPU path)",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {n}^3 Float32 tile, box {box:g} Mpc/h, seed "
                               f"{cfg['seed']}, {len(radii)} top-hat filters "
                               f"{cfg['rmin']}-{cfg['rmax']} Mpc/h, threshold 1.5, nbuff "
                               f"{nbuff}, hunt radius {nhunt} cells ({stab.nshells} shells), "
                               "z = 0, psi1 + psi2, no lapd",
                   "method": f"shell scan with exact early exit, schedule '{args.order}'",
                   "schedule_ablation_shell_ms": ablation,
                   "autotune": tuning,
                   "variant": args.variant, "candidates": npk, "halos": nhalo,
                   "peak_find_ms": round(t_pf, 3), "shell_analysis_ms": round(t_sa, 3),
                   "tile_peak_find": tile_pf,
                   "records": records,
                   "merge": merge,
                   "finalize": finalize,
                   "field_synthesis": synth_t,
                   "sub_workloads": subs,
                   "oracle_check": oracle_check,
                   "shell_only_peaks_per_s": round(npk / (t_sa * 1e-3), 1) if t_sa > 0 else None,
                   "mean_cells_gathered": round(float(cells.mean()), 1) if npk else 0.0,
                   "mean_cells_reference_extent": round(float(ref_cells.mean()), 1)
                   if npk else 0.0,
                   "stage_ms": {k: round(v, 3) for k, v in stage_ms.items()},
                   "kernels_GBs": {k: round(v["GBs"], 1) for k, v in kern.items()},
                   "parallelism": "1 GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


RESULT_FIELDS = ("RTHL", "Fbarx", "e_v", "p_v", "strain", "eigs", "Srb", "Sbar", "Sbar2",
                 "gradpk", "gradpkf", "gradpkrf", "d2F", "zvir_half", "nshells_used",
                 "crossing_shell")


def same_results(a, b, npk):
    """Every result field of two shell analyses equal bit for bit (NaN where both are NaN)."""
    import torch

    for k in RESULT_FIELDS:
        x, y = a[k][:npk], b[k][:npk]
        if x.dtype.is_floating_point:
            x, y = torch.nan_to_num(x, nan=7.0), torch.nan_to_num(y, nan=7.0)
        if not torch.equal(x, y):
            return False
    return True


def schedule_ablation(args, ctx, fields, pf, r, npk, Rf_tab, ir2_tab, stab, ct, alatt, fcrit,
                      nbuff, mask, net, dev):
    """Shell analysis alone (validation, repack, schedule, gather + radial logic) under each
    schedule, on the step's candidates: crossing (delta-only pass for the exact extents),
    plain (ir2min, Morton), learned (the bracket network; its verification rates), perfect
    (the exact extents of the step given as the bracket: the best any bracket can do here,
    no pass of its own).  Each timed on a warm second call."""
    import torch

    from huntian_hip import api

    Rf = Rf_tab[pf["filt"].long()]
    ir2 = ir2_tab[pf["filt"].long()]
    extent = r["nshells_used"][:npk].clone()
    out = {}
    for o in ("crossing", "plain", "learned", "perfect"):
        if o == "learned" and net is None:
            continue
        res2 = api.alloc_results(npk, dev)
        kw = dict(order="given", extent=extent) if o == "perfect" else dict(order=o, bracket=net)
        for rep in range(2):  # first call warms the schedule's buffers
            mask.zero_()
            ctx.set_profiling(True)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            r2 = api.shell_analyse(fields, pf["i"], pf["j"], pf["k"], Rf, ir2, stab, ct,
                                   alatt=alatt, ZZon=1.0, fcrit=fcrit, nbuff=nbuff,
                                   variant=args.variant, early_exit=True, mask=mask,
                                   results=res2, **kw)
            e1.record()
            torch.cuda.synchronize()
            prof = ctx.profile()
            ctx.set_profiling(False)
        ent = dict(ms=round(e0.elapsed_time(e1), 3), same_results=same_results(r2, r, npk),
                   stage_ms={k: round(v[0], 3) for k, v in prof.items() if v[1]})
        if o == "learned":
            fl = r2["bracket"][:npk]
            halo = r["RTHL"][:npk] > 0
            ent["bracket"] = dict(
                gather_verified=round(float(((fl & 1) > 0).float().mean()), 4),
                walk_verified_of_halos=round(float(((fl & 2) > 0)[halo].float().mean()), 4),
                both_verified_of_halos=round(float(((fl & 3) == 3)[halo].float().mean()), 4),
                collapse_logit_right=round(float(((fl & 4) > 0).float().mean()), 4))
        out[o] = ent
        del r2, res2
        log(f"schedule {o}: {ent}")
    if "learned" in out and "plain" in out:
        out["learned_over_plain_speedup"] = round(out["plain"]["ms"] / out["learned"]["ms"], 3)
    if "perfect" in out:
        out["perfect_over_plain_speedup"] = round(out["plain"]["ms"] / out["perfect"]["ms"], 3)
    out["note"] = ("shell analysis alone on the step's candidates, same results required; "
                   "plain = exact early exit in (ir2min, Morton) order; perfect = the step's "
                   "own extents given as the bracket")
    return out


def host_info():
    """Cores this process may use, nproc, the machine's CPU count and model."""
    import subprocess

    info = dict(affinity=len(os.sched_getaffinity(0)), cpu_count=os.cpu_count())
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True,
                                           timeout=10).stdout.strip())
    except Exception:
        info["nproc"] = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return info


def cpu_baseline_and_check(args, fields, pf, r, npk, radii, alatt, nhunt, ct, fcrit, nbuff):
    """(cpu_baseline, oracle_check).  The baseline: the oracle's FP64 restatement of
    RadialShell.analyse_peak (variant A, full hunt radius, no early exit), OpenMP over peaks
    on every core this process may use, on a uniform random sample of the step's candidates
    sized to about --cpu-seconds.  The check: the oracle in the GPU's variant with early exit
    on the same sample against the GPU's results of the step (halo flags, worst relative
    error per field)."""
    import numpy as np

    from huntian_hip import api
    from oracle import oracle as O

    host = host_info()
    cores = host["affinity"]
    hf = {k: v.cpu().numpy() for k, v in fields.items()}
    pi = pf["i"].cpu().numpy()
    pj = pf["j"].cpu().numpy()
    pk_ = pf["k"].cpu().numpy()
    filt = pf["filt"].cpu().numpy()
    Rf_h = np.array(radii)[filt]
    ir2_h = np.array([api.ir2min_of(R, alatt) for R in radii], np.int32)[filt]
    t_oracle = O.shell_tables(nhunt)
    rng = np.random.default_rng(0)
    order = rng.permutation(npk)
    done, elapsed, nsamp = 0, 0.0, max(cores * 4, 64)
    while elapsed < args.cpu_seconds and done < npk:
        sel = order[done:done + nsamp]
        t1 = time.time()
        O.analyse_peaks(hf, pi[sel], pj[sel], pk_[sel], Rf_h[sel], ir2_h[sel], t_oracle, ct,
                        alatt=alatt, ZZon=1.0, fcrit=fcrit, nbuff=nbuff, variant="A",
                        early_exit=False, nthreads=cores)
        dt = time.time() - t1
        elapsed += dt
        done += len(sel)
        rate = len(sel) / max(dt, 1e-9)
        nsamp = int(min(npk - done, max(16, rate * (args.cpu_seconds - elapsed))))
        if nsamp <= 0:
            break
    cpu = dict(value=round(done / elapsed, 2), unit="peaks/s", cores=cores, kind="port",
               host=host,
               sample=f"{done} of {npk} candidates drawn uniformly (seed 0); FP64 oracle of "
                      "RadialShell.analyse_peak (variant A, full hunt radius, no early exit), "
                      f"OpenMP over peaks on {cores} threads; {elapsed:.1f} s")
    log(f"cpu baseline: {cpu}")
    # the check, on the same sample and a larger one (early exit makes it cheap)
    nchk = min(npk, max(done, 200_000))
    sel = np.sort(order[:nchk])
    t1 = time.time()
    ref = O.analyse_peaks(hf, pi[sel], pj[sel], pk_[sel], Rf_h[sel], ir2_h[sel], t_oracle, ct,
                          alatt=alatt, ZZon=1.0, fcrit=fcrit, nbuff=nbuff,
                          variant=args.variant, early_exit=True, nthreads=cores)
    t_chk = time.time() - t1
    import torch

    sel_t = torch.as_tensor(sel, device=r["RTHL"].device)
    g = {k: v[sel_t].cpu().numpy() for k, v in r.items() if k in ref}
    g["strain"] = np.swapaxes(g["strain"], 1, 2)
    halo_g, halo_r = g["RTHL"] > 0, ref["RTHL"] > 0
    worst = {}
    both = halo_g & halo_r
    for k in ("RTHL", "Fbarx", "e_v", "p_v", "strain", "eigs", "Srb", "Sbar", "Sbar2", "gradpk",
              "gradpkf", "gradpkrf", "d2F", "zvir_half"):
        a = g[k][both].reshape(int(both.sum()), -1)
        b = ref[k][both].reshape(int(both.sum()), -1)
        fin = np.isfinite(a) & np.isfinite(b)
        scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-30)
        err = np.where(fin, np.abs(a - b) / np.maximum(np.abs(b), 1e-9 * scale), 0.0)
        worst[k] = float(err.max()) if err.size else 0.0
    check = dict(candidates=int(nchk), halos_gpu=int(halo_g.sum()), halos_oracle=int(halo_r.sum()),
                 halo_flags_equal=bool(np.array_equal(halo_g, halo_r)),
                 nshells_used_equal=bool(np.array_equal(g["nshells_used"], ref["nshells_used"])),
                 worst_rel_error={k: float(f"{v:.3g}") for k, v in worst.items()},
                 oracle=f"FP64 oracle, variant {args.variant}, early exit, {nchk} sampled "
                        f"candidates (seed 0), {t_chk:.1f} s on {cores} threads")
    log(f"oracle check: {check}")
    return cpu, check


def time_field_synthesis(n, box, nbuff, dev):
    """One tile's fields as the tile driver builds them (src/MultiResolution.jl:881-1010):
    the tile's Hoffman-Ribak residual noise (Threefry stream of a N = 4096 global grid less
    the block means of its M = 256 coarse grid, tile (2, 2, 2)), delta_self and psi1_self from
    one zero-padded (2n)^3 isolated convolution of it, the coarse delta and psi1 interpolated
    onto the tile, then the periodic 2LPT and Laplacian of the combined delta.  Each piece
    timed once after a warm call; the coarse grid's downsample (once per run in the driver)
    timed once; the coarse fields interpolated are synthetic."""
    import torch
    from huntian_hip import api, synth

    pk = synth.PowerSpectrum().file
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    M, N, nsub, seed = 256, 4096, n - 2 * nbuff, 7
    coarse = [torch.randn((M, M, M), generator=g, device=dev) for _ in range(4)]
    kernels = [(0, 0, 0), (1, 1, 0), (1, 2, 0), (1, 3, 0)]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn, warm=True):
        if warm:
            fn()  # warm (plans, scratch)
        torch.cuda.synchronize()
        ev[0].record()
        out = fn()
        ev[1].record()
        torch.cuda.synchronize()
        return out, round(ev[0].elapsed_time(ev[1]), 3)

    coarse_noise, t_down = timed(lambda: api.downsample_noise(N, M, seed, return_device=True),
                                 warm=False)
    residual, t_res = timed(lambda: api.generate_extended_residual(
        2, 2, 2, nsub, n, N, seed, coarse_noise, M, 0, return_device=True))
    del coarse_noise
    selfs, t_conv = timed(lambda: api.isolated_convolve_gpu_multi(
        residual, pk, box, n, kernels=kernels, return_device=True))
    del residual

    def interp_all():
        return [api.interpolate_to_tile_gpu(c, 2, 2, 2, nsub, n, N, M, return_device=True)
                for c in coarse]
    longs, t_interp = timed(interp_all)
    delta_tile = selfs[0] + longs[0]
    del selfs, longs
    psi2, t_2lpt = timed(lambda: api.compute_2lpt_gpu(delta_tile, n, box, return_device=True))
    del psi2
    lap, t_lap = timed(lambda: api.compute_laplacian_gpu(delta_tile, n, box,
                                                         return_device=True))
    del lap, delta_tile
    api.context(0).release_scratch()
    torch.cuda.empty_cache()
    return dict(residual_ms=t_res, isolated_convolve_ms=t_conv, interpolate_ms=t_interp,
                compute_2lpt_ms=t_2lpt, laplacian_ms=t_lap,
                total_ms=round(t_res + t_conv + t_interp + t_2lpt + t_lap, 3),
                downsample_once_per_run_ms=t_down,
                note=f"{n}^3 tile of a {N}^3 grid: Threefry residual noise less the block "
                     f"means of the {M}^3 coarse grid, isolated convolution of delta + psi1 "
                     f"x3 over the {2 * n}^3 padded box (1 R2C + 4 C2R), 4 tri-cubic "
                     f"interpolations from {M}^3, 2LPT (2 R2C + 9 C2R at {n}^3), Laplacian "
                     f"(1 + 1); the coarse downsample (all {N}^3 cells) runs once per run")


def batched_result(args, cfg_name):
    """Uneven tiles: shell analysis of all candidates of 64 tiles, (a) one library call per
    tile as the reference drivers do, (b) one persistent launch over all tiles, (c) the same
    recorded as a hipGraph plan and replayed.  Peak finding runs per tile, untimed.  Returns
    the timings and whether the three give the same results."""
    import torch

    from huntian_hip import api, synth
    from huntian_hip.homeltab import WEBSKY_TABLE, read_homeltab

    cfg = CONFIGS[cfg_name]
    dev = torch.device("cuda", 0)
    specs = synth.batched_specs(cfg_name)
    nt = len(specs)
    sides = [s for s, _, _ in specs]
    nbuffs = [b for _, b, _ in specs]
    radii = synth.filter_radii(cfg["rmin"], cfg["rmax"], cfg["nfilt"])
    alatt, nhunt = cfg["alatt"], cfg["nhunt"]
    ct = read_homeltab(WEBSKY_TABLE)
    fcrit = api.fsc_of_z(ct, 0.0)
    stab = api.build_shell_tables(nhunt)
    Rf_tab = torch.tensor(radii, dtype=torch.float64, device=dev)
    ir2_tab = torch.tensor([api.ir2min_of(R, alatt) for R in radii], dtype=torch.int32,
                           device=dev)
    tiles, counts = [], []
    t0 = time.time()
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
                          ir2min=ir2_tab[f], mask=mask, nbuff=nb))
        counts.append(len(pf["i"]))
    torch.cuda.synchronize()
    total = int(sum(counts))
    log(f"{nt} tiles, sides {sorted(set(int(s) for s in sides))}, {total} candidates "
        f"(per tile {min(counts)}..{max(counts)}), setup {time.time() - t0:.1f} s")
    kw = dict(alatt=alatt, ZZon=1.0, fcrit=fcrit, variant=args.variant)
    nbuff_all = int(min(nbuffs))  # one mask buffer width for the batched call
    per = [None] * nt
    batch = [None]

    def per_tile():
        for q, t in enumerate(tiles):
            per[q] = api.shell_analyse(t["fields"], t["pi"], t["pj"], t["pk"], t["Rf"],
                                       t["ir2min"], stab, ct, nbuff=nbuff_all, mask=t["mask"],
                                       **kw)

    def batched():
        batch[0] = api.shell_analyse_tiles(tiles, stab, ct, nbuff=nbuff_all, **kw)

    plan = api.ShellPlan(tiles, stab, ct, nbuff=nbuff_all, **kw)
    modes = {"per_tile_calls": per_tile, "one_launch": batched, "hipgraph_plan": plan.launch}
    ms = {}
    steps = max(1, args.steps)
    for name, fn in modes.items():
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms[name] = e0.elapsed_time(e1) / steps
        log(f"{name}: {ms[name]:.2f} ms per pass")
    same = all(same_results(per[q], batch[0][q], counts[q]) and
               same_results(per[q], plan.results[q], counts[q]) for q in range(nt))
    plan.close()
    best = min(ms, key=ms.get)
    res = dict(
        workload=f"{nt} tiles, sides from {list(cfg['sides'])}, nbuff "
                 f"{cfg['nbuff'][0]}-{cfg['nbuff'][1]}, seeds {cfg['seeds'][0]}-"
                 f"{cfg['seeds'][1]}, {cfg['nfilt']} top-hat filters {cfg['rmin']}-"
                 f"{cfg['rmax']} Mpc/h, hunt radius {nhunt}",
        candidates=total, candidates_per_tile=[min(counts), max(counts)],
        ms_per_pass={k: round(v, 3) for k, v in ms.items()},
        best_mode=best, peaks_per_s=round(total / (ms[best] * 1e-3), 1),
        speedup_vs_per_tile_calls=round(ms["per_tile_calls"] / ms[best], 3),
        same_results=bool(same))
    del tiles, plan, per, batch
    torch.cuda.empty_cache()
    return res


def run_batched(args, cfg):
    """--config batched64: the batched workload as the JSON line."""
    res = batched_result(args, args.config)
    out = {
        "metric": "peaks resolved/s, 64 uneven tiles, shell analysis (batched)",
        "value": res["peaks_per_s"], "unit": "peaks/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["ms_per_pass"][res["best_mode"]],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (sums of f32 inputs and f32 products, as the reference CPU path)",
        "data": "synthetic",
        "config": dict(res, parallelism="1 GPU"),
    }
    print(json.dumps(out), flush=True)


def tile_result(args, name, seed=None, net=None, ablate=False):
    """One tile workload end to end on the GPU: peak finding over the pre-smoothed fields,
    then the shell analysis of its candidates (crossing schedule), each timed on a warm call;
    with ablate, the shell analysis under every schedule (schedule_ablation)."""
    import torch

    from huntian_hip import api, synth
    from huntian_hip.homeltab import WEBSKY_TABLE, read_homeltab

    dev = torch.device("cuda", 0)
    t0 = time.time()
    w = synth.workload_tile(name, dev, seed=seed)
    n, nbuff, nhunt, alatt = w["n"], w["nbuff"], w["nhunt"], w["alatt"]
    radii = w["radii"]
    ct = read_homeltab(WEBSKY_TABLE)
    fcrit = api.fsc_of_z(ct, 0.0)
    stab = api.build_shell_tables(nhunt)
    Rf_tab = torch.tensor(radii, dtype=torch.float64, device=dev)
    ir2_tab = torch.tensor([api.ir2min_of(R, alatt) for R in radii], dtype=torch.int32,
                           device=dev)
    mask = api.fortran_empty((n, n, n), torch.int8, dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev[0].record()
        out = fn()
        ev[1].record()
        torch.cuda.synchronize()
        return out, ev[0].elapsed_time(ev[1])

    def pfind():
        mask.zero_()
        return api.peak_find(w["smoothed"], [1.5] * len(radii), nbuff, mask,
                             capacity=(n - 2 * nbuff) ** 3)
    pf, t_pf = timed(pfind)
    npk = len(pf["i"])
    Rf = Rf_tab[pf["filt"].long()]
    ir2 = ir2_tab[pf["filt"].long()]
    res = api.alloc_results(npk, dev)

    def shells():
        mask.zero_()
        return api.shell_analyse(w["fields"], pf["i"], pf["j"], pf["k"], Rf, ir2, stab, ct,
                                 alatt=alatt, ZZon=1.0, fcrit=fcrit, nbuff=nbuff,
                                 variant=args.variant, mask=mask, results=res)
    r, t_sa = timed(shells)
    out = dict(workload=f"{name}: {n}^3, box {w['box']:g} Mpc/h, seed {w['seed']}, "
                        f"{len(radii)} top-hat filters {w['rmin']}-{w['rmax']} Mpc/h, nbuff "
                        f"{nbuff}, hunt radius {nhunt} ({stab.nshells} shells)",
               candidates=npk, halos=int((r["RTHL"][:npk] > 0).sum()),
               peak_find_ms=round(t_pf, 3), shell_analysis_ms=round(t_sa, 3),
               peaks_per_s=round(npk / ((t_pf + t_sa) * 1e-3), 1),
               shell_only_peaks_per_s=round(npk / (t_sa * 1e-3), 1))
    if ablate:
        ctx = api.context(0)
        out["schedule_ablation_shell_ms"] = schedule_ablation(
            args, ctx, w["fields"], pf, r, npk, Rf_tab, ir2_tab, stab, ct, alatt, fcrit,
            nbuff, mask, net, dev)
    log(f"{name}: {out} ({time.time() - t0:.1f} s)")
    del w, pf, r, res, mask
    api.context(0).release_scratch()
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()

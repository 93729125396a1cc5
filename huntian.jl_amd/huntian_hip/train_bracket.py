"""Train the bracket network on the GPU from the plain method's own results.

    python -m huntian_hip.train_bracket --out huntian_hip/data/bracket_mlp.npz

Training tiles: the 128^3 workload's tiles (box 256 Mpc/h, 2 Mpc/h cells, BBKS spectrum,
sigma8 0.81, 16 top-hat filters 1.5..20 Mpc/h, nbuff 16, hunt radius 15 cells) with seeds
1..8; evaluation seed 99.  Each tile runs the filter bank, peak finding and the plain shell
analysis; the targets are each candidate's RTHL shell, gather extent and collapse flag, the
inputs its 18 inner-shell features (hnt_bracket_features).  Everything runs through the
library.  The evaluation runs the learned schedule with the trained network on the seed-99
tile and reports how often the bracket verified.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(msg):
    print(f"[train_bracket] {msg}", file=sys.stderr, flush=True)


def tile_samples(n, box, seed, nbuff, nhunt, radii, stab, ct, fcrit, device):
    """(features [np, 18], targets [np, 3], extent [np], peaks, tile inputs) of one tile."""
    import torch

    from . import api, bracket, synth

    alatt = box / n
    fields, dk = synth.fields_torch(n, box, seed, device)
    sm = [synth.smooth_tophat_torch(dk, n, box, R) for R in radii]
    del dk
    mask = api.fortran_empty((n, n, n), torch.int8, device)
    mask.zero_()
    pf = api.peak_find(sm, [1.5] * len(radii), nbuff, mask, capacity=(n - 2 * nbuff) ** 3)
    del sm
    npk = len(pf["i"])
    Rf = torch.tensor(radii, dtype=torch.float64, device=device)[pf["filt"].long()]
    ir2 = torch.tensor([api.ir2min_of(R, alatt) for R in radii], dtype=torch.int32,
                       device=device)[pf["filt"].long()]
    mask.zero_()
    r = api.shell_analyse(fields, pf["i"], pf["j"], pf["k"], Rf, ir2, stab, ct, alatt=alatt,
                          ZZon=1.0, fcrit=fcrit, nbuff=nbuff, mask=mask)
    x = bracket.features(fields, pf["i"], pf["j"], pf["k"], Rf, stab, alatt)
    t = bracket.targets(r, stab, npk)
    inputs = dict(fields=fields, Rf=Rf, ir2=ir2, alatt=alatt, mask=mask, result=r)
    return x, t, r["nshells_used"][:npk].clone(), pf, inputs


def wave_efficiency(cells, key_order):
    """Mean over wavefronts of 64 consecutive peaks (in key_order) of mean/max gather cells."""
    c = cells[key_order]
    ng = (len(c) + 63) // 64
    pad = np.zeros(ng * 64, np.float64)
    pad[:len(c)] = c
    g = pad.reshape(ng, 64)
    return float(c.sum() / (g.max(1).sum() * 64))


def morton(i, j, k):
    def spread(v):
        x = v.astype(np.uint64) & np.uint64(0x1FFFFF)
        for sh, m in ((32, 0x1F00000000FFFF), (16, 0x1F0000FF0000FF), (8, 0x100F00F00F00F00F),
                      (4, 0x10C30C30C30C30C3), (2, 0x1249249249249249)):
            x = (x | (x << np.uint64(sh))) & np.uint64(m)
        return x
    return spread(i - 1) | (spread(j - 1) << np.uint64(1)) | (spread(k - 1) << np.uint64(2))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--box", type=float, default=256.0)
    ap.add_argument("--nbuff", type=int, default=16)
    ap.add_argument("--nhunt", type=int, default=15)
    ap.add_argument("--rmin", type=float, default=1.5)
    ap.add_argument("--rmax", type=float, default=20.0)
    ap.add_argument("--nfilt", type=int, default=16)
    ap.add_argument("--seeds", type=str, default="1-8")
    ap.add_argument("--eval-seed", type=int, default=99)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", required=True)
    ap.add_argument("--summary", default=None)
    args = ap.parse_args(argv)

    import torch

    from . import api, bracket, synth
    from .homeltab import WEBSKY_TABLE, read_homeltab

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(0)
    a, b = (int(v) for v in args.seeds.split("-"))
    radii = synth.filter_radii(args.rmin, args.rmax, args.nfilt)
    stab = api.build_shell_tables(args.nhunt)
    ct = read_homeltab(WEBSKY_TABLE)
    fcrit = api.fsc_of_z(ct, 0.0)
    xs, ts = [], []
    t0 = time.time()
    for seed in range(a, b + 1):
        x, t, _, _, _ = tile_samples(args.n, args.box, seed, args.nbuff, args.nhunt, radii,
                                     stab, ct, fcrit, dev)
        xs.append(x)
        ts.append(t)
        log(f"seed {seed}: {len(x)} samples ({time.time() - t0:.1f} s)")
    x = torch.cat(xs)
    t = torch.cat(ts)
    del xs, ts
    net = bracket.BracketNet(seed=0, device=dev)
    net.set_norm_from(x)
    t1 = time.time()
    hist = net.fit(x, t, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
    torch.cuda.synchronize()
    train_s = time.time() - t1
    net.save(args.out)
    log(f"trained on {len(x)} samples in {train_s:.1f} s; saved {args.out}")

    # evaluation tile
    xe, te, used, pf, inp = tile_samples(args.n, args.box, args.eval_seed, args.nbuff,
                                         args.nhunt, radii, stab, ct, fcrit, dev)
    pred = net.forward(xe)
    p = pred.cpu().numpy()
    tt = te.cpu().numpy()
    cells_tab = np.concatenate([[0], np.cumsum(stab.shell_count)])
    cells = cells_tab[used.cpu().numpy()]
    mort = morton(pf["i"].cpu().numpy(), pf["j"].cpu().numpy(), pf["k"].cpu().numpy())
    ir2 = np.array([api.ir2min_of(R, args.box / args.n) for R in radii])[pf["filt"].cpu().numpy()]
    stop_pred = np.clip(np.rint(np.expm1(p[:, 1])), 1, stab.nshells - 1)
    orders = {
        "plain (ir2min, Morton)": np.lexsort((mort, ir2)),
        "learned (predicted extent, Morton)": np.lexsort((mort, stop_pred)),
        "crossing (exact extent, Morton)": np.lexsort((mort, used.cpu().numpy())),
    }
    eff = {k: wave_efficiency(cells, o) for k, o in orders.items()}
    collapse_acc = float(((p[:, 2] > 0) == (tt[:, 2] > 0.5)).mean())
    # the learned schedule on the evaluation tile: results must equal the plain method's,
    # and the bracket flags say how often the bracket verified
    inp["mask"].zero_()
    rl = api.shell_analyse(inp["fields"], pf["i"], pf["j"], pf["k"], inp["Rf"], inp["ir2"],
                           stab, ct, alatt=inp["alatt"], ZZon=1.0, fcrit=fcrit,
                           nbuff=args.nbuff, mask=inp["mask"], order="learned", bracket=net)
    r0 = inp["result"]
    same = all(torch.equal(torch.nan_to_num(rl[k], nan=7.0), torch.nan_to_num(r0[k], nan=7.0))
               for k in ("RTHL", "Fbarx", "strain", "e_v", "p_v", "nshells_used"))
    fl = rl["bracket"][:len(xe)].cpu().numpy()
    halo = rl["RTHL"][:len(xe)].cpu().numpy() > 0
    lo = np.clip(np.rint(np.expm1(p[:, 0])), 1, stab.nshells - 1)
    lo_true = np.rint(np.expm1(tt[:, 0]))
    summary = dict(
        train_samples=int(len(x)), eval_samples=int(len(xe)), epochs=args.epochs,
        train_seconds=round(train_s, 2), loss_per_epoch=[round(h, 5) for h in hist],
        eval_mae_log_rthl_shell=float(np.abs(p[:, 0] - tt[:, 0]).mean()),
        eval_mae_log_extent=float(np.abs(p[:, 1] - tt[:, 1]).mean()),
        eval_rthl_shell_within_2=float((np.abs(lo - lo_true)[halo] <= 2).mean()),
        eval_extent_within_2=float((np.abs(stop_pred - used.cpu().numpy()) <= 2).mean()),
        eval_collapse_accuracy=collapse_acc,
        eval_bracket_gather_verified=float((fl & 1).astype(bool).mean()),
        eval_bracket_walk_verified_of_halos=float(((fl & 2) > 0)[halo].mean()),
        eval_bracket_both_verified_of_halos=float(((fl & 3) == 3)[halo].mean()),
        eval_learned_results_equal_plain=bool(same),
        eval_wave_efficiency=eff,
        tiles=f"{args.n}^3, box {args.box:g} Mpc/h, nbuff {args.nbuff}, hunt radius "
              f"{args.nhunt}, {args.nfilt} filters {args.rmin:g}-{args.rmax:g} Mpc/h, seeds "
              f"{args.seeds}, eval seed {args.eval_seed}")
    log(json.dumps(summary))
    if args.summary:
        with open(args.summary, "w") as fh:
            json.dump(summary, fh, indent=1)
    return summary


if __name__ == "__main__":
    main()

"""Comparison helpers of the GPU parity tests: the library's results against the oracle's."""
import numpy as np

FLOAT_KEYS = ("RTHL", "Fbarx", "e_v", "p_v", "Srb", "d2F", "zvir_half", "eigs", "Sbar",
              "Sbar2", "gradpk", "gradpkf", "gradpkrf", "strain")

# the bars the library is held to (the FP64 path differs from the FP64 oracle only where the
# device and host log10 / pow differ by an ulp) and the looser ones the product promises
# for reduced-precision sums
RTOL_FP64 = 1e-10
BARS = {"RTHL": 1e-5, "Fbarx": 1e-5, "strain": 1e-4, "e_v": 1e-4, "p_v": 1e-4}


def to_host(r, sel=None):
    """Device result dict -> numpy, strain as [p, L, K] like the oracle's."""
    out = {}
    for k, v in r.items():
        if v is None:
            continue
        a = v.cpu().numpy()
        out[k] = a[sel] if sel is not None else a
    out["strain"] = np.swapaxes(out["strain"], 1, 2)
    return out


def rel_errors(gpu, ref, keys=FLOAT_KEYS):
    """Worst relative error per field over peaks that collapse on both sides; entries tiny
    against the rest of their vector are compared absolutely.  NaN patterns must agree."""
    both = (gpu["RTHL"] > 0) & (ref["RTHL"] > 0)
    worst = {}
    if not both.any():
        return {k: 0.0 for k in keys}
    for k in keys:
        a = gpu[k][both].reshape(int(both.sum()), -1)
        b = ref[k][both].reshape(int(both.sum()), -1)
        assert np.array_equal(np.isnan(a), np.isnan(b)), (k, "NaN pattern")
        a = np.where(np.isnan(a), 0.0, a)
        b = np.where(np.isnan(b), 0.0, b)
        if a.size == 0:
            worst[k] = 0.0
            continue
        vscale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-30)
        err = np.abs(a - b) / np.maximum(np.abs(b), 1e-9 * vscale)
        worst[k] = float(err.max())
    return worst


def bar_errors(gpu, ref):
    """Worst errors, over peaks that collapse on both sides, in the measure the required bars
    are stated in: RTHL and Fbarx relative; a strain entry relative to itself or, if it is
    smaller, to 1e-3 of the peak's largest entry (off-diagonal entries can be ~0); e_v and
    p_v relative to themselves or, near 0, absolutely below 1e-3 (p ranges over [-e, e])."""
    both = (gpu["RTHL"] > 0) & (ref["RTHL"] > 0)
    out = {}
    if not both.any():
        return {**{k: 0.0 for k in BARS}, **{k + "_over_bar": 0 for k in BARS}}
    for k in BARS:
        a = gpu[k][both].reshape(int(both.sum()), -1)
        b = ref[k][both].reshape(int(both.sum()), -1)
        if k in ("e_v", "p_v"):
            floor = 1e-3
        else:
            floor = 1e-3 * np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-30) \
                if k == "strain" else 1e-300
        err = np.abs(a - b) / np.maximum(np.abs(b), floor)
        per_peak = err.max(axis=1) if err.size else np.zeros(0)
        out[k] = float(per_peak.max()) if per_peak.size else 0.0
        out[k + "_over_bar"] = int((per_peak > BARS[k]).sum())
    return out


def compare(gpu, ref, label, rtol=RTOL_FP64):
    """Same halo set and every float field within rtol; prints and returns the worst errors."""
    n = len(ref["RTHL"])
    assert np.array_equal(gpu["RTHL"][:n] > 0, ref["RTHL"] > 0), label
    worst = rel_errors({k: v[:n] for k, v in gpu.items()}, ref)
    bad = {k: v for k, v in worst.items() if v > rtol}
    exact = sum(np.array_equal(gpu[k][:n], ref[k], equal_nan=True) for k in FLOAT_KEYS)
    print(f"{label}: {n} peaks, {int((ref['RTHL'] > 0).sum())} halos, {exact}/{len(FLOAT_KEYS)} "
          f"fields bit-identical, worst rel {max(worst.values()):.2e} "
          f"({max(worst, key=worst.get)})", flush=True)
    assert not bad, (label, bad)
    return worst

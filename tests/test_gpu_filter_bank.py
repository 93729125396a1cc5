"""hnt_peak_find_tile: filter-bank smoothing (one real-to-complex FFT, then window x spectrum
and a complex-to-real FFT per filter) fused with the peak test.

The reference's smoothed values come from FFTW / cuFFT, whose last bits no restatement can
reproduce; the smoothed fields are therefore checked against an FP64 numpy FFT with the same
Float32-rounded window (Filters.jl:72-99, ext/CUDAExt.jl:2831-2871) within a tolerance, and
the peaks exactly against the oracle's find_peaks loop run on the very fields the tile call
produced (and against hnt_peak_find on them)."""
import numpy as np
import pytest

from huntian_hip import api, synth

pytestmark = pytest.mark.gpu


def window(kind, k, R):
    x = k * R
    if kind == 1:
        w = synth.tophat_w(x)
    elif kind == 0:
        w = np.exp(-x * x * 0.125)
    else:
        w = np.exp(-x * x * 0.125) * k * k
    w = w.astype(np.float32).astype(np.float64)
    return np.where(k == 0, 1.0, w)  # k = 0 left unchanged, for every window


def smooth_ref(delta, alatt, R, kind):
    """FP64 smoothing of a column-major periodic tile of box n * alatt."""
    n = delta.shape[0]
    fk = np.fft.rfftn(np.ascontiguousarray(delta.T).astype(np.float64))
    KZ, KY, KX, _ = synth._kgrid(n, n * alatt)
    k = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    return np.asfortranarray(np.fft.irfftn(fk * window(kind, k, R), s=(n, n, n), axes=(0, 1, 2)).T)


@pytest.fixture(scope="module")
def tile():
    n, alatt = 64, 2.0
    delta, _ = synth.gaussian_field(n, n * alatt, 4242)
    return delta, alatt


def _dev(a, dtype=None):
    import torch

    return api.to_device_field(a, dtype=dtype or torch.float32)


def _outs(n, k):
    import torch

    return [api.fortran_empty((n, n, n), torch.float32, "cuda") for _ in range(k)]


@pytest.mark.parametrize("wsmooth,odd", [(1, False), (0, False), (3, False), (1, True)])
def test_smoothed_fields_match_fp64_fft(tile, wsmooth, odd):
    """Smoothed fields within 2e-5 of the FP64 FFT (relative to their rms); an odd n has a
    negative frequency for every |k| along y and z (no Nyquist row)."""
    import torch

    delta, alatt = tile
    if odd:
        delta, _ = synth.gaussian_field(45, 45 * alatt, 77)
    n = delta.shape[0]
    radii = [8.0, 4.0, 2.0]
    outs = _outs(n, len(radii))
    mask = _dev(np.zeros((n, n, n), np.int8), torch.int8)
    d = _dev(delta)
    before = d.clone()
    api.peak_find_tile(d, radii, [1e30] * len(radii), 2, mask, alatt=alatt, wsmooth=wsmooth,
                       smoothed_out=outs)
    assert torch.equal(d, before)  # the caller's delta is not touched
    for R, o in zip(radii, outs):
        ref = smooth_ref(delta, alatt, R, wsmooth)
        got = o.cpu().numpy()
        err = np.abs(got - ref).max() / np.sqrt(np.mean(ref ** 2))
        assert err < 2e-5, (R, err)


@pytest.mark.parametrize("nbuff,seed", [(3, 1), (9, 2)])
def test_tile_peaks_equal_oracle_on_its_fields(oracle, tile, nbuff, seed):
    """Peaks, order, values and mask equal the oracle's filter loop over the smoothed fields
    the call produced, and hnt_peak_find on those fields; also under a capacity cut."""
    import torch

    delta, alatt = tile
    n = delta.shape[0]
    radii = [12.0, 7.0, 4.0, 2.5, 1.5]
    fcrits = [float(np.sqrt(np.mean(smooth_ref(delta, alatt, R, 1) ** 2))) for R in radii]
    rng = np.random.default_rng(seed)
    mask0 = np.zeros((n, n, n), np.int8, order="F")
    mask0[rng.random((n, n, n)) < 0.03] = 1
    outs = _outs(n, len(radii))
    md = _dev(mask0, torch.int8)
    got = api.peak_find_tile(_dev(delta), radii, fcrits, nbuff, md, alatt=alatt,
                             capacity=n ** 3, smoothed_out=outs)
    fields = [np.asfortranarray(o.cpu().numpy()) for o in outs]
    mref = mask0.copy(order="F")
    ref, tot = oracle.peak_find_multi(fields, fcrits, nbuff, mref)
    assert got["total"] == tot > 100
    assert len(set(ref["filt"])) == len(radii)
    for k in ("i", "j", "k", "filt", "delta"):
        assert np.array_equal(got[k].cpu().numpy(), ref[k]), k
    assert np.array_equal(md.cpu().numpy(), mref)
    # hnt_peak_find on the same device fields
    md2 = _dev(mask0, torch.int8)
    two = api.peak_find(outs, fcrits, nbuff, md2, capacity=n ** 3)
    for k in ("i", "j", "k", "filt", "delta"):
        assert torch.equal(two[k], got[k]), k
    assert torch.equal(md2, md)
    # capacity cut: the first `cap` peaks in (filter, k, j, i) order are written and masked;
    # total stays the number of peak cells (the oracle, like the reference, counts a cell
    # again in a later filter when its first peak fell beyond the capacity and went unmasked)
    cap = tot // 2
    md3 = _dev(mask0, torch.int8)
    cut = api.peak_find_tile(_dev(delta), radii, fcrits, nbuff, md3, alatt=alatt, capacity=cap)
    mref3 = mask0.copy(order="F")
    ref3, tot3 = oracle.peak_find_multi(fields, fcrits, nbuff, mref3, cap=cap)
    assert cut["total"] == tot <= tot3 and len(cut["i"]) == cap
    for k in ("i", "j", "k", "filt"):
        assert np.array_equal(cut[k].cpu().numpy(), ref3[k]), k
    assert np.array_equal(md3.cpu().numpy(), mref3)


@pytest.mark.parametrize("nbuff", [1, 5, 16])
def test_inner_cube_transforms_give_the_same_peaks(tile, nbuff):
    """Without smoothed_out the inverse transforms write only the inner cube and its one-cell
    halo; peaks, values, d2Rf and the mask equal those of the call that writes whole fields."""
    import torch

    delta, alatt = tile
    n = delta.shape[0]
    radii = [12.0, 7.0, 4.0, 2.5]
    fcrits = [0.6 * float(np.sqrt(np.mean(smooth_ref(delta, alatt, R, 1) ** 2))) for R in radii]
    mask0 = np.zeros((n, n, n), np.int8, order="F")
    mask0[np.random.default_rng(nbuff).random((n, n, n)) < 0.02] = 1
    res = []
    for whole in (True, False):
        md = _dev(mask0, torch.int8)
        r = api.peak_find_tile(_dev(delta), radii, fcrits, nbuff, md, alatt=alatt,
                               with_lapd=True, capacity=n ** 3,
                               smoothed_out=_outs(n, len(radii)) if whole else None)
        res.append((r, md))
    (a, ma), (b, mb) = res
    assert a["total"] == b["total"] > 50
    for k in ("i", "j", "k", "filt", "delta", "d2Rf"):
        assert torch.equal(a[k], b[k]), k
    assert torch.equal(ma, mb)


def test_d2Rf_is_the_k2_gaussian_at_each_peak(tile):
    """With the Laplacian output (ioutshear >= 1), d2Rf of a peak of filter f is the
    window-3 smoothing at radius Rf[f] sampled at the peak cell."""
    import torch

    delta, alatt = tile
    n = delta.shape[0]
    radii = [9.0, 5.0, 2.5]
    fcrits = [0.5 * float(np.sqrt(np.mean(smooth_ref(delta, alatt, R, 1) ** 2))) for R in radii]
    md = _dev(np.zeros((n, n, n), np.int8), torch.int8)
    got = api.peak_find_tile(_dev(delta), radii, fcrits, 4, md, alatt=alatt, with_lapd=True,
                             capacity=n ** 3)
    i = got["i"].cpu().numpy() - 1
    j = got["j"].cpu().numpy() - 1
    k = got["k"].cpu().numpy() - 1
    f = got["filt"].cpu().numpy()
    d2 = got["d2Rf"].cpu().numpy()
    assert len(d2) > 50 and np.all(d2 != 0)
    for q, R in enumerate(radii):
        sel = f == q
        lap = smooth_ref(delta, alatt, R, 3)
        scale = np.sqrt(np.mean(lap ** 2))
        err = np.abs(d2[sel] - lap[i[sel], j[sel], k[sel]]).max() / scale
        assert err < 2e-5, (R, err)


def test_tile_mirror_returns_peak_candidates(tile):
    """peak_find_tile_gpu mirror: positions x = xbx + alatt (i - (n+1)/2), Rsmooth per
    filter, FcRf the smoothed value, mask with every peak set."""
    delta, alatt = tile
    n = delta.shape[0]
    filters = [(1, 1.0, 8.0), (2, 1.0, 3.0)]
    fcrits = [0.4, 0.6]
    peaks, Rf, FcRf, d2Rf, mask = api.peak_find_tile_gpu(
        delta, filters, fcrits, np.zeros((n, n, n), np.int8), 10.0, -4.0, 2.5, alatt, 3, 1, 1)
    assert len(peaks) > 20
    np.testing.assert_array_equal(peaks["x"], 10.0 + alatt * (peaks["i"] - 0.5 * (n + 1)))
    np.testing.assert_array_equal(peaks["z"], 2.5 + alatt * (peaks["k"] - 0.5 * (n + 1)))
    assert set(np.unique(Rf)) <= {8.0, 3.0}
    np.testing.assert_array_equal(FcRf, peaks["delta"])
    assert np.all(mask[peaks["i"] - 1, peaks["j"] - 1, peaks["k"] - 1] == 1)
    assert mask.sum() == len(peaks) and np.all(d2Rf != 0)


def test_tile_rejects_bad_arguments(tile):
    import torch

    from huntian_hip.abi import HuntianError

    delta, alatt = tile
    n = delta.shape[0]
    md = _dev(np.zeros((n, n, n), np.int8), torch.int8)
    with pytest.raises(HuntianError):
        api.peak_find_tile(_dev(delta), [2.0], [1.0], 2, md, alatt=alatt, wsmooth=2)
    with pytest.raises(HuntianError):
        api.peak_find_tile(_dev(delta), [2.0], [1.0], 0, md, alatt=alatt)
    with pytest.raises(HuntianError):
        api.peak_find_tile(_dev(delta), [2.0], [1.0], 2, md, alatt=-1.0)


# ---------------------------------------------------------------- the 1024^3 tile

@pytest.fixture(scope="module")
def tile1024():
    """A 1024^3 tile of the headline workload's size (2 Mpc/h cells, buffer 64): Gaussian
    white noise, whose smoothing at a few cells gives peaks everywhere; three filters with
    thresholds at three times each smoothed field's rms keep the peak lists near 10^5."""
    import torch

    n, alatt, nbuff = 1024, 2.0, 64
    g = torch.Generator(device="cuda")
    g.manual_seed(1024)
    d = api.fortran_empty((n, n, n), torch.float32, "cuda")
    d.copy_(torch.randn((n, n, n), generator=g, device="cuda").permute(2, 1, 0))
    radii = [12.0, 6.0, 3.0]
    print("tile1024: field ready", flush=True)
    return d, alatt, nbuff, radii


def _rel_l2(a, b):
    import torch

    num = den = 0.0
    for z in range(0, a.shape[2], 128):  # slabs bound the float64 temporaries
        x = a[:, :, z:z + 128].double()
        y = b[:, :, z:z + 128].double()
        num += float(torch.sum((x - y) ** 2))
        den += float(torch.sum(y ** 2))
    return (num / den) ** 0.5


def test_tile1024_library_fft_cropped_and_oracle(oracle, tile1024):
    """At 1024^3 (the 1024-thread column transforms, the window table read in the first
    inverse pass, the inverse transform cropped to the inner cube): the smoothed fields
    within relative L2 2e-6 of hipFFT's (tuning fft = 1); the cropped call (no smoothed_out)
    gives the same peaks, values, d2Rf and mask as the whole-field call; and the peak list and
    mask equal the oracle's filter loop over the smoothed fields, streamed to the host one
    filter at a time with the mask carried."""
    import torch

    d, alatt, nbuff, radii = tile1024
    n = d.shape[0]
    ctx = api.context(0)
    zero = api.fortran_empty((n, n, n), torch.int8, "cuda")
    zero.zero_()
    # thresholds from a first pass (library FFT, whole fields)
    outs = _outs(n, len(radii))
    md = zero.clone()
    api.peak_find_tile(d, radii, [1e30] * len(radii), nbuff, md, alatt=alatt,
                       smoothed_out=outs)
    fcrits = [3.0 * float(torch.sqrt(torch.mean(o.double() ** 2))) for o in outs]
    cap = 1 << 24
    md_whole = zero.clone()
    whole = api.peak_find_tile(d, radii, fcrits, nbuff, md_whole, alatt=alatt,
                               with_lapd=True, capacity=cap, smoothed_out=outs)
    print(f"tile1024: whole-field call, {whole['total']} peaks", flush=True)
    assert 10_000 < whole["total"] < cap
    md_crop = zero.clone()
    crop = api.peak_find_tile(d, radii, fcrits, nbuff, md_crop, alatt=alatt, with_lapd=True,
                              capacity=cap)
    assert crop["total"] == whole["total"]
    for k in ("i", "j", "k", "filt", "delta", "d2Rf"):
        assert torch.equal(crop[k], whole[k]), k
    assert torch.equal(md_crop, md_whole)
    del md_crop, crop
    # hipFFT: same transforms in full, its own rounding
    hip_outs = _outs(n, len(radii))
    try:
        ctx.set_tuning(fft=1)
        api.peak_find_tile(d, radii, fcrits, nbuff, zero.clone(), alatt=alatt, capacity=cap,
                           smoothed_out=hip_outs)
    finally:
        ctx.set_tuning(fft=0)
    for R, a, b in zip(radii, outs, hip_outs):
        err = _rel_l2(a, b)
        print(f"tile1024: R={R} library vs hipFFT rel L2 {err:.2e}", flush=True)
        assert err <= 2e-6, (R, err)
    del hip_outs
    torch.cuda.empty_cache()
    # the oracle over the library's smoothed fields, one filter at a time, mask carried
    mref = np.zeros((n, n, n), np.int8, order="F")
    parts = []
    for f, (o, fc) in enumerate(zip(outs, fcrits)):
        field = np.asfortranarray(o.cpu().numpy())
        ref, tot = oracle.peak_find_multi([field], [fc], nbuff, mref, cap=cap, nthreads=16)
        ref["filt"] = ref["filt"] + f
        parts.append(ref)
        del field
        print(f"tile1024: oracle filter {f}: {tot} peaks", flush=True)
    ref = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    assert len(ref["i"]) == whole["total"]
    for k in ("i", "j", "k", "filt", "delta"):
        assert np.array_equal(whole[k].cpu().numpy(), ref[k]), k
    assert np.array_equal(md_whole.cpu().numpy(), mref)

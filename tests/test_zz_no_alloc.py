"""hnt_shell_analyse_noalloc (all scratch in the library's fixed device array) against
hnt_shell_analyse, both called through the C ABI on the same inputs.

The two run the same kernels over the same peaks, only with buffers from another place and
possibly fewer resident wavefronts, so every result field, the gather extents and the mask
must be identical bit for bit.  Two calls in a row, the second on a smaller tile, show that
the fixed array's leftovers from the first call (profile slots, queue counters, tables) do
not leak into the second."""
import ctypes as C

import numpy as np
import pytest
import torch

from fieldgen import bumps_field
from huntian_hip import abi, api, synth

pytestmark = pytest.mark.gpu


def _grf_tile(n, box, seed, nbuff, radii):
    """A Gaussian random field tile with 1LPT displacements, its smoothed fields' peaks."""
    delta, dk = synth.gaussian_field(n, box, seed)
    f = {"delta": delta}
    f.update(synth.lpt_fields(dk, n, box))
    sm = [api.to_device_field(synth.smooth_tophat(dk, n, box, R)) for R in radii]
    mask = api.to_device_field(np.zeros((n, n, n), np.int8), dtype=torch.int8)
    pf = api.peak_find(sm, [1.5] * len(radii), nbuff, mask, capacity=n ** 3)
    return f, pf


def _call(entry, fields, pi, pj, pk, Rf, ir2, tables, ct, *, alatt, nbuff, fcrit, mask,
          order="crossing", variant="C"):
    ctx = api.context()
    dev = torch.device("cuda", ctx.device)
    fs, keep_f = api._fields_struct(fields)
    b, keep_b = api._batch_struct(pi, pj, pk, Rf, ir2, 1.0, fcrit, None, None, dev)
    prm = api._params(alatt, nbuff, variant, True, order, None, "fp64")
    r = api.alloc_results(len(pi), dev)
    rs = api._results_struct(r)
    cs, keep_t = api._ctab_struct(ct)
    ts = tables.struct()
    m = abi.as_ptr(api._ptr(mask), C.c_int8)
    st = getattr(ctx.lib, entry)(ctx.handle, C.byref(fs), C.byref(b), C.byref(ts), C.byref(cs),
                                 C.byref(prm), C.byref(rs), m)
    msg = abi.last_error(ctx.lib)
    torch.cuda.synchronize()
    del keep_f, keep_b, keep_t
    return st, msg, {k: v[:len(pi)].cpu().numpy() for k, v in r.items()}


def _same(a, b, label):
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), (label, k)


def test_noalloc_matches_shell_analyse_twice(websky):
    fcrit = api.fsc_of_z(websky, 0.0)
    cases = [
        # (n, box, seed, nbuff, nhunt, order): the second call is smaller in every respect
        (64, 128.0, 6060, 11, 10, "crossing"),
        (40, 80.0, 77, 8, 7, "plain"),
    ]
    for n, box, seed, nbuff, nhunt, order in cases:
        print(f"noalloc case n={n} seed={seed} nhunt={nhunt} order={order}", flush=True)
        alatt = box / n
        radii = synth.filter_radii(1.5, 12.0, 6)
        f, pf = _grf_tile(n, box, seed, nbuff, radii)
        Rf = torch.tensor(radii, dtype=torch.float64, device="cuda")[pf["filt"].long()]
        ir2 = torch.tensor([api.ir2min_of(R, alatt) for R in radii], dtype=torch.int32,
                           device="cuda")[pf["filt"].long()]
        tables = api.build_shell_tables(nhunt)
        outs, masks = [], []
        for entry in ("hnt_shell_analyse", "hnt_shell_analyse_noalloc"):
            mask = api.to_device_field(np.zeros((n, n, n), np.int8), dtype=torch.int8)
            st, msg, r = _call(entry, f, pf["i"], pf["j"], pf["k"], Rf, ir2, tables, websky,
                               alatt=alatt, nbuff=nbuff, fcrit=fcrit, mask=mask, order=order)
            assert st == abi.HNT_OK, (entry, msg)
            outs.append(r)
            masks.append(mask.cpu().numpy())
        npk = len(pf["i"])
        nh = int((outs[0]["RTHL"] > 0).sum())
        print(f"  {npk} peaks, {nh} halos", flush=True)
        assert npk > 100 and nh > 10
        _same(outs[0], outs[1], f"n={n}")
        assert np.array_equal(masks[0], masks[1])


def test_noalloc_bumps_with_lapd_and_boundary_peaks(fortran_table):
    """Peaks whose balls leave the grid, lapd and 2LPT fields: the bounds-checked gather."""
    n = 48
    f = bumps_field(n, 11, nbumps=24, margin=3)
    rng = np.random.default_rng(5)
    npk = 300
    pi, pj, pk = (rng.integers(1, n + 1, npk).astype(np.int32) for _ in range(3))
    tables = api.build_shell_tables(9)
    outs = []
    for entry in ("hnt_shell_analyse", "hnt_shell_analyse_noalloc"):
        mask = api.to_device_field(np.zeros((n, n, n), np.int8), dtype=torch.int8)
        st, msg, r = _call(entry, f, pi, pj, pk, 3.0, 4, tables, fortran_table, alatt=1.0,
                           nbuff=2, fcrit=1.0, mask=mask, variant="A")
        assert st == abi.HNT_OK, (entry, msg)
        outs.append((r, mask.cpu().numpy()))
    _same(outs[0][0], outs[1][0], "bumps")
    assert np.array_equal(outs[0][1], outs[1][1])


def test_noalloc_rejects_what_does_not_fit(websky):
    """A 256^3 tile needs 512 MiB of cell records alone: a clear error before any gather."""
    n = 256
    zero = torch.zeros((n, n, n), dtype=torch.float32, device="cuda")
    f = {"delta": zero, "etax": zero, "etay": zero, "etaz": zero}
    pi = np.array([128], np.int32)
    st, msg, _ = _call("hnt_shell_analyse_noalloc", f, pi, pi, pi, 3.0, 4,
                       api.build_shell_tables(5), websky, alatt=1.0, nbuff=2, fcrit=1.0,
                       mask=None)
    assert st == abi.HNT_ERR_CAPACITY, (st, msg)
    assert "256 MiB" in msg, msg

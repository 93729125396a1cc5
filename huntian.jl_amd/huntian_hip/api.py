"""Host-side mirror of Huntian.jl's GPU plugin API for the halo-finding path.

Each function keeps the name and argument meaning of the stub the reference declares
(src/Huntian.jl:98-412) and that ext/CUDAExt.jl implements, and calls libhuntian_hip.so
through its C ABI.  Device memory comes from PyTorch (ROCm); the library never sees a
torch type.  Arrays follow the reference layout: shape (n1, n2, n3) indexed [i, j, k],
column-major, peak indices 1-based.

There is no CPU fallback: every call fails loudly if the HIP library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi
from .homeltab import make_ctab

_ctx = None


def _torch():
    import torch

    return torch


def _require_gpu():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("huntian_hip needs a HIP GPU (torch.cuda.is_available() is False)")
    return torch


class Context:
    """One hnt_ctx on one device (the reference binds one task per device,
    set_cuda_device!, src/Huntian.jl:414-424)."""

    def __init__(self, device: int = 0):
        torch = _require_gpu()
        self.lib = abi.load()
        self.device = device
        torch.cuda.set_device(device)
        h = C.c_void_p()
        abi.check(self.lib, self.lib.hnt_ctx_create(device, C.byref(h)))
        self.handle = h
        self._stream_ptr = -1  # nothing bound yet (0 is the default stream)

    def bind_stream(self, stream=None):
        """Run library work on `stream` (a torch.cuda.Stream) or torch's current stream."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        ptr = s.cuda_stream
        if ptr != self._stream_ptr:
            abi.check(self.lib, self.lib.hnt_ctx_set_stream(self.handle, C.c_void_p(ptr)))
            self._stream_ptr = ptr

    def set_tuning(self, **knobs):
        """Launch-shape knobs (brick_log2, waves_per_cu, gang = lanes per peak: 1, 16 or
        64): speed only, never results; fft = 0 (the library's FFTs) or 1 (hipFFT) for the
        field-synthesis entries, which differ by the transforms' Float32 rounding;
        fft_direct = 1 (default) or 0: the 1024- and 2048-point passes load into LDS directly
        or through registers; tricubic_tab = 1 (default) or 0: the interpolation's weight
        products tabulated per row or formed per cell (both speed only, same bits)."""
        for k, v in knobs.items():
            abi.check(self.lib, self.lib.hnt_ctx_set_tuning(self.handle, abi.TUNING[k], int(v)))

    def release_scratch(self):
        """Free the device scratch and FFT plans kept between calls (hnt_ctx_release_scratch)."""
        abi.check(self.lib, self.lib.hnt_ctx_release_scratch(self.handle))

    def set_profiling(self, on: bool = True):
        abi.check(self.lib, self.lib.hnt_ctx_set_profiling(self.handle, int(bool(on))))

    def profile(self) -> dict:
        """Summed device milliseconds and call counts per stage since set_profiling."""
        n = len(abi.STAGES)
        ms = (C.c_double * n)()
        calls = (C.c_int64 * n)()
        abi.check(self.lib, self.lib.hnt_ctx_get_profile(self.handle, ms, calls, n))
        return {s: (ms[q], calls[q]) for q, s in enumerate(abi.STAGES)}

    def close(self):
        if self.handle:
            self.lib.hnt_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def context(device: int = 0) -> Context:
    global _ctx
    if _ctx is None or _ctx.device != device:
        _ctx = Context(device)
    _ctx.bind_stream()
    return _ctx


# ---------------------------------------------------------------- layout helpers

def fortran_empty(shape, dtype, device="cuda"):
    """torch tensor of `shape` with column-major strides (the reference layout)."""
    torch = _torch()
    return torch.empty(tuple(shape)[::-1], dtype=dtype, device=device).permute(
        *range(len(shape) - 1, -1, -1))


def to_device_field(a, dtype=None, device="cuda"):
    """numpy (any order) or torch array -> column-major device tensor."""
    torch = _torch()
    dtype = dtype or torch.float32
    if isinstance(a, torch.Tensor):
        t = a.to(device=device, dtype=dtype)
    else:
        t = torch.from_numpy(np.asarray(a))
        t = t.to(device=device, dtype=dtype)
    if t.dim() == 3 and not t.permute(2, 1, 0).is_contiguous():
        out = fortran_empty(t.shape, dtype, device)
        out.copy_(t)
        t = out
    return t


def _ptr(t):
    return 0 if t is None else t.data_ptr()


def _dev1(a, dtype, device="cuda"):
    torch = _torch()
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    arr = np.ascontiguousarray(a)
    if not arr.flags.writeable:  # torch warns on read-only numpy arrays
        arr = arr.copy()
    return torch.as_tensor(arr, device=device).to(dtype).contiguous()


# ---------------------------------------------------------------- tables

@dataclass
class ShellTables:
    """build_shell_tables(rmax) (src/ShellAnalysisGPU.jl:28-71); shell_start 0-based."""
    rmax: int
    di: np.ndarray
    dj: np.ndarray
    dk: np.ndarray
    shell_start: np.ndarray
    shell_count: np.ndarray
    shell_r2: np.ndarray

    @property
    def nshells(self):
        return len(self.shell_r2)

    @property
    def ncells(self):
        return len(self.di)

    def struct(self) -> abi.ShellTables:
        p = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
        return abi.ShellTables(self.rmax, self.nshells, self.ncells, p(self.di), p(self.dj),
                               p(self.dk), p(self.shell_start), p(self.shell_count),
                               p(self.shell_r2))


def build_shell_tables(rmax: int) -> ShellTables:
    lib = abi.load()
    nc = C.c_int64()
    ns = C.c_int32()
    abi.check(lib, lib.hnt_shell_table_size(rmax, C.byref(nc), C.byref(ns)))
    arr = {k: np.zeros(nc.value, np.int32) for k in ("di", "dj", "dk")}
    arr.update({k: np.zeros(ns.value, np.int32)
                for k in ("shell_start", "shell_count", "shell_r2")})
    abi.check(lib, lib.hnt_build_shell_tables(
        rmax, *[arr[k].ctypes.data_as(C.POINTER(C.c_int32))
                for k in ("di", "dj", "dk", "shell_start", "shell_count", "shell_r2")]))
    return ShellTables(rmax=rmax, **arr)


def _ctab_struct(ct: dict):
    tab = np.ascontiguousarray(np.asarray(ct["table"], np.float32).ravel(order="F"))
    s = abi.CTab(ct["nx"], ct["ny"], ct["nz"], ct["x1"], ct["x2"], ct["y1"], ct["y2"], ct["z1"],
                 ct["z2"], ct.get("out_val", -1.0), tab.ctypes.data_as(C.POINTER(C.c_float)))
    return s, tab


def fsc_of_z(ct: dict, z: float) -> float:
    """RadialShell.fsc_of_z (src/HaloFinder/RadialShell.jl:181-195)."""
    lib = abi.load()
    s, keep = _ctab_struct(ct)
    v = lib.hnt_fsc_of_z(C.byref(s), float(z))
    del keep
    return v


def ir2min_of(Rf: float, alatt: float) -> int:
    """src/MultiResolution.jl:575-576."""
    import math

    return min(int(math.floor((1.75 * Rf / alatt) ** 2)),
               int(math.floor((40.0 / alatt - 1) ** 2)))


# ---------------------------------------------------------------- peak finding

def peak_find(smoothed, fcrits, nbuff: int, mask, capacity: int | None = None):
    """Multi-filter peak finding on already-smoothed device fields (hnt_peak_find).

    smoothed: list of column-major (n, n, n) float32 device tensors, in processing order.
    mask: column-major (n, n, n) int8 device tensor, updated in place.
    Returns dict of device tensors i, j, k (1-based), filt, delta and the total count."""
    torch = _require_gpu()
    ctx = context(smoothed[0].device.index or 0)
    n = smoothed[0].shape[0]
    for f in smoothed:
        assert f.shape == (n, n, n) and f.dtype == torch.float32 and f.permute(2, 1, 0).is_contiguous()
    assert mask.shape == (n, n, n) and mask.dtype == torch.int8 and mask.permute(2, 1, 0).is_contiguous()
    m = max(n - 2 * nbuff, 0)
    if capacity is None:
        capacity = max(1024, m ** 3 // 64)
    dev = smoothed[0].device
    out = {k: torch.empty(max(capacity, 1), dtype=torch.int32, device=dev)
           for k in ("i", "j", "k", "filt")}
    out["delta"] = torch.empty(max(capacity, 1), dtype=torch.float32, device=dev)
    pk = abi.Peaks(capacity, *[abi.as_ptr(out[k].data_ptr(), C.c_int32)
                               for k in ("i", "j", "k", "filt")],
                   abi.as_ptr(out["delta"].data_ptr(), C.c_float), 0, 0)
    ptrs = (C.c_void_p * len(smoothed))(*[f.data_ptr() for f in smoothed])
    fc = np.ascontiguousarray(np.asarray(fcrits, np.float32))
    st = ctx.lib.hnt_peak_find(ctx.handle, ptrs, fc.ctypes.data_as(C.POINTER(C.c_float)),
                               len(smoothed), n, nbuff,
                               abi.as_ptr(mask.data_ptr(), C.c_int8), C.byref(pk))
    if st != abi.HNT_ERR_CAPACITY:
        abi.check(ctx.lib, st)
    # On HNT_ERR_CAPACITY the first `capacity` peaks (in filter, k, j, i order) are returned
    # and masked, like the reference GPU method that drops slots beyond max_peaks
    # (ext/CUDAExt.jl:2921-2926, 3020-3021); "total" tells how many were found.
    cnt = int(pk.count)
    res = {k: v[:cnt] for k, v in out.items()}
    res["total"] = int(pk.total)
    return res


def peak_find_tile(delta, radii, fcrits, nbuff: int, mask, *, alatt: float, wsmooth: int = 1,
                   with_lapd: bool = False, capacity: int | None = None, smoothed_out=None):
    """Filter-bank smoothing and peak finding of one tile on the device (hnt_peak_find_tile).

    delta: column-major (n, n, n) float32 device tensor (left unchanged); radii (Mpc/h) and
    fcrits in processing order; mask: column-major (n, n, n) int8 device tensor, updated.
    with_lapd: also return d2Rf, the k^2-Gaussian smoothing sampled at each peak.
    smoothed_out: optional list of column-major float32 device tensors that receive the
    smoothed fields (None entries are skipped).
    Returns dict of device tensors i, j, k (1-based), filt, delta (FcRf), d2Rf, and total."""
    torch = _require_gpu()
    ctx = context(delta.device.index or 0)
    n = delta.shape[0]
    assert delta.shape == (n, n, n) and delta.dtype == torch.float32
    assert delta.permute(2, 1, 0).is_contiguous()
    assert mask.shape == (n, n, n) and mask.dtype == torch.int8 and mask.permute(2, 1, 0).is_contiguous()
    nf = len(radii)
    assert len(fcrits) == nf
    m = max(n - 2 * nbuff, 0)
    if capacity is None:
        capacity = max(1024, m ** 3 // 64)
    dev = delta.device
    out = {k: torch.empty(max(capacity, 1), dtype=torch.int32, device=dev)
           for k in ("i", "j", "k", "filt")}
    out["delta"] = torch.empty(max(capacity, 1), dtype=torch.float32, device=dev)
    d2 = torch.zeros(max(capacity, 1), dtype=torch.float32, device=dev)
    pk = abi.Peaks(capacity, *[abi.as_ptr(out[k].data_ptr(), C.c_int32)
                               for k in ("i", "j", "k", "filt")],
                   abi.as_ptr(out["delta"].data_ptr(), C.c_float), 0, 0)
    rf = np.ascontiguousarray(np.asarray(radii, np.float64))
    fc = np.ascontiguousarray(np.asarray(fcrits, np.float32))
    bank = abi.FilterBank(nf, rf.ctypes.data_as(C.POINTER(C.c_double)),
                          fc.ctypes.data_as(C.POINTER(C.c_float)), int(wsmooth), float(alatt))
    so = None
    if smoothed_out is not None:
        assert len(smoothed_out) == nf
        for t in smoothed_out:
            assert t is None or (t.shape == (n, n, n) and t.dtype == torch.float32
                                 and t.permute(2, 1, 0).is_contiguous())
        so = (C.c_void_p * nf)(*[0 if t is None else t.data_ptr() for t in smoothed_out])
    st = ctx.lib.hnt_peak_find_tile(ctx.handle, abi.as_ptr(delta.data_ptr(), C.c_float), n,
                                    int(nbuff), C.byref(bank),
                                    abi.as_ptr(mask.data_ptr(), C.c_int8), C.byref(pk),
                                    abi.as_ptr(d2.data_ptr() if with_lapd else 0, C.c_float),
                                    so)
    if st != abi.HNT_ERR_CAPACITY:
        abi.check(ctx.lib, st)
    cnt = int(pk.count)
    res = {k: v[:cnt] for k, v in out.items()}
    res["d2Rf"] = d2[:cnt]
    res["total"] = int(pk.total)
    return res


def peak_find_tile_gpu(delta_tile, filters, fcrits, tile_masks_in, xbx, ybx, zbx, alatt,
                       nbuff, wsmooth, ioutshear, threads=256, max_peaks=2_000_000):
    """Mirror of peak_find_tile_gpu (src/Huntian.jl:387-412, ext/CUDAExt.jl:2938-3060).

    filters: list of (id, fcrit, Rf) in processing order (largest Rf first in the drivers).
    Smoothing (one rfft, then window + irfft per filter), the peak test and, with
    ioutshear >= 1, the k^2-Gaussian sample at each peak all run in libhuntian_hip
    (hnt_peak_find_tile).  Returns (peaks, Rf, FcRf, d2Rf, mask) with peaks a numpy
    structured array of PeakCandidate fields (i, j, k, ipp, x, y, z, delta, Rsmooth).
    Unlike the CUDA method the peak order is deterministic (filter, then k, j, i)."""
    torch = _require_gpu()
    n = int(delta_tile.shape[0])
    d = to_device_field(delta_tile)
    mask = to_device_field(tile_masks_in, dtype=torch.int8)
    radii = [float(f[2]) for f in filters]
    res = peak_find_tile(d, radii, fcrits, int(nbuff), mask, alatt=float(alatt),
                         wsmooth=int(wsmooth), with_lapd=ioutshear >= 1,
                         capacity=int(max_peaks))
    cnt = len(res["i"])
    i = res["i"].cpu().numpy().astype(np.int64)
    j = res["j"].cpu().numpy().astype(np.int64)
    k = res["k"].cpu().numpy().astype(np.int64)
    fidx = res["filt"].cpu().numpy()
    dl = res["delta"].cpu().numpy()
    cen = 0.5 * (n + 1)
    peaks = np.zeros(cnt, dtype=[("i", "i8"), ("j", "i8"), ("k", "i8"), ("ipp", "i8"),
                                 ("x", "f8"), ("y", "f8"), ("z", "f8"), ("delta", "f4"),
                                 ("Rsmooth", "f8")])
    Rf = np.array([radii[f] for f in fidx], np.float64)
    peaks["i"], peaks["j"], peaks["k"] = i, j, k
    peaks["ipp"] = i + (j - 1) * n + (k - 1) * n * n
    peaks["x"] = float(xbx) + float(alatt) * (i - cen)
    peaks["y"] = float(ybx) + float(alatt) * (j - cen)
    peaks["z"] = float(zbx) + float(alatt) * (k - cen)
    peaks["delta"] = dl
    peaks["Rsmooth"] = Rf
    d2Rf = res["d2Rf"].cpu().numpy() if ioutshear >= 1 else np.zeros(cnt, np.float32)
    return peaks, Rf, dl.copy(), d2Rf, np.asfortranarray(mask.cpu().numpy())


# ---------------------------------------------------------------- shell analysis

def _fields_struct(fields: dict):
    torch = _torch()
    d = fields["delta"]
    n1, n2, n3 = d.shape
    keep = {}
    ptrs = []
    for key in ("delta", "etax", "etay", "etaz", "eta2x", "eta2y", "eta2z", "lapd"):
        a = fields.get(key)
        if a is None:
            ptrs.append(abi.as_ptr(0, C.c_float))
            continue
        t = to_device_field(a, torch.float32, d.device if isinstance(d, torch.Tensor) else "cuda")
        keep[key] = t
        ptrs.append(abi.as_ptr(t.data_ptr(), C.c_float))
    return abi.Fields(n1, n2, n3, *ptrs), keep


def alloc_results(npk: int, device="cuda", schedule_extent=False):
    """Result arrays of npk peaks; schedule_extent adds that optional output (the crossing
    pass's extents, HNT_ORDER_CROSSING), which stays unset otherwise."""
    torch = _torch()
    r = {k: torch.empty(max(npk, 1), dtype=torch.float64, device=device)
         for k in abi.RESULT_SCALARS}
    r.update({k: torch.empty((max(npk, 1), 3), dtype=torch.float64, device=device)
              for k in abi.RESULT_VEC3})
    r["strain"] = torch.empty((max(npk, 1), 3, 3), dtype=torch.float64, device=device)
    for k in abi.RESULT_INT:
        r[k] = torch.zeros(max(npk, 1), dtype=torch.int32, device=device)
    if schedule_extent:
        r["schedule_extent"] = torch.zeros(max(npk, 1), dtype=torch.int32, device=device)
    return r


def _results_struct(r):
    vals = []
    for name, _ in abi.Results._fields_:
        t = r.get(name)
        ct = C.c_int32 if name in abi.RESULT_I32 else C.c_double
        vals.append(abi.as_ptr(_ptr(t), ct))
    return abi.Results(*vals)


def _batch_struct(pi, pj, pk, Rf, ir2min, ZZon, fcrit, ZZon_pp, fcrit_pp, dev):
    """abi.Batch over device copies of the peak arrays (kept alive by the returned list)."""
    torch = _torch()
    npk = int(len(pi))
    pi_ = _dev1(pi, torch.int32, dev)
    pj_ = _dev1(pj, torch.int32, dev)
    pk_ = _dev1(pk, torch.int32, dev)
    Rf_ = _dev1(np.broadcast_to(np.asarray(Rf, np.float64), (npk,)) if not isinstance(
        Rf, torch.Tensor) else Rf, torch.float64, dev)
    ir_ = _dev1(np.broadcast_to(np.asarray(ir2min, np.int32), (npk,)) if not isinstance(
        ir2min, torch.Tensor) else ir2min, torch.int32, dev)
    zz = _dev1(ZZon_pp, torch.float64, dev) if ZZon_pp is not None else None
    fc = _dev1(fcrit_pp, torch.float64, dev) if fcrit_pp is not None else None
    b = abi.Batch(npk, abi.as_ptr(_ptr(pi_), C.c_int32), abi.as_ptr(_ptr(pj_), C.c_int32),
                  abi.as_ptr(_ptr(pk_), C.c_int32), abi.as_ptr(_ptr(Rf_), C.c_double),
                  abi.as_ptr(_ptr(ir_), C.c_int32), abi.as_ptr(_ptr(zz), C.c_double),
                  abi.as_ptr(_ptr(fc), C.c_double), float(ZZon), float(fcrit))
    return b, [pi_, pj_, pk_, Rf_, ir_, zz, fc]


def _params(alatt, nbuff, variant, early_exit, order, bracket, precision, extent=None,
            rmax2rs=0.0):
    net = None
    if order == "learned":
        if bracket is None:
            raise ValueError("order='learned' needs a bracket network")
        net = bracket.struct()
    if order == "given" and extent is None:
        raise ValueError("order='given' needs the extents (extent=...)")
    return abi.Params(float(alatt), int(nbuff), float(rmax2rs), abi.VARIANTS[variant],
                      int(bool(early_exit)),
                      abi.ORDERS[order], abi.PRECISIONS[precision],
                      C.pointer(net) if net is not None else None,
                      abi.as_ptr(_ptr(extent), C.c_int32))


def shell_analyse(fields: dict, pi, pj, pk, Rf, ir2min, tables: ShellTables, ct: dict, *,
                  alatt: float, ZZon: float, fcrit: float, nbuff: int = 0, variant="C",
                  early_exit: bool = True, mask=None, ZZon_pp=None, fcrit_pp=None,
                  results=None, order: str = "crossing", bracket=None,
                  precision: str = "fp64", extent=None, rmax2rs: float = 0.0):
    """hnt_shell_analyse over one batch of peaks (device arrays or numpy).  Returns a dict of
    device tensors: scalars [np], vectors [np, 3], strain [np, 3, 3] ([p, L, K]).

    order: wavefront schedule ("crossing", "plain", "learned" with `bracket`, a
    huntian_hip.bracket.BracketNet, or "given" with `extent`, a device int32 [np] of gather
    extents); it changes the speed, never a result.
    precision: per-shell sums in "fp64" (the reference CPU arithmetic) or "fp32_kahan".
    rmax2rs > 0 (variants "B" and "C"): each peak uses only its shells with
    r <= rmax2rs Rf / alatt, all of them gathered."""
    torch = _require_gpu()
    ctx = context()
    fs, keep_f = _fields_struct(fields)
    npk = int(len(pi))
    dev = torch.device("cuda", ctx.device)
    b, keep_b = _batch_struct(pi, pj, pk, Rf, ir2min, ZZon, fcrit, ZZon_pp, fcrit_pp, dev)
    ext = _dev1(extent, torch.int32, dev) if extent is not None else None
    prm = _params(alatt, nbuff, variant, early_exit, order, bracket, precision, ext, rmax2rs)
    r = results if results is not None else alloc_results(npk, dev)
    rs = _results_struct(r)
    cs, keep_t = _ctab_struct(ct)
    ts = tables.struct()
    m = abi.as_ptr(_ptr(mask), C.c_int8)
    st = ctx.lib.hnt_shell_analyse(ctx.handle, C.byref(fs), C.byref(b), C.byref(ts),
                                   C.byref(cs), C.byref(prm), C.byref(rs), m)
    abi.check(ctx.lib, st)
    del keep_f, keep_t, keep_b, ext
    if npk == 0:
        return {k: v[:0] for k, v in r.items()}
    return r


def _tiles_structs(tiles, ZZon, fcrit, dev, schedule_extent=False):
    nt = len(tiles)
    FA = (abi.Fields * nt)()
    BA = (abi.Batch * nt)()
    RA = (abi.Results * nt)()
    MA = (C.POINTER(C.c_int8) * nt)()
    keep, out = [], []
    for q, t in enumerate(tiles):
        fs, kf = _fields_struct(t["fields"])
        b, kb = _batch_struct(t["pi"], t["pj"], t["pk"], t["Rf"], t["ir2min"], ZZon, fcrit,
                              t.get("ZZon_pp"), t.get("fcrit_pp"), dev)
        r = alloc_results(int(len(t["pi"])), dev, schedule_extent)
        FA[q], BA[q], RA[q] = fs, b, _results_struct(r)
        MA[q] = abi.as_ptr(_ptr(t.get("mask")), C.c_int8)
        keep += [kf, kb]
        out.append(r)
    return FA, BA, RA, MA, keep, out


def shell_analyse_tiles(tiles, tables: ShellTables, ct: dict, *, alatt: float, ZZon: float,
                        fcrit: float, nbuff: int = 0, variant="C", early_exit: bool = True,
                        order: str = "crossing", bracket=None, precision: str = "fp64",
                        rmax2rs: float = 0.0, schedule_extent: bool = False):
    """hnt_shell_analyse_tiles: every tile in one persistent launch.

    tiles: list of dicts with keys fields, pi, pj, pk, Rf, ir2min and optionally mask,
    ZZon_pp, fcrit_pp.  Returns one result dict (as shell_analyse) per tile, with the
    crossing pass's extents as "schedule_extent" if asked."""
    torch = _require_gpu()
    ctx = context()
    dev = torch.device("cuda", ctx.device)
    FA, BA, RA, MA, keep, out = _tiles_structs(tiles, ZZon, fcrit, dev, schedule_extent)
    prm = _params(alatt, nbuff, variant, early_exit, order, bracket, precision,
                  rmax2rs=rmax2rs)
    cs, keep_t = _ctab_struct(ct)
    ts = tables.struct()
    st = ctx.lib.hnt_shell_analyse_tiles(ctx.handle, len(tiles), FA, BA, C.byref(ts),
                                         C.byref(cs), C.byref(prm), RA, MA)
    abi.check(ctx.lib, st)
    del keep, keep_t
    return [{k: v[:len(t["pi"])] for k, v in r.items()} for r, t in zip(out, tiles)]


class ShellPlan:
    """hnt_plan_create / hnt_plan_launch: the tiles' analysis recorded once as a hipGraph.

    Construction runs the analysis (results are valid afterwards); launch() replays it over
    the current contents of the same field and peak arrays.  `results` holds one result
    dict per tile; `inputs` the device arrays the plan reads (update them in place)."""

    def __init__(self, tiles, tables: ShellTables, ct: dict, *, alatt: float, ZZon: float,
                 fcrit: float, nbuff: int = 0, variant="C", early_exit: bool = True,
                 order: str = "crossing", bracket=None, precision: str = "fp64",
                 rmax2rs: float = 0.0):
        torch = _require_gpu()
        self.ctx = context()
        dev = torch.device("cuda", self.ctx.device)
        FA, BA, RA, MA, keep, out = _tiles_structs(tiles, ZZon, fcrit, dev)
        prm = _params(alatt, nbuff, variant, early_exit, order, bracket, precision,
                      rmax2rs=rmax2rs)
        cs, keep_t = _ctab_struct(ct)
        ts = tables.struct()
        h = C.c_void_p()
        abi.check(self.ctx.lib, self.ctx.lib.hnt_plan_create(
            self.ctx.handle, len(tiles), FA, BA, C.byref(ts), C.byref(cs), C.byref(prm), RA, MA,
            C.byref(h)))
        self.handle = h
        self._keep = (FA, BA, RA, MA, keep, keep_t, out, tables, bracket, tiles)
        self.inputs = keep
        self.results = [{k: v[:len(t["pi"])] for k, v in r.items()} for r, t in zip(out, tiles)]

    def launch(self):
        context(self.ctx.device)  # run on torch's current stream
        abi.check(self.ctx.lib, self.ctx.lib.hnt_plan_launch(self.handle))

    def close(self):
        if self.handle:
            self.ctx.lib.hnt_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _host_results(r, npk):
    out = {}
    for k, v in r.items():
        if v is None:
            continue
        a = v[:npk].cpu().numpy()
        out[k] = a
    return out


def cosmo_tables_struct(growth, chi2z=None):
    """abi.CosmoTables over host arrays: growth = (atab, D_atab[, dlnD]) from
    cosmology.dlinear_tables; chi2z = (zs, chis, z_max) from cosmology.build_chi_to_z or
    None.  Returns (struct, arrays to keep alive)."""
    at = np.ascontiguousarray(np.asarray(growth[0], np.float64))
    Dt = np.ascontiguousarray(np.asarray(growth[1], np.float64))
    keep = [at, Dt]
    f64 = C.POINTER(C.c_double)
    if chi2z is not None:
        zs = np.ascontiguousarray(np.asarray(chi2z[0], np.float64))
        ch = np.ascontiguousarray(np.asarray(chi2z[1], np.float64))
        keep += [zs, ch]
        t = abi.CosmoTables(len(at), at.ctypes.data_as(f64), Dt.ctypes.data_as(f64), len(zs),
                            zs.ctypes.data_as(f64), ch.ctypes.data_as(f64), float(chi2z[2]))
    else:
        t = abi.CosmoTables(len(at), at.ctypes.data_as(f64), Dt.ctypes.data_as(f64), 0,
                            None, None, 0.0)
    if len(growth) > 2:
        ft = np.ascontiguousarray(np.asarray(growth[2], np.float64))
        keep.append(ft)
        t.dlnD_atab = ft.ctypes.data_as(f64)
    return t, keep


def pack_params(*, alatt, n, xbx=0.0, ybx=0.0, zbx=0.0, ZZon=1.0, ievol=0, obs=(0.0, 0.0, 0.0),
                z_max=0.0, ilpt=2, Om=0.31, OL=0.69, ioutshear=0):
    return abi.PackParams(float(alatt), int(n), float(xbx), float(ybx), float(zbx), float(ZZon),
                          int(ievol), (C.c_double * 3)(*[float(v) for v in obs]), float(z_max),
                          int(ilpt), float(Om), float(OL), int(ioutshear))


def pack_records(pi, pj, pk, Rf, results: dict, FcRf, *, growth, chi2z=None, d2Rf=None,
                 capacity: int | None = None, **params):
    """Halo records of one tile on the device (hnt_pack_records; the record loop of
    src/MultiResolution.jl:1159-1219, layouts of src/IO/Catalog.jl).

    pi, pj, pk, Rf: the peaks of the shell analysis (device or host); results: its result
    dict (device FP64 tensors, as from shell_analyse); FcRf / d2Rf: float32 per peak (the
    peak finder's delta and d2Rf).  growth / chi2z: host tables (cosmology module).
    params: alatt, n, xbx, ybx, zbx, ZZon, ievol, obs, z_max, ilpt, Om, OL, ioutshear.
    Returns a device float32 tensor (nhalo, 11 or 33) in peak order."""
    torch = _require_gpu()
    dev = results["RTHL"].device
    ctx = context(dev.index or 0)
    npk = int(len(pi))
    pi_, pj_, pk_ = (_dev1(x, torch.int32, dev) for x in (pi, pj, pk))
    Rf_ = _dev1(Rf if isinstance(Rf, torch.Tensor) else
                np.broadcast_to(np.asarray(Rf, np.float64), (npk,)), torch.float64, dev)
    keep = [pi_, pj_, pk_, Rf_]
    # the packer reads pi, pj, pk and Rf only
    b = abi.Batch(npk, abi.as_ptr(_ptr(pi_), C.c_int32), abi.as_ptr(_ptr(pj_), C.c_int32),
                  abi.as_ptr(_ptr(pk_), C.c_int32), abi.as_ptr(_ptr(Rf_), C.c_double),
                  abi.as_ptr(0, C.c_int32), abi.as_ptr(0, C.c_double), abi.as_ptr(0, C.c_double),
                  1.0, 1.0)
    fc = _dev1(FcRf, torch.float32, dev)
    d2 = _dev1(d2Rf, torch.float32, dev) if d2Rf is not None else None
    nf = 33 if int(params.get("ioutshear", 0)) >= 1 else 11
    if capacity is None:
        capacity = npk
    out = torch.empty((max(int(capacity), 1), nf), dtype=torch.float32, device=dev)
    t, tkeep = cosmo_tables_struct(growth, chi2z)
    pp = pack_params(**params)
    cnt = C.c_int64(0)
    st = ctx.lib.hnt_pack_records(ctx.handle, C.byref(b), C.byref(_results_struct(results)),
                                  abi.as_ptr(_ptr(fc), C.c_float),
                                  abi.as_ptr(_ptr(d2), C.c_float), C.byref(t), C.byref(pp),
                                  abi.as_ptr(out.data_ptr(), C.c_float), int(capacity),
                                  C.byref(cnt))
    if st != abi.HNT_ERR_CAPACITY:
        abi.check(ctx.lib, st)
    del keep, tkeep
    return out[:min(int(cnt.value), int(capacity))]


def finalize_eulerian(records, growth, obs=(0.0, 0.0, 0.0), *, Om, OL, ievol=1, z_out=0.0,
                      chi2z=None, out=None):
    """Eulerian positions and km/s velocities of a raw catalog on the device
    (hnt_finalize_eulerian; Merger.finalize_eulerian, src/Merger/Merger.jl:107-166).

    records: (N, 11) or (N, 33) float32 rows, device or host; growth: the three
    cosmology.dlinear_tables columns; chi2z: cosmology.build_chi_to_z(cosmo, z_max=10.0)
    for ievol = 1 (the reference's table).  out: a device float32 tensor of the records'
    shape (records itself for an in-place update); a new one by default.  Returns out."""
    torch = _require_gpu()
    dev = records.device if isinstance(records, torch.Tensor) and records.is_cuda else \
        torch.device("cuda", torch.cuda.current_device())
    rec = torch.as_tensor(records, dtype=torch.float32).to(dev).contiguous()
    if rec.ndim != 2 or rec.shape[1] not in (11, 33):
        raise ValueError(f"records must be (N, 11) or (N, 33), got {tuple(rec.shape)}")
    if len(growth) < 3:
        raise ValueError("growth must hold (atab, D_atab, dlnD_dlnatab)")
    ctx = context(dev.index or 0)
    if out is None:
        out = torch.empty_like(rec)
    elif (not out.is_cuda or out.dtype != torch.float32 or out.shape != rec.shape
          or not out.is_contiguous()):
        raise ValueError("out must be a contiguous device float32 tensor of the records' shape")
    t, keep = cosmo_tables_struct(growth, chi2z if int(ievol) == 1 else None)
    o = (C.c_double * 3)(*[float(v) for v in obs])
    abi.check(ctx.lib, ctx.lib.hnt_finalize_eulerian(
        ctx.handle, abi.as_ptr(rec.data_ptr(), C.c_float), int(rec.shape[0]), int(rec.shape[1]),
        C.byref(t), float(Om), float(OL), o, int(ievol), float(z_out),
        abi.as_ptr(out.data_ptr(), C.c_float)))
    del keep
    return out


def merge_catalog(records, *, volume_reduction: bool = False, fortran_ties: bool = False):
    """Lagrangian exclusion, then optionally volume reduction, of a halo catalog on the
    device (hnt_merge_catalog; Merger.merge_catalog, src/Merger/Merger.jl:23-89).

    records: (N, 11) or (N, 33) float32 HaloRecord / ExtHaloRecord rows, on the device or
    the host.  Returns the survivors in catalog order as a device float32 tensor with RTHL
    replaced by the reduced radius (unchanged without volume_reduction)."""
    torch = _require_gpu()
    dev = records.device if isinstance(records, torch.Tensor) and records.is_cuda else \
        torch.device("cuda", torch.cuda.current_device())
    rec = torch.as_tensor(records, dtype=torch.float32).to(dev).contiguous()
    if rec.ndim != 2 or rec.shape[1] not in (11, 33):
        raise ValueError(f"records must be (N, 11) or (N, 33), got {tuple(rec.shape)}")
    ctx = context(dev.index or 0)
    out = torch.empty((max(rec.shape[0], 1), rec.shape[1]), dtype=torch.float32, device=dev)
    cnt = C.c_int64(0)
    abi.check(ctx.lib, ctx.lib.hnt_merge_catalog(
        ctx.handle, abi.as_ptr(rec.data_ptr(), C.c_float), int(rec.shape[0]),
        int(rec.shape[1]), int(bool(volume_reduction)), int(bool(fortran_ties)),
        abi.as_ptr(out.data_ptr(), C.c_float), C.byref(cnt)))
    return out[:int(cnt.value)]


def analyse_peaks_gpu_cuda_multirf(delta, etax, etay, etaz, eta2x, eta2y, eta2z, batches,
                                   stab: ShellTables, ct_table, X1, X2, Y1, Y2, Z1, Z2, alatt,
                                   ZZon, *, fcrit_override=None, ct=None, rmax2rs=0.0,
                                   threads=128, ct_out_val=-1.0, lapd=None, mask=None,
                                   nbuff=0, variant="C"):
    """Mirror of analyse_peaks_gpu_cuda_multirf (src/Huntian.jl:259-281,
    ext/CUDAExt.jl:2085-2344).

    batches: list of (peaks_i, peaks_j, peaks_k, Rf, ir2min[, ZZon_pp, fcrit_pp]).
    The fields are uploaded once and every batch runs in ONE library call (all peaks carry
    their own Rf and ir2min).  Returns (results, mask): results[b] is a dict of host
    Float32 arrays with the fields of the reference's NamedTuple; mask is the updated host
    Int8 array or None.  `threads` is accepted for signature compatibility and ignored."""
    torch = _require_gpu()
    ctd = make_ctab(ct_table, X1, X2, Y1, Y2, Z1, Z2, ct_out_val)
    if fcrit_override is not None:
        fcrit = float(fcrit_override)
    elif ct is not None:
        fcrit = fsc_of_z(ct, float(ZZon) - 1.0)
    else:
        fcrit = 1.686
    fields = dict(delta=delta, etax=etax, etay=etay, etaz=etaz, eta2x=eta2x, eta2y=eta2y,
                  eta2z=eta2z, lapd=lapd)
    n = int(delta.shape[0])
    keys = ("pi", "pj", "pk", "Rf", "ir2", "zz", "fc")
    cat = {k: [] for k in keys}
    sizes = []
    has_pp = any(len(b) >= 7 for b in batches)
    for b in batches:
        npb = len(b[0])
        sizes.append(npb)
        cat["pi"].append(np.asarray(b[0], np.int32))
        cat["pj"].append(np.asarray(b[1], np.int32))
        cat["pk"].append(np.asarray(b[2], np.int32))
        cat["Rf"].append(np.full(npb, float(b[3])))
        cat["ir2"].append(np.full(npb, int(b[4]), np.int32))
        if has_pp:
            cat["zz"].append(np.asarray(b[5], np.float64) if len(b) >= 7 else np.full(npb, float(ZZon)))
            cat["fc"].append(np.asarray(b[6], np.float64) if len(b) >= 7 else np.full(npb, fcrit))
    cc = {k: (np.concatenate(v) if v else None) for k, v in cat.items()}
    mask_d = to_device_field(mask, dtype=torch.int8) if mask is not None else None
    tot = int(sum(sizes))
    r = shell_analyse(fields, cc["pi"], cc["pj"], cc["pk"], cc["Rf"], cc["ir2"], stab, ctd,
                      alatt=alatt, ZZon=ZZon, fcrit=fcrit, nbuff=nbuff, variant=variant,
                      early_exit=True, mask=mask_d,
                      ZZon_pp=cc["zz"] if has_pp else None,
                      fcrit_pp=cc["fc"] if has_pp else None, rmax2rs=rmax2rs)
    host = _host_results(r, tot)
    results = []
    off = 0
    for npb in sizes:
        sl = slice(off, off + npb)
        d = {k: host[k][sl].astype(np.float32) for k in abi.RESULT_SCALARS}
        for k in abi.RESULT_VEC3:
            d[k] = np.ascontiguousarray(host[k][sl].T.astype(np.float32))  # (3, npeaks)
        d["strain_final"] = np.ascontiguousarray(
            np.transpose(host["strain"][sl], (1, 2, 0)).astype(np.float32))  # (3,3,npeaks)
        results.append(d)
        off += npb
    mask_out = None
    if mask_d is not None:
        mask_out = np.asfortranarray(mask_d.cpu().numpy())
    return results, mask_out


def analyse_peak_gpu_cuda(delta, etax, etay, etaz, eta2x, eta2y, eta2z, peaks_i, peaks_j,
                          peaks_k, stab, ct_table, X1, X2, Y1, Y2, Z1, Z2, alatt, ir2min,
                          ZZon, Rfclvi, *, fcrit_override=None, ct=None, rmax2rs=0.0,
                          threads=128, ct_out_val=-1.0, lapd=None, mask=None, nbuff=0,
                          variant="C"):
    """Mirror of analyse_peak_gpu_cuda (src/Huntian.jl:219-257, ext/CUDAExt.jl:1914-2078):
    one batch, scalar ZZon / fcrit; returns the reference's NamedTuple fields plus mask."""
    res, m = analyse_peaks_gpu_cuda_multirf(
        delta, etax, etay, etaz, eta2x, eta2y, eta2z,
        [(peaks_i, peaks_j, peaks_k, Rfclvi, ir2min)], stab, ct_table, X1, X2, Y1, Y2, Z1, Z2,
        alatt, ZZon, fcrit_override=fcrit_override, ct=ct, rmax2rs=rmax2rs, threads=threads,
        ct_out_val=ct_out_val, lapd=lapd, mask=mask, nbuff=nbuff, variant=variant)
    out = dict(res[0])
    out["mask"] = m
    return out


# ---------------------------------------------------------------- primitive harnesses

def shell_gather_full_gpu(delta, etax, etay, etaz, eta2x, eta2y, eta2z, peaks_i, peaks_j,
                          peaks_k, stab: ShellTables, threads=128, lapd=None, return_lapd=False):
    """Mirror of shell_gather_full_gpu (src/Huntian.jl:146-166): per-shell sums, Float64,
    shaped like the reference's (nshells, npeaks) / (3, nshells, npeaks) /
    (3, 3, nshells, npeaks) host arrays."""
    torch = _require_gpu()
    ctx = context()
    fs, keep = _fields_struct(dict(delta=delta, etax=etax, etay=etay, etaz=etaz, eta2x=eta2x,
                                   eta2y=eta2y, eta2z=eta2z, lapd=lapd))
    npk = len(peaks_i)
    ns = stab.nshells
    dev = torch.device("cuda", ctx.device)
    pi_ = _dev1(peaks_i, torch.int32, dev)
    pj_ = _dev1(peaks_j, torch.int32, dev)
    pk_ = _dev1(peaks_k, torch.int32, dev)
    F = torch.empty((max(npk, 1), ns), dtype=torch.float64, device=dev)
    N = torch.empty((max(npk, 1), ns), dtype=torch.int32, device=dev)
    L = torch.empty_like(F)
    S = torch.empty((max(npk, 1), ns, 3), dtype=torch.float64, device=dev)
    S2 = torch.empty_like(S)
    G = torch.empty_like(S)
    Gf = torch.empty_like(S)
    SR = torch.empty((max(npk, 1), ns, 3, 3), dtype=torch.float64, device=dev)  # [.., K, L]
    ts = stab.struct()
    d = lambda t: abi.as_ptr(t.data_ptr(), C.c_double)  # noqa: E731
    abi.check(ctx.lib, ctx.lib.hnt_shell_gather_full(
        ctx.handle, C.byref(fs), npk, abi.as_ptr(pi_.data_ptr(), C.c_int32),
        abi.as_ptr(pj_.data_ptr(), C.c_int32), abi.as_ptr(pk_.data_ptr(), C.c_int32),
        C.byref(ts), d(F), abi.as_ptr(N.data_ptr(), C.c_int32), d(S), d(S2), d(G), d(Gf),
        d(SR), d(L)))
    del keep
    h = lambda t: t[:npk].cpu().numpy()  # noqa: E731
    Fh, Nh, Sh, S2h, Gh, Gfh, SRh, Lh = map(h, (F, N, S, S2, G, Gf, SR, L))
    out = (Fh.T, Nh.T, np.transpose(Sh, (2, 1, 0)), np.transpose(S2h, (2, 1, 0)),
           np.transpose(Gh, (2, 1, 0)), np.transpose(Gfh, (2, 1, 0)),
           np.transpose(SRh, (3, 2, 1, 0)))  # SR -> [L, K, s, p]
    return out + (Lh.T,) if return_lapd else out


def eig3_symmetric_batch_gpu(mats, threads=128):
    """Mirror of eig3_symmetric_batch_gpu (src/Huntian.jl:168-180): mats (3, 3, N) ->
    ascending eigenvalues (3, N), Float64."""
    torch = _require_gpu()
    ctx = context()
    m = np.asarray(mats, np.float64)
    N = m.shape[2]
    flat = np.ascontiguousarray(np.transpose(m, (2, 1, 0)).reshape(N, 9))  # per matrix col-major
    dm = torch.as_tensor(flat, device="cuda")
    out = torch.empty((max(N, 1), 3), dtype=torch.float64, device="cuda")
    abi.check(ctx.lib, ctx.lib.hnt_eig3_symmetric_batch(
        ctx.handle, abi.as_ptr(dm.data_ptr(), C.c_double), N,
        abi.as_ptr(out.data_ptr(), C.c_double)))
    return out[:N].cpu().numpy().T.copy()


def interp3_trilinear_gpu(table, X1, X2, Y1, Y2, Z1, Z2, xs, ys, zs, out_val=-1.0):
    """Mirror of interp3_trilinear_gpu (src/Huntian.jl:182-195), Float64."""
    torch = _require_gpu()
    ctx = context()
    ctd = make_ctab(table, X1, X2, Y1, Y2, Z1, Z2, out_val)
    cs, keep = _ctab_struct(ctd)
    x = torch.as_tensor(np.asarray(xs, np.float64), device="cuda")
    y = torch.as_tensor(np.asarray(ys, np.float64), device="cuda")
    z = torch.as_tensor(np.asarray(zs, np.float64), device="cuda")
    n = len(x)
    out = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
    abi.check(ctx.lib, ctx.lib.hnt_interp3_trilinear(
        ctx.handle, C.byref(cs), abi.as_ptr(x.data_ptr(), C.c_double),
        abi.as_ptr(y.data_ptr(), C.c_double), abi.as_ptr(z.data_ptr(), C.c_double), n,
        abi.as_ptr(out.data_ptr(), C.c_double)))
    del keep
    return out[:n].cpu().numpy()


def kernel_strain_gpu(rad, Gshell, Gshellf, SRshell, mp, mlow, mupp, wRnor, aRnor, hlatt_1,
                      hlatt_2):
    """Mirror of kernel_strain_gpu (src/Huntian.jl:197-217).  Gshell, Gshellf: (3, ns);
    SRshell: (3, 3, ns); 1-based mp/mlow/mupp.  Returns (Ebar 3x3, grad 3, gradf 3)."""
    ctx = context()
    rad = np.ascontiguousarray(rad, np.float64)
    ns = len(rad)
    G = np.ascontiguousarray(np.asarray(Gshell, np.float64).T)       # [s, L]
    Gf = np.ascontiguousarray(np.asarray(Gshellf, np.float64).T)
    SR = np.ascontiguousarray(np.transpose(np.asarray(SRshell, np.float64), (2, 0, 1)))  # [s, L, K]
    out = np.zeros(12)
    d = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    abi.check(ctx.lib, ctx.lib.hnt_kernel_strain(ctx.handle, d(rad), d(G), d(Gf), d(SR), ns,
                                                 mp, mlow, mupp, wRnor, aRnor, hlatt_1,
                                                 hlatt_2, d(out)))
    E = np.array([[out[0], out[1], out[2]], [out[1], out[3], out[4]], [out[2], out[4], out[5]]])
    return E, out[6:9].copy(), out[9:12].copy()


# ---------------------------------------------------------------- tile field synthesis

def pk_table(pk, boxsize_local: float, n: int, n_input: int, pk_table_n: int = 8192):
    """The sqrt(P(k) dk^3 n2^3) table of isolated_convolve_gpu (ext/CUDAExt.jl:2489-2497):
    pk_table_n points uniform in ln k from ln(dk/2) to ln(1.5 sqrt(3) (n2/2) dk), n2 = 2
    n_input, dk = 2 pi / (n2 boxsize_local / n).  Returns (float32 values, log_k_min, d_log_k)
    with both scalars rounded to float32, as the reference passes them."""
    n2 = 2 * int(n_input)
    dx = float(boxsize_local) / int(n)
    dk = 2.0 * np.pi / (n2 * dx)
    k_max = np.sqrt(3.0) * (n2 / 2) * dk
    log_k = np.linspace(np.log(dk * 0.5), np.log(k_max * 1.5), int(pk_table_n))
    norm = dk ** 3 * float(n2) ** 3
    kk = np.exp(log_k)
    try:
        pv = np.asarray(pk(kk), dtype=np.float64)
        if pv.shape != kk.shape:
            raise ValueError
    except Exception:
        pv = np.array([float(pk(k)) for k in kk])
    vals = np.sqrt(pv * norm).astype(np.float32)
    return vals, np.float32(log_k[0]), np.float32(log_k[1] - log_k[0])


def _field_out(n, dev, k=1):
    torch = _torch()
    return [fortran_empty((n, n, n), torch.float32, dev) for _ in range(k)]


def _host_or_device(outs, return_device):
    if return_device:
        return outs
    return [o.cpu().numpy() for o in outs]


def _fptrs(outs):
    return (C.POINTER(C.c_float) * len(outs))(*[abi.as_ptr(o.data_ptr(), C.c_float)
                                               for o in outs])


def isolated_convolve_gpu_multi(noise, pk, boxsize_local, n, *, kernels, nshell=0,
                                pk_table_n=8192, threads=256, return_device=False):
    """Mirror of isolated_convolve_gpu_multi (ext/CUDAExt.jl:2538-2622): one zero-padded
    forward transform of the noise, then per (kernel_fn_id, dim1, dim2) the transfer, the
    Hermitian DC/Nyquist planes, the inverse transform and the n^3 crop (hnt_isolated_convolve).
    noise: (n_input,)*3 indexed [i, j, k], host or device.  Returns one array per kernel
    (host numpy, or column-major device tensors with return_device)."""
    torch = _require_gpu()
    x = to_device_field(noise)
    n_input = int(x.shape[0])
    dev = x.device
    ctx = context(dev.index or 0)
    vals, lkmin, dlk = pk_table(pk, boxsize_local, n, n_input, pk_table_n)
    tab = abi.PkTable(abi.as_ptr(vals.ctypes.data, C.c_float), len(vals), float(lkmin),
                      float(dlk))
    ks = list(kernels)
    tr = (abi.Transfer * len(ks))(*[abi.Transfer(int(a), int(b), int(c)) for a, b, c in ks])
    outs = _field_out(int(n), dev, len(ks))
    abi.check(ctx.lib, ctx.lib.hnt_isolated_convolve(
        ctx.handle, abi.as_ptr(x.data_ptr(), C.c_float), n_input, int(n),
        float(boxsize_local), C.byref(tab), tr, len(ks), int(nshell), _fptrs(outs)))
    del threads
    return _host_or_device(outs, return_device)


def isolated_convolve_gpu(noise, pk, boxsize_local, n, *, kernel_fn_id=0, dim1=0, dim2=0,
                          nshell=0, pk_table_n=8192, threads=256, return_device=False):
    """Mirror of isolated_convolve_gpu (ext/CUDAExt.jl:2475-2536)."""
    return isolated_convolve_gpu_multi(noise, pk, boxsize_local, n,
                                       kernels=[(kernel_fn_id, dim1, dim2)], nshell=nshell,
                                       pk_table_n=pk_table_n, threads=threads,
                                       return_device=return_device)[0]


def compute_2lpt_gpu(delta_tile, nmesh, boxsize_local, *, threads=256, return_device=False):
    """Mirror of compute_2lpt_gpu (ext/CUDAExt.jl:2724-2791): the three periodic 2LPT
    displacements of the tile's delta (hnt_compute_2lpt)."""
    _require_gpu()
    x = to_device_field(delta_tile)
    if tuple(x.shape) != (int(nmesh),) * 3:
        raise ValueError(f"delta_tile must be {nmesh}^3, got {tuple(x.shape)}")
    ctx = context(x.device.index or 0)
    outs = _field_out(int(nmesh), x.device, 3)
    abi.check(ctx.lib, ctx.lib.hnt_compute_2lpt(ctx.handle, abi.as_ptr(x.data_ptr(), C.c_float),
                                                int(nmesh), float(boxsize_local), _fptrs(outs)))
    return _host_or_device(outs, return_device)


def compute_laplacian_gpu(delta_tile, nmesh, boxsize_local, *, threads=256,
                          return_device=False):
    """Mirror of compute_laplacian_gpu (ext/CUDAExt.jl:2793-2825): irfft(k^2 rfft(delta))."""
    _require_gpu()
    x = to_device_field(delta_tile)
    if tuple(x.shape) != (int(nmesh),) * 3:
        raise ValueError(f"delta_tile must be {nmesh}^3, got {tuple(x.shape)}")
    ctx = context(x.device.index or 0)
    out = _field_out(int(nmesh), x.device)[0]
    abi.check(ctx.lib, ctx.lib.hnt_compute_laplacian(
        ctx.handle, abi.as_ptr(x.data_ptr(), C.c_float), int(nmesh), float(boxsize_local),
        abi.as_ptr(out.data_ptr(), C.c_float)))
    return out if return_device else out.cpu().numpy()


def isolated_poisson_psi_gpu(dS, boxsize_local, *, threads=256):
    """Mirror of isolated_poisson_psi_gpu (ext/CUDAExt.jl:2679-2701): a tuple of three
    column-major device tensors."""
    _require_gpu()
    x = to_device_field(dS)
    n = int(x.shape[0])
    ctx = context(x.device.index or 0)
    outs = _field_out(n, x.device, 3)
    abi.check(ctx.lib, ctx.lib.hnt_isolated_poisson_psi(
        ctx.handle, abi.as_ptr(x.data_ptr(), C.c_float), n, float(boxsize_local), _fptrs(outs)))
    return tuple(outs)


def isolated_short_pot_gpu(delta, boxsize_local, rs):
    """Mirror of isolated_short_pot_gpu (ext/CUDAExt.jl:2703-2722): a column-major device
    tensor."""
    _require_gpu()
    x = to_device_field(delta)
    n = int(x.shape[0])
    ctx = context(x.device.index or 0)
    out = _field_out(n, x.device)[0]
    abi.check(ctx.lib, ctx.lib.hnt_isolated_short_pot(
        ctx.handle, abi.as_ptr(x.data_ptr(), C.c_float), n, float(boxsize_local), float(rs),
        abi.as_ptr(out.data_ptr(), C.c_float)))
    return out


def interpolate_to_tile_gpu(coarse, it, jt, kt, nsub, nmesh, N, M, *, threads=256,
                            return_device=False):
    """Mirror of interpolate_to_tile_gpu (ext/CUDAExt.jl:3142-3176): tri-cubic Catmull-Rom
    interpolation of the periodic M^3 coarse grid onto tile (it, jt, kt) (1-based)."""
    _require_gpu()
    x = to_device_field(coarse)
    if tuple(x.shape) != (int(M),) * 3:
        raise ValueError(f"coarse must be {M}^3, got {tuple(x.shape)}")
    ctx = context(x.device.index or 0)
    out = _field_out(int(nmesh), x.device)[0]
    abi.check(ctx.lib, ctx.lib.hnt_interpolate_to_tile(
        ctx.handle, abi.as_ptr(x.data_ptr(), C.c_float), int(M), int(N), int(it), int(jt),
        int(kt), int(nsub), int(nmesh), abi.as_ptr(out.data_ptr(), C.c_float)))
    return out if return_device else out.cpu().numpy()


# ---------------------------------------------------------------- tile white noise

def generate_tile_noise(it, jt, kt, nsub, nmesh, N, seed, *, return_device=False):
    """Mirror of _generate_tile_noise (src/MultiResolution.jl:35-57) on the device
    (hnt_tile_noise): the Threefry Gaussian stream of the global N^3 grid over tile
    (it, jt, kt)'s nmesh^3 cells (1-based tile indices, wrapped mod N), Float32."""
    torch = _require_gpu()
    ctx = context()
    out = _field_out(int(nmesh), torch.device("cuda", ctx.device))[0]
    abi.check(ctx.lib, ctx.lib.hnt_tile_noise(
        ctx.handle, int(it), int(jt), int(kt), int(nsub), int(nmesh), int(N), int(seed),
        abi.as_ptr(out.data_ptr(), C.c_float)))
    return out if return_device else out.cpu().numpy()


def downsample_noise(N, M, seed, *, return_device=False):
    """Mirror of _downsample_noise (src/MultiResolution.jl:64-87) on the device
    (hnt_downsample_noise): each coarse cell the FP64 sum of its (N/M)^3 fine Gaussians over
    sqrt((N/M)^3), Float32 [M^3]."""
    torch = _require_gpu()
    ctx = context()
    out = _field_out(int(M), torch.device("cuda", ctx.device))[0]
    abi.check(ctx.lib, ctx.lib.hnt_downsample_noise(
        ctx.handle, int(N), int(M), int(seed), abi.as_ptr(out.data_ptr(), C.c_float)))
    return out if return_device else out.cpu().numpy()


def generate_extended_residual(it, jt, kt, nsub, nmesh, N, seed, coarse_noise, M, nshell, *,
                               return_device=False):
    """Mirror of _generate_extended_residual (src/MultiResolution.jl:190-227) on the device
    (hnt_extended_residual): the Hoffman-Ribak residual fine - block mean of coarse_noise
    (the M^3 output of downsample_noise) over the (nmesh + 2 nshell)^3 region around tile
    (it, jt, kt), Float32."""
    _require_gpu()
    c = to_device_field(coarse_noise)
    if tuple(c.shape) != (int(M),) * 3:
        raise ValueError(f"coarse_noise must be {M}^3, got {tuple(c.shape)}")
    ctx = context(c.device.index or 0)
    side = int(nmesh) + 2 * int(nshell)
    out = _field_out(side, c.device)[0]
    abi.check(ctx.lib, ctx.lib.hnt_extended_residual(
        ctx.handle, int(it), int(jt), int(kt), int(nsub), int(nmesh), int(N), int(seed),
        abi.as_ptr(c.data_ptr(), C.c_float), int(M), int(nshell),
        abi.as_ptr(out.data_ptr(), C.c_float)))
    return out if return_device else out.cpu().numpy()

"""TEST INFRASTRUCTURE: ctypes wrapper of the FP64 C oracle (libhuntian_oracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, and
only as the checker / CPU baseline.  The product (huntian_hip) never imports it.

The oracle restates the reference CPU path (Huntian.jl src/HaloFinder/PeakFind.jl,
src/HaloFinder/RadialShell.jl, src/ShellAnalysisGPU.jl,
src/EllipsoidalCollapse/CollapseTable.jl); see huntian_oracle.c for line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "huntian.jl_amd"))

from huntian_hip import abi  # noqa: E402  (struct layouts only)

LIB = os.path.join(HERE, "libhuntian_oracle.so")

_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i8p = C.POINTER(C.c_int8)

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.hro_shell_cells.argtypes = [C.c_int32]
        L.hro_shell_cells.restype = C.c_int64
        L.hro_shell_count.argtypes = [C.c_int32]
        L.hro_shell_count.restype = C.c_int32
        L.hro_build_shell_tables.argtypes = [C.c_int32] + [_i32p] * 6
        L.hro_build_shell_tables.restype = C.c_int32
        L.hro_hRinteg.argtypes = [C.c_double, C.c_double]
        L.hro_hRinteg.restype = C.c_double
        L.hro_atab4.argtypes = [_f64p]
        L.hro_kernel_strain.argtypes = [_f64p, _f64p, _f64p, _f64p, C.c_int32, C.c_int32,
                                        C.c_int32, C.c_double, C.c_double, C.c_double,
                                        C.c_double, _f64p]
        L.hro_eig3.argtypes = [_f64p, _f64p]
        L.hro_interp3.argtypes = [C.POINTER(abi.CTab), C.c_double, C.c_double, C.c_double]
        L.hro_interp3.restype = C.c_double
        L.hro_fsc_of_z.argtypes = [C.POINTER(abi.CTab), C.c_double]
        L.hro_fsc_of_z.restype = C.c_double
        L.hro_find_peaks.argtypes = [_f32p, _i8p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                     C.c_float, C.c_int64, _i32p, _i32p, _i32p, _f32p]
        L.hro_find_peaks.restype = C.c_int64
        L.hro_peak_find_multi.argtypes = [C.POINTER(_f32p), _f32p, C.c_int32, C.c_int32,
                                          C.c_int32, _i8p, C.c_int64, _i32p, _i32p, _i32p,
                                          _i32p, _f32p]
        L.hro_peak_find_multi.restype = C.c_int64
        L.hro_peak_find_multi_par.argtypes = L.hro_peak_find_multi.argtypes + [C.c_int32]
        L.hro_peak_find_multi_par.restype = C.c_int64
        L.hro_shell_gather_full.argtypes = [C.POINTER(abi.Fields), C.c_int64, _i32p, _i32p,
                                            _i32p, C.POINTER(abi.ShellTables)] + [_f64p, _i32p] + [_f64p] * 6
        L.hro_analyse_peaks.argtypes = [C.POINTER(abi.Fields), C.POINTER(abi.Batch),
                                        C.POINTER(abi.ShellTables), C.POINTER(abi.CTab),
                                        C.POINTER(abi.Params), C.POINTER(abi.Results), _i8p,
                                        C.c_int32]
        L.hro_analyse_peaks.restype = C.c_int
        L.hro_pack_records.argtypes = [C.c_int64, _i32p, _i32p, _i32p, _f64p,
                                       C.POINTER(abi.Results), _f32p, _f32p,
                                       C.POINTER(abi.CosmoTables), C.POINTER(abi.PackParams),
                                       _f32p, C.c_int64]
        L.hro_pack_records.restype = C.c_int64
        L.hro_merge_catalog.argtypes = [_f32p, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                        _f32p]
        L.hro_merge_catalog.restype = C.c_int64
        L.hro_merge_passes.argtypes = [_f32p, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                       C.POINTER(C.c_uint8), _f64p]
        L.hro_merge_passes.restype = C.c_int32
        L.hro_sphere_overlap.argtypes = [C.c_double, C.c_double, C.c_double, _f64p, _f64p]
        L.hro_growth_f.argtypes = [C.POINTER(abi.CosmoTables), C.c_double]
        L.hro_growth_f.restype = C.c_double
        L.hro_finalize_eulerian.argtypes = [_f32p, C.c_int64, C.c_int32,
                                            C.POINTER(abi.CosmoTables), C.c_double, C.c_double,
                                            _f64p, C.c_int32, C.c_double, _f32p]
        L.hro_growth_D.argtypes = [C.POINTER(abi.CosmoTables), C.c_double]
        L.hro_growth_D.restype = C.c_double
        L.hro_chi_to_z.argtypes = [C.POINTER(abi.CosmoTables), C.c_double]
        L.hro_chi_to_z.restype = C.c_double
        L.hro_pow_m3.argtypes = [C.c_double]
        L.hro_pow_m3.restype = C.c_double
        L.hro_threefry2x64_20.argtypes = [C.c_uint64] * 4 + [C.POINTER(C.c_uint64)]
        L.hro_threefry_gauss.argtypes = [C.c_uint64, C.c_uint64]
        L.hro_threefry_gauss.restype = C.c_double
        L.hro_tile_noise.argtypes = [C.c_int32] * 5 + [C.c_int64, C.c_uint64, _f32p]
        L.hro_downsample_noise.argtypes = [C.c_int64, C.c_int32, C.c_uint64, _f32p]
        L.hro_extended_residual.argtypes = ([C.c_int32] * 5 +
                                            [C.c_int64, C.c_uint64, _f32p, C.c_int32,
                                             C.c_int32, _f32p])
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else C.cast(None, C.POINTER(ct))


# ---------------------------------------------------------------- shells / small kernels

def shell_tables(rmax: int) -> dict:
    L = lib()
    nc = L.hro_shell_cells(rmax)
    ns = L.hro_shell_count(rmax)
    t = {k: np.zeros(nc, np.int32) for k in ("di", "dj", "dk")}
    t.update({k: np.zeros(ns, np.int32) for k in ("shell_start", "shell_count", "shell_r2")})
    got = L.hro_build_shell_tables(rmax, *[_p(t[k], C.c_int32) for k in
                                           ("di", "dj", "dk", "shell_start", "shell_count",
                                            "shell_r2")])
    assert got == ns
    t["rmax"] = rmax
    t["nshells"] = ns
    t["ncells"] = nc
    return t


def tables_struct(t: dict) -> abi.ShellTables:
    return abi.ShellTables(t["rmax"], t["nshells"], t["ncells"],
                           *[_p(t[k], C.c_int32) for k in ("di", "dj", "dk", "shell_start",
                                                           "shell_count", "shell_r2")])


def ctab_struct(ct: dict) -> abi.CTab:
    tab = np.ascontiguousarray(ct["table"].ravel(order="F"), dtype=np.float32)
    ct["_flat"] = tab  # keep alive
    return abi.CTab(ct["nx"], ct["ny"], ct["nz"], ct["x1"], ct["x2"], ct["y1"], ct["y2"],
                    ct["z1"], ct["z2"], ct.get("out_val", -1.0), _p(tab, C.c_float))


def hRinteg(u0, u12):
    return lib().hro_hRinteg(float(u0), float(u12))


def atab4():
    out = np.zeros(1101)
    lib().hro_atab4(_p(out, C.c_double))
    return out


def kernel_strain(rad, G, Gf, SR, mp, mlow, mupp, wRnor, aRnor, h1=1.0, h2=1.0):
    """rad [ns]; G, Gf [ns,3]; SR [ns,3,3] (L,K); 1-based mp/mlow/mupp.
    Returns (Ebar 3x3, grad 3, gradf 3)."""
    rad = np.ascontiguousarray(rad, np.float64)
    G = np.ascontiguousarray(G, np.float64)
    Gf = np.ascontiguousarray(Gf, np.float64)
    SR = np.ascontiguousarray(SR, np.float64)
    out = np.zeros(12)
    lib().hro_kernel_strain(_p(rad, C.c_double), _p(G, C.c_double), _p(Gf, C.c_double),
                            _p(SR, C.c_double), mp, mlow, mupp, wRnor, aRnor, h1, h2,
                            _p(out, C.c_double))
    e = out
    E = np.array([[e[0], e[1], e[2]], [e[1], e[3], e[4]], [e[2], e[4], e[5]]])
    return E, out[6:9].copy(), out[9:12].copy()


def eig3(A):
    A = np.asfortranarray(np.asarray(A, np.float64).reshape(3, 3))
    flat = np.ascontiguousarray(A.ravel(order="F"))
    lam = np.zeros(3)
    lib().hro_eig3(_p(flat, C.c_double), _p(lam, C.c_double))
    return lam


def interp3(ct: dict, x, y, z):
    s = ctab_struct(ct)
    return lib().hro_interp3(C.byref(s), float(x), float(y), float(z))


def fsc_of_z(ct: dict, z: float) -> float:
    s = ctab_struct(ct)
    return lib().hro_fsc_of_z(C.byref(s), float(z))


# ---------------------------------------------------------------- peak finding

def find_peaks(delta, mask, nbuff, fcrit, max_peaks=2_000_000):
    """One filter, as PeakFind.find_peaks.  delta/mask are (n1,n2,n3) Fortran-ordered
    arrays; mask is updated in place.  Returns (i, j, k, delta) 1-based arrays."""
    d = np.asfortranarray(delta, dtype=np.float32)
    assert mask.flags.f_contiguous and mask.dtype == np.int8
    n1, n2, n3 = d.shape
    cap = max(1, min(max_peaks, d.size))
    pi = np.zeros(cap, np.int32); pj = np.zeros(cap, np.int32); pk = np.zeros(cap, np.int32)
    pd = np.zeros(cap, np.float32)
    np_ = lib().hro_find_peaks(_p(d, C.c_float), _p(mask, C.c_int8), n1, n2, n3, nbuff,
                               float(fcrit), max_peaks, _p(pi, C.c_int32), _p(pj, C.c_int32),
                               _p(pk, C.c_int32), _p(pd, C.c_float))
    return pi[:np_], pj[:np_], pk[:np_], pd[:np_]


def peak_find_multi(fields, fcrits, nbuff, mask, cap=None, nthreads=1):
    """The filter loop of find_peaks scans, mask carried.  nthreads > 1 splits each filter's
    scan over k-planes (same output, hro_peak_find_multi_par)."""
    fs = [np.asfortranarray(f, dtype=np.float32) for f in fields]
    n = fs[0].shape[0]
    if cap is None:
        cap = max(1, min((n - 2 * nbuff) ** 3, 1 << 26))
    arr = (_f32p * len(fs))(*[_p(f, C.c_float) for f in fs])
    fc = np.asarray(fcrits, np.float32)
    out = {k: np.zeros(cap, np.int32) for k in ("i", "j", "k", "filt")}
    out["delta"] = np.zeros(cap, np.float32)
    args = (arr, _p(fc, C.c_float), len(fs), n, nbuff, _p(mask, C.c_int8), cap,
            _p(out["i"], C.c_int32), _p(out["j"], C.c_int32), _p(out["k"], C.c_int32),
            _p(out["filt"], C.c_int32), _p(out["delta"], C.c_float))
    if nthreads > 1:
        tot = lib().hro_peak_find_multi_par(*args, int(nthreads))
        assert tot >= 0, "oracle peak_find_multi_par: out of host memory"
    else:
        tot = lib().hro_peak_find_multi(*args)
    cnt = min(tot, cap)
    return {k: v[:cnt] for k, v in out.items()}, tot


# ---------------------------------------------------------------- shell analysis

def fields_struct(flds: dict) -> abi.Fields:
    d = flds["delta"]
    n1, n2, n3 = d.shape
    keep = {}
    ptrs = []
    for k in ("delta", "etax", "etay", "etaz", "eta2x", "eta2y", "eta2z", "lapd"):
        a = flds.get(k)
        if a is None:
            ptrs.append(C.cast(None, _f32p))
        else:
            a = np.asfortranarray(a, dtype=np.float32)
            keep[k] = a
            ptrs.append(_p(a, C.c_float))
    s = abi.Fields(n1, n2, n3, *ptrs)
    s._keep = keep
    return s


def shell_gather_full(flds: dict, pi, pj, pk, t: dict):
    fs = fields_struct(flds)
    ts = tables_struct(t)
    npk = len(pi)
    ns = t["nshells"]
    pi = np.ascontiguousarray(pi, np.int32); pj = np.ascontiguousarray(pj, np.int32)
    pk = np.ascontiguousarray(pk, np.int32)
    F = np.zeros((npk, ns)); n = np.zeros((npk, ns), np.int32); lp = np.zeros((npk, ns))
    S = np.zeros((npk, ns, 3)); S2 = np.zeros((npk, ns, 3)); G = np.zeros((npk, ns, 3))
    Gf = np.zeros((npk, ns, 3)); SR = np.zeros((npk, ns, 3, 3))  # SR[..., K, L] (L fastest)
    lib().hro_shell_gather_full(C.byref(fs), npk, _p(pi, C.c_int32), _p(pj, C.c_int32),
                                _p(pk, C.c_int32), C.byref(ts), _p(F, C.c_double),
                                _p(n, C.c_int32), _p(S, C.c_double), _p(S2, C.c_double),
                                _p(G, C.c_double), _p(Gf, C.c_double), _p(SR, C.c_double),
                                _p(lp, C.c_double))
    SR = np.swapaxes(SR, -1, -2)  # -> SR[..., L, K]
    return dict(F=F, n=n, S=S, S2=S2, G=G, Gf=Gf, SR=SR, lapd=lp)


def alloc_results(npk: int, with_used=True):
    r = {k: np.zeros(npk) for k in abi.RESULT_SCALARS}
    r.update({k: np.zeros((npk, 3)) for k in abi.RESULT_VEC3})
    r["strain"] = np.zeros((npk, 3, 3))  # [p, K, L] column-major per peak
    for k in abi.RESULT_INT:
        r[k] = np.zeros(npk, np.int32) if with_used else None
    return r


def results_struct(r: dict) -> abi.Results:
    vals = []
    for name, _ in abi.Results._fields_:
        a = r.get(name)
        if a is None:
            vals.append(C.cast(None, _i32p if name in abi.RESULT_I32 else _f64p))
        else:
            vals.append(_p(a, C.c_int32 if name in abi.RESULT_I32 else C.c_double))
    return abi.Results(*vals)


def analyse_peaks(flds: dict, pi, pj, pk, Rf, ir2min, t: dict, ct: dict, *, alatt, ZZon,
                  fcrit, nbuff=0, variant="A", early_exit=False, mask=None, nthreads=1,
                  ZZon_pp=None, fcrit_pp=None, rmax2rs=0.0):
    """Variant "A" = RadialShell.analyse_peak, "B" = analyse_peak_gpu, "C" = B + zvir_half.
    Returns a dict of per-peak arrays (strain as [p, L, K])."""
    npk = len(pi)
    fs = fields_struct(flds)
    ts = tables_struct(t)
    cs = ctab_struct(ct)
    pi = np.ascontiguousarray(pi, np.int32); pj = np.ascontiguousarray(pj, np.int32)
    pk = np.ascontiguousarray(pk, np.int32)
    Rf = np.ascontiguousarray(np.broadcast_to(np.asarray(Rf, np.float64), (npk,)))
    ir2 = np.ascontiguousarray(np.broadcast_to(np.asarray(ir2min, np.int32), (npk,)))
    zz = np.ascontiguousarray(ZZon_pp, np.float64) if ZZon_pp is not None else None
    fc = np.ascontiguousarray(fcrit_pp, np.float64) if fcrit_pp is not None else None
    b = abi.Batch(npk, _p(pi, C.c_int32), _p(pj, C.c_int32), _p(pk, C.c_int32),
                  _p(Rf, C.c_double), _p(ir2, C.c_int32), _p(zz, C.c_double),
                  _p(fc, C.c_double), float(ZZon), float(fcrit))
    prm = abi.Params(float(alatt), int(nbuff), float(rmax2rs), abi.VARIANTS[variant],
                     int(bool(early_exit)))
    r = alloc_results(npk)
    rs = results_struct(r)
    if mask is not None:
        assert mask.flags.f_contiguous and mask.dtype == np.int8
    st = lib().hro_analyse_peaks(C.byref(fs), C.byref(b), C.byref(ts), C.byref(cs),
                                 C.byref(prm), C.byref(rs), _p(mask, C.c_int8), nthreads)
    if st != 0:
        raise ValueError(f"oracle analyse_peaks failed with code {st}")
    r["strain"] = np.swapaxes(r["strain"], 1, 2)  # -> [p, L, K]
    return r


# ---------------------------------------------------------------- halo records

def cosmo_tables(growth, chi2z=None):
    """abi.CosmoTables over host arrays (kept alive in the returned list); growth = (atab,
    D_atab[, dlnD_dlnatab])."""
    at = np.ascontiguousarray(growth[0], np.float64)
    Dt = np.ascontiguousarray(growth[1], np.float64)
    keep = [at, Dt]
    if chi2z is None:
        t = abi.CosmoTables(len(at), _p(at, C.c_double), _p(Dt, C.c_double), 0, None, None, 0.0)
    else:
        zs = np.ascontiguousarray(chi2z[0], np.float64)
        ch = np.ascontiguousarray(chi2z[1], np.float64)
        keep += [zs, ch]
        t = abi.CosmoTables(len(at), _p(at, C.c_double), _p(Dt, C.c_double), len(zs),
                            _p(zs, C.c_double), _p(ch, C.c_double), float(chi2z[2]))
    if len(growth) > 2:
        ft = np.ascontiguousarray(growth[2], np.float64)
        keep.append(ft)
        t.dlnD_atab = _p(ft, C.c_double)
    return t, keep


def growth_f(growth, a):
    """Dlinear_ab(a)[2], the growth rate dlnD/dlna."""
    t, keep = cosmo_tables(growth)
    return lib().hro_growth_f(C.byref(t), float(a))


def finalize_eulerian(records, growth, obs=(0.0, 0.0, 0.0), *, Om, OL, ievol=1, z_out=0.0,
                      chi2z=None):
    """Merger.finalize_eulerian on an (N, 11 or 33) float32 record array."""
    rec = np.ascontiguousarray(records, np.float32)
    n, nf = rec.shape
    out = np.zeros_like(rec)
    t, keep = cosmo_tables(growth, chi2z if ievol == 1 else None)
    o = np.ascontiguousarray(obs, np.float64)
    lib().hro_finalize_eulerian(_p(rec, C.c_float), n, nf, C.byref(t), float(Om), float(OL),
                                _p(o, C.c_double), int(ievol), float(z_out), _p(out, C.c_float))
    return out


def growth_D(growth, a):
    t, keep = cosmo_tables(growth)
    return lib().hro_growth_D(C.byref(t), float(a))


def chi_to_z(chi2z, chi, growth=((0.1, 0.2, 0.3, 0.4), (1.0, 1.0, 1.0, 1.0))):
    t, keep = cosmo_tables(growth, chi2z)
    return lib().hro_chi_to_z(C.byref(t), float(chi))


def pack_records(pi, pj, pk, Rf, r: dict, FcRf, *, growth, chi2z=None, d2Rf=None,
                 capacity=None, alatt, n, xbx=0.0, ybx=0.0, zbx=0.0, ZZon=1.0, ievol=0,
                 obs=(0.0, 0.0, 0.0), z_max=0.0, ilpt=2, Om=0.31, OL=0.69, ioutshear=0):
    """The record loop of MultiResolution.jl:1159-1219 on host arrays; r holds FP64 results
    (strain symmetric).  Returns (records (min(nhalo, capacity), 11 or 33) float32, nhalo)."""
    npk = len(pi)
    pi = np.ascontiguousarray(pi, np.int32)
    pj = np.ascontiguousarray(pj, np.int32)
    pk = np.ascontiguousarray(pk, np.int32)
    Rf = np.ascontiguousarray(np.broadcast_to(np.asarray(Rf, np.float64), (npk,)))
    fc = np.ascontiguousarray(FcRf, np.float32)
    d2 = np.ascontiguousarray(d2Rf, np.float32) if d2Rf is not None else None
    rr = {k: np.ascontiguousarray(v) if v is not None else None for k, v in r.items()}
    rs = results_struct(rr)
    t, keep = cosmo_tables(growth, chi2z)
    pp = abi.PackParams(float(alatt), int(n), float(xbx), float(ybx), float(zbx), float(ZZon),
                        int(ievol), (C.c_double * 3)(*[float(v) for v in obs]), float(z_max),
                        int(ilpt), float(Om), float(OL), int(ioutshear))
    nf = 33 if ioutshear >= 1 else 11
    cap = npk if capacity is None else int(capacity)
    out = np.zeros((max(cap, 1), nf), np.float32)
    nrec = lib().hro_pack_records(npk, _p(pi, C.c_int32), _p(pj, C.c_int32), _p(pk, C.c_int32),
                                  _p(Rf, C.c_double), C.byref(rs), _p(fc, C.c_float),
                                  _p(d2, C.c_float), C.byref(t), C.byref(pp),
                                  _p(out, C.c_float), cap)
    return out[:min(nrec, cap)], int(nrec)


# ---------------------------------------------------------------- catalog merge

def merge_catalog(records, volume_reduction=False, fortran_ties=False):
    """Merger.merge_catalog on an (N, 11 or 33) float32 record array: (survivors, count)."""
    rec = np.ascontiguousarray(records, np.float32)
    n, nf = rec.shape
    out = np.zeros((max(n, 1), nf), np.float32)
    k = lib().hro_merge_catalog(_p(rec, C.c_float), n, nf, int(bool(volume_reduction)),
                                int(bool(fortran_ties)), _p(out, C.c_float))
    _merge_status(k)
    return out[:k], int(k)


def _merge_status(k):
    if k == -1:
        raise MemoryError("oracle merge_catalog: allocation failed")
    if k == -2:
        raise ValueError("oracle merge_catalog: empty domain")


def merge_passes(records, volume_reduction=False, fortran_ties=False):
    """The two passes alone: (survived bool [N], new radii float64 [N])."""
    rec = np.ascontiguousarray(records, np.float32)
    n, nf = rec.shape
    surv = np.zeros(max(n, 1), np.uint8)
    newr = np.zeros(max(n, 1), np.float64)
    _merge_status(lib().hro_merge_passes(_p(rec, C.c_float), n, nf, int(bool(volume_reduction)),
                                         int(bool(fortran_ties)), _p(surv, C.c_uint8),
                                         _p(newr, C.c_double)))
    return surv[:n].astype(bool), newr[:n]


def sphere_overlap(d, r1, r2):
    v1, v2 = C.c_double(), C.c_double()
    lib().hro_sphere_overlap(float(d), float(r1), float(r2), C.byref(v1), C.byref(v2))
    return v1.value, v2.value


# ---------------------------------------------------------------- tile white noise

def threefry2x64_20(c0, c1, k0, k1):
    out = (C.c_uint64 * 2)()
    lib().hro_threefry2x64_20(c0, c1, k0, k1, out)
    return int(out[0]), int(out[1])


def threefry_gauss(seed, idx0):
    return lib().hro_threefry_gauss(int(seed), int(idx0))


def tile_noise(it, jt, kt, nsub, nmesh, N, seed):
    out = np.empty((nmesh, nmesh, nmesh), np.float32, order="F")
    lib().hro_tile_noise(it, jt, kt, nsub, nmesh, int(N), int(seed), _p(out, C.c_float))
    return out


def downsample_noise(N, M, seed):
    out = np.empty((M, M, M), np.float32, order="F")
    lib().hro_downsample_noise(int(N), int(M), int(seed), _p(out, C.c_float))
    return out


def extended_residual(it, jt, kt, nsub, nmesh, N, seed, coarse, M, nshell):
    side = nmesh + 2 * nshell
    out = np.empty((side, side, side), np.float32, order="F")
    c = np.asfortranarray(coarse, np.float32)
    lib().hro_extended_residual(it, jt, kt, nsub, nmesh, int(N), int(seed), _p(c, C.c_float),
                                int(M), int(nshell), _p(out, C.c_float))
    return out

"""Pins the oracle's peak finder: the nine known-answer cases of test/test_peakfind.jl and a
brute-force 26-neighbour scan written with numpy (independent of the oracle's loops)."""
import numpy as np
import pytest


def F(n, fill=0.0):
    return np.full((n, n, n), fill, dtype=np.float32, order="F")


def M(n):
    return np.zeros((n, n, n), dtype=np.int8, order="F")


def test_single_peak(oracle):
    d = F(11)
    d[5, 5, 5] = 2.0
    m = M(11)
    i, j, k, v = oracle.find_peaks(d, m, 1, 1.0)
    assert (list(i), list(j), list(k)) == ([6], [6], [6]) and v[0] == 2.0
    assert m[5, 5, 5] == 1  # mask mutation


def test_below_threshold_and_gradient(oracle):
    assert len(oracle.find_peaks(F(11, 0.5), M(11), 1, 1.0)[0]) == 0
    ii, jj, kk = np.meshgrid(*[np.arange(1, 12)] * 3, indexing="ij")
    g = np.asfortranarray((ii + jj + kk).astype(np.float32))
    assert len(oracle.find_peaks(g, M(11), 1, 1.0)[0]) == 0


def test_mask_and_buffer(oracle):
    d = F(11)
    d[5, 5, 5] = 2.0
    m = M(11)
    m[5, 5, 5] = 1
    assert len(oracle.find_peaks(d, m, 1, 1.0)[0]) == 0
    d = F(11)
    d[0, 5, 5] = 2.0
    assert len(oracle.find_peaks(d, M(11), 1, 1.0)[0]) == 0


def test_scan_order_and_flat_index(oracle):
    d = F(15)
    d[3, 3, 3] = 2.0
    d[11, 11, 11] = 3.0
    i, j, k, _ = oracle.find_peaks(d, M(15), 1, 1.0)
    assert list(zip(i, j, k)) == [(4, 4, 4), (12, 12, 12)]
    d = F(11)
    d[2, 4, 6] = 2.0
    i, j, k, _ = oracle.find_peaks(d, M(11), 1, 1.0)
    assert i[0] + (j[0] - 1) * 11 + (k[0] - 1) * 121 == 3 + 4 * 11 + 6 * 121


def test_max_peaks(oracle):
    d = F(15)
    for (a, b, c) in [(3, 3, 3), (3, 3, 11), (11, 11, 3), (11, 11, 11)]:
        d[a, b, c] = 2.0
    assert len(oracle.find_peaks(d, M(15), 1, 1.0, max_peaks=2)[0]) == 2


def test_ties_are_peaks(oracle):
    # `ff < neighbour` rejects; equal neighbours do not: a flat plateau is all peaks
    d = F(7, 2.0)
    i, _, _, _ = oracle.find_peaks(d, M(7), 1, 1.0)
    assert len(i) == 5 ** 3


def brute_force(fields, fcrits, nbuff, mask0):
    """numpy restatement: cell is a peak of the first filter where it is unmasked,
    >= fcrit and no 26-neighbour is larger."""
    n = fields[0].shape[0]
    sl = slice(nbuff, n - nbuff)
    claimed = np.zeros((n, n, n), bool)
    out = []
    for f, (d, fc) in enumerate(zip(fields, fcrits)):
        nb = np.full(d.shape, -np.inf, np.float32)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    nb[1:-1, 1:-1, 1:-1] = np.maximum(
                        nb[1:-1, 1:-1, 1:-1],
                        d[1 + dx:n - 1 + dx, 1 + dy:n - 1 + dy, 1 + dz:n - 1 + dz])
        ok = np.zeros((n, n, n), bool)
        ok[sl, sl, sl] = True
        ok &= (mask0 != 1) & ~claimed & ~(d < fc) & ~(d < nb)
        idx = np.argwhere(ok)  # C order of (i, j, k) -> sort to (k, j, i)
        idx = idx[np.lexsort((idx[:, 0], idx[:, 1], idx[:, 2]))]
        for (a, b, c) in idx:
            out.append((a + 1, b + 1, c + 1, f, d[a, b, c]))
        claimed |= ok
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_multi_filter_vs_brute_force(oracle, seed):
    rng = np.random.default_rng(seed)
    n, nbuff = 24, 2
    fields = [np.asfortranarray(rng.standard_normal((n, n, n)).astype(np.float32) * s)
              for s in (1.0, 1.3, 0.8)]
    fields[1][5:8, 5:8, 5:8] = 3.0  # a plateau: ties
    fc = [0.5, 1.0, 0.2]
    mask0 = M(n)
    mask0[rng.random((n, n, n)) < 0.05] = 1
    mask = mask0.copy(order="F")
    got, tot = oracle.peak_find_multi(fields, fc, nbuff, mask)
    ref = brute_force(fields, fc, nbuff, mask0)
    assert tot == len(ref) > 20
    assert list(zip(got["i"], got["j"], got["k"], got["filt"], got["delta"])) == \
        [tuple(r) for r in ref]
    exp_mask = mask0.copy(order="F")
    for (a, b, c, _, _) in ref:
        exp_mask[a - 1, b - 1, c - 1] = 1
    assert np.array_equal(mask, exp_mask)


@pytest.mark.parametrize("cap", [None, 150])
def test_threaded_scan_equals_serial_scan(oracle, cap):
    """The k-plane-threaded filter loop (used on the large tiles) returns the serial scan's
    peaks, total and mask, also under a capacity cut inside a filter (the cut peaks are
    counted, neither written nor masked, and a later filter may count their cells again)."""
    rng = np.random.default_rng(9)
    n, nbuff = 40, 3
    fields = [np.asfortranarray(rng.standard_normal((n, n, n)).astype(np.float32) * s)
              for s in (1.0, 1.2, 0.9, 1.1)]
    fields[2][10:14, 10:14, 10:14] = 4.0  # plateau ties
    fields[3][rng.random((n, n, n)) < 0.003] = np.nan
    fc = [1.0, 0.8, 0.5, 0.7]
    mask0 = M(n)
    mask0[rng.random((n, n, n)) < 0.04] = 1
    m_ser, m_par = mask0.copy(order="F"), mask0.copy(order="F")
    ser, tot_s = oracle.peak_find_multi(fields, fc, nbuff, m_ser, cap=cap)
    par, tot_p = oracle.peak_find_multi(fields, fc, nbuff, m_par, cap=cap, nthreads=4)
    assert tot_s == tot_p and len(ser["i"]) > 100
    for k in ser:
        assert np.array_equal(ser[k], par[k], equal_nan=True), k
    assert np.array_equal(m_ser, m_par)

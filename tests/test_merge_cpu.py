"""Catalog merge: the oracle's restatement of Merger.merge_catalog (src/Merger/Merger.jl:23-89,
src/Merger/Exclusion.jl) pinned by the reference's own cases (test/test_merger.jl), the
textbook lens volume of two spheres, a hash-free brute-force greedy exclusion, an all-pairs
volume reduction, and the summation order of the reduction checked against an explicit sum.
Catalogs are synthetic: the reference's MersenneTwister draws are replaced by numpy draws
of the same distributions."""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest

# the C library's cbrt, as the oracle calls it (numpy's vectorised cbrt differs in the last
# place for about half of all arguments)
_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.cbrt.restype = ctypes.c_double
_libm.cbrt.argtypes = [ctypes.c_double]


def mk(x, y, z, R, vel=(0, 0, 0), vel2=(0, 0, 0), od=0.0):
    return [x, y, z, *vel, R, *vel2, od]


def cat(rows):
    return np.array(rows, np.float32).reshape(-1, 11)


def lens_volume(d, r1, r2):
    """Volume of the intersection of two spheres (the lens), textbook closed form."""
    return (math.pi * (r1 + r2 - d) ** 2 *
            (d * d + 2 * d * r2 - 3 * r2 * r2 + 2 * d * r1 + 6 * r1 * r2 - 3 * r1 * r1) / (12 * d))


def test_sphere_overlap_reference_cases_and_lens_volume(oracle):
    so = oracle.sphere_overlap
    assert so(10.0, 2.0, 3.0) == (0.0, 0.0)   # apart
    assert so(5.0, 2.0, 3.0) == (0.0, 0.0)    # touching
    assert so(0.5, 3.0, 1.0) == (0.0, 0.0)    # one inside the other
    assert so(0.0, 2.0, 2.0) == (0.0, 0.0)
    v1, v2 = so(2.0, 2.0, 2.0)                # equal spheres at d = r: caps of height r/2
    assert v1 == v2 and v1 == pytest.approx(math.pi / 3 * 1.0 * 5.0, rel=1e-15)
    v1, v2 = so(3.0, 3.0, 1.0)
    assert 0 < v1 < v2
    # the two caps make up the lens, for equal and unequal spheres
    for d, r1, r2 in [(3.0, 2.0, 2.0), (1.0, 2.0, 2.0), (4.0, 3.0, 2.0), (2.5, 3.0, 1.0),
                      (0.7, 0.5, 0.4)]:
        v1, v2 = so(d, r1, r2)
        assert v1 + v2 == pytest.approx(lens_volume(d, r1, r2), rel=1e-13)
    d, r = 1.3, 1.0                            # equal spheres: pi/12 (4r + d)(2r - d)^2
    assert sum(so(d, r, r)) == pytest.approx(math.pi / 12 * (4 * r + d) * (2 * r - d) ** 2,
                                             rel=1e-14)


def test_reference_merge_cases(oracle):
    # test_merger.jl "merge_catalog -- synthetic HaloRecords"
    h = cat([mk(0, 0, 0, 5, od=1.0), mk(1, 1, 1, 1, od=0.5), mk(7, 0, 0, 3, od=0.8),
             mk(50, 50, 50, 2, od=0.3)])
    out, n = oracle.merge_catalog(h)
    assert n == 3 and np.array_equal(out, h[[0, 2, 3]])
    # chain: only the largest survives
    h = cat([mk(0, 0, 0, 10), mk(1, 0, 0, 5), mk(1.5, 0, 0, 2)])
    out, n = oracle.merge_catalog(h)
    assert n == 1 and out[0, 6] == 10
    # no overlaps, single halo, fields preserved
    h = cat([mk(0, 0, 0, 1), mk(100, 0, 0, 1), mk(0, 100, 0, 1)])
    assert np.array_equal(oracle.merge_catalog(h)[0], h)
    h = cat([mk(0, 0, 0, 5, vel=(1, 2, 3), od=1.0)])
    assert np.array_equal(oracle.merge_catalog(h)[0], h)
    h = cat([mk(0, 0, 0, 1, (10, 20, 30), (4, 5, 6), 0.5), mk(50, 0, 0, 2, (-1, -2, -3),
                                                             (-4, -5, -6), 0.8)])
    assert np.array_equal(oracle.merge_catalog(h)[0], h)
    # equal radii: the first in (-r, x, y, z) order kills the other
    h = cat([mk(0, 0, 0, 2), mk(0.5, 0, 0, 2)])
    out, n = oracle.merge_catalog(h)
    assert n == 1 and out[0, 0] == 0
    # ExtHaloRecord: every field kept
    ext = np.arange(33, dtype=np.float32)[None, :] * 0.01
    ext[0, :3] = 0
    ext[0, 6] = 5
    assert np.array_equal(oracle.merge_catalog(ext)[0], ext)
    # empty input
    out, n = oracle.merge_catalog(np.zeros((0, 11), np.float32))
    assert n == 0 and out.shape == (0, 11)


def test_reference_volume_reduction_cases(oracle):
    h = cat([mk(50, 50, 50, 3.0), mk(54, 50, 50, 2.0)])
    v1, v2 = oracle.sphere_overlap(4.0, 3.0, 2.0)
    assert np.array_equal(oracle.merge_catalog(h)[0][:, 6], [3, 2])
    out, _ = oracle.merge_catalog(h, volume_reduction=True)
    want = [np.cbrt(27 - 3 * v1 / (4 * math.pi)), np.cbrt(8 - 3 * v2 / (4 * math.pi))]
    np.testing.assert_allclose(out[:, 6], want, rtol=1e-6)
    # equal radii: reduced once, or twice with fortran_ties
    h = cat([mk(50, 50, 50, 2.0), mk(53, 50, 50, 2.0)])
    v1t, _ = oracle.sphere_overlap(3.0, 2.0, 2.0)
    out, _ = oracle.merge_catalog(h, volume_reduction=True)
    np.testing.assert_allclose(out[:, 6], np.cbrt(8 - 3 * v1t / (4 * math.pi)), rtol=1e-6)
    out, _ = oracle.merge_catalog(h, volume_reduction=True, fortran_ties=True)
    np.testing.assert_allclose(out[:, 6], np.cbrt(8 - 6 * v1t / (4 * math.pi)), rtol=1e-6)
    # far apart: unchanged
    h = cat([mk(0, 0, 0, 2.0), mk(100, 0, 0, 3.0)])
    assert np.array_equal(oracle.merge_catalog(h, volume_reduction=True)[0], h)


def brute_exclusion(rec):
    """Greedy exclusion without a hash: walk the halos in (-r, x, y, z) order (stable), each
    survivor kills the survivors of no larger radius whose centres lie inside it."""
    x, y, z, r = (rec[:, c].astype(np.float64) for c in (0, 1, 2, 6))
    order = np.lexsort((z, y, x, -r))
    alive = np.ones(len(rec), bool)
    for i in order:
        if not alive[i]:
            continue
        dx, dy, dz = x - x[i], y - y[i], z - z[i]
        dist = np.sqrt(dx * dx + dy * dy + dz * dz)
        kill = alive & (r <= r[i]) & (dist < r[i])
        kill[i] = False
        alive &= ~kill
    return alive


def tie_heavy(n, box, seed):
    rng = np.random.default_rng(seed)
    radii = np.array([1.0, 1.4142135, 1.7320508, 2.0, 2.236068], np.float32)
    rec = np.zeros((n, 11), np.float32)
    rec[:, :3] = rng.uniform(0, box, (n, 3))
    rec[:, 6] = rng.choice(radii, n)
    rec[:, 10] = np.arange(n)
    return rec


def test_exclusion_matches_brute_force_and_ignores_input_order(oracle):
    rec = tie_heavy(4000, 60.0, 11)
    surv, _ = oracle.merge_passes(rec)
    assert np.array_equal(surv, brute_exclusion(rec))
    assert 1000 < surv.sum() < 4000
    out, _ = oracle.merge_catalog(rec)
    key = lambda c: sorted(map(tuple, c[:, [0, 1, 2, 6]].tolist()))
    ref = key(out)
    for s in range(3):
        perm = np.random.default_rng(s).permutation(len(rec))
        assert key(oracle.merge_catalog(rec[perm])[0]) == ref
    # continuous radii, and halos spread over a larger box (a different hash grid)
    rng = np.random.default_rng(2)
    rec = np.zeros((3000, 11), np.float32)
    rec[:, :3] = rng.uniform(-500, 500, (3000, 3)) * [1, 0.2, 0.05]
    rec[:, 6] = rng.uniform(1, 12, 3000)
    assert np.array_equal(oracle.merge_passes(rec)[0], brute_exclusion(rec))


def test_volume_reduction_matches_all_pairs(oracle):
    """test_merger.jl's random-set check: after exclusion, every overlapping survivor pair
    cuts each member's own cap; radii from the summed caps (any order, rtol)."""
    rng = np.random.default_rng(7)
    rec = np.zeros((600, 11), np.float32)
    rec[:, :3] = rng.uniform(0, 100, (600, 3))
    rec[:, 6] = 1 + 4 * rng.uniform(size=600)
    ex, _ = oracle.merge_catalog(rec)
    red, _ = oracle.merge_catalog(rec, volume_reduction=True)
    dV = np.zeros(len(ex))
    for a in range(len(ex)):
        for b in range(a + 1, len(ex)):
            p, q = ex[a].astype(np.float64), ex[b].astype(np.float64)
            d = math.sqrt(((p[:3] - q[:3]) ** 2).sum())
            if d < p[6] + q[6]:
                va, vb = oracle.sphere_overlap(d, p[6], q[6])
                dV[a] += va
                dV[b] += vb
    r3 = ex[:, 6].astype(np.float64) ** 3 - 3 * dV / (4 * math.pi)
    keep = r3 > 0
    assert (dV > 0).sum() > 50
    assert len(red) == keep.sum()
    np.testing.assert_array_equal(red[:, :6], ex[keep, :6])
    np.testing.assert_allclose(red[:, 6], np.cbrt(r3[keep]), rtol=1e-5)


def test_volume_reduction_summation_order(oracle):
    """The reduction adds to a halo first the caps cut by larger halos, in their rank order,
    then its own caps against smaller halos in its cell-scan order (Exclusion.jl:172-189).
    For a halo with two larger and two smaller overlapping neighbours, the FP64 radius must
    equal cbrt of the volume less that ordered sum, and the case is chosen so that each of
    three other orders (own caps first, larger halos reversed, interleaved) gives other bits."""
    rng = np.random.default_rng(42)
    pi43 = 4 * math.pi / 3
    for trial in range(3000):
        c = np.array([50.0, 50.0, 50.0])
        rows = [mk(*c, 3.0)]
        # two larger halos (radii 4.5 and 4.0) and two smaller (2.0, 1.5), all overlapping
        # the centre halo without containing or being contained
        dirs = rng.normal(size=(4, 3))
        dirs /= np.linalg.norm(dirs, axis=1)[:, None]
        for R, dist, u in zip((4.5, 4.0, 2.0, 1.5), (5.0, 4.8, 3.6, 3.2), dirs):
            rows.append(mk(*(c + dist * u * rng.uniform(0.97, 1.0)), R))
        rec = cat(rows)
        surv, newr = oracle.merge_passes(rec, volume_reduction=True)
        if not surv.all():
            continue
        p = rec[:, :3].astype(np.float64)
        r = rec[:, 6].astype(np.float64)

        def cap(a, b):  # cap that halo a loses to halo b
            d = math.sqrt(((p[a] - p[b]) ** 2).sum())
            return oracle.sphere_overlap(d, r[a], r[b])[0] if d < r[a] + r[b] else 0.0

        incoming = [cap(0, 1), cap(0, 2)]  # halos 1 and 2 are larger: ranks 0 and 1
        own = [cap(0, 3), cap(0, 4)]
        if min(incoming + own) <= 0.0:
            continue
        # own caps in the reference's cell order: the hash cells of halos 3 and 4 relative
        # to halo 0, scanned by dz, then dy, then dx; same cell: higher index first
        s = 0.0
        for v in incoming:
            s += v
        ox = rec[:, :3].min(0).astype(np.float64) - 2 * r.max()
        L = (rec[:, :3].max(0).astype(np.float64) + 2 * r.max() - ox).max()
        cs = L / 16
        cell = np.clip(np.floor((p - ox) / cs).astype(int) + 1, 1, 16)
        k3 = (cell[3, 2], cell[3, 1], cell[3, 0], -3)
        k4 = (cell[4, 2], cell[4, 1], cell[4, 0], -4)
        for v in (own if k3 < k4 else own[::-1]):
            s += v
        want = _libm.cbrt((pi43 * (r[0] * r[0] * r[0]) - s) * 3 / (4 * math.pi))
        others = set()
        for seq in ([own[0], own[1], incoming[0], incoming[1]],
                    [incoming[1], incoming[0], own[0], own[1]],
                    [incoming[0], own[0], incoming[1], own[1]]):
            t = 0.0
            for v in seq:
                t += v
            others.add(_libm.cbrt((pi43 * (r[0] * r[0] * r[0]) - t) * 3 / (4 * math.pi)))
        if want in others:
            continue  # some other order rounds alike: try another layout
        assert newr[0] == want, trial
        return
    pytest.fail("no layout separated the summation orders")


def test_empty_domain_is_an_error(oracle):
    with pytest.raises(ValueError):
        oracle.merge_catalog(cat([mk(1, 2, 3, 0.0)]))

"""The oracle's tile white noise: Threefry-2x64-20 (RandomField._threefry_gaussian,
src/InitialConditions/RandomField.jl:45-53), _generate_tile_noise, _downsample_noise and
_generate_extended_residual (src/MultiResolution.jl:35-87, 190-227).

Pins: the Random123 known-answer vector of threefry2x64_20 (counter and key zero); the
synthetic-input generator's independent implementation on random words; the Box-Muller pair
and its even/odd split written from the definition; the reference's own RandomField tests
(statistics, skewness / kurtosis, a region equal to the full field, decomposition
independence); the coarse grid as block sums of the full field; the residual's zero block
means (what makes T * residual short-range) and its definition fine - mean."""
import math

import numpy as np
import pytest


def test_threefry_known_answer(oracle):
    # Random123 kat_vectors: threefry2x64_20, counter {0, 0}, key {0, 0}
    assert oracle.threefry2x64_20(0, 0, 0, 0) == (0xC2B6E3A8C2C69865, 0x6F81ED42F350084D)


def test_threefry_matches_the_independent_numpy_form(oracle):
    from huntian_hip import synth

    rng = np.random.default_rng(1)
    for _ in range(50):
        c0, c1, k0, k1 = (int(v) for v in rng.integers(0, 2**64, 4, dtype=np.uint64))
        a = oracle.threefry2x64_20(c0, c1, k0, k1)
        b = synth.threefry2x64_20(np.uint64(c0), np.uint64(c1), np.uint64(k0), np.uint64(k1))
        assert a == (int(b[0]), int(b[1]))


def test_gaussian_pair_from_its_definition(oracle):
    seed = 12345
    for idx0 in (0, 1, 2, 3, 1000, 1001, 2**40 + 7):
        w0, w1 = oracle.threefry2x64_20(idx0 >> 1, 0, seed, 0)
        u1 = ((w0 >> 11) + 1) * 2.0 ** -53
        u2 = ((w1 >> 11) + 1) * 2.0 ** -53
        r = math.sqrt(-2.0 * math.log(u1))
        g = r * (math.cos if idx0 % 2 == 0 else math.sin)(2.0 * math.pi * u2)
        assert oracle.threefry_gauss(seed, idx0) == g


def test_noise_statistics(oracle):
    # test/test_randomfield.jl:107-128: mean 0, variance 1, skewness 0, kurtosis 3
    x = oracle.tile_noise(1, 1, 1, 96, 96, 96, 42).astype(np.float64).ravel()
    assert abs(x.mean()) < 0.01 and abs(x.var() - 1.0) < 0.01
    z = (x - x.mean()) / x.std()
    assert abs((z ** 3).mean()) < 0.02 and abs((z ** 4).mean() - 3.0) < 0.05


def test_tiles_are_wrapped_regions_of_the_global_field(oracle):
    N, nsub, nmesh = 32, 16, 24  # nbuff 4: edge tiles wrap
    full = oracle.tile_noise(1, 1, 1, N, N, N, 7)  # nsub = nmesh = N: the whole grid
    for (it, jt, kt) in ((1, 1, 1), (2, 1, 2), (2, 2, 2), (1, 2, 1)):
        t = oracle.tile_noise(it, jt, kt, nsub, nmesh, N, 7)
        idx = [(np.arange(nmesh) + (c - 1) * nsub - (nmesh - nsub) // 2) % N
               for c in (it, jt, kt)]
        assert np.array_equal(t, full[np.ix_(*idx)])


def test_coarse_is_the_block_sum_over_sqrt_block3(oracle):
    N, M, seed = 32, 8, 99
    full = np.array([[[oracle.threefry_gauss(seed, i + N * j + N * N * k) for k in range(N)]
                      for j in range(N)] for i in range(N)])  # FP64, [i, j, k]
    b = N // M
    want = full.reshape(M, b, M, b, M, b).sum(axis=(1, 3, 5)) / math.sqrt(b ** 3)
    got = oracle.downsample_noise(N, M, seed)
    np.testing.assert_allclose(got, want.astype(np.float32), rtol=2e-6, atol=2e-7)
    # Float32 of the same sums: the fine field is the tile noise's FP64 source
    assert np.array_equal(full.astype(np.float32), oracle.tile_noise(1, 1, 1, N, N, N, seed))


@pytest.mark.parametrize("nshell", [0, 3])
def test_residual_is_fine_minus_block_mean(oracle, nshell):
    N, M, seed = 32, 8, 5
    coarse = oracle.downsample_noise(N, M, seed)
    it, jt, kt, nsub, nmesh = 2, 1, 2, 16, 20
    res = oracle.extended_residual(it, jt, kt, nsub, nmesh, N, seed, coarse, M, nshell)
    side = nmesh + 2 * nshell
    full = oracle.tile_noise(1, 1, 1, N, N, N, seed)
    off = [(c - 1) * nsub - (nmesh - nsub) // 2 - nshell for c in (it, jt, kt)]
    idx = [(np.arange(side) + o) % N for o in off]
    fine = full[np.ix_(*idx)]
    blk = [ix // (N // M) for ix in idx]
    mean = coarse[np.ix_(*blk)] * np.float32(1.0 / math.sqrt((N // M) ** 3))
    assert np.array_equal(res, fine - mean)


def test_residual_has_zero_block_means(oracle):
    # a tile made of whole coarse blocks: every block of the residual sums to 0 up to the
    # Float32 roundings of the fine values and of the mean
    N, M, seed = 48, 12, 3
    b = N // M
    coarse = oracle.downsample_noise(N, M, seed)
    res = oracle.extended_residual(1, 1, 1, N, N, N, seed, coarse, M, 0).astype(np.float64)
    sums = res.reshape(M, b, M, b, M, b).sum(axis=(1, 3, 5))
    assert np.abs(sums).max() < 1e-4
    # against the fine field's own block sums (order 8 in magnitude)
    fine = oracle.tile_noise(1, 1, 1, N, N, N, seed).astype(np.float64)
    assert np.abs(fine.reshape(M, b, M, b, M, b).sum(axis=(1, 3, 5))).max() > 1.0

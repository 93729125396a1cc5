"""GPU primitive entry points vs the oracle (and vs LAPACK / scipy).  These run the same
FP64 operation sequences as the oracle, so agreement is bit for bit."""
import numpy as np
import pytest

from fieldgen import bumps_field
from huntian_hip import api
from huntian_hip.homeltab import make_ctab

pytestmark = pytest.mark.gpu


def test_eig3_batch_bitwise(oracle):
    from test_oracle_kernels import _reference_matrices

    mats = _reference_matrices()
    got = api.eig3_symmetric_batch_gpu(np.stack(mats, axis=2))
    for q, A in enumerate(mats):
        ref = oracle.eig3(A)
        assert np.array_equal(got[:, q], ref), (q, got[:, q], ref)
        np.testing.assert_allclose(got[:, q], np.linalg.eigvalsh(A), rtol=0,
                                   atol=4e-15 * (np.abs(ref).max() + 1e-300))


def test_interp3_bitwise(oracle):
    xs, ys, zs = np.linspace(-2, 2, 21), np.linspace(0, 3, 17), np.linspace(1, 5, 13)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    t = (np.sin(X) * Y + Z * 0.1).astype(np.float32)
    rng = np.random.default_rng(3)
    px = np.concatenate([rng.uniform(-2.2, 2.2, 400), [-2.0, 2.0, 2.0 + 1e-15]])
    py = np.concatenate([rng.uniform(-0.1, 3.1, 400), [0.0, 3.0, 1.0]])
    pz = np.concatenate([rng.uniform(0.9, 5.1, 400), [1.0, 5.0, 2.0]])
    got = api.interp3_trilinear_gpu(t, -2.0, 2.0, 0.0, 3.0, 1.0, 5.0, px, py, pz, -1.0)
    ct = make_ctab(t, -2.0, 2.0, 0.0, 3.0, 1.0, 5.0)
    ref = np.array([oracle.interp3(ct, a, b, c) for a, b, c in zip(px, py, pz)])
    assert np.array_equal(got, ref)
    assert got[-1] == -1.0 and np.sum(got == -1.0) > 10


@pytest.mark.parametrize("win", [(5, 3, 7), (10, 7, 13), (15, 10, 18), (3, 2, 5), (1, 2, 8),
                                 (1, 2, 11), (7, 4, 10)])
def test_kernel_strain_bitwise(oracle, win):
    rng = np.random.default_rng(2718)
    ns = 20
    rad = np.sqrt(np.arange(ns, dtype=np.float64))
    G, Gf = rng.standard_normal((3, ns)), rng.standard_normal((3, ns))
    SR = rng.standard_normal((3, 3, ns))
    for (wR, aR, h1, h2) in [(1.0, 2.5, 1.0, 1.0), (0.5, 1.0, 0.7, 1.3)]:
        g = api.kernel_strain_gpu(rad, G, Gf, SR, *win, wR, aR, h1, h2)
        o = oracle.kernel_strain(rad, G.T, Gf.T, np.transpose(SR, (2, 0, 1)), *win, wR, aR, h1, h2)
        for a, b in zip(g, o):
            assert np.array_equal(a, b), (win, a, b)


@pytest.mark.parametrize("with_2lpt,with_lapd", [(True, True), (False, False)])
def test_shell_gather_full_bitwise(oracle, with_2lpt, with_lapd):
    f = bumps_field(32, 314159, nbumps=3, margin=5)
    if not with_2lpt:
        for k in ("eta2x", "eta2y", "eta2z"):
            f[k] = None
    if not with_lapd:
        f["lapd"] = None
    stab = api.build_shell_tables(8)
    t = oracle.shell_tables(8)
    pi, pj, pk = [16, 2, 31, 8, 1], [16, 30, 3, 8, 32], [16, 14, 20, 31, 1]
    F, N, S, S2, G, Gf, SR, L = api.shell_gather_full_gpu(
        f["delta"], f["etax"], f["etay"], f["etaz"], f["eta2x"], f["eta2y"], f["eta2z"], pi, pj,
        pk, stab, lapd=f["lapd"], return_lapd=True)
    o = oracle.shell_gather_full(f, pi, pj, pk, t)
    assert np.array_equal(N.T, o["n"])
    assert np.array_equal(F.T, o["F"])
    assert np.array_equal(np.transpose(S, (2, 1, 0)), o["S"])
    assert np.array_equal(np.transpose(S2, (2, 1, 0)), o["S2"])
    assert np.array_equal(np.transpose(G, (2, 1, 0)), o["G"])
    assert np.array_equal(np.transpose(Gf, (2, 1, 0)), o["Gf"])
    assert np.array_equal(np.transpose(SR, (3, 2, 0, 1)), o["SR"])
    assert np.array_equal(L.T, o["lapd"])
    if not with_2lpt:
        assert np.all(S2 == 0)

"""CPU checks of the bracket network's restatement (tests/mlp_ref.py) and host helpers:
finite differences pin the FP64 backward pass, known values pin the bfloat16 rounding."""
import os
import re

import numpy as np

import mlp_ref
from huntian_hip import abi, bracket

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "huntian_abi.h")


def test_parameter_layout_matches_the_header():
    src = open(HEADER).read()
    expr = re.search(r"#define HNT_MLP_NPARAM (.+)", src).group(1)
    assert eval(expr) == abi.MLP_NPARAM == len(bracket.init_params(0))
    p = bracket.init_params(1)
    assert np.array_equal(bracket.flatten(bracket.unflatten(p)), p)
    d = mlp_ref.unflatten(p)
    assert d["W1"].shape == (18, 128) and d["W4"].shape == (128, 3)
    assert len(bracket.FEATURE_NAMES) == abi.MLP_INPUTS


def test_bf16_rounding_known_values():
    one = 1.0
    ulp = 2.0 ** -7  # bfloat16 spacing just above 1
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 2 + 2 ** -20, -2.5,
                     3.4e38, np.inf], np.float32)
    want = np.array([one, one, one + 2 * ulp, one + ulp, -2.5, np.inf, np.inf])
    got = mlp_ref.bf16_round(vals)
    assert np.array_equal(got, want)
    assert np.isnan(mlp_ref.bf16_round(np.array([np.nan], np.float32))[0])


def test_fp64_gradient_against_finite_differences():
    rng = np.random.default_rng(0)
    p = bracket.init_params(2).astype(np.float64)
    p += rng.normal(0, 0.05, p.shape)
    norm = np.concatenate([rng.normal(0, 0.3, 18), rng.uniform(0.5, 2, 18)])
    x = rng.normal(0, 1, (16, 18))
    t = np.stack([rng.uniform(0, 2, 16), rng.uniform(0, 2, 16), rng.integers(0, 2, 16)], 1)
    _, g = mlp_ref.loss_and_grad(p, norm, x, t)
    blocks = bracket.unflatten(np.arange(len(p)))
    idx = np.concatenate([rng.choice(b.ravel(), min(6, b.size), replace=False)
                          for b in blocks.values()])
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (mlp_ref.loss_and_grad(pp, norm, x, t)[0] - mlp_ref.loss_and_grad(pm, norm, x, t)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])


def test_features_restatement_on_a_symmetric_field():
    """A radially symmetric delta and psi1 = -x (pure compression): the shell means are the
    profile, e = p = 0, the mean over 12 shells is the volume average."""
    n = 21
    c = 11
    ii, jj, kk = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), np.arange(1, n + 1),
                             indexing="ij")
    r2 = (ii - c) ** 2 + (jj - c) ** 2 + (kk - c) ** 2
    f = {"delta": np.asfortranarray((5.0 / (1.0 + r2)).astype(np.float32)),
         "etax": np.asfortranarray(-(ii - c).astype(np.float32)),
         "etay": np.asfortranarray(-(jj - c).astype(np.float32)),
         "etaz": np.asfortranarray(-(kk - c).astype(np.float32))}
    from oracle import oracle as O

    t = O.shell_tables(5)
    got = mlp_ref.features_ref(f, [c], [c], [c], [4.0], 2.0, t)[0]
    r2s = t["shell_r2"][:12]
    np.testing.assert_allclose(got[:12], 5.0 / (1.0 + r2s), rtol=1e-6)
    # psi1 = -x gives T = -sum d d^T, negative definite: F <= 0 -> e = p = 0
    assert np.all(got[12:16] == 0.0)
    w = t["shell_count"][:12]
    np.testing.assert_allclose(got[16], np.sum(w * 5.0 / (1.0 + r2s)) / w.sum(), rtol=1e-6)
    assert got[17] == 2.0

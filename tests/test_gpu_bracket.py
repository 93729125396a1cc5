"""The learned bracket on the GPU: BF16 matrix-core forward and backward against the CPU
restatement (tests/mlp_ref.py), features against a brute-force loop, training, and the
learned schedule leaving every shell-analysis result unchanged."""
import numpy as np
import pytest

import mlp_ref
from huntian_hip import api, bracket, synth

pytestmark = pytest.mark.gpu


def _net(seed=3):
    rng = np.random.default_rng(seed)
    p = bracket.init_params(seed)
    d = bracket.unflatten(p.copy())
    for b in ("b1", "b2", "b3", "b4"):
        d[b][:] = rng.normal(0, 0.1, d[b].shape)
    p = bracket.flatten(d)
    norm = np.concatenate([rng.normal(0, 0.5, 18), rng.uniform(0.5, 2.0, 18)]).astype(np.float32)
    return p, norm


def _batch(B, seed=4):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1.5, (B, 18)).astype(np.float32)
    t = np.stack([rng.uniform(0, 3, B), rng.uniform(1, 4, B), rng.integers(0, 2, B)],
                 1).astype(np.float32)
    return x, t


@pytest.mark.parametrize("B", [1, 45, 300])
def test_forward_matches_cpu_restatement(B):
    import torch

    p, norm = _net()
    x, _ = _batch(B)
    net = bracket.BracketNet(p, norm)
    y = net.forward(torch.tensor(x, device="cuda")).cpu().numpy().astype(np.float64)
    yb, _ = mlp_ref.forward(p, norm, x, bf16=True)
    y64, _ = mlp_ref.forward(p, norm, x, bf16=False)
    scale = np.abs(y64).max() + 1e-3
    err_b = np.abs(y - yb).max() / scale
    err_64 = np.abs(y - y64).max() / scale
    print(f"B={B}: max err vs bf16-rounded FP64 {err_b:.2e}, vs pure FP64 {err_64:.2e}")
    assert err_b < 1e-2
    assert err_64 < 5e-2


def test_gradient_matches_cpu_restatement():
    import torch

    p, norm = _net(5)
    x, t = _batch(256, 6)
    net = bracket.BracketNet(p, norm)
    g, loss = net.gradient(torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda"))
    g = g.cpu().numpy().astype(np.float64)
    lb, gb = mlp_ref.loss_and_grad(p, norm, x, t, bf16=True)
    l64, g64 = mlp_ref.loss_and_grad(p, norm, x, t, bf16=False)
    rel_b = np.linalg.norm(g - gb) / np.linalg.norm(gb)
    rel_64 = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    # per-block agreement too, so that a wrong layer cannot hide behind a large one
    blocks = bracket.unflatten(np.arange(len(g)))
    worst = max(np.linalg.norm(g[idx.ravel()] - gb[idx.ravel()]) /
                (np.linalg.norm(gb[idx.ravel()]) + 1e-12) for idx in blocks.values())
    print(f"grad rel err vs bf16-rounded FP64 {rel_b:.2e} (worst block {worst:.2e}), "
          f"vs pure FP64 {rel_64:.2e}; loss {loss:.6f} vs {lb:.6f}")
    assert abs(loss - lb) / abs(lb) < 1e-3
    assert rel_b < 1e-2 and worst < 3e-2
    assert rel_64 < 5e-2


def test_training_reduces_the_loss():
    import torch

    rng = np.random.default_rng(8)
    x = rng.normal(0, 1, (8192, 18)).astype(np.float32)
    t = np.stack([np.tanh(x[:, 0] + 0.5 * x[:, 1]) + 1.0, 0.3 * x[:, 2] ** 2 + x[:, 16],
                  (x[:, 3] + x[:, 4] > 0)], 1).astype(np.float32)
    xd, td = torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda")
    net = bracket.BracketNet(seed=1)
    net.set_norm_from(xd)
    _, l0 = net.gradient(xd, td)
    hist = net.fit(xd, td, epochs=15, batch=512, lr=3e-3)
    _, l1 = net.gradient(xd, td)
    print(f"loss {l0:.4f} -> {l1:.4f}; epochs {['%.3f' % h for h in hist]}")
    assert net.step == 15 * (8192 // 512)
    assert l1 < 0.3 * l0


@pytest.fixture(scope="module")
def cosmo48():
    n, box = 48, 96.0
    delta, dk = synth.gaussian_field(n, box, 4242)
    f = {"delta": delta}
    f.update(synth.lpt_fields(dk, n, box))
    radii = synth.filter_radii(1.5, 12.0, 6)
    sm = [synth.smooth_tophat(dk, n, box, R) for R in radii]
    return f, radii, sm


def test_features_match_brute_force(oracle, cosmo48):
    f, radii, sm = cosmo48
    pk, tot = oracle.peak_find_multi(sm, [1.5] * len(radii), 1,
                                     np.zeros((48, 48, 48), np.int8, order="F"))
    Rf = np.array([radii[q] for q in pk["filt"]])
    t = oracle.shell_tables(6)
    stab = api.build_shell_tables(6)
    sel = np.arange(0, len(pk["i"]), max(1, len(pk["i"]) // 60))  # incl. boundary peaks
    got = bracket.features(f, pk["i"][sel], pk["j"][sel], pk["k"][sel], Rf[sel], stab,
                           2.0).cpu().numpy()
    ref = mlp_ref.features_ref(f, pk["i"][sel], pk["j"][sel], pk["k"][sel], Rf[sel], 2.0, t)
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-6)
    assert (pk["i"][sel] <= 6).any() or (pk["i"][sel] >= 43).any()


@pytest.mark.parametrize("order", ["plain", "learned"])
def test_schedules_do_not_change_results(oracle, websky, cosmo48, order):
    import torch

    f, radii, sm = cosmo48
    nbuff = 7
    pk, tot = oracle.peak_find_multi(sm, [1.5] * len(radii), nbuff,
                                     np.zeros((48, 48, 48), np.int8, order="F"))
    Rf = np.array([radii[q] for q in pk["filt"]])
    ir2 = np.array([api.ir2min_of(r, 2.0) for r in Rf], np.int32)
    stab = api.build_shell_tables(14)
    fcrit = api.fsc_of_z(websky, 0.0)
    net = None
    if order == "learned":
        p, norm = _net(11)
        net = bracket.BracketNet(p, norm)
    out = {}
    for o in ("crossing", order):
        mask = api.to_device_field(np.zeros((48, 48, 48), np.int8), dtype=torch.int8)
        r = api.shell_analyse(f, pk["i"], pk["j"], pk["k"], Rf, ir2, stab, websky, alatt=2.0,
                              ZZon=1.0, fcrit=fcrit, nbuff=nbuff, mask=mask, order=o,
                              bracket=net)
        out[o] = {k: v.cpu().numpy() for k, v in r.items()}
        out[o]["mask"] = mask.cpu().numpy()
    assert (out["crossing"]["RTHL"] > 0).sum() > 10
    for k in out["crossing"]:
        if k != "bracket":
            assert np.array_equal(out["crossing"][k], out[order][k], equal_nan=True), k
    if order == "learned":
        # the verification flags follow from the network's own outputs
        x = bracket.features(f, pk["i"], pk["j"], pk["k"], Rf, stab, 2.0)
        y = net.forward(x).cpu().numpy()
        m = stab.nshells - 1
        dec = lambda v: np.clip(np.floor(np.maximum(np.exp(v) - 1.0, 1.0) + 0.5), 1, m)
        hi = dec(y[:, 1])
        fl = out[order]["bracket"]
        bit0 = (out["crossing"]["nshells_used"] <= hi + 2)
        bit2 = (y[:, 2] > 0) == (out["crossing"]["RTHL"] > 0)
        assert np.mean((fl & 1).astype(bool) == bit0) > 0.99
        assert np.mean(((fl & 4) > 0) == bit2) > 0.99
        assert np.all(fl[out["crossing"]["RTHL"] <= 0] & 2 == 0)
    else:
        assert np.all(out[order]["bracket"] == 0)


def test_given_extents_order_like_the_crossing_pass(oracle, websky, cosmo48):
    """order='given' with the crossing run's own extents (the perfect bracket): same results,
    no pass of the library's own."""
    import torch

    f, radii, sm = cosmo48
    nbuff = 7
    pk, tot = oracle.peak_find_multi(sm, [1.5] * len(radii), nbuff,
                                     np.zeros((48, 48, 48), np.int8, order="F"))
    Rf = np.array([radii[q] for q in pk["filt"]])
    ir2 = np.array([api.ir2min_of(r, 2.0) for r in Rf], np.int32)
    stab = api.build_shell_tables(14)
    fcrit = api.fsc_of_z(websky, 0.0)
    kw = dict(alatt=2.0, ZZon=1.0, fcrit=fcrit, nbuff=nbuff)
    r0 = api.shell_analyse(f, pk["i"], pk["j"], pk["k"], Rf, ir2, stab, websky, **kw)
    ext = r0["nshells_used"].clone()
    r1 = api.shell_analyse(f, pk["i"], pk["j"], pk["k"], Rf, ir2, stab, websky, order="given",
                           extent=ext, **kw)
    for k in r0:
        assert torch.equal(torch.nan_to_num(r0[k], nan=7.0), torch.nan_to_num(r1[k], nan=7.0)), k
    with pytest.raises(ValueError):
        api.shell_analyse(f, pk["i"], pk["j"], pk["k"], Rf, ir2, stab, websky, order="given",
                          **kw)


@pytest.mark.timeout(300)
def test_trained_bracket_on_its_evaluation_tile(websky):
    """The committed network (trained on the 128^3 workload's tiles, seeds 1-8) on the seed-99
    tile of that workload: the learned schedule gives exactly the crossing schedule's
    results; the share of peaks whose bracket verified is printed."""
    import torch

    dev = torch.device("cuda", 0)
    w = synth.workload_tile("tile128", dev, seed=99)
    n, nbuff, nhunt, alatt = w["n"], w["nbuff"], w["nhunt"], w["alatt"]
    mask = api.fortran_empty((n, n, n), torch.int8, dev)
    mask.zero_()
    pf = api.peak_find(w["smoothed"], [1.5] * len(w["radii"]), nbuff, mask,
                       capacity=(n - 2 * nbuff) ** 3)
    f = pf["filt"].long()
    Rf = torch.tensor(w["radii"], dtype=torch.float64, device=dev)[f]
    ir2 = torch.tensor([api.ir2min_of(R, alatt) for R in w["radii"]], dtype=torch.int32,
                       device=dev)[f]
    stab = api.build_shell_tables(nhunt)
    fcrit = api.fsc_of_z(websky, 0.0)
    net = bracket.BracketNet.load(bracket.DEFAULT_WEIGHTS, device=dev)
    out = {}
    for o in ("crossing", "learned"):
        mask.zero_()
        r = api.shell_analyse(w["fields"], pf["i"], pf["j"], pf["k"], Rf, ir2, stab, websky,
                              alatt=alatt, ZZon=1.0, fcrit=fcrit, nbuff=nbuff, mask=mask,
                              order=o, bracket=net)
        out[o] = ({k: v.clone() for k, v in r.items()}, mask.clone())
    for k in out["crossing"][0]:
        if k != "bracket":
            assert torch.equal(torch.nan_to_num(out["crossing"][0][k], nan=7.0),
                               torch.nan_to_num(out["learned"][0][k], nan=7.0)), k
    assert torch.equal(out["crossing"][1], out["learned"][1])
    fl = out["learned"][0]["bracket"].cpu().numpy()
    halo = out["crossing"][0]["RTHL"].cpu().numpy() > 0
    print(f"seed 99, {len(fl)} candidates, {halo.sum()} halos: gather verified "
          f"{(fl & 1).astype(bool).mean():.3f}, walk verified {((fl & 2) > 0)[halo].mean():.3f} "
          f"of halos, collapse logit right {((fl & 4) > 0).mean():.3f}", flush=True)
    assert len(fl) > 5000 and (fl & 1).astype(bool).mean() > 0.5

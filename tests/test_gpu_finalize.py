"""hnt_finalize_eulerian (device) against the oracle's Merger.finalize_eulerian
(src/Merger/Merger.jl:107-166): random catalogs of HaloRecord and ExtHaloRecord rows in
snapshot (ievol = 0) and lightcone (ievol = 1) mode, bit for bit, in place and out of place,
the records of a realistic tile after the merge, and the rejected arguments."""
import numpy as np
import pytest

from huntian_hip import api, cosmology
from huntian_hip.abi import HuntianError

pytestmark = pytest.mark.gpu

COSMO = cosmology.CosmologyParams(0.31, 0.049, 0.69, 0.68)


@pytest.fixture(scope="module")
def tabs():
    return dict(growth=cosmology.dlinear_tables(COSMO),
                chi2z=cosmology.build_chi_to_z(COSMO, z_max=10.0))


def catalog(n, nf, seed, half=3000.0):
    rng = np.random.default_rng(seed)
    r = rng.normal(0.0, 1.0, (n, nf)).astype(np.float32) * 3
    r[:, :3] = rng.uniform(-half, half, (n, 3))
    r[:, 3:6] = rng.normal(0.0, 5.0, (n, 3))
    r[:, 6] = rng.uniform(1.0, 10.0, n)
    r[:, 7:10] = rng.normal(0.0, 1.0, (n, 3))
    return r


def run(oracle, tabs, rec, obs, ievol, z_out, inplace=False):
    import torch

    print(f"finalize: N={len(rec)} nf={rec.shape[1]} ievol={ievol} z_out={z_out} "
          f"inplace={inplace}", flush=True)
    kw = dict(Om=COSMO.Om, OL=COSMO.OL, ievol=ievol, z_out=z_out, chi2z=tabs["chi2z"])
    ref = oracle.finalize_eulerian(rec, tabs["growth"], obs, **kw)
    if inplace:
        d = torch.as_tensor(rec, device="cuda").clone()
        got = api.finalize_eulerian(d, tabs["growth"], obs, out=d, **kw)
        assert got.data_ptr() == d.data_ptr()
    else:
        got = api.finalize_eulerian(rec, tabs["growth"], obs, **kw)
    got = got.cpu().numpy()
    np.testing.assert_array_equal(got, ref)
    return got


@pytest.mark.parametrize("nf", [11, 33])
@pytest.mark.parametrize("ievol,z_out", [(0, 0.0), (0, 1.7), (1, 0.0)])
def test_random_catalogs_match_the_oracle(oracle, tabs, nf, ievol, z_out):
    rec = catalog(200_003, nf, seed=nf * 10 + ievol)
    got = run(oracle, tabs, rec, (-500.0, 100.0, 50.0), ievol, z_out)
    assert np.all(got[:, 7:10] == 0.0)
    assert np.array_equal(got[:, 6], rec[:, 6]) and np.array_equal(got[:, 10:], rec[:, 10:])


def test_in_place_and_the_observer_itself(oracle, tabs):
    rec = catalog(5000, 33, seed=7)
    rec[0, :3] = 0.0
    rec[0, 3:6] = 0.0
    rec[0, 7:10] = 0.0                # at the observer: r = 0, a = 1
    got = run(oracle, tabs, rec, (0.0, 0.0, 0.0), 1, 0.0, inplace=True)
    assert tuple(got[0, :6]) == (0.0,) * 6


def test_records_of_a_realistic_tile(oracle, gpu_tile, cosmo_tables):
    """The packed records of a tile, merged, then finalized: the reference's order of
    calls (merge_catalog, then finalize_eulerian)."""
    t = gpu_tile
    pf = t["pf"]
    growth = cosmo_tables["growth"]
    recs = api.pack_records(pf["i"], pf["j"], pf["k"], t["Rf"], t["r"], pf["delta"],
                            growth=growth, d2Rf=t["d2"], alatt=t["alatt"], n=t["n"],
                            ZZon=1.0, ilpt=2, Om=COSMO.Om, OL=COSMO.OL, ioutshear=1)
    merged = api.merge_catalog(recs)
    assert merged.shape[0] > 10
    tb = dict(growth=growth, chi2z=cosmology.build_chi_to_z(cosmo_tables["cosmo"], z_max=10.0))
    got = run(oracle, tb, merged.cpu().numpy(), (0.0, 0.0, -2000.0), 1, 0.0)
    assert np.any(got[:, 3:6] != 0.0)


def test_empty_catalog_and_bad_arguments(tabs):
    import torch

    g = tabs["growth"]
    kw = dict(Om=COSMO.Om, OL=COSMO.OL)
    assert api.finalize_eulerian(np.zeros((0, 11), np.float32), g, ievol=0, **kw).shape == (0, 11)
    with pytest.raises(ValueError):
        api.finalize_eulerian(np.zeros((4, 12), np.float32), g, ievol=0, **kw)
    with pytest.raises(ValueError):
        api.finalize_eulerian(np.zeros((4, 11), np.float32), g[:2], ievol=0, **kw)
    with pytest.raises(HuntianError, match="chi -> z"):
        api.finalize_eulerian(np.zeros((4, 11), np.float32), g, ievol=1, **kw)
    with pytest.raises(HuntianError, match="beyond the growth table"):
        api.finalize_eulerian(np.zeros((4, 11), np.float32), g, ievol=0, z_out=5000.0, **kw)
    torch.cuda.synchronize()

"""The crossing pass of the crossing schedule (HNT_ORDER_CROSSING) gives each peak its exact
gather extent: with FP64 sums the extent it finds (hnt_results.schedule_extent) equals the
extent the analysis then gathers (nshells_used), and the schedule changes no result.  Cases:
every candidate of the 128^3 workload; uneven tiles in one call with peaks on the tile faces
(shells that leave the tile); a hunt radius of 3 cells (no crossing within it)."""
import numpy as np
import pytest
import torch

from huntian_hip import api, synth

pytestmark = pytest.mark.gpu

FIELDS = ("RTHL", "Fbarx", "e_v", "p_v", "strain", "Sbar", "gradpkf", "zvir_half",
          "nshells_used")


def _check(r, plain, label):
    ext = r["schedule_extent"]
    assert torch.equal(ext, r["nshells_used"]), label
    for k in FIELDS:
        assert torch.equal(r[k], plain[k]), (label, k)
    e = ext.cpu().numpy()
    print(f"{label}: {len(e)} peaks, extents {e.min()}..{e.max()}", flush=True)
    return e


def test_tile128_every_candidate(websky):
    dev = torch.device("cuda", 0)
    w = synth.workload_tile("tile128", dev)
    n, nbuff, alatt = w["n"], w["nbuff"], w["alatt"]
    mask = api.fortran_empty((n, n, n), torch.int8, dev)
    mask.zero_()
    pf = api.peak_find(w["smoothed"], [1.5] * len(w["radii"]), nbuff, mask,
                       capacity=(n - 2 * nbuff) ** 3)
    f = pf["filt"].long()
    Rf = torch.tensor(w["radii"], dtype=torch.float64, device=dev)[f]
    ir2 = torch.tensor([api.ir2min_of(R, alatt) for R in w["radii"]], dtype=torch.int32,
                       device=dev)[f]
    stab = api.build_shell_tables(w["nhunt"])
    fcrit = api.fsc_of_z(websky, 0.0)
    npk = len(pf["i"])

    def run(order):
        r = api.alloc_results(npk, dev, schedule_extent=True)
        api.shell_analyse(w["fields"], pf["i"], pf["j"], pf["k"], Rf, ir2, stab, websky,
                          alatt=alatt, ZZon=1.0, fcrit=fcrit, nbuff=nbuff, results=r,
                          order=order)
        return r

    plain = run("plain")
    assert int(plain["schedule_extent"].abs().sum()) == 0  # other orders leave it as it is
    ext = _check(run("crossing"), plain, "tile128")
    assert npk > 5000 and ext.min() < ext.max()
    del w, plain
    torch.cuda.empty_cache()


def _face_tile(oracle, n, seed):
    box = 2.0 * n
    delta, dk = synth.gaussian_field(n, box, seed)
    f = {"delta": delta}
    f.update(synth.lpt_fields(dk, n, box))
    radii = synth.filter_radii(1.5, 8.0, 4)
    sm = [synth.smooth_tophat(dk, n, box, R) for R in radii]
    pk, _ = oracle.peak_find_multi(sm, [1.0] * len(radii), 0,
                                   np.zeros((n, n, n), np.int8, order="F"))
    Rf = np.array([radii[q] for q in pk["filt"]])
    ir2 = np.array([api.ir2min_of(r, 2.0) for r in Rf], np.int32)
    return dict(fields=f, pi=pk["i"], pj=pk["j"], pk=pk["k"], Rf=Rf, ir2min=ir2, n=n)


def _near_faces(t, d):
    """Peaks within d cells of a tile face, whose shells beyond d leave the tile."""
    n = t["n"]
    c = np.stack([t["pi"], t["pj"], t["pk"]])
    return int(((c <= d) | (c > n - d)).any(axis=0).sum())


@pytest.mark.parametrize("nhunt", [15, 3])
def test_uneven_tiles_with_face_peaks(oracle, websky, nhunt):
    tiles = [_face_tile(oracle, n, s) for n, s in ((40, 1), (56, 2), (33, 3), (48, 4))]
    faces = sum(_near_faces(t, nhunt) for t in tiles)
    stab = api.build_shell_tables(nhunt)
    fcrit = api.fsc_of_z(websky, 0.0)
    kw = dict(alatt=2.0, ZZon=1.0, fcrit=fcrit, nbuff=0)
    crossing = api.shell_analyse_tiles(tiles, stab, websky, order="crossing",
                                       schedule_extent=True, **kw)
    plain = api.shell_analyse_tiles(tiles, stab, websky, order="plain", **kw)
    top = 0
    for q, (r, rp) in enumerate(zip(crossing, plain)):
        top = max(top, int(_check(r, rp, f"nhunt {nhunt} tile {q}").max()))
    print(f"nhunt {nhunt}: {faces} peaks within {nhunt} cells of a face, largest extent {top}",
          flush=True)
    assert faces > 0

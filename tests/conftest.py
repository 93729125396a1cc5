import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "huntian.jl_amd"))
sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and libhuntian_hip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def websky():
    from huntian_hip import WEBSKY_TABLE, read_homeltab

    return read_homeltab(WEBSKY_TABLE)


@pytest.fixture(scope="session")
def fortran_table():
    from huntian_hip import read_homeltab

    return read_homeltab(os.path.join(REPO, "tests", "golden", "HomelTab_fortran.dat"))


@pytest.fixture(scope="session")
def gpu_tile(websky):
    """A realistic tile analysed on the GPU: a 64^3 Gaussian field in a 128 Mpc/h box,
    eight top-hat filters, peak finding and shell analysis (nbuff 11, nhunt 10)."""
    import numpy as np
    import torch
    from huntian_hip import api, synth

    n, box, nbuff, nhunt = 64, 128.0, 11, 10
    alatt = box / n
    delta, dk = synth.gaussian_field(n, box, 6060)
    f = {"delta": delta}
    f.update(synth.lpt_fields(dk, n, box))
    radii = synth.filter_radii(1.5, 16.0, 8)
    sm = [api.to_device_field(synth.smooth_tophat(dk, n, box, R)) for R in radii]
    mask = api.to_device_field(np.zeros((n, n, n), np.int8), dtype=torch.int8)
    pf = api.peak_find(sm, [1.5] * len(radii), nbuff, mask, capacity=n ** 3)
    Rf = torch.tensor(radii, dtype=torch.float64, device="cuda")[pf["filt"].long()]
    ir2 = torch.tensor([api.ir2min_of(R, alatt) for R in radii], dtype=torch.int32,
                       device="cuda")[pf["filt"].long()]
    mask.zero_()
    r = api.shell_analyse(f, pf["i"], pf["j"], pf["k"], Rf, ir2, api.build_shell_tables(nhunt),
                          websky, alatt=alatt, ZZon=1.0, fcrit=api.fsc_of_z(websky, 0.0),
                          nbuff=nbuff, mask=mask)
    assert (r["RTHL"] > 0).sum().item() > 30
    d2 = torch.randn(len(pf["i"]), device="cuda")
    return dict(n=n, alatt=alatt, pf=pf, Rf=Rf, r=r, d2=d2)


@pytest.fixture(scope="session")
def cosmo_tables():
    """Growth and chi -> z tables of a flat Lambda-CDM cosmology (Om 0.31, h 0.68)."""
    from huntian_hip import cosmology

    c = cosmology.CosmologyParams(0.31, 0.049, 0.69, 0.68)
    return dict(cosmo=c, growth=cosmology.dlinear_tables(c),
                chi2z=cosmology.build_chi_to_z(c, z_max=1.0, npts=500))

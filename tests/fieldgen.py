"""Small synthetic inputs shared by the tests (seeded, Fortran-ordered numpy arrays)."""
import numpy as np


def fortran_zeros(n, dtype=np.float32):
    return np.zeros((n, n, n), dtype=dtype, order="F")


def fortran_golden_peak():
    """The single synthetic peak of the reference's Fortran comparison
    (test/test_radialshell_fortran.jl:26-51): 142^3 grid, peak at (71,71,71)."""
    n = 142
    delta, etax, etay, etaz = (fortran_zeros(n) for _ in range(4))
    c = 70  # 0-based index of cell 71
    delta[c, c, c] = 10
    for d in [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]:
        delta[c + d[0], c + d[1], c + d[2]] = 5
    etax[c, c, c] = 0.1
    etax[c - 1, c, c] = 0.05
    etax[c + 1, c, c] = 0.15
    for d in [(0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]:
        etax[c + d[0], c + d[1], c + d[2]] = 0.1
    etay[c, c, c] = 0.2
    etay[c - 1, c, c] = 0.2
    etay[c + 1, c, c] = 0.2
    etay[c, c - 1, c] = 0.1
    etay[c, c + 1, c] = 0.3
    etay[c, c, c - 1] = 0.2
    etay[c, c, c + 1] = 0.2
    etaz[c, c, c] = 0.3
    for d in [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0)]:
        etaz[c + d[0], c + d[1], c + d[2]] = 0.3
    etaz[c, c, c - 1] = 0.15
    etaz[c, c, c + 1] = 0.45
    return dict(delta=delta, etax=etax, etay=etay, etaz=etaz)


# Fortran get_homel outputs for that peak (test/test_radialshell_fortran.jl:8-24)
FORTRAN_GOLDEN = dict(
    fcrit=1.7161525487899780,
    RTHL=1.7320507764816284,
    Srb=1.9020472764968872,
    Fbarx=2.3230779170989990,
    Sbar=(2.5925926864147186e-02, 5.1851853728294373e-02, 7.7777773141860962e-02),
    Sbar2=(0.0, 0.0, 0.0),
    strain=np.diag([3.8717970252037048e-01, 7.7435940504074097e-01, 1.1615388393402100e+00]),
    gradpk=(0.0, 0.0, 0.0),
    gradpkrf=(0.0, 0.0, 0.0),
    e_v=1.6666662693023682e-01,
    p_v=-5.1315236504478889e-08,
    zvir_half=-1.0,
)


def bumps_field(n, seed, nbumps=20, noise=0.05, margin=10):
    """Gaussian bumps on noise plus displacement fields, as the reference's end-to-end GPU
    test builds them (test/test_shell_gpu.jl:678-716)."""
    rng = np.random.default_rng(seed)
    delta = (noise * rng.standard_normal((n, n, n))).astype(np.float32)
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for _ in range(nbumps):
        c = rng.integers(margin, n - margin, size=3)
        amp = 1.0 + 2.0 * rng.random()
        s2 = 15.0 + 10.0 * rng.random()
        d2 = (ii - c[0]) ** 2 + (jj - c[1]) ** 2 + (kk - c[2]) ** 2
        delta += (amp * np.exp(-d2 / s2)).astype(np.float32)
    f = {"delta": np.asfortranarray(delta)}
    for name, sc in (("etax", 0.3), ("etay", 0.3), ("etaz", 0.3), ("eta2x", 0.15),
                     ("eta2y", 0.15), ("eta2z", 0.15), ("lapd", 0.05)):
        f[name] = np.asfortranarray((sc * rng.standard_normal((n, n, n))).astype(np.float32))
    return f


def nontrivial_ctab(make_ctab):
    """The nontrivial 11^3 collapse table of test/test_shell_gpu.jl:724-736."""
    i = np.arange(11)[:, None, None]
    j = np.arange(11)[None, :, None]
    k = np.arange(11)[None, None, :]
    t = (0.5 + 0.3 * np.sin(i * 0.3) + 0.2 * np.cos(j * 0.4) + 0.1 * k * 0.1).astype(np.float32)
    return make_ctab(t, -2.0, 2.0, 0.0, 1.0, 0.0, 2.0)

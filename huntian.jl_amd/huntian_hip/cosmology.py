"""Host-side cosmology tables the record packing needs, as the reference driver builds them
once per run (src/MultiResolution.jl:719-743).

  dlinear_tables  -- Dlinear_tables (src/Cosmology/Cosmology.jl:133-284): RK4 integration of
                     the linear growth ODE in ln a from a = 1e-7 in 2000 steps, tabulated
                     for 1/1001 <= a <= 2 and normalised to D(a = 1) = 1
  chi             -- comoving distance in Mpc/h (Cosmology.jl:19-25)
  build_chi_to_z  -- the chi -> z lookup table (Cosmology.jl:337-350)

The growth ODE is restated with the reference's operation order (x^3 as x*x*x, the same
left-to-right sums); exp/log/sqrt come from Python's libm.  The integral of chi uses scipy's
adaptive quadrature instead of QuadGK, both converged to well below the table's use."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class CosmologyParams:
    """CosmologyParams (Cosmology.jl:5-12)."""
    Om: float   # total matter
    OB: float   # baryons
    OL: float   # dark energy
    h: float
    ns: float = 0.965
    s8: float = 0.81


def _deriv(y, a, h, Omnr, Omvac, Omcurv, Omer, OmX, OmB, Omhdm, fhdmclus, Omnrd, taud):
    """_deriv_Dlin! (Cosmology.jl:136-159) with ldecay = 0 (Omerd = 0)."""
    Omerd = 0.0
    a3 = a * a * a
    hub = Omer / a + Omnr + Omvac * a3 + Omcurv * a + Omerd / a + Omnrd
    d1 = 1.0 / (h * math.sqrt(hub / a3))
    fb = y[2]
    q = Omer / a + Omnr + Omvac * a3 + Omcurv * a + Omerd / a + Omnrd
    qs = 0.5 / q
    q_val = qs * (2.0 * Omer / a + Omnr - 2.0 * Omvac * a3 + 2.0 * Omerd / a + Omnrd)
    fnr = (OmX + OmB + fhdmclus * Omhdm) / Omnr
    if OmB != 0.0 and a <= 1.0e-3:
        fnr = (OmX + fhdmclus * Omhdm) / Omnr
    fer = 0.0
    qf = qs * (2.0 * Omer / a * fer + Omnr * fnr + Omnrd * fnr)
    return [d1, fb, 3.0 * qf - fb * fb - (1.0 - q_val) * fb, Omnrd * a * d1 / taud, d1 / a]


def dlinear_tables(cosmo: CosmologyParams):
    """(atab, D_atab, dlnD_dlnatab): scale factors ascending, D(a)/a with D(1) = 1, and
    dlnD/dlna (Dlinear_tables, Cosmology.jl:161-284)."""
    Omnr, Omvac = cosmo.Om, cosmo.OL
    Omcurv = 1.0 - cosmo.Om - cosmo.OL
    Omer, OmX, Omhdm, fhdmclus, taud, Omnrd = 0.0, cosmo.Om - cosmo.OB, 0.0, 0.0, 1.0, 0.0
    amaxtab, amintab_target, ainit, nmax = 1.0 / 0.5, 1.0 / 1001.0, 1.0e-7, 2000
    a = ainit
    alogn = math.log(ainit)
    amlog = math.log(amaxtab)
    hh = (amlog - alogn) / nmax
    t0 = 1.0 / cosmo.h / math.sqrt(Omnr)
    aeq = 0.0
    yaeq = math.sqrt(a + aeq)
    yaeq0 = math.sqrt(aeq)
    fb = 0.25 * (math.sqrt(24.0 * (OmX + 0.0 * cosmo.OB + fhdmclus * Omhdm) / Omnr + 1.0) - 1.0)
    y = [t0 * (2.0 / 3.0 * (yaeq * yaeq * yaeq - yaeq0 * yaeq0 * yaeq0) - 2.0 * aeq * (yaeq - yaeq0)),
         math.log(a), fb, 0.0, t0 * 2.0 * (yaeq - yaeq0)]
    args = (cosmo.h, Omnr, Omvac, Omcurv, Omer, OmX, cosmo.OB, Omhdm, fhdmclus, Omnrd, taud)
    at, Dt, ft = [], [], []
    lnorm, anorm, bnorm, ip0 = 1, 0.001, 1.0, 0
    for _ in range(nmax):
        w1 = _deriv(y, a, *args)
        a_half = math.exp(alogn + hh * 0.5)
        w2 = _deriv([y[q] + hh * w1[q] * 0.5 for q in range(5)], a_half, *args)
        w3 = _deriv([y[q] + hh * w2[q] * 0.5 for q in range(5)], a_half, *args)
        alogn_new = alogn + hh
        a_new = math.exp(alogn_new)
        w4 = _deriv([y[q] + hh * w3[q] for q in range(5)], a_new, *args)
        y = [y[q] + hh / 6.0 * (w1[q] + w4[q] + 2.0 * (w2[q] + w3[q])) for q in range(5)]
        alogn, a = alogn_new, a_new
        if lnorm == 1 and a >= anorm:
            lnorm = 0
            bnorm = math.exp(y[1]) / a
        if amintab_target <= a <= amaxtab:
            at.append(a)
            Dt.append(math.exp(y[1]) / a)
            ft.append(y[2])
            if a >= 1.0 and ip0 == 0:
                ip0 = len(at)  # 1-based index of the first a >= 1
    if ip0 > 1:
        ainterp = (at[ip0 - 1] - 1.0) / (at[ip0 - 1] - at[ip0 - 2])
        D_agrowth = (Dt[ip0 - 1] * (1.0 - ainterp) + Dt[ip0 - 2] * ainterp) / bnorm
        fac = 1.0 / (D_agrowth * bnorm)
    else:
        fac = 1.0 / bnorm
    return np.array(at), np.array([d * fac for d in Dt]), np.array(ft)


def chi(z: float, cosmo: CosmologyParams) -> float:
    """Comoving distance in Mpc/h, (c / 100 km/s/Mpc) * int_0^z dz' / E(z') (flat, no
    radiation; Cosmology.jl:15-25)."""
    from scipy.integrate import quad

    val, _ = quad(lambda zp: 1.0 / math.sqrt(cosmo.Om * (1.0 + zp) ** 3 + cosmo.OL), 0.0,
                  float(z), epsabs=0.0, epsrel=1e-13, limit=200)
    return (2.998e5 / 100.0) * val


def build_chi_to_z(cosmo: CosmologyParams, z_max: float = 10.0, npts: int = 2000):
    """(zs, chis, z_max): zs uniform on [0, z_max] and their comoving distances
    (build_chi_to_z, Cosmology.jl:343-350)."""
    zs = np.linspace(0.0, z_max, npts)
    return zs, np.array([chi(z, cosmo) for z in zs]), float(z_max)

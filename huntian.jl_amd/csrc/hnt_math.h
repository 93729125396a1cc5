// hnt_math.h -- FP64 scalar math of the radial-shell collapse, shared by host and device code.
//
// Every function restates a piece of the reference CPU path with the same operation order,
// so that the GPU result equals the FP64 CPU result bit for bit wherever only +, -, *, /
// and sqrt are involved (all correctly rounded on gfx950 and on the host).  The library is
// compiled with -ffp-contract=off for that reason.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HNT_HD __host__ __device__ inline

namespace hnt {

constexpr double kPi = 3.14159265358979323846;

// x^-3 for a Float64 x as Julia evaluates the literal power `x^-3` (Base.literal_pow of an
// AbstractFloat forwards to x^n with an Int64 n; n = -3 takes the compensated
// square-and-multiply of its integer-power routine): r = 1/x with its residual
// rlo = -(x r - 1) r, then r^3 carried as a double-double (two_mul products with their
// fma-exact low parts) and summed once at the end.  The reference runs on hosts with FMA,
// where its muladd steps are fused; fma() here does the same on host and device.
HNT_HD double pow_m3(double x) {
    const double r = 1.0 / x;
    const double rlo = isfinite(x) ? -fma(x, r, -1.0) * r : 0.0;
    // y = r (ylo = rlo); then x2 = r * r with its low part and the cross term 2 r rlo
    const double y = r, ylo = rlo;
    const double x2 = r * r;
    const double x2lo = fma(r, r, -x2) + r * 2.0 * rlo;
    if (!isfinite(x2)) return x2 * y;
    return fma(x2, y, fma(y, x2lo, x2 * ylo));
}

// hRinteg: integral of the SPH kernel (RadialShell.jl:95-121).
HNT_HD double hRinteg(double u0, double u12) {
    double x0 = fabs(u0);
    if (x0 >= 2.0) return 0.0;
    const double one7 = 1.0 / 7.0;
    double result = 1.6 * u12 - 6.4 * one7;
    double x02 = x0 * x0;
    double x04 = x02 * x02;
    if (x0 >= 1.0) {
        double h1 = x04 * (2.0 - x0 * (2.4 - x0 * (1.0 - x0 * one7)));
        double h2 = x02 * (4.0 - x0 * (4.0 - x0 * (1.5 - x0 * 0.2)));
        result = result - h2 * u12 + h1;
    } else {
        result = result - 0.2 * (u12 - one7);
        if (x0 > 0.0) {
            double h1 = x04 * (1.0 - x02 * (1.0 - 3.0 * x0 * one7));
            double h2 = x02 * (2.0 - x02 * (1.5 - 0.6 * x0));
            result = result - h2 * u12 + h1;
        }
    }
    return result;
}

HNT_HD double cube(double x) { return x * x * x; }
HNT_HD double pow5(double x) {
    double x2 = x * x;
    return x2 * x2 * x;
}

// Ascending eigenvalues of the symmetric matrix (a11 a12 a13; . a22 a23; . . a33) by cyclic
// Jacobi rotations.  Replaces LAPACK eigvals(Symmetric) (RadialShell.jl:215-220) with a
// fixed, deterministic sequence of correctly rounded operations.
HNT_HD void eig3_sym(double a11, double a12, double a13, double a22, double a23, double a33,
                     double lam[3]) {
    double a00 = a11, a01 = a12, a02 = a13, b11 = a22, a12v = a23, a22v = a33;
    // a[p][q] held in scalars: d0 d1 d2 on the diagonal, o01 o02 o12 off the diagonal
    double d0 = a00, d1 = b11, d2 = a22v, o01 = a01, o02 = a02, o12 = a12v;
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = fabs(o01) + fabs(o02) + fabs(o12);
        if (off == 0.0) break;
        for (int r = 0; r < 3; ++r) {
            // (p, q, o) = (0,1,2), (0,2,1), (1,2,0)
            double apq = r == 0 ? o01 : (r == 1 ? o02 : o12);
            if (apq == 0.0) continue;
            double app = r == 2 ? d1 : d0;
            double aqq = r == 0 ? d1 : d2;
            if (fabs(apq) <= 0x1p-60 * (fabs(app) + fabs(aqq))) {
                if (r == 0) o01 = 0.0;
                else if (r == 1) o02 = 0.0;
                else o12 = 0.0;
                continue;
            }
            double theta = (aqq - app) / (2.0 * apq);
            double t;
            if (fabs(theta) > 1e150) {
                t = 0.5 / theta;
            } else {
                t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0.0) t = -t;
            }
            double c = 1.0 / sqrt(t * t + 1.0);
            double s = t * c;
            double tau = s / (1.0 + c);
            double npp = app - t * apq;
            double nqq = aqq + t * apq;
            // off-diagonal entries coupling the third index o with p and q
            double aop, aoq;
            if (r == 0) { aop = o02; aoq = o12; }       // o = 2: a[2][0], a[2][1]
            else if (r == 1) { aop = o01; aoq = o12; }  // o = 1: a[1][0], a[1][2]
            else { aop = o01; aoq = o02; }              // o = 0: a[0][1], a[0][2]
            double nop = aop - s * (aoq + tau * aop);
            double noq = aoq + s * (aop - tau * aoq);
            if (r == 0) { d0 = npp; d1 = nqq; o01 = 0.0; o02 = nop; o12 = noq; }
            else if (r == 1) { d0 = npp; d2 = nqq; o02 = 0.0; o01 = nop; o12 = noq; }
            else { d1 = npp; d2 = nqq; o12 = 0.0; o01 = nop; o02 = noq; }
        }
    }
    double x = d0, y = d1, z = d2, tmp;
    if (y < x) { tmp = x; x = y; y = tmp; }
    if (z < y) { tmp = y; y = z; z = tmp; }
    if (y < x) { tmp = x; x = y; y = tmp; }
    lam[0] = x;
    lam[1] = y;
    lam[2] = z;
}

// Frho = sum of eigenvalues; e_v, p_v as in RadialShell.jl:557-562.
HNT_HD void ev_from_lam(const double lam[3], double* Frho, double* e, double* p) {
    double F = lam[0] + lam[1] + lam[2];
    *Frho = F;
    if (F > 0.0) {
        *e = 0.5 * (lam[2] - lam[0]) / F;
        *p = 0.5 * (lam[2] + lam[0] - 2.0 * lam[1]) / F;
    } else {
        *e = 0.0;
        *p = 0.0;
    }
}

struct CTabView {
    const float* table;
    int nx, ny, nz;
    double x1, x2, y1, y2, z1, z2, out_val;
};

// Trilinear lookup on the uniform axes of the collapse table, out_val outside the closed
// box (CollapseTable.jl:119-138).
HNT_HD double interp3(const CTabView& c, double x, double y, double z) {
    if (!(x >= c.x1 && x <= c.x2 && y >= c.y1 && y <= c.y2 && z >= c.z1 && z <= c.z2))
        return c.out_val;
    int nx = c.nx, ny = c.ny, nz = c.nz;
    double fx = (nx > 1) ? (x - c.x1) / ((c.x2 - c.x1) / (double)(nx - 1)) : 0.0;
    double fy = (ny > 1) ? (y - c.y1) / ((c.y2 - c.y1) / (double)(ny - 1)) : 0.0;
    double fz = (nz > 1) ? (z - c.z1) / ((c.z2 - c.z1) / (double)(nz - 1)) : 0.0;
    int ix = (int)floor(fx), iy = (int)floor(fy), iz = (int)floor(fz);
    if (ix > nx - 2) ix = nx - 2;
    if (iy > ny - 2) iy = ny - 2;
    if (iz > nz - 2) iz = nz - 2;
    if (ix < 0) ix = 0;
    if (iy < 0) iy = 0;
    if (iz < 0) iz = 0;
    double tx = fx - (double)ix, ty = fy - (double)iy, tz = fz - (double)iz;
    double ox = 1.0 - tx, oy = 1.0 - ty, oz = 1.0 - tz;
    int ix1 = (nx > 1) ? ix + 1 : ix, iy1 = (ny > 1) ? iy + 1 : iy, iz1 = (nz > 1) ? iz + 1 : iz;
    const float* T = c.table;
    auto tv = [&](int a, int b, int d) -> double {
        return (double)T[a + (int64_t)nx * (b + (int64_t)ny * d)];
    };
    double c00 = tv(ix, iy, iz) * ox + tv(ix1, iy, iz) * tx;
    double c10 = tv(ix, iy1, iz) * ox + tv(ix1, iy1, iz) * tx;
    double c01 = tv(ix, iy, iz1) * ox + tv(ix1, iy, iz1) * tx;
    double c11 = tv(ix, iy1, iz1) * ox + tv(ix1, iy1, iz1) * tx;
    double c0 = c00 * oy + c10 * ty;
    double c1 = c01 * oy + c11 * ty;
    return c0 * oz + c1 * tz;
}

// First index m in [lo, hi] with v[m] > x for increasing v; hi + 1 if there is none.
HNT_HD int first_gt(const double* v, int lo, int hi, double x) {
    int a = lo, b = hi + 1;
    while (a < b) {
        int mid = (a + b) >> 1;
        if (v[mid] > x) b = mid;
        else a = mid + 1;
    }
    return a;
}

// First index m in [lo, hi] with v[m] >= x for increasing v; hi + 1 if there is none.
HNT_HD int first_ge(const double* v, int lo, int hi, double x) {
    int a = lo, b = hi + 1;
    while (a < b) {
        int mid = (a + b) >> 1;
        if (v[mid] >= x) b = mid;
        else a = mid + 1;
    }
    return a;
}

}  // namespace hnt

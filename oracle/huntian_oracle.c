/*
 * huntian_oracle.c -- TEST INFRASTRUCTURE (the parity checker and the CPU baseline leg of
 * bench.py; never linked into libhuntian_hip.so).
 *
 * FP64 restatement of Huntian.jl's halo-finding hot path in plain C, following the
 * reference CPU code operation by operation:
 *   src/HaloFinder/PeakFind.jl          find_peaks
 *   src/HaloFinder/RadialShell.jl       precompute_shells, hRinteg, atab4, kernel_strain,
 *                                       fsc_of_z, normalize_strain, get_evals, analyse_peak
 *   src/ShellAnalysisGPU.jl             build_shell_tables, analyse_peak_gpu
 *   src/EllipsoidalCollapse/CollapseTable.jl   read_homeltab layout, trilinear lookup
 * Arithmetic notes (they matter for bit-level agreement with the Julia code):
 *   - Julia promotes Float32 field values to Float64 when adding them to Float64 sums, but
 *     a product Float32 * Int32 (delta * di, eta * dk) is a Float32 product; it is rounded
 *     to Float32 before the Float64 add.  That is reproduced here.
 *   - x^2 and x^3 with literal exponents are x*x and x*x*x in Julia.  x^5 and x^(1/3) go
 *     through Julia's pow; here they are x2*x2*x and C pow (agreement to ~1 ulp, not bits).
 *   - Build with -ffp-contract=off: Julia does not contract a*b+c into an FMA.
 * Eigenvalues: the reference calls LAPACK (eigvals(Symmetric)); this file uses a cyclic
 * Jacobi iteration (same code as the GPU library), pinned against numpy.linalg.eigvalsh.
 */
#include "huntian_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.14159265358979323846

/* ------------------------------------------------------------------ shells */

/* RadialShell.precompute_shells (RadialShell.jl:151-166): offsets generated dk outer, dj,
 * di inner, then a stable sort by r^2 -- done here as a counting sort, which is stable. */
int64_t hro_shell_cells(int32_t rmax) {
    int64_t c = 0;
    int32_t r2max = rmax * rmax;
    for (int32_t dk = -rmax; dk <= rmax; ++dk)
        for (int32_t dj = -rmax; dj <= rmax; ++dj)
            for (int32_t di = -rmax; di <= rmax; ++di)
                if (di * di + dj * dj + dk * dk <= r2max) ++c;
    return c;
}

int32_t hro_shell_count(int32_t rmax) {
    int32_t r2max = rmax * rmax;
    char* seen = (char*)calloc((size_t)r2max + 1, 1);
    for (int32_t dk = -rmax; dk <= rmax; ++dk)
        for (int32_t dj = -rmax; dj <= rmax; ++dj)
            for (int32_t di = -rmax; di <= rmax; ++di) {
                int32_t r2 = di * di + dj * dj + dk * dk;
                if (r2 <= r2max) seen[r2] = 1;
            }
    int32_t s = 0;
    for (int32_t v = 0; v <= r2max; ++v) s += seen[v];
    free(seen);
    return s;
}

/* ShellAnalysisGPU.build_shell_tables (ShellAnalysisGPU.jl:44-71), shell_start 0-based. */
int32_t hro_build_shell_tables(int32_t rmax, int32_t* di_o, int32_t* dj_o, int32_t* dk_o,
                               int32_t* shell_start, int32_t* shell_count, int32_t* shell_r2) {
    int32_t r2max = rmax * rmax;
    int64_t* pos = (int64_t*)calloc((size_t)r2max + 2, sizeof(int64_t));
    for (int32_t dk = -rmax; dk <= rmax; ++dk)
        for (int32_t dj = -rmax; dj <= rmax; ++dj)
            for (int32_t di = -rmax; di <= rmax; ++di) {
                int32_t r2 = di * di + dj * dj + dk * dk;
                if (r2 <= r2max) pos[r2 + 1]++;
            }
    int32_t ns = 0;
    for (int32_t v = 0; v <= r2max; ++v) {
        if (pos[v + 1] > 0) {
            shell_start[ns] = (int32_t)pos[v];
            shell_count[ns] = (int32_t)pos[v + 1];
            shell_r2[ns] = v;
            ++ns;
        }
        pos[v + 1] += pos[v];
    }
    for (int32_t dk = -rmax; dk <= rmax; ++dk)
        for (int32_t dj = -rmax; dj <= rmax; ++dj)
            for (int32_t di = -rmax; di <= rmax; ++di) {
                int32_t r2 = di * di + dj * dj + dk * dk;
                if (r2 <= r2max) {
                    int64_t q = pos[r2]++;
                    di_o[q] = di;
                    dj_o[q] = dj;
                    dk_o[q] = dk;
                }
            }
    free(pos);
    return ns;
}

/* ------------------------------------------------------------------ kernels */

/* hRinteg (RadialShell.jl:95-121) */
double hro_hRinteg(double u0, double u12) {
    double x0 = fabs(u0);
    if (x0 >= 2.0) return 0.0;
    double one7 = 1.0 / 7.0;
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

/* _build_atab4 (RadialShell.jl:125-142): cubic-spline SPH kernel tabulated in u = r^2 */
void hro_atab4(double* akk) {
    double ca = 3.0 / (2.0 * PI_D);
    double cb = 1.0 / (4.0 * PI_D);
    double two_thirds = 2.0 / 3.0;
    for (int i = 1; i <= 1100; ++i) {
        double u = (double)(i - 1) * 0.01;
        double ru = sqrt(u);
        double v;
        if (ru < 1.0) {
            v = ca * (two_thirds - u + 0.5 * u * ru);
        } else if (ru <= 2.0) {
            double t = 2.0 - ru;
            v = cb * (t * t * t);
        } else {
            v = 0.0;
        }
        akk[i - 1] = v;
    }
    akk[1100] = 0.0;
}

/* kernel_strain (RadialShell.jl:225-274) on 1-based profile arrays (index m). */
static void kstrain1(const double* rad, const double* G, const double* Gf, const double* SR,
                     const double* akk, int mp, int mlow, int mupp, double wRnor,
                     double aRnor, double h1, double h2, double E[9], double g[3],
                     double gf[3]) {
    for (int q = 0; q < 9; ++q) E[q] = 0.0;
    for (int q = 0; q < 3; ++q) { g[q] = 0.0; gf[q] = 0.0; }
    if (mp > 1) {
        double rmp = rad[mp];
        for (int m1 = mlow; m1 <= mupp; ++m1) {
            double u0 = h1 * (rmp - rad[m1]);
            double u02 = u0 * u0;
            if (u02 < 4.0) {
                double u12 = h2 * (rmp * rmp + rad[m1] * rad[m1]);
                double d = u12 - u02;
                double wt = hro_hRinteg(u0, u12) / (d * d);
                for (int L = 0; L < 3; ++L) {
                    g[L] += wt * G[m1 * 3 + L];
                    gf[L] += wt * Gf[m1 * 3 + L];
                    for (int K = 0; K < 3; ++K)
                        E[L * 3 + K] -= 0.5 * wt * (SR[m1 * 9 + L * 3 + K] + SR[m1 * 9 + K * 3 + L]);
                }
            }
        }
        double scale = wRnor / rmp;
        for (int q = 0; q < 9; ++q) E[q] *= scale;
        for (int q = 0; q < 3; ++q) { g[q] *= scale; gf[q] *= scale; }
    } else {
        for (int m1 = mlow; m1 <= mupp; ++m1) {
            double u0 = h1 * rad[m1];
            double con = 100.0 * u0 * u0;
            double fl = floor(con);
            /* the reference indexes a 1101-entry table; the CUDA port skips out-of-table
             * entries (ext/CUDAExt.jl:185), which is what happens here */
            if (!(fl >= 0.0 && fl < 1100.0)) continue;
            int icon = (int)fl;
            double diff = con - (double)icon;
            double aww = akk[icon] + diff * (akk[icon + 1] - akk[icon]);
            double wt = 0.5 * aww / rad[m1];
            for (int L = 0; L < 3; ++L) {
                g[L] += wt * G[m1 * 3 + L];
                gf[L] += wt * Gf[m1 * 3 + L];
                for (int K = 0; K < 3; ++K)
                    E[L * 3 + K] -= 0.5 * wt * (SR[m1 * 9 + L * 3 + K] + SR[m1 * 9 + K * 3 + L]);
            }
        }
        for (int q = 0; q < 9; ++q) E[q] *= aRnor;
        for (int q = 0; q < 3; ++q) { g[q] *= aRnor; gf[q] *= aRnor; }
    }
}

void hro_kernel_strain(const double* rad, const double* G, const double* Gf, const double* SR,
                       int32_t mp, int32_t mlow, int32_t mupp, double wRnor, double aRnor,
                       double hlatt_1, double hlatt_2, double* out) {
    double akk[1101];
    hro_atab4(akk);
    double E[9], g[3], gf[3];
    /* public arrays are 0-based (shell m at m-1); shift to the 1-based internal form */
    kstrain1(rad - 1, G - 3, Gf - 3, SR - 9, akk, mp, mlow, mupp, wRnor, aRnor, hlatt_1,
             hlatt_2, E, g, gf);
    out[0] = E[0]; out[1] = E[1]; out[2] = E[2];
    out[3] = E[4]; out[4] = E[5]; out[5] = E[8];
    out[6] = g[0]; out[7] = g[1]; out[8] = g[2];
    out[9] = gf[0]; out[10] = gf[1]; out[11] = gf[2];
}

/* Symmetric 3x3 eigenvalues, ascending, by cyclic Jacobi rotations in FP64.
 * A: column-major, upper triangle (a11 a12 a13 a22 a23 a33) used, as Symmetric(mat). */
static void eig3_sym(double a11, double a12, double a13, double a22, double a23, double a33,
                     double lam[3]) {
    double a[3][3] = {{a11, a12, a13}, {a12, a22, a23}, {a13, a23, a33}};
    static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        if (off == 0.0) break;
        for (int r = 0; r < 3; ++r) {
            int p = P[r], q = Q[r];
            double apq = a[p][q];
            if (apq == 0.0) continue;
            double app = a[p][p], aqq = a[q][q];
            /* negligible against both diagonal entries: drop it */
            if (fabs(apq) <= 0x1p-60 * (fabs(app) + fabs(aqq))) {
                a[p][q] = 0.0;
                a[q][p] = 0.0;
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
            a[p][p] = app - t * apq;
            a[q][q] = aqq + t * apq;
            a[p][q] = 0.0;
            a[q][p] = 0.0;
            int o = 3 - p - q; /* the third index */
            double aop = a[o][p], aoq = a[o][q];
            double nop = aop - s * (aoq + tau * aop);
            double noq = aoq + s * (aop - tau * aoq);
            a[o][p] = nop; a[p][o] = nop;
            a[o][q] = noq; a[q][o] = noq;
        }
    }
    double x = a[0][0], y = a[1][1], z = a[2][2], tmp;
    if (y < x) { tmp = x; x = y; y = tmp; }
    if (z < y) { tmp = y; y = z; z = tmp; }
    if (y < x) { tmp = x; x = y; y = tmp; }
    lam[0] = x; lam[1] = y; lam[2] = z;
}

void hro_eig3(const double* A, double* lam) {
    eig3_sym(A[0], A[3], A[6], A[4], A[7], A[8], lam);
}

/* CollapseTableInterp + interpolate (CollapseTable.jl:119-138): trilinear on the uniform
 * axes range(X1,X2,Nx) x ..., out_val outside the closed box. */
double hro_interp3(const hnt_ctab* ct, double x, double y, double z) {
    if (!(x >= ct->x1 && x <= ct->x2 && y >= ct->y1 && y <= ct->y2 && z >= ct->z1 &&
          z <= ct->z2))
        return ct->out_val;
    int nx = ct->nx, ny = ct->ny, nz = ct->nz;
    double fx = (nx > 1) ? (x - ct->x1) / ((ct->x2 - ct->x1) / (double)(nx - 1)) : 0.0;
    double fy = (ny > 1) ? (y - ct->y1) / ((ct->y2 - ct->y1) / (double)(ny - 1)) : 0.0;
    double fz = (nz > 1) ? (z - ct->z1) / ((ct->z2 - ct->z1) / (double)(nz - 1)) : 0.0;
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
    const float* T = ct->table;
#define TV(a, b, c) ((double)T[(a) + (int64_t)nx * ((b) + (int64_t)ny * (c))])
    double c00 = TV(ix, iy, iz) * ox + TV(ix1, iy, iz) * tx;
    double c10 = TV(ix, iy1, iz) * ox + TV(ix1, iy1, iz) * tx;
    double c01 = TV(ix, iy, iz1) * ox + TV(ix1, iy, iz1) * tx;
    double c11 = TV(ix, iy1, iz1) * ox + TV(ix1, iy1, iz1) * tx;
#undef TV
    double c0 = c00 * oy + c10 * ty;
    double c1 = c01 * oy + c11 * ty;
    return c0 * oz + c1 * tz;
}

/* fsc_of_z (RadialShell.jl:181-195) */
double hro_fsc_of_z(const hnt_ctab* ct, double z) {
    double target = 1.0 + z;
    double lo = ct->x1, hi = ct->x2;
    for (int it = 0; it < 50; ++it) {
        double mid = 0.5 * (lo + hi);
        double zvir = hro_interp3(ct, mid, 0.0, 0.0);
        if (zvir > target) hi = mid;
        else lo = mid;
    }
    return pow(10.0, lo);
}

/* ------------------------------------------------------------------ peak finding */

/* find_peaks (PeakFind.jl:15-78), 1-based indices, column-major fields. */
int64_t hro_find_peaks(const float* delta, int8_t* mask, int32_t n1, int32_t n2, int32_t n3,
                       int32_t nbuff, float fcrit, int64_t max_peaks,
                       int32_t* pi, int32_t* pj, int32_t* pk, float* pdelta) {
    int64_t np = 0;
    int64_t s1 = n1, s2 = (int64_t)n1 * n2;
    for (int32_t k = nbuff + 1; k <= n3 - nbuff; ++k)
        for (int32_t j = nbuff + 1; j <= n2 - nbuff; ++j)
            for (int32_t i = nbuff + 1; i <= n1 - nbuff; ++i) {
                int64_t c = (i - 1) + (j - 1) * s1 + (k - 1) * s2;
                if (mask[c] == 1) continue;
                float ff = delta[c];
                if (ff < fcrit) continue;
                int is_max = 1;
                for (int kk = -1; kk <= 1 && is_max; ++kk)
                    for (int jj = -1; jj <= 1 && is_max; ++jj)
                        for (int ii = -1; ii <= 1; ++ii)
                            if (ff < delta[c + ii + jj * s1 + kk * s2]) { is_max = 0; break; }
                if (!is_max) continue;
                if (pi) { pi[np] = i; pj[np] = j; pk[np] = k; pdelta[np] = ff; }
                np++;
                mask[c] = 1;
                if (np >= max_peaks) return np;
            }
    return np;
}

/* The drivers' filter loop (MultiResolution.jl:1046-1066, Pipeline.jl:164-198): one
 * find_peaks scan per filter in the given order, the mask carried over.  Returns the
 * total; entries beyond cap are counted but not written (and their cells not masked,
 * as the GPU reference drops slots beyond max_peaks, ext/CUDAExt.jl:2921-2926). */
int64_t hro_peak_find_multi(const float* const* fields, const float* fcrit, int32_t nfilt,
                            int32_t n, int32_t nbuff, int8_t* mask, int64_t cap,
                            int32_t* pi, int32_t* pj, int32_t* pk, int32_t* filt, float* pdelta) {
    int64_t total = 0;
    int64_t s1 = n, s2 = (int64_t)n * n;
    for (int32_t f = 0; f < nfilt; ++f) {
        const float* delta = fields[f];
        for (int32_t k = nbuff + 1; k <= n - nbuff; ++k)
            for (int32_t j = nbuff + 1; j <= n - nbuff; ++j)
                for (int32_t i = nbuff + 1; i <= n - nbuff; ++i) {
                    int64_t c = (i - 1) + (j - 1) * s1 + (k - 1) * s2;
                    if (mask[c] == 1) continue;
                    float ff = delta[c];
                    if (ff < fcrit[f]) continue;
                    int is_max = 1;
                    for (int kk = -1; kk <= 1 && is_max; ++kk)
                        for (int jj = -1; jj <= 1 && is_max; ++jj)
                            for (int ii = -1; ii <= 1; ++ii)
                                if (ff < delta[c + ii + jj * s1 + kk * s2]) { is_max = 0; break; }
                    if (!is_max) continue;
                    if (total < cap) {
                        pi[total] = i; pj[total] = j; pk[total] = k;
                        filt[total] = f; pdelta[total] = ff;
                        mask[c] = 1;
                    }
                    total++;
                }
    }
    return total;
}

/* The same filter loop with each filter's scan split over k-planes between threads.  A
 * cell's test reads the mask only at that cell and a scan visits each cell once, so the
 * planes are independent within a filter: each thread counts the peaks of its planes, the
 * counts give every plane its offset in the (k, j, i) order of the serial scan, and a second
 * pass writes the peaks and masks those that fit under cap.  Same outputs as
 * hro_peak_find_multi. */
static int peak_at(const float* delta, const int8_t* mask, int64_t c, int64_t s1, int64_t s2,
                   float fc) {
    if (mask[c] == 1) return 0;
    float ff = delta[c];
    if (ff < fc) return 0;
    for (int kk = -1; kk <= 1; ++kk)
        for (int jj = -1; jj <= 1; ++jj)
            for (int ii = -1; ii <= 1; ++ii)
                if (ff < delta[c + ii + jj * s1 + kk * s2]) return 0;
    return 1;
}

int64_t hro_peak_find_multi_par(const float* const* fields, const float* fcrit, int32_t nfilt,
                                int32_t n, int32_t nbuff, int8_t* mask, int64_t cap,
                                int32_t* pi, int32_t* pj, int32_t* pk, int32_t* filt,
                                float* pdelta, int32_t nthreads) {
    const int64_t s1 = n, s2 = (int64_t)n * n;
    const int32_t k0 = nbuff + 1, k1 = n - nbuff;
    if (k1 < k0) return 0;
    const int32_t np = k1 - k0 + 1;
    int64_t* off = (int64_t*)calloc((size_t)np + 1, sizeof(int64_t));
    if (!off) return -1;
    if (nthreads < 1) nthreads = 1;
    int64_t total = 0;
    for (int32_t f = 0; f < nfilt; ++f) {
        const float* delta = fields[f];
        const float fc = fcrit[f];
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
#endif
        for (int32_t q = 0; q < np; ++q) {
            const int32_t k = k0 + q;
            int64_t cnt = 0;
            for (int32_t j = nbuff + 1; j <= n - nbuff; ++j)
                for (int32_t i = nbuff + 1; i <= n - nbuff; ++i)
                    cnt += peak_at(delta, mask, (i - 1) + (j - 1) * s1 + (k - 1) * s2, s1, s2,
                                   fc);
            off[q + 1] = cnt;
        }
        off[0] = total;
        for (int32_t q = 0; q < np; ++q) off[q + 1] += off[q];
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
#endif
        for (int32_t q = 0; q < np; ++q) {
            const int32_t k = k0 + q;
            int64_t o = off[q];
            if (o >= cap) continue;
            for (int32_t j = nbuff + 1; j <= n - nbuff && o < cap; ++j)
                for (int32_t i = nbuff + 1; i <= n - nbuff && o < cap; ++i) {
                    const int64_t c = (i - 1) + (j - 1) * s1 + (k - 1) * s2;
                    if (!peak_at(delta, mask, c, s1, s2, fc)) continue;
                    pi[o] = i; pj[o] = j; pk[o] = k; filt[o] = f; pdelta[o] = delta[c];
                    mask[c] = 1;
                    ++o;
                }
        }
        total = off[np];
    }
    free(off);
    return total;
}

/* ------------------------------------------------------------------ shell gather */

void hro_shell_gather_full(const hnt_fields* f, int64_t npeaks, const int32_t* pi,
                           const int32_t* pj, const int32_t* pk, const hnt_shell_tables* t,
                           double* F, int32_t* n, double* S, double* S2, double* G, double* Gf,
                           double* SR, double* lapd) {
    int32_t n1 = f->n1, n2 = f->n2, n3 = f->n3;
    int64_t s1 = n1, s2 = (int64_t)n1 * n2;
    int32_t ns = t->nshells;
    for (int64_t p = 0; p < npeaks; ++p) {
        int32_t ci = pi[p], cj = pj[p], ck = pk[p];
        for (int32_t s = 0; s < ns; ++s) {
            double fs = 0.0, lp = 0.0, sv[3] = {0, 0, 0}, s2v[3] = {0, 0, 0}, gv[3] = {0, 0, 0},
                   gfv[3] = {0, 0, 0}, srv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            int32_t cnt = 0;
            for (int32_t c = 0; c < t->shell_count[s]; ++c) {
                int64_t q = (int64_t)t->shell_start[s] + c;
                int32_t di = t->di[q], dj = t->dj[q], dk = t->dk[q];
                int32_t i1 = ci + di, i2 = cj + dj, i3 = ck + dk;
                if (i1 < 1 || i1 > n1 || i2 < 1 || i2 > n2 || i3 < 1 || i3 > n3) continue;
                cnt++;
                int64_t idx = (i1 - 1) + (i2 - 1) * s1 + (i3 - 1) * s2;
                int64_t idxT = (i3 - 1) + (i1 - 1) * s1 + (i2 - 1) * s2;
                float d = f->delta[idx], dT = f->delta[idxT];
                float e[3] = {f->etax[idx], f->etay[idx], f->etaz[idx]};
                float e2[3] = {0.0f, 0.0f, 0.0f};
                if (f->eta2x) { e2[0] = f->eta2x[idx]; e2[1] = f->eta2y[idx]; e2[2] = f->eta2z[idx]; }
                float ix[3] = {(float)di, (float)dj, (float)dk};
                fs += (double)d;
                for (int L = 0; L < 3; ++L) {
                    sv[L] += (double)e[L];
                    s2v[L] += (double)e2[L];
                    gv[L] += (double)(d * ix[L]);
                    gfv[L] += (double)(dT * ix[L]);
                    for (int K = 0; K < 3; ++K) srv[L * 3 + K] += (double)(e[L] * ix[K]);
                }
                if (f->lapd) lp += (double)f->lapd[idx];
            }
            int64_t o = p * ns + s;
            F[o] = fs;
            n[o] = cnt;
            lapd[o] = lp;
            for (int L = 0; L < 3; ++L) {
                S[o * 3 + L] = sv[L];
                S2[o * 3 + L] = s2v[L];
                G[o * 3 + L] = gv[L];
                Gf[o * 3 + L] = gfv[L];
                /* SR component (L, K) at L + 3K: column-major 3x3 as the reference arrays */
                for (int K = 0; K < 3; ++K) SR[o * 9 + L + 3 * K] = srv[L * 3 + K];
            }
        }
    }
}

/* ------------------------------------------------------------------ analyse_peak */

typedef struct {
    double RTHL, Srb, Sbar[3], Sbar2[3], strain[9], gradpk[3], gradpkrf[3], Fbarx, e_v, p_v,
        zvir_half, gradpkf[3], d2F, eigs[3];
} res_t;

static void no_collapse(res_t* r) {
    memset(r, 0, sizeof(*r));
    r->RTHL = -1.0;
    r->zvir_half = -1.0;
}

typedef struct {
    double *rad, *Fbar, *S, *S2, *G, *Gf, *SR, *lapA, *lapB;
    int64_t *nsh, *cntA, *cntB;
} prof_t;

static int prof_alloc(prof_t* w, int32_t ns) {
    size_t m = (size_t)ns + 2;
    w->rad = (double*)malloc(m * sizeof(double));
    w->Fbar = (double*)malloc(m * sizeof(double));
    w->S = (double*)malloc(3 * m * sizeof(double));
    w->S2 = (double*)malloc(3 * m * sizeof(double));
    w->G = (double*)malloc(3 * m * sizeof(double));
    w->Gf = (double*)malloc(3 * m * sizeof(double));
    w->SR = (double*)malloc(9 * m * sizeof(double));
    w->lapA = (double*)malloc(m * sizeof(double));
    w->lapB = (double*)malloc(m * sizeof(double));
    w->nsh = (int64_t*)malloc(m * sizeof(int64_t));
    w->cntA = (int64_t*)malloc(m * sizeof(int64_t));
    w->cntB = (int64_t*)malloc(m * sizeof(int64_t));
    return w->rad && w->Fbar && w->S && w->S2 && w->G && w->Gf && w->SR && w->lapA &&
           w->lapB && w->nsh && w->cntA && w->cntB;
}

static void prof_free(prof_t* w) {
    free(w->rad); free(w->Fbar); free(w->S); free(w->S2); free(w->G); free(w->Gf);
    free(w->SR); free(w->lapA); free(w->lapB); free(w->nsh); free(w->cntA); free(w->cntB);
}

static inline double cube(double x) { return x * x * x; }
static inline double pow5(double x) { double x2 = x * x; return x2 * x2 * x; }

static inline void ev_from_lam(const double lam[3], double* Frho, double* e, double* p) {
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

static inline void eig_full(const double E[9], double lam[3]) {
    eig3_sym(E[0], E[1], E[2], E[4], E[5], E[8], lam);
}

/* One peak: gather (RadialShell.jl:318-446 == ShellAnalysisGPU.jl:122-229) and the radial
 * logic (RadialShell.jl:449-865 for A, ShellAnalysisGPU.jl:231-535 for B/C). */
static void analyse_one(const hnt_fields* f, const hnt_shell_tables* t, const hnt_ctab* ct,
                        const double* akk, const hnt_params* prm, int32_t ci, int32_t cj,
                        int32_t ck, double Rf, int32_t ir2min, double ZZon, double fcrit,
                        prof_t* w, res_t* out, int8_t* mask, int32_t* nused,
                        int32_t* ncross) {
    const int variant = prm->variant;
    const double alatt = prm->alatt;
    const int32_t n1 = f->n1, n2 = f->n2, n3 = f->n3;
    const int64_t s1 = n1, s2 = (int64_t)n1 * n2;
    const int32_t NS = t->nshells;
    int m = NS - 1; /* last finalised shell (the last shell never is) */
    if (prm->rmax2rs > 0.0) {
        /* B and C: only the leading shells with r^2 <= (rmax2rs Rf / alatt)^2
         * (ShellAnalysisGPU.jl:90-97, ext/CUDAExt.jl:2177-2185) */
        double lim = prm->rmax2rs * Rf / alatt;
        lim = lim * lim;
        int cnt = 0;
        for (int s = 1; s <= NS - 1; ++s) {
            if (!((double)t->shell_r2[s - 1] <= lim)) break;
            cnt = s;
        }
        m = cnt;
        if (m < 2) { /* ShellAnalysisGPU.jl:99 */
            no_collapse(out);
            *nused = 0;
            *ncross = 0;
            return;
        }
    }
    /* hlatt = 1 (RadialShell.jl:296-300) */
    const double aRnor = 3.0 * 1.0 / alatt;
    const double wRnor = 3.0 * (1.0 / (4.0 * PI_D)) / alatt;
    double* rad = w->rad;
    for (int s = 1; s <= NS; ++s) rad[s] = sqrt((double)t->shell_r2[s - 1]);

    /* ---- mrf: shell closest to (Rf/alatt)^2, with each variant's index lag ---- */
    double rf = Rf / alatt;
    double rf2 = rf * rf;
    int mrf = 1;
    {
        double rmrf = 10.0;
        int mrfi = 1;
        if (variant == HNT_VARIANT_A) {
            /* RadialShell.jl:450-464: cell loop over jp = 2..npart (all shells) */
            int32_t ir2p = t->shell_r2[1];
            for (int s = 2; s <= NS; ++s) {
                int32_t ir2 = t->shell_r2[s - 1];
                double dist = fabs((double)ir2 - rf2);
                for (int32_t c = 0; c < t->shell_count[s - 1]; ++c) {
                    if (dist < rmrf) {
                        mrf = mrfi;
                        rmrf = rmrf < dist ? rmrf : dist;
                    }
                    if (ir2 != ir2p) {
                        mrfi += 1;
                        ir2p = ir2;
                    }
                }
            }
        } else {
            /* ShellAnalysisGPU.jl:234-254 */
            for (int s = 2; s <= m; ++s) {
                double dist = fabs((double)t->shell_r2[s - 1] - rf2);
                for (int32_t c = 0; c < t->shell_count[s - 1]; ++c)
                    if (dist < rmrf) { mrf = mrfi; rmrf = dist; }
                if (s < m && t->shell_r2[s] != t->shell_r2[s - 1]) mrfi += 1;
            }
        }
    }
    double rlow_rf = rad[mrf] - 2.0;
    if (rlow_rf < 0.0) rlow_rf = 0.0;
    int mlow_rf = mrf;
    for (int m1 = mrf - 1; m1 >= 1; --m1)
        if (rad[m1] > rlow_rf) mlow_rf = m1;
    double rupp_rf = rad[mrf] + 2.0;
    int mupp_rf = mrf;
    for (int m1 = mrf + 1; m1 <= m; ++m1)
        if (rad[m1] < rupp_rf) mupp_rf = m1;

    /* ---- gather ---- */
    int64_t c0 = (ci - 1) + (cj - 1) * s1 + (ck - 1) * s2;
    w->Fbar[1] = (double)f->delta[c0];
    w->nsh[1] = 1;
    w->S[3 + 0] = (double)f->etax[c0];
    w->S[3 + 1] = (double)f->etay[c0];
    w->S[3 + 2] = (double)f->etaz[c0];
    for (int L = 0; L < 3; ++L) {
        w->S2[3 + L] = f->eta2x ? (double)(L == 0 ? f->eta2x[c0] : L == 1 ? f->eta2y[c0]
                                                                          : f->eta2z[c0])
                                : 0.0;
        w->G[3 + L] = 0.0;
        w->Gf[3 + L] = 0.0;
    }
    for (int q = 0; q < 9; ++q) w->SR[9 + q] = 0.0;
    double lapA = 0.0, lapB = 0.0;
    int64_t cntA = 0, cntB = 0;
    int oobA = 0;
    if (f->lapd) { lapB = 0.0 + (double)f->lapd[c0]; cntB = 1; }
    w->lapA[1] = lapA; w->cntA[1] = cntA; w->lapB[1] = lapB; w->cntB[1] = cntB;

    double dFbarp = w->Fbar[1];
    int past_m0 = 0, crossed = 0;
    double rstop = 0.0;
    int stop = m;     /* last shell the GPU gather needs (early-exit rule) */
    int stop_set = 0;
    int last = m;     /* last shell gathered here */
    for (int s = 2; s <= m; ++s) {
        int32_t st = t->shell_start[s - 1], nc = t->shell_count[s - 1];
        double Fs = 0.0, sv[3] = {0, 0, 0}, s2v[3] = {0, 0, 0}, gv[3] = {0, 0, 0},
               gfv[3] = {0, 0, 0}, srv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        int64_t nin = 0;
        for (int32_t c = 0; c < nc; ++c) {
            int32_t di = t->di[st + c], dj = t->dj[st + c], dk = t->dk[st + c];
            int32_t i1 = ci + di, i2 = cj + dj, i3 = ck + dk;
            if (i1 < 1 || i1 > n1 || i2 < 1 || i2 > n2 || i3 < 1 || i3 > n3) {
                oobA = 1;
                continue;
            }
            nin++;
            int64_t idx = (i1 - 1) + (i2 - 1) * s1 + (i3 - 1) * s2;
            int64_t idxT = (i3 - 1) + (i1 - 1) * s1 + (i2 - 1) * s2;
            float d = f->delta[idx];
            float dT = f->delta[idxT];
            float e[3] = {f->etax[idx], f->etay[idx], f->etaz[idx]};
            float ix[3] = {(float)di, (float)dj, (float)dk};
            Fs += (double)d;
            for (int L = 0; L < 3; ++L) {
                sv[L] += (double)e[L];
                gv[L] += (double)(d * ix[L]);
                gfv[L] += (double)(dT * ix[L]);
                for (int K = 0; K < 3; ++K) srv[L * 3 + K] += (double)(e[L] * ix[K]);
            }
            if (f->eta2x) {
                s2v[0] += (double)f->eta2x[idx];
                s2v[1] += (double)f->eta2y[idx];
                s2v[2] += (double)f->eta2z[idx];
            }
            if (f->lapd) {
                double v = (double)f->lapd[idx];
                lapB += v; cntB++;
                if (!oobA) { lapA += v; cntA++; }
            }
        }
        w->nsh[s] = nin;
        double dFbar = nin > 0 ? Fs / (double)nin : 0.0;
        double rad3p = cube(rad[s - 1]);
        double rad3 = cube(rad[s]);
        w->Fbar[s] = (rad3p * w->Fbar[s - 1] + 0.5 * (dFbarp + dFbar) * (rad3 - rad3p)) / rad3;
        dFbarp = dFbar;
        for (int L = 0; L < 3; ++L) {
            w->S[s * 3 + L] = sv[L];
            w->S2[s * 3 + L] = s2v[L];
            w->G[s * 3 + L] = gv[L] / rad[s];
            w->Gf[s * 3 + L] = gfv[L] / rad[s];
        }
        for (int q = 0; q < 9; ++q) w->SR[s * 9 + q] = srv[q] / rad[s];
        w->lapA[s] = lapA; w->cntA[s] = cntA; w->lapB[s] = lapB; w->cntB[s] = cntB;

        /* early-exit rule.  Let s_c be the first shell past ir2min whose Fbar drops below
         * fcrit.  The radial logic then has m0 <= s_c; it reads Fbar up to s_c (s_c + 1 for
         * the formation-redshift radius when RTHL rounds up to rad[m0]) and the kernel_strain
         * inputs of shells with rad < rad[s_c + 1] + 2, plus the filter-scale window up to
         * mupp_rf.  The gather stops before the first shell beyond all of these.  The
         * reference states the same rule with a Float32 margin (ext/CUDAExt.jl:774-781,
         * 896-917); here Fbar is the exact FP64 value both sides use. */
        if (!past_m0 && t->shell_r2[s - 1] > ir2min) past_m0 = 1;
        if (past_m0 && !crossed && w->Fbar[s] < fcrit) {
            crossed = s;
            rstop = rad[s + 1] + 2.0;
        }
        if (!stop_set && crossed && rad[s + 1] >= rstop && s >= mupp_rf) {
            stop = s;
            stop_set = 1;
            /* with rmax2rs > 0 every shell up to the limit is gathered, as the reference
             * GPU path passes no shell_r2 to its gather then (ext/CUDAExt.jl:2235-2236) */
            if (prm->early_exit && !(prm->rmax2rs > 0.0)) { last = s; break; }
        }
    }
    if (nused) *nused = stop;
    if (ncross) *ncross = crossed;
    /* shells not gathered must never be read: poison them */
    for (int s = last + 1; s <= m + 1 && s <= NS; ++s) {
        w->Fbar[s] = NAN;
        for (int L = 0; L < 3; ++L) {
            w->S[s * 3 + L] = NAN; w->S2[s * 3 + L] = NAN;
            w->G[s * 3 + L] = NAN; w->Gf[s * 3 + L] = NAN;
        }
        for (int q = 0; q < 9; ++q) w->SR[s * 9 + q] = NAN;
        w->lapA[s] = NAN; w->lapB[s] = NAN; w->nsh[s] = -1;
    }

    /* ---- m0: first shell with r^2 > ir2min (else the last) ---- */
    int m0 = 1;
    for (int s = 2; s <= m; ++s) {
        m0 = s;
        if (t->shell_r2[s - 1] > ir2min) break;
    }
    double akkE[9], g3[3], gf3[3];
    kstrain1(rad, w->G, w->Gf, w->SR, akk, mrf, mlow_rf, mupp_rf, wRnor, aRnor, 1.0, 1.0,
             akkE, g3, gf3);
    double gradpkrf[3] = {g3[0], g3[1], g3[2]};

    /* ---- fcrit crossing (RadialShell.jl:485-527) ---- */
    int mupp;
    double* Fbar = w->Fbar;
    if (Fbar[m0] >= fcrit) {
        int found = 0;
        for (int mm = m0; mm <= m; ++mm)
            if (Fbar[mm] < fcrit) { m0 = mm; found = 1; break; }
        if (!found) {
            mupp = m;
        } else {
            double rupp_m0 = rad[m0] + 2.0;
            mupp = m0;
            for (int mm = m0 + 1; mm <= m; ++mm)
                if (rad[mm] < rupp_m0) mupp = mm;
        }
    } else {
        int mstart = m0 - 1 > 1 ? m0 - 1 : 1;
        int found = 0;
        for (int mp = mstart; mp >= 1; --mp) {
            if (Fbar[mp] >= fcrit) { found = 1; break; }
            m0 = mp;
        }
        if (!found) { no_collapse(out); return; }
        double rupp_m0 = rad[m0] + 2.0;
        mupp = m0;
        for (int mm = m0 + 1; mm <= m; ++mm)
            if (rad[mm] < rupp_m0) mupp = mm;
    }

    /* ---- inward walk (RadialShell.jl:529-686) ---- */
    double rlow = rad[m0] - 2.0;
    if (rlow < 0.0) rlow = 0.0;
    int mlow = m0;
    for (int m1 = m0 - 1; m1 >= 1; --m1)
        if (rad[m1] > rlow) mlow = m1;

    double E_m0[9], gk_m0[3], gkf_m0[3], lam[3];
    kstrain1(rad, w->G, w->Gf, w->SR, akk, m0, mlow, mupp, wRnor, aRnor, 1.0, 1.0, E_m0, gk_m0,
             gkf_m0);
    double Frho_raw, e_v_c, p_v_c;
    eig_full(E_m0, lam);
    ev_from_lam(lam, &Frho_raw, &e_v_c, &p_v_c);
    /* zvir at m0 is -1 in the reference (Fortran reads an unset value); ZZon > -1 is
     * required by the caller, so the walk always runs */
    double zvir1 = -1.0, zvir1p = -1.0;
    double Frhoh = Frho_raw;
    double Ebar_last[9], strain_mat[9], gk_last[3], gk_prev[3], gkf_last[3], gkf_prev[3];
    memcpy(Ebar_last, E_m0, sizeof(E_m0));
    memcpy(strain_mat, E_m0, sizeof(E_m0));
    memcpy(gk_last, gk_m0, sizeof(gk_m0));
    memcpy(gk_prev, gk_m0, sizeof(gk_m0));
    memcpy(gkf_last, gkf_m0, sizeof(gkf_m0));
    memcpy(gkf_prev, gkf_m0, sizeof(gkf_m0));
    if (m0 == 1) { no_collapse(out); return; }
    int collapsed = 0;
    for (int mp = m0 - 1; mp >= 1; --mp) {
        double rupp_mp = rad[mp] + 2.0;
        double rlow_mp = rad[mp] - 2.0;
        if (rlow_mp < 0.0) rlow_mp = 0.0;
        if (rad[mupp] > rupp_mp) {
            int mupp_new = mp;
            for (int m1 = mp + 1; m1 <= mupp; ++m1)
                if (rad[m1] < rupp_mp) mupp_new = m1;
            mupp = mupp_new;
        }
        if (mlow > 1 && rad[mlow - 1] > rlow_mp) {
            int mlownew = mlow;
            for (int m1 = mlow - 1; m1 >= 1; --m1)
                if (rad[m1] > rlow_mp) mlownew = m1;
            mlow = mlownew;
        }
        double E_mp[9], gk_mp[3], gkf_mp[3];
        kstrain1(rad, w->G, w->Gf, w->SR, akk, mp, mlow, mupp, wRnor, aRnor, 1.0, 1.0, E_mp,
                 gk_mp, gkf_mp);
        double Fr, e_mp, p_mp;
        eig_full(E_mp, lam);
        ev_from_lam(lam, &Fr, &e_mp, &p_mp);
        double Frho_mp = Fbar[mp];
        double zv = -1.0;
        if (Frho_mp > 0.0) {
            double poe = e_mp < 1e-5 ? 0.0 : p_mp / e_mp;
            zv = hro_interp3(ct, log10(Frho_mp), e_mp, poe);
        }
        if (zv >= ZZon) {
            m0 = mp + 1;
            zvir1p = zv;
            memcpy(Ebar_last, E_mp, sizeof(E_mp));
            memcpy(gk_last, gk_mp, sizeof(gk_mp));
            memcpy(gkf_last, gkf_mp, sizeof(gkf_mp));
            Frhoh = Fr;
            collapsed = 1;
            break;
        }
        zvir1 = zv;
        memcpy(gk_prev, gk_mp, sizeof(gk_mp));
        memcpy(gkf_prev, gkf_mp, sizeof(gkf_mp));
        memcpy(strain_mat, E_mp, sizeof(E_mp));
        Frhoh = Fr;
        memcpy(Ebar_last, E_mp, sizeof(E_mp));
        memcpy(gk_last, gk_mp, sizeof(gk_mp));
        memcpy(gkf_last, gkf_mp, sizeof(gkf_mp));
    }
    if (!collapsed) { no_collapse(out); return; }

    /* ---- RTHL and the interpolated fields (RadialShell.jl:688-736) ---- */
    double dZvir = zvir1p - zvir1;
    double RTHL;
    if (zvir1 > 0.0 && dZvir != 0.0) {
        double RTHL3 = cube(rad[m0 - 1]) + (cube(rad[m0]) - cube(rad[m0 - 1])) * (zvir1p - ZZon) / dZvir;
        RTHL = pow(RTHL3, 1.0 / 3.0);
    } else {
        RTHL = rad[m0 - 1];
    }
    if (RTHL <= 0.0) { no_collapse(out); return; }
    double RTHL3 = cube(RTHL);
    double RTHL5 = RTHL3 * RTHL * RTHL;
    double rad3p = cube(rad[m0 - 1]);
    double rad3 = cube(rad[m0]);
    double drad3 = rad3 - rad3p;
    double dFb = Fbar[m0] - Fbar[m0 - 1];
    double Fbarx, gk_fin[3], gkf_fin[3];
    if (zvir1 > 0.0 && drad3 != 0.0) {
        double frac = (RTHL3 - rad3p) / drad3;
        Fbarx = Fbar[m0 - 1] + frac * dFb;
        for (int q = 0; q < 9; ++q) strain_mat[q] = Ebar_last[q] + frac * (strain_mat[q] - Ebar_last[q]);
        for (int L = 0; L < 3; ++L) {
            gk_fin[L] = gk_last[L] + frac * (gk_prev[L] - gk_last[L]);
            gkf_fin[L] = gkf_last[L] + frac * (gkf_prev[L] - gkf_last[L]);
        }
    } else {
        Fbarx = Fbar[m0 - 1];
        memcpy(strain_mat, Ebar_last, sizeof(Ebar_last));
        memcpy(gk_fin, gk_prev, sizeof(gk_prev));
        memcpy(gkf_fin, gkf_prev, sizeof(gkf_prev));
    }
    /* normalize_strain (RadialShell.jl:207-210) then the final eigenvalues */
    double tr = strain_mat[0] + strain_mat[4] + strain_mat[8];
    double sn[9];
    if (tr != 0.0) {
        double sc = Fbarx / tr;
        for (int q = 0; q < 9; ++q) sn[q] = strain_mat[q] * sc;
    } else {
        memcpy(sn, strain_mat, sizeof(sn));
    }
    double lamf[3], Frhoc, e_f, p_f;
    eig_full(sn, lamf);
    ev_from_lam(lamf, &Frhoc, &e_f, &p_f);
    double strain_fin[9];
    if (variant == HNT_VARIANT_A || Frhoh != 0.0) {
        double sc2 = Fbarx / Frhoh;
        for (int q = 0; q < 9; ++q) strain_fin[q] = Ebar_last[q] * sc2;
    } else {
        memcpy(strain_fin, Ebar_last, sizeof(strain_fin));
    }

    /* ---- Sbar, Sbar2 over shells 1..m0-1 (RadialShell.jl:738-754) ---- */
    double Sb[3] = {0, 0, 0}, Sb2[3] = {0, 0, 0};
    int64_t nSbar = 0;
    for (int m1 = 1; m1 <= m0 - 1; ++m1) {
        if (w->nsh[m1] > 0) {
            for (int L = 0; L < 3; ++L) {
                Sb[L] += w->S[m1 * 3 + L];
                Sb2[L] += w->S2[m1 * 3 + L];
            }
            nSbar += w->nsh[m1];
        }
    }
    if (nSbar > 0)
        for (int L = 0; L < 3; ++L) {
            Sb[L] /= (double)nSbar;
            Sb2[L] /= (double)nSbar;
        }

    /* ---- Srb (RadialShell.jl:756-771) ---- */
    double Srb = 0.0, rad5p = 0.0;
    if (m0 > 2)
        for (int mp = 2; mp <= m0 - 1; ++mp) {
            double rad5 = pow5(rad[mp]);
            Srb += 0.5 * (Fbar[mp - 1] + Fbar[mp]) * (rad5 - rad5p);
            rad5p = rad5;
        }
    if (zvir1 > 0.0 && dZvir != 0.0) Srb += 0.5 * (Fbar[m0 - 1] + Fbarx) * (RTHL5 - rad5p);
    if (Fbarx * RTHL5 > 0.0) Srb /= (Fbarx * RTHL5);

    /* ---- last shell within RTHL: mask and d2F ---- */
    int sR = 1;
    for (int s = 1; s <= NS; ++s) {
        if (sqrt((double)t->shell_r2[s - 1]) > RTHL) break;
        sR = s;
    }
    if (mask) {
        for (int s = 1; s <= sR; ++s) {
            int32_t st = t->shell_start[s - 1], nc = t->shell_count[s - 1];
            for (int32_t c = 0; c < nc; ++c) {
                int32_t i1 = ci + t->di[st + c], i2 = cj + t->dj[st + c], i3 = ck + t->dk[st + c];
                if (i1 < 1 || i1 > n1 || i2 < 1 || i2 > n2 || i3 < 1 || i3 > n3) continue;
                if (i1 > prm->nbuff && i1 <= n1 - prm->nbuff && i2 > prm->nbuff &&
                    i2 <= n2 - prm->nbuff && i3 > prm->nbuff && i3 <= n3 - prm->nbuff) {
                    int64_t idx = (i1 - 1) + (i2 - 1) * s1 + (i3 - 1) * s2;
                    __atomic_store_n(&mask[idx], (int8_t)1, __ATOMIC_RELAXED);
                }
            }
        }
    }
    double d2F = 0.0;
    if (f->lapd) {
        if (variant == HNT_VARIANT_A) {
            /* centre excluded, stop at the first out-of-bounds cell (RadialShell.jl:784-799) */
            if (w->cntA[sR] > 0) d2F = w->lapA[sR] / (double)w->cntA[sR];
        } else {
            /* centre included, out-of-bounds cells skipped (ShellAnalysisGPU.jl:495-511) */
            if (w->cntB[sR] > 0) d2F = w->lapB[sR] / (double)w->cntB[sR];
        }
    }

    /* ---- formation redshift zvir_half (RadialShell.jl:801-859) ---- */
    double zvir_half = -1.0;
    if (variant == HNT_VARIANT_B) {
        zvir_half = 0.0;
    } else if (RTHL / 1.0 >= 3.0) {
        for (int jj = 1; jj <= 10; ++jj) {
            double rcur = RTHL * pow((double)(jj - 1) / 9.0 * 0.5 + 0.5, 1.0 / 3.0);
            double rupp_rc = rcur + 2.0;
            int mupp_rc = 1;
            for (int m1 = 2; m1 <= m - 1; ++m1) {
                if (rad[m1] > rupp_rc) { mupp_rc = m1; break; }
                else if (rad[m1] == 0.0) { mupp_rc = m1 - 1; break; }
                mupp_rc = m1;
            }
            double rlow_rc = rcur - 2.0;
            if (rlow_rc < 0.0) rlow_rc = 0.0;
            int mlow_rc = 1;
            for (int m1 = 1; m1 <= m - 1; ++m1)
                if (rad[m1] > rlow_rc) { mlow_rc = m1; break; }
            int m0_rc = 1;
            for (int m1 = 1; m1 <= m - 1; ++m1) {
                if (rad[m1] > rcur) { m0_rc = m1; break; }
                m0_rc = m1;
            }
            double E_rc[9], gr[3], gfr[3];
            kstrain1(rad, w->G, w->Gf, w->SR, akk, m0_rc, mlow_rc, mupp_rc, wRnor, aRnor, 1.0,
                     1.0, E_rc, gr, gfr);
            double Fr, e_rc, p_rc;
            eig_full(E_rc, lam);
            ev_from_lam(lam, &Fr, &e_rc, &p_rc);
            double Frhoc_rc = Fbar[m0_rc];
            double poe = e_rc < 1e-5 ? 0.0 : p_rc / e_rc;
            double zr = -1.0;
            if (Frhoc_rc > 0.0) zr = hro_interp3(ct, log10(Frhoc_rc), e_rc, poe);
            double zs = zr - 1.0;
            if (zs > zvir_half) zvir_half = zs;
        }
    }

    out->RTHL = RTHL;
    out->Srb = Srb;
    for (int L = 0; L < 3; ++L) {
        out->Sbar[L] = Sb[L];
        out->Sbar2[L] = Sb2[L];
        out->gradpk[L] = gk_fin[L];
        out->gradpkf[L] = gkf_fin[L];
        out->gradpkrf[L] = gradpkrf[L];
        out->eigs[L] = lamf[L];
    }
    /* strain as a column-major 3x3 (symmetric) */
    for (int L = 0; L < 3; ++L)
        for (int K = 0; K < 3; ++K) out->strain[L + 3 * K] = strain_fin[L * 3 + K];
    out->Fbarx = Fbarx;
    out->e_v = e_f;
    out->p_v = p_f;
    out->zvir_half = zvir_half;
    out->d2F = d2F;
}

static void store(hnt_results* r, int64_t p, const res_t* v) {
    r->RTHL[p] = v->RTHL;
    r->Fbarx[p] = v->Fbarx;
    r->e_v[p] = v->e_v;
    r->p_v[p] = v->p_v;
    r->Srb[p] = v->Srb;
    r->d2F[p] = v->d2F;
    r->zvir_half[p] = v->zvir_half;
    for (int q = 0; q < 9; ++q) r->strain[p * 9 + q] = v->strain[q];
    for (int L = 0; L < 3; ++L) {
        r->eigs[p * 3 + L] = v->eigs[L];
        r->Sbar[p * 3 + L] = v->Sbar[L];
        r->Sbar2[p * 3 + L] = v->Sbar2[L];
        r->gradpk[p * 3 + L] = v->gradpk[L];
        r->gradpkf[p * 3 + L] = v->gradpkf[L];
        r->gradpkrf[p * 3 + L] = v->gradpkrf[L];
    }
}

int hro_analyse_peaks(const hnt_fields* f, const hnt_batch* b, const hnt_shell_tables* t,
                      const hnt_ctab* ct, const hnt_params* p, hnt_results* r, int8_t* mask,
                      int32_t nthreads) {
    if (!f || !b || !t || !ct || !p || !r) return 1;
    if (f->n1 != f->n2 || f->n2 != f->n3) return 2;
    if (t->nshells < 3) return 3;
    if (p->variant < 0 || p->variant > 2) return 5;
    /* rmax2rs > 0: the r^2 limit of variants B and C; variant A limits by a cell count
     * (RadialShell.jl), not restated here */
    if (!(p->rmax2rs >= 0.0) || (p->rmax2rs > 0.0 && p->variant == HNT_VARIANT_A)) return 4;
    /* variant B skips a peak whose cell the mask already holds beyond floor(Rf/alatt
     * rmax2rs) (ShellAnalysisGPU.jl:112-117).  The mask is read as it is at the call: the
     * peaks of one call are analysed together, so one peak's halo does not skip another of
     * the same call (a caller wanting that analyses one peak per call) */
    int8_t* skip = NULL;
    if (p->rmax2rs > 0.0 && p->variant == HNT_VARIANT_B && mask) {
        skip = (int8_t*)calloc((size_t)(b->npeaks > 0 ? b->npeaks : 1), 1);
        if (!skip) return 8;
        const int64_t s1 = f->n1, s2 = (int64_t)f->n1 * f->n2;
        for (int64_t q = 0; q < b->npeaks; ++q) {
            const int64_t c = (b->pi[q] - 1) + (b->pj[q] - 1) * s1 + (b->pk[q] - 1) * s2;
            skip[q] = (double)mask[c] > floor(b->Rf[q] / p->alatt * p->rmax2rs);
        }
    }
    double akk[1101];
    hro_atab4(akk);
    int bad = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(| : bad)
#endif
    {
        prof_t w;
        if (!prof_alloc(&w, t->nshells)) {
            bad |= 8;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 8)
#endif
            for (int64_t q = 0; q < b->npeaks; ++q) {
                res_t v;
                double ZZ = b->ZZon ? b->ZZon[q] : b->ZZon_scalar;
                double fc = b->fcrit ? b->fcrit[q] : b->fcrit_scalar;
                if (!(ZZ > -1.0)) { bad |= 16; no_collapse(&v); store(r, q, &v); continue; }
                if (skip && skip[q]) {
                    no_collapse(&v);
                    store(r, q, &v);
                    if (r->nshells_used) r->nshells_used[q] = 0;
                    if (r->crossing_shell) r->crossing_shell[q] = 0;
                    continue;
                }
                int32_t nu = 0, nx = 0;
                analyse_one(f, t, ct, akk, p, b->pi[q], b->pj[q], b->pk[q], b->Rf[q],
                            b->ir2min[q], ZZ, fc, &w, &v, mask, &nu, &nx);
                store(r, q, &v);
                if (r->nshells_used) r->nshells_used[q] = nu;
                if (r->crossing_shell) r->crossing_shell[q] = nx;
            }
            prof_free(&w);
        }
    }
    free(skip);
    return bad;
}

/* ------------------------------------------------------------------ halo records */

/* _hunt (Cosmology.jl:116-130), 1-based jlo over xx[0..n) */
static int hunt_1(const double* xx, int n, double x) {
    int ascnd = xx[n - 1] > xx[0];
    int jlo = 0, jhi = n + 1;
    while (jhi - jlo > 1) {
        int jm = (jhi + jlo) / 2;
        if ((x > xx[jm - 1]) == ascnd) jlo = jm;
        else jhi = jm;
    }
    return jlo;
}

/* Dlinear_ab(ai)[1] (Cosmology.jl:287-307) with _lagrint_2col's first column (:88-110);
 * Julia indexes xa = atab[kkl:kkl+nlint-1] (1-based). */
double hro_growth_D(const hnt_cosmo_tables* t, double ai) {
    const double* at = t->atab;
    const double* Dt = t->D_atab;
    int ntab = t->ntab;
    if (ai > at[0] && ai < at[ntab - 1]) {
        int kk = hunt_1(at, ntab, ai);
        int nlint, kkl;
        if (kk < 2 || kk >= ntab) {
            nlint = 2;
            kkl = kk - 1 > 1 ? kk - 1 : 1;
        } else {
            nlint = 4;
            kkl = kk - 2;
        }
        if (kkl < 1) return NAN; /* the reference's index would fall before its table */
        double xa1 = at[kkl - 1], xa2 = at[kkl], x = ai;
        double D_a;
        if (nlint == 4) {
            double xa3 = at[kkl + 1], xa4 = at[kkl + 2];
            double a1 = (x - xa2) / (xa1 - xa2) * (x - xa3) / (xa1 - xa3) * (x - xa4) / (xa1 - xa4);
            double a2 = (x - xa1) / (xa2 - xa1) * (x - xa3) / (xa2 - xa3) * (x - xa4) / (xa2 - xa4);
            double a3 = (x - xa1) / (xa3 - xa1) * (x - xa2) / (xa3 - xa2) * (x - xa4) / (xa3 - xa4);
            double a4 = (x - xa1) / (xa4 - xa1) * (x - xa2) / (xa4 - xa2) * (x - xa3) / (xa4 - xa3);
            D_a = a1 * Dt[kkl - 1] + a2 * Dt[kkl] + a3 * Dt[kkl + 1] + a4 * Dt[kkl + 2];
        } else {
            double a1 = (x - xa2) / (xa1 - xa2);
            double a2 = 1.0 - a1;
            D_a = a1 * Dt[kkl - 1] + a2 * Dt[kkl];
        }
        return D_a * ai;
    } else if (ai >= at[ntab - 1]) {
        return Dt[ntab - 1] * ai;
    }
    return Dt[0] * ai;
}

/* Dlinear_ab(ai)[2] = f (Cosmology.jl:287-307): _lagrint_2col's second column with the same
 * weights; the table's end values outside it. */
double hro_growth_f(const hnt_cosmo_tables* t, double ai) {
    const double* at = t->atab;
    const double* ft = t->dlnD_atab;
    int ntab = t->ntab;
    if (ai > at[0] && ai < at[ntab - 1]) {
        int kk = hunt_1(at, ntab, ai);
        int nlint, kkl;
        if (kk < 2 || kk >= ntab) {
            nlint = 2;
            kkl = kk - 1 > 1 ? kk - 1 : 1;
        } else {
            nlint = 4;
            kkl = kk - 2;
        }
        if (kkl < 1) return NAN; /* the reference's index would fall before its table */
        double xa1 = at[kkl - 1], xa2 = at[kkl], x = ai;
        if (nlint == 4) {
            double xa3 = at[kkl + 1], xa4 = at[kkl + 2];
            double a1 = (x - xa2) / (xa1 - xa2) * (x - xa3) / (xa1 - xa3) * (x - xa4) / (xa1 - xa4);
            double a2 = (x - xa1) / (xa2 - xa1) * (x - xa3) / (xa2 - xa3) * (x - xa4) / (xa2 - xa4);
            double a3 = (x - xa1) / (xa3 - xa1) * (x - xa2) / (xa3 - xa2) * (x - xa4) / (xa3 - xa4);
            double a4 = (x - xa1) / (xa4 - xa1) * (x - xa2) / (xa4 - xa2) * (x - xa3) / (xa4 - xa3);
            return a1 * ft[kkl - 1] + a2 * ft[kkl] + a3 * ft[kkl + 1] + a4 * ft[kkl + 2];
        }
        double a1 = (x - xa2) / (xa1 - xa2);
        double a2 = 1.0 - a1;
        return a1 * ft[kkl - 1] + a2 * ft[kkl];
    } else if (ai >= at[ntab - 1]) {
        return ft[ntab - 1];
    }
    return ft[0];
}

/* chi_to_z (Cosmology.jl:357-372) */
double hro_chi_to_z(const hnt_cosmo_tables* t, double chi_val) {
    const double* zs = t->chi2z_z;
    const double* chis = t->chi2z_chi;
    int n = t->nchi;
    if (chi_val <= chis[0]) return 0.0;
    if (chi_val >= chis[n - 1]) return t->chi2z_zmax;
    int lo = 1, hi = n;
    while (hi - lo > 1) {
        int mid = (int)(((unsigned)lo + (unsigned)hi) >> 1);
        if (chis[mid - 1] <= chi_val) lo = mid;
        else hi = mid;
    }
    double frac = (chi_val - chis[lo - 1]) / (chis[hi - 1] - chis[lo - 1]);
    return zs[lo - 1] + frac * (zs[hi - 1] - zs[lo - 1]);
}

int64_t hro_pack_records(int64_t npeaks, const int32_t* pi, const int32_t* pj,
                         const int32_t* pk, const double* Rf, const hnt_results* r,
                         const float* FcRf, const float* d2Rf, const hnt_cosmo_tables* t,
                         const hnt_pack_params* pp, float* out, int64_t capacity) {
    const int nf = pp->ioutshear >= 1 ? 33 : 11;
    const double coef2_base = -3.0 / 7.0;
    const double cen1 = 0.5 * (pp->n + 1);  /* peak_find_tile_gpu cell centres */
    int64_t nrec = 0;
    for (int64_t idx = 0; idx < npeaks; ++idx) {
        float RTHL32 = (float)r->RTHL[idx];
        if (RTHL32 <= 0.0f) continue;
        double px = pp->xbx + pp->alatt * ((double)pi[idx] - cen1);
        double py = pp->ybx + pp->alatt * ((double)pj[idx] - cen1);
        double pz = pp->zbx + pp->alatt * ((double)pk[idx] - cen1);
        double ZZon_pk = pp->ZZon;
        if (pp->ievol == 1) {
            double dx = px - pp->obs[0], dy = py - pp->obs[1], dz = pz - pp->obs[2];
            double rr = sqrt(dx * dx + dy * dy + dz * dz);
            double z_pk = hro_chi_to_z(t, rr);
            ZZon_pk = 1.0 + z_pk;
            if (z_pk > pp->z_max) continue;
        }
        double a_pk = 1.0 / ZZon_pk;
        double D_pk = hro_growth_D(t, a_pk);
        float D32 = (float)D_pk;
        float RTHL_phys = (float)((double)RTHL32 * pp->alatt);
        float S1 = (float)r->Sbar[idx * 3] * D32, S2 = (float)r->Sbar[idx * 3 + 1] * D32,
              S3 = (float)r->Sbar[idx * 3 + 2] * D32;
        float T1 = 0.0f, T2 = 0.0f, T3 = 0.0f;
        if (pp->ilpt >= 2) {
            double a3 = a_pk * a_pk * a_pk;
            double Om_a = (pp->Om / a3) / (pp->Om / a3 + pp->OL);
            float coef = (float)(coef2_base * pow(Om_a, -1.0 / 143) * (D_pk * D_pk));
            T1 = (float)r->Sbar2[idx * 3] * coef;
            T2 = (float)r->Sbar2[idx * 3 + 1] * coef;
            T3 = (float)r->Sbar2[idx * 3 + 2] * coef;
        }
        if (nrec < capacity) {
            float* o = out + nrec * nf;
            o[0] = (float)px; o[1] = (float)py; o[2] = (float)pz;
            o[3] = S1; o[4] = S2; o[5] = S3;
            o[6] = RTHL_phys;
            o[7] = T1; o[8] = T2; o[9] = T3;
            o[10] = (float)r->Fbarx[idx];
            if (nf == 33) {
                const double* st = r->strain + idx * 9;
                o[11] = (float)r->e_v[idx]; o[12] = (float)r->p_v[idx];
                o[13] = (float)st[0]; o[14] = (float)st[4]; o[15] = (float)st[8];
                o[16] = (float)st[5]; o[17] = (float)st[2]; o[18] = (float)st[1];
                o[19] = (float)r->d2F[idx];
                o[20] = (float)r->zvir_half[idx];
                for (int L = 0; L < 3; ++L) o[21 + L] = (float)r->gradpk[idx * 3 + L];
                for (int L = 0; L < 3; ++L) o[24 + L] = (float)r->gradpkf[idx * 3 + L];
                o[27] = (float)Rf[idx];
                o[28] = FcRf[idx];
                o[29] = d2Rf ? d2Rf[idx] : 0.0f;
                for (int L = 0; L < 3; ++L) o[30 + L] = (float)r->gradpkrf[idx * 3 + L];
            }
        }
        nrec++;
    }
    return nrec;
}

/* ------------------------------------------------------------------ catalog merge */

void hro_sphere_overlap(double d, double r1, double r2, double* v1, double* v2) {
    *v1 = *v2 = 0.0;
    if (d >= r1 + r2) return;
    if (d <= fabs(r1 - r2)) return;
    double h1 = (r2 - r1 + d) * (r2 + r1 - d) / (2.0 * d);
    double h2 = (r1 - r2 + d) * (r1 + r2 - d) / (2.0 * d);
    *v1 = (PI_D / 3.0) * (h1 * h1) * (3.0 * r1 - h1);
    *v2 = (PI_D / 3.0) * (h2 * h2) * (3.0 * r2 - h2);
}

static const double* mg_x;
static const double* mg_y;
static const double* mg_z;
static const double* mg_r;

/* (-r, x, y, z), then catalog index: the stable sortperm of Merger.jl:59 */
static int mg_cmp(const void* pa, const void* pb) {
    int64_t a = *(const int64_t*)pa, b = *(const int64_t*)pb;
    double ka[4] = {-mg_r[a], mg_x[a], mg_y[a], mg_z[a]};
    double kb[4] = {-mg_r[b], mg_x[b], mg_y[b], mg_z[b]};
    for (int q = 0; q < 4; ++q) {
        if (ka[q] < kb[q]) return -1;
        if (ka[q] > kb[q]) return 1;
        /* Julia's isless orders -0.0 before 0.0 */
        if (ka[q] == 0.0 && kb[q] == 0.0 && signbit(ka[q]) != signbit(kb[q]))
            return signbit(ka[q]) ? -1 : 1;
    }
    return a < b ? -1 : (a > b ? 1 : 0);
}

static int mg_cell(double p, double o, double cs, int nc) {
    double f = floor((p - o) / cs);
    long c = (long)f + 1;
    return c < 1 ? 1 : (c > nc ? nc : (int)c);
}

/* The two passes of _merge_impl (Merger.jl:47-79): survivors and the new radii in FP64.
 * Returns 0, -1 if memory runs out, -2 for an empty domain (the reference's floor of NaN). */
int32_t hro_merge_passes(const float* rec, int64_t N, int32_t nf, int32_t reduce,
                         int32_t fortran_ties, uint8_t* surv, double* newr) {
    if (N == 0) return 0;
    double* x = malloc(sizeof(double) * N);
    double* y = malloc(sizeof(double) * N);
    double* z = malloc(sizeof(double) * N);
    double* r = malloc(sizeof(double) * N);
    int64_t* order = malloc(sizeof(int64_t) * N);
    int64_t* rank = malloc(sizeof(int64_t) * N);
    int32_t* ll = calloc(N, sizeof(int32_t));
    double* dV = calloc(N, sizeof(double));
    int32_t status = 0;
    int32_t* hoc = NULL;
    if (!x || !y || !z || !r || !order || !rank || !ll || !dV) {
        status = -1;
        goto done;
    }
    double rmax = 0.0, mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t m = 0; m < N; ++m) {
        x[m] = rec[m * nf]; y[m] = rec[m * nf + 1]; z[m] = rec[m * nf + 2]; r[m] = rec[m * nf + 6];
        order[m] = m;
        surv[m] = 1;
        if (r[m] > rmax) rmax = r[m];
        double p[3] = {x[m], y[m], z[m]};
        for (int c = 0; c < 3; ++c) {
            if (p[c] < mn[c]) mn[c] = p[c];
            if (p[c] > mx[c]) mx[c] = p[c];
        }
    }
    mg_x = x; mg_y = y; mg_z = z; mg_r = r;
    qsort(order, (size_t)N, sizeof(int64_t), mg_cmp);
    for (int64_t q = 0; q < N; ++q) rank[order[q]] = q;
    double pad = 2.0 * rmax;
    double dmin[3] = {mn[0] - pad, mn[1] - pad, mn[2] - pad};
    double L = 0.0;
    for (int c = 0; c < 3; ++c) {
        double e = (mx[c] + pad) - dmin[c];
        if (e > L) L = e;
    }
    if (!(L > 0.0)) {
        status = -2;
        goto done;
    }
    int nc = (int)nearbyint(cbrt((double)N));
    nc = nc < 16 ? 16 : (nc > 1024 ? 1024 : nc);
    double cs = L / nc;
    hoc = calloc((size_t)nc * nc * nc, sizeof(int32_t));
    if (!hoc) {
        status = -1;
        goto done;
    }
#define HOC(ix, iy, iz) hoc[((ix) - 1) + (int64_t)nc * (((iy) - 1) + (int64_t)nc * ((iz) - 1))]
    for (int64_t m = 0; m < N; ++m) {  /* build_hash: linked lists, newest first */
        int ix = mg_cell(x[m], dmin[0], cs, nc), iy = mg_cell(y[m], dmin[1], cs, nc),
            iz = mg_cell(z[m], dmin[2], cs, nc);
        ll[m] = HOC(ix, iy, iz);
        HOC(ix, iy, iz) = (int32_t)(m + 1);
    }
    /* lagrangian_exclusion! */
    for (int64_t q = 0; q < N; ++q) {
        int64_t i = order[q];
        if (!surv[i]) continue;
        double ri = r[i], xi = x[i], yi = y[i], zi = z[i];
        int dcell = (int)ceil(ri / cs);
        int ci = mg_cell(xi, dmin[0], cs, nc), cj = mg_cell(yi, dmin[1], cs, nc),
            ck = mg_cell(zi, dmin[2], cs, nc);
        for (int dz = -dcell; dz <= dcell; ++dz)
            for (int dy = -dcell; dy <= dcell; ++dy)
                for (int dx = -dcell; dx <= dcell; ++dx) {
                    int jx = dx + ci, jy = dy + cj, jz = dz + ck;
                    if (jx < 1 || jx > nc || jy < 1 || jy > nc || jz < 1 || jz > nc) continue;
                    int32_t j1 = HOC(jx, jy, jz);
                    while (j1 > 0) {
                        int64_t j = j1 - 1;
                        if (j != i && surv[j]) {
                            double rj = r[j];
                            if (ri >= rj) {
                                double ex = x[j] - xi, ey = y[j] - yi, ez = z[j] - zi;
                                double dist = sqrt(ex * ex + ey * ey + ez * ez);
                                if (dist < ri) surv[j] = 0;
                            }
                        }
                        j1 = ll[j];
                    }
                }
    }
    /* volume_reduction! */
    for (int64_t m = 0; m < N; ++m) newr[m] = r[m];
    if (reduce) {
        for (int64_t q = 0; q < N; ++q) {
            int64_t i = order[q];
            if (!surv[i]) continue;
            double ri = r[i], xi = x[i], yi = y[i], zi = z[i];
            int dcell = (int)ceil(2.0 * ri / cs);
            int ci = mg_cell(xi, dmin[0], cs, nc), cj = mg_cell(yi, dmin[1], cs, nc),
                ck = mg_cell(zi, dmin[2], cs, nc);
            for (int dz = -dcell; dz <= dcell; ++dz)
                for (int dy = -dcell; dy <= dcell; ++dy)
                    for (int dx = -dcell; dx <= dcell; ++dx) {
                        int jx = dx + ci, jy = dy + cj, jz = dz + ck;
                        if (jx < 1 || jx > nc || jy < 1 || jy > nc || jz < 1 || jz > nc) continue;
                        int32_t j1 = HOC(jx, jy, jz);
                        while (j1 > 0) {
                            int64_t j = j1 - 1;
                            if (j != i && surv[j] && rank[j] > rank[i]) {
                                double rj = r[j];
                                double ex = x[j] - xi, ey = y[j] - yi, ez = z[j] - zi;
                                double d = sqrt(ex * ex + ey * ey + ez * ez);
                                if (d < ri + rj) {
                                    double v1, v2;
                                    hro_sphere_overlap(d, ri, rj, &v1, &v2);
                                    double w = (fortran_ties && rj == ri) ? 2.0 : 1.0;
                                    dV[i] += w * v1;
                                    dV[j] += w * v2;
                                }
                            }
                            j1 = ll[j];
                        }
                    }
        }
        for (int64_t m = 0; m < N; ++m) {
            if (!surv[m]) continue;
            double vol_new = 4.0 * PI_D / 3.0 * (r[m] * r[m] * r[m]) - dV[m];
            if (vol_new <= 0.0) {
                surv[m] = 0;
                newr[m] = 0.0;
            } else {
                newr[m] = cbrt(vol_new * 3.0 / (4.0 * PI_D));
            }
        }
    }
#undef HOC
done:
    free(x); free(y); free(z); free(r); free(order); free(rank); free(ll); free(dV); free(hoc);
    return status;
}

int64_t hro_merge_catalog(const float* rec, int64_t N, int32_t nf, int32_t reduce,
                          int32_t fortran_ties, float* out) {
    if (N == 0) return 0;
    uint8_t* surv = malloc(N);
    double* newr = malloc(sizeof(double) * N);
    if (!surv || !newr) return -1;
    int32_t st = hro_merge_passes(rec, N, nf, reduce, fortran_ties, surv, newr);
    int64_t nout = 0;
    for (int64_t m = 0; st == 0 && m < N; ++m) {
        if (!surv[m]) continue;
        for (int c = 0; c < nf; ++c) out[nout * nf + c] = rec[m * nf + c];
        out[nout * nf + 6] = (float)newr[m];
        nout++;
    }
    free(surv); free(newr);
    return st == 0 ? nout : st;
}

/* ------------------------------------------------------------------ finalize_eulerian */

/* Julia's `x^-3` for a Float64 x: Base.literal_pow forwards to x^n (Int64 n), whose
 * integer-power routine (base/math.jl, power by squaring) takes r = 1/x with the residual
 * rlo = -fma(x, r, -1) r and carries the product r * r * r as a double-double: y = r
 * (low part rlo); r^2 by two_mul with its low part plus the cross term 2 r rlo; the result
 * fma(r^2, y, fma(y, lo(r^2), r^2 * lo(y))).  muladd is fused on FMA hosts, so fma here. */
double hro_pow_m3(double x) {
    double r = 1.0 / x;
    double rlo = isfinite(x) ? -fma(x, r, -1.0) * r : 0.0;
    double y = r, ylo = rlo;
    double x2 = r * r;
    double x2lo = fma(r, r, -x2) + r * 2.0 * rlo;
    if (!isfinite(x2)) return x2 * y;
    return fma(x2, y, fma(y, x2lo, x2 * ylo));
}

/* Merger.finalize_eulerian (src/Merger/Merger.jl:123-166): Eulerian position q + d1 + d2 and
 * velocity a 100 E(a) f(a) (d1 + 2 d2), E = sqrt(Om a^-3 + OL) with a^-3 as Julia's literal
 * power evaluates it (hro_pow_m3); the
 * 2LPT fields zeroed, the rest kept.  a from the Eulerian distance to obs (ievol = 1) or
 * 1/(1 + z_out). */
void hro_finalize_eulerian(const float* rec, int64_t N, int32_t nf, const hnt_cosmo_tables* t,
                           double Om, double OL, const double* obs, int32_t ievol, double z_out,
                           float* out) {
    double a_snap = 1.0 / (1.0 + z_out);
    for (int64_t i = 0; i < N; ++i) {
        const float* h = rec + i * nf;
        float* o = out + i * nf;
        double d1x = h[3], d1y = h[4], d1z = h[5];
        double d2x = h[7], d2y = h[8], d2z = h[9];
        double ex = (double)h[0] + d1x + d2x, ey = (double)h[1] + d1y + d2y,
               ez = (double)h[2] + d1z + d2z;
        double a = a_snap;
        if (ievol == 1) {
            double rx = ex - obs[0], ry = ey - obs[1], rz = ez - obs[2];
            double r = sqrt(rx * rx + ry * ry + rz * rz);
            a = r > 0.0 ? 1.0 / (1.0 + hro_chi_to_z(t, r)) : 1.0;
        }
        double E = sqrt(Om * hro_pow_m3(a) + OL);
        double f = hro_growth_f(t, a);
        double vfac = a * 100.0 * E * f;
        float keep[33];
        for (int q = 0; q < nf; ++q) keep[q] = h[q];
        o[0] = (float)ex;
        o[1] = (float)ey;
        o[2] = (float)ez;
        o[3] = (float)(vfac * (d1x + 2.0 * d2x));
        o[4] = (float)(vfac * (d1y + 2.0 * d2y));
        o[5] = (float)(vfac * (d1z + 2.0 * d2z));
        o[6] = keep[6];
        o[7] = o[8] = o[9] = 0.0f;
        for (int q = 10; q < nf; ++q) o[q] = keep[q];
    }
}

/* ------------------------------------------------------------------ tile white noise */

/* Threefry-2x64 with 20 rounds (Salmon, Moraes, Dror, Shaw, SC'11, the Random123 generator
 * RandomField.jl calls as threefry((seed, 0), (pair, 0), Val(20)), key first): key words
 * k0, k1 and the parity word 0x1BD11BDAA9FC1A22 ^ k0 ^ k1; each round adds word 1 into
 * word 0, rotates word 1 left by 16 42 12 31 16 32 24 21 (round mod 8) and xors word 0 into
 * it; after rounds 4, 8, .., 20 the key word s mod 3 is added to word 0 and key word
 * (s + 1) mod 3 plus s to word 1 (s = 1 .. 5); the words start as counter + key. */
void hro_threefry2x64_20(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1, uint64_t* out) {
    static const int rot[8] = {16, 42, 12, 31, 16, 32, 24, 21};
    const uint64_t ks[3] = {k0, k1, 0x1BD11BDAA9FC1A22ull ^ k0 ^ k1};
    uint64_t x0 = c0 + ks[0], x1 = c1 + ks[1];
    for (int r = 0; r < 20; ++r) {
        x0 += x1;
        x1 = (x1 << rot[r % 8]) | (x1 >> (64 - rot[r % 8]));
        x1 ^= x0;
        if (r % 4 == 3) {
            int s = (r + 1) / 4;
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3] + (uint64_t)s;
        }
    }
    out[0] = x0;
    out[1] = x1;
}

/* _threefry_gaussian (RandomField.jl:45-53) for the cell of 0-based global index idx0:
 * g1 for an even index, g2 for an odd one, of pair idx0 >> 1. */
double hro_threefry_gauss(uint64_t seed, uint64_t idx0) {
    uint64_t w[2];
    hro_threefry2x64_20(idx0 >> 1, 0, seed, 0, w);
    double u1 = ((double)(w[0] >> 11) + 1.0) * 1.1102230246251565e-16; /* (x + 1) 2^-53 */
    double u2 = ((double)(w[1] >> 11) + 1.0) * 1.1102230246251565e-16;
    double r = sqrt(-2.0 * log(u1));
    double th = (2.0 * 3.141592653589793) * u2; /* 2 pi in Float64 */
    return (idx0 % 2 == 0) ? r * cos(th) : r * sin(th);
}

static int64_t mod1_(int64_t x, int64_t n) { /* Julia's mod1 */
    int64_t m = (x - 1) % n;
    if (m < 0) m += n;
    return m + 1;
}

/* _generate_tile_noise (MultiResolution.jl:35-57) and _generate_extended_residual
 * (:190-227): the side^3 region from global cell (i0, j0, k0), 1-based, wrapped; with
 * coarse, Float32 fine value less Float32 coarse block value times Float32(1/sqrt(block^3)). */
static void region_noise(int64_t i0, int64_t j0, int64_t k0, int64_t side, int64_t N,
                         uint64_t seed, const float* coarse, int64_t M, float* out) {
    int64_t block = coarse ? N / M : 1;
    double b = (double)block;
    float inv = (float)(1.0 / sqrt(b * b * b));
#pragma omp parallel for schedule(static)
    for (int64_t lk = 1; lk <= side; ++lk)
        for (int64_t lj = 1; lj <= side; ++lj)
            for (int64_t li = 1; li <= side; ++li) {
                int64_t gi = mod1_(i0 + li - 1, N), gj = mod1_(j0 + lj - 1, N);
                int64_t gk = mod1_(k0 + lk - 1, N);
                int64_t idx = gi + (gj - 1) * N + (gk - 1) * N * N;
                float fine = (float)hro_threefry_gauss(seed, (uint64_t)(idx - 1));
                int64_t o = (li - 1) + (lj - 1) * side + (lk - 1) * side * side;
                if (!coarse) {
                    out[o] = fine;
                } else {
                    int64_t ci = (gi - 1) / block + 1, cj = (gj - 1) / block + 1;
                    int64_t ck = (gk - 1) / block + 1;
                    float mean = coarse[(ci - 1) + (cj - 1) * M + (ck - 1) * M * M] * inv;
                    out[o] = fine - mean;
                }
            }
}

void hro_tile_noise(int32_t it, int32_t jt, int32_t kt, int32_t nsub, int32_t nmesh, int64_t N,
                    uint64_t seed, float* out) {
    int64_t nbuff = (nmesh - nsub) / 2;
    region_noise((int64_t)(it - 1) * nsub + 1 - nbuff, (int64_t)(jt - 1) * nsub + 1 - nbuff,
                 (int64_t)(kt - 1) * nsub + 1 - nbuff, nmesh, N, seed, NULL, 1, out);
}

void hro_extended_residual(int32_t it, int32_t jt, int32_t kt, int32_t nsub, int32_t nmesh,
                           int64_t N, uint64_t seed, const float* coarse, int32_t M,
                           int32_t nshell, float* out) {
    int64_t nbuff = (nmesh - nsub) / 2;
    region_noise((int64_t)(it - 1) * nsub + 1 - nbuff - nshell,
                 (int64_t)(jt - 1) * nsub + 1 - nbuff - nshell,
                 (int64_t)(kt - 1) * nsub + 1 - nbuff - nshell, nmesh + 2 * nshell, N, seed,
                 coarse, M, out);
}

/* _downsample_noise (MultiResolution.jl:64-87): FP64 block sums in (dk, dj, di) order. */
void hro_downsample_noise(int64_t N, int32_t M, uint64_t seed, float* coarse) {
    int64_t block = N / M;
    double b = (double)block;
    double norm = sqrt(b * b * b);
#pragma omp parallel for schedule(static)
    for (int64_t K = 1; K <= M; ++K)
        for (int64_t J = 1; J <= M; ++J)
            for (int64_t I = 1; I <= M; ++I) {
                double val = 0.0;
                for (int64_t dk = 0; dk < block; ++dk)
                    for (int64_t dj = 0; dj < block; ++dj)
                        for (int64_t di = 0; di < block; ++di) {
                            int64_t gi = (I - 1) * block + di + 1, gj = (J - 1) * block + dj + 1;
                            int64_t gk = (K - 1) * block + dk + 1;
                            int64_t idx = gi + (gj - 1) * N + (gk - 1) * N * N;
                            val += hro_threefry_gauss(seed, (uint64_t)(idx - 1));
                        }
                coarse[(I - 1) + (J - 1) * M + (K - 1) * M * M] = (float)(val / norm);
            }
}

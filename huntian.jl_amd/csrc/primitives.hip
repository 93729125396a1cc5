// primitives.hip -- single-primitive entry points of the path.  The reference exposes the
// same primitives as GPU test harnesses (src/Huntian.jl:98-217): the full shell gather,
// the symmetric 3x3 eigen-solve, the collapse-table lookup and one kernel_strain call.
#include "hnt_internal.h"

namespace {

using namespace hnt;

// Per-shell sums of every shell for every peak, one thread per (peak, shell), cells in
// table order: the arithmetic of ShellAnalysisGPU.analyse_peak_gpu (ShellAnalysisGPU.jl:153-195).
__global__ void gather_full_kernel(const float* __restrict__ delta, const float* __restrict__ ex,
                                   const float* __restrict__ ey, const float* __restrict__ ez,
                                   const float* __restrict__ e2x, const float* __restrict__ e2y,
                                   const float* __restrict__ e2z, const float* __restrict__ lapd,
                                   int n1, int n2, int n3, int64_t np,
                                   const int32_t* __restrict__ pi, const int32_t* __restrict__ pj,
                                   const int32_t* __restrict__ pk, const int4* __restrict__ cells,
                                   const int32_t* __restrict__ start,
                                   const int32_t* __restrict__ count, int ns,
                                   double* __restrict__ F, int32_t* __restrict__ N,
                                   double* __restrict__ S, double* __restrict__ S2,
                                   double* __restrict__ G, double* __restrict__ Gf,
                                   double* __restrict__ SR, double* __restrict__ L) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np * ns) return;
    const int64_t p = q / ns;
    const int s = (int)(q - p * ns);
    const int ci = pi[p], cj = pj[p], ck = pk[p];
    const int64_t s1 = n1, s2 = (int64_t)n1 * n2;
    double fs = 0.0, lp = 0.0, sv[3] = {0, 0, 0}, tv[3] = {0, 0, 0}, gv[3] = {0, 0, 0},
           hv[3] = {0, 0, 0}, rv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int cnt = 0;
    for (int c = 0; c < count[s]; ++c) {
        const int4 o = cells[start[s] + c];
        const int i1 = ci + o.x, i2 = cj + o.y, i3 = ck + o.z;
        if (i1 < 1 || i1 > n1 || i2 < 1 || i2 > n2 || i3 < 1 || i3 > n3) continue;
        ++cnt;
        const int64_t idx = (i1 - 1) + (i2 - 1) * s1 + (i3 - 1) * s2;
        const int64_t idxT = (i3 - 1) + (i1 - 1) * s1 + (i2 - 1) * s2;
        const float d = delta[idx], dT = delta[idxT];
        const float e[3] = {ex[idx], ey[idx], ez[idx]};
        const float ix[3] = {(float)o.x, (float)o.y, (float)o.z};
        fs += (double)d;
        for (int a = 0; a < 3; ++a) {
            sv[a] += (double)e[a];
            gv[a] += (double)(d * ix[a]);
            hv[a] += (double)(dT * ix[a]);
            for (int b = 0; b < 3; ++b) rv[a * 3 + b] += (double)(e[a] * ix[b]);
        }
        if (e2x) {
            tv[0] += (double)e2x[idx];
            tv[1] += (double)e2y[idx];
            tv[2] += (double)e2z[idx];
        }
        if (lapd) lp += (double)lapd[idx];
    }
    F[q] = fs;
    N[q] = cnt;
    L[q] = lp;
    for (int a = 0; a < 3; ++a) {
        S[q * 3 + a] = sv[a];
        S2[q * 3 + a] = tv[a];
        G[q * 3 + a] = gv[a];
        Gf[q * 3 + a] = hv[a];
        for (int b = 0; b < 3; ++b) SR[q * 9 + a + 3 * b] = rv[a * 3 + b];
    }
}

__global__ void eig3_kernel(const double* __restrict__ mats, int64_t n, double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double* A = mats + q * 9;  // column-major: upper triangle (1,2) at 3, (1,3) at 6, (2,3) at 7
    double lam[3];
    eig3_sym(A[0], A[3], A[6], A[4], A[7], A[8], lam);
    out[q * 3] = lam[0];
    out[q * 3 + 1] = lam[1];
    out[q * 3 + 2] = lam[2];
}

__global__ void interp3_kernel(CTabView ct, const double* __restrict__ x,
                               const double* __restrict__ y, const double* __restrict__ z,
                               int64_t n, double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    out[q] = interp3(ct, x[q], y[q], z[q]);
}

// One kernel_strain evaluation on 1-based shell arrays (RadialShell.jl:225-274); SR is the
// 3x3 per shell with (L, K) at L*3 + K.
__global__ void kstrain_kernel(const double* __restrict__ rad, const double* __restrict__ G,
                               const double* __restrict__ Gf, const double* __restrict__ SR,
                               const double* __restrict__ akk, int mp, int mlow, int mupp,
                               double wRnor, double aRnor, double h1, double h2,
                               double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double E[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, g[3] = {0, 0, 0}, gf[3] = {0, 0, 0};
    auto add = [&](int m1, double wt) {
        for (int L = 0; L < 3; ++L) {
            g[L] += wt * G[m1 * 3 + L];
            gf[L] += wt * Gf[m1 * 3 + L];
            for (int K = 0; K < 3; ++K)
                E[L * 3 + K] -= 0.5 * wt * (SR[m1 * 9 + L * 3 + K] + SR[m1 * 9 + K * 3 + L]);
        }
    };
    if (mp > 1) {
        const double rmp = rad[mp];
        for (int m1 = mlow; m1 <= mupp; ++m1) {
            const double u0 = h1 * (rmp - rad[m1]);
            const double u02 = u0 * u0;
            if (u02 < 4.0) {
                const double u12 = h2 * (rmp * rmp + rad[m1] * rad[m1]);
                const double d = u12 - u02;
                add(m1, hRinteg(u0, u12) / (d * d));
            }
        }
        const double sc = wRnor / rmp;
        for (int q = 0; q < 9; ++q) E[q] *= sc;
        for (int q = 0; q < 3; ++q) { g[q] *= sc; gf[q] *= sc; }
    } else {
        for (int m1 = mlow; m1 <= mupp; ++m1) {
            const double u0 = h1 * rad[m1];
            const double con = 100.0 * u0 * u0;
            const double fl = floor(con);
            if (!(fl >= 0.0 && fl < 1100.0)) continue;
            const int icon = (int)fl;
            const double diff = con - (double)icon;
            const double aww = akk[icon] + diff * (akk[icon + 1] - akk[icon]);
            add(m1, 0.5 * aww / rad[m1]);
        }
        for (int q = 0; q < 9; ++q) E[q] *= aRnor;
        for (int q = 0; q < 3; ++q) { g[q] *= aRnor; gf[q] *= aRnor; }
    }
    out[0] = E[0]; out[1] = E[1]; out[2] = E[2];
    out[3] = E[4]; out[4] = E[5]; out[5] = E[8];
    out[6] = g[0]; out[7] = g[1]; out[8] = g[2];
    out[9] = gf[0]; out[10] = gf[1]; out[11] = gf[2];
}

}  // namespace

using namespace hnt;

extern "C" {

hnt_status hnt_shell_gather_full(hnt_ctx* ctx, const hnt_fields* f, int64_t npeaks,
                                 const int32_t* pi, const int32_t* pj, const int32_t* pk,
                                 const hnt_shell_tables* tables, double* F, int32_t* n,
                                 double* S, double* S2, double* G, double* Gf, double* SR,
                                 double* lapd) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(f && tables && f->delta && f->etax && f->etay && f->etaz,
                "hnt_shell_gather_full: NULL argument");
    HNT_REQUIRE(f->n1 == f->n2 && f->n2 == f->n3,
                "hnt_shell_gather_full: the transposed read needs a cubic grid");
    HNT_REQUIRE(npeaks >= 0, "hnt_shell_gather_full: npeaks < 0");
    HNT_TRY(upload_tables(ctx, tables));
    if (npeaks == 0) return HNT_OK;
    HNT_REQUIRE(pi && pj && pk && F && n && S && S2 && G && Gf && SR && lapd,
                "hnt_shell_gather_full: NULL array");
    const int ns = tables->nshells;
    const int64_t tot = npeaks * ns;
    hipLaunchKernelGGL(gather_full_kernel, dim3((unsigned)((tot + 127) / 128)), dim3(128), 0,
                       ctx->stream, f->delta, f->etax, f->etay, f->etaz, f->eta2x, f->eta2y,
                       f->eta2z, f->lapd, f->n1, f->n2, f->n3, npeaks, pi, pj, pk,
                       (const int4*)ctx->tables.cells.p, (const int32_t*)ctx->tables.start.p,
                       (const int32_t*)ctx->tables.count.p, ns, F, n, S, S2, G, Gf, SR, lapd);
    return finish_call(ctx, "hnt_shell_gather_full");
}

hnt_status hnt_eig3_symmetric_batch(hnt_ctx* ctx, const double* mats, int64_t n, double* out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(n >= 0, "hnt_eig3_symmetric_batch: n < 0");
    if (n == 0) return HNT_OK;
    HNT_REQUIRE(mats && out, "hnt_eig3_symmetric_batch: NULL array");
    hipLaunchKernelGGL(eig3_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, ctx->stream,
                       mats, n, out);
    return finish_call(ctx, "hnt_eig3_symmetric_batch");
}

hnt_status hnt_interp3_trilinear(hnt_ctx* ctx, const hnt_ctab* ctab, const double* x,
                                 const double* y, const double* z, int64_t n, double* out) {
    HNT_TRY(begin_call(ctx));
    HNT_TRY(upload_ctab(ctx, ctab));
    HNT_REQUIRE(n >= 0, "hnt_interp3_trilinear: n < 0");
    if (n == 0) return HNT_OK;
    HNT_REQUIRE(x && y && z && out, "hnt_interp3_trilinear: NULL array");
    CTabView v{(const float*)ctx->ctab.table.p, ctab->nx, ctab->ny, ctab->nz, ctab->x1, ctab->x2,
               ctab->y1, ctab->y2, ctab->z1, ctab->z2, ctab->out_val};
    hipLaunchKernelGGL(interp3_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0,
                       ctx->stream, v, x, y, z, n, out);
    return finish_call(ctx, "hnt_interp3_trilinear");
}

hnt_status hnt_kernel_strain(hnt_ctx* ctx, const double* rad, const double* G, const double* Gf,
                             const double* SR, int32_t nshells, int32_t mp, int32_t mlow,
                             int32_t mupp, double wRnor, double aRnor, double hlatt_1,
                             double hlatt_2, double* out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(rad && G && Gf && SR && out, "hnt_kernel_strain: NULL array");
    HNT_REQUIRE(nshells >= 1 && mp >= 1 && mp <= nshells && mlow >= 1 && mupp <= nshells,
                "hnt_kernel_strain: shell indices out of 1..%d", nshells);
    HNT_TRY(ensure_akk(ctx));
    // device copies, 1-based (index 0 unused)
    const size_t ns1 = (size_t)nshells + 1;
    const size_t bytes = sizeof(double) * (ns1 * (1 + 3 + 3 + 9) + 12);
    HNT_TRY(reserve(ctx, ctx->small, bytes));
    double* d = (double*)ctx->small.p;
    double* dr = d;
    double* dG = dr + ns1;
    double* dGf = dG + 3 * ns1;
    double* dSR = dGf + 3 * ns1;
    double* dout = dSR + 9 * ns1;
    hipStream_t st = ctx->stream;
    HNT_HIP(hipMemcpyAsync(dr + 1, rad, sizeof(double) * nshells, hipMemcpyHostToDevice, st));
    HNT_HIP(hipMemcpyAsync(dG + 3, G, sizeof(double) * 3 * nshells, hipMemcpyHostToDevice, st));
    HNT_HIP(hipMemcpyAsync(dGf + 3, Gf, sizeof(double) * 3 * nshells, hipMemcpyHostToDevice, st));
    HNT_HIP(hipMemcpyAsync(dSR + 9, SR, sizeof(double) * 9 * nshells, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(kstrain_kernel, dim3(1), dim3(64), 0, st, dr, dG, dGf, dSR,
                       (const double*)ctx->akk.p, mp, mlow, mupp, wRnor, aRnor, hlatt_1,
                       hlatt_2, dout);
    HNT_HIP(hipMemcpyAsync(out, dout, sizeof(double) * 12, hipMemcpyDeviceToHost, st));
    return finish_call(ctx, "hnt_kernel_strain");
}

}  // extern "C"

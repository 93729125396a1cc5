// records.hip -- halo records of a tile on the device (hnt_pack_records).
//
// Replaces the record loop of the tile driver after the shell analysis
// (src/MultiResolution.jl:1159-1219), with the record layouts of src/IO/Catalog.jl:4-46, the
// growth factor Dlinear_ab (src/Cosmology/Cosmology.jl:287-307, with _hunt and _lagrint_2col
// :70-130), omega_m_a (:50) and the lightcone redshift chi_to_z / peak_redshift (:357-385).
// The host loop pushes a record for every peak with RTHL > 0 (and z <= z_max in lightcone
// mode) in peak order; here rec_flag_kernel marks those peaks, a scan gives each its slot
// and rec_write_kernel writes the record there.  Double-precision steps keep the reference's
// operation order (no FMA contraction), so records equal the host loop's bit for bit except
// where device and host pow() differ in the last place (the 2LPT factor).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "hnt_internal.h"

namespace {

using namespace hnt;

struct RecArgs {
    const int32_t* pi;
    const int32_t* pj;
    const int32_t* pk;
    const double* Rf;
    hnt_results res;
    const float* FcRf;
    const float* d2Rf;
    const double* atab;
    const double* Dtab;
    int ntab;
    const double* chiz;
    const double* chid;
    int nchi;
    double chi_zmax;
    hnt_pack_params pp;
    int64_t npeaks;
    int nf;  // floats per record: 11 or 33
};

// _hunt (Cosmology.jl:116-130): bisection for x in the ascending or descending table xx[0..n),
// returning the 1-based jlo with xx[jlo] <= x < xx[jlo + 1] (0: below, n: above).
__device__ __host__ int hunt(const double* xx, int n, double x) {
    const bool ascnd = xx[n - 1] > xx[0];
    int jlo = 0, jhi = n + 1;
    while (jhi - jlo > 1) {
        const int jm = (jhi + jlo) / 2;
        if ((x > xx[jm - 1]) == ascnd) jlo = jm;
        else jhi = jm;
    }
    return jlo;
}

// Dlinear_ab(ai)[1] = D(ai) (Cosmology.jl:287-307): Lagrange interpolation of D(a)/a over 4
// table points (2 at the table ends), times ai; D/a held constant outside the table.
__device__ double growth_D(const double* atab, const double* Dtab, int ntab, double ai) {
    if (ai > atab[0] && ai < atab[ntab - 1]) {
        const int kk = hunt(atab, ntab, ai);
        // 1-based first point kkl; the caller has checked that kk == 2 (which the reference
        // would index out of its table) cannot occur
        int nlint, kkl;
        if (kk < 2 || kk >= ntab) {
            nlint = 2;
            kkl = kk - 1 > 1 ? kk - 1 : 1;
        } else {
            nlint = 4;
            kkl = kk - 2;
        }
        const double* xa = atab + (kkl - 1);
        const double* v = Dtab + (kkl - 1);
        const double x = ai;
        double D_a;
        if (nlint == 4) {
            const double a1 = (x - xa[1]) / (xa[0] - xa[1]) * (x - xa[2]) / (xa[0] - xa[2]) *
                              (x - xa[3]) / (xa[0] - xa[3]);
            const double a2 = (x - xa[0]) / (xa[1] - xa[0]) * (x - xa[2]) / (xa[1] - xa[2]) *
                              (x - xa[3]) / (xa[1] - xa[3]);
            const double a3 = (x - xa[0]) / (xa[2] - xa[0]) * (x - xa[1]) / (xa[2] - xa[1]) *
                              (x - xa[3]) / (xa[2] - xa[3]);
            const double a4 = (x - xa[0]) / (xa[3] - xa[0]) * (x - xa[1]) / (xa[3] - xa[1]) *
                              (x - xa[2]) / (xa[3] - xa[2]);
            D_a = a1 * v[0] + a2 * v[1] + a3 * v[2] + a4 * v[3];
        } else {
            const double a1 = (x - xa[1]) / (xa[0] - xa[1]);
            const double a2 = 1.0 - a1;
            D_a = a1 * v[0] + a2 * v[1];
        }
        return D_a * ai;
    }
    if (ai >= atab[ntab - 1]) return Dtab[ntab - 1] * ai;
    return Dtab[0] * ai;
}

// Dlinear_ab(ai)[2] = f = dlnD/dlna (Cosmology.jl:287-307): the same 4-point (2 at the table
// ends) Lagrange weights as growth_D over the growth-rate column; end values outside the table.
__device__ double growth_f(const double* atab, const double* ftab, int ntab, double ai) {
    if (ai > atab[0] && ai < atab[ntab - 1]) {
        const int kk = hunt(atab, ntab, ai);
        int nlint, kkl;
        if (kk < 2 || kk >= ntab) {
            nlint = 2;
            kkl = kk - 1 > 1 ? kk - 1 : 1;
        } else {
            nlint = 4;
            kkl = kk - 2;
        }
        const double* xa = atab + (kkl - 1);
        const double* v = ftab + (kkl - 1);
        const double x = ai;
        if (nlint == 4) {
            const double a1 = (x - xa[1]) / (xa[0] - xa[1]) * (x - xa[2]) / (xa[0] - xa[2]) *
                              (x - xa[3]) / (xa[0] - xa[3]);
            const double a2 = (x - xa[0]) / (xa[1] - xa[0]) * (x - xa[2]) / (xa[1] - xa[2]) *
                              (x - xa[3]) / (xa[1] - xa[3]);
            const double a3 = (x - xa[0]) / (xa[2] - xa[0]) * (x - xa[1]) / (xa[2] - xa[1]) *
                              (x - xa[3]) / (xa[2] - xa[3]);
            const double a4 = (x - xa[0]) / (xa[3] - xa[0]) * (x - xa[1]) / (xa[3] - xa[1]) *
                              (x - xa[2]) / (xa[3] - xa[2]);
            return a1 * v[0] + a2 * v[1] + a3 * v[2] + a4 * v[3];
        }
        const double a1 = (x - xa[1]) / (xa[0] - xa[1]);
        const double a2 = 1.0 - a1;
        return a1 * v[0] + a2 * v[1];
    }
    if (ai >= atab[ntab - 1]) return ftab[ntab - 1];
    return ftab[0];
}

// chi_to_z (Cosmology.jl:357-372): bisection in the comoving-distance table, then linear
// interpolation in z.
__device__ double chi_to_z(const double* zs, const double* chis, int n, double zmax,
                           double chi) {
    if (chi <= chis[0]) return 0.0;
    if (chi >= chis[n - 1]) return zmax;
    int lo = 1, hi = n;  // 1-based
    while (hi - lo > 1) {
        const int mid = (int)(((unsigned)lo + (unsigned)hi) >> 1);
        if (chis[mid - 1] <= chi) lo = mid;
        else hi = mid;
    }
    const double frac = (chi - chis[lo - 1]) / (chis[hi - 1] - chis[lo - 1]);
    return zs[lo - 1] + frac * (zs[hi - 1] - zs[lo - 1]);
}

struct PeakPos {
    double x, y, z;
};

__device__ PeakPos peak_position(const RecArgs& a, int64_t p) {
    // cell centres as in peak_find_tile_gpu (ext/CUDAExt.jl:3036-3042)
    const double cen1 = 0.5 * (a.pp.n + 1);
    return {a.pp.xbx + a.pp.alatt * ((double)a.pi[p] - cen1),
            a.pp.ybx + a.pp.alatt * ((double)a.pj[p] - cen1),
            a.pp.zbx + a.pp.alatt * ((double)a.pk[p] - cen1)};
}

// 1 + z of peak p, or 0 if its record is dropped (no collapse, or beyond z_max).
__device__ double record_ZZon(const RecArgs& a, int64_t p, const PeakPos& q) {
    // `r.RTHL[idx] <= 0 && continue` (MultiResolution.jl:1169): a NaN RTHL keeps its record
    if ((float)a.res.RTHL[p] <= 0.0f) return 0.0;
    if (!a.pp.ievol) return a.pp.ZZon;
    // peak_redshift (Cosmology.jl:376-385)
    const double dx = q.x - a.pp.obs[0], dy = q.y - a.pp.obs[1], dz = q.z - a.pp.obs[2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double z = chi_to_z(a.chiz, a.chid, a.nchi, a.chi_zmax, r);
    if (z > a.pp.z_max) return 0.0;
    return 1.0 + z;
}

constexpr int REC_BLOCK = 256;

__global__ __launch_bounds__(REC_BLOCK) void rec_flag_kernel(RecArgs a,
                                                             int32_t* __restrict__ keep) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npeaks) return;
    keep[p] = record_ZZon(a, p, peak_position(a, p)) > 0.0 ? 1 : 0;
}

// One record per peak, built in LDS at the peak's rank among the workgroup's kept peaks, then
// the workgroup's records -- consecutive in the output -- are written as one contiguous run
// (the per-peak rows would otherwise be 44- or 132-byte strided stores).

__global__ __launch_bounds__(REC_BLOCK) void rec_write_kernel(RecArgs a,
                                                              const int32_t* __restrict__ keep,
                                                              const int64_t* __restrict__ slot,
                                                              float* __restrict__ out,
                                                              int64_t capacity) {
    __shared__ float rows[REC_BLOCK * 33];
    const int64_t p0 = (int64_t)blockIdx.x * REC_BLOCK;
    const int64_t p = p0 + threadIdx.x;
    const int64_t s0 = slot[p0];  // slot of the workgroup's first kept peak
    const int64_t pe = min(p0 + REC_BLOCK, a.npeaks) - 1;
    const int64_t nkeep = slot[pe] + keep[pe] - s0;
    if (p < a.npeaks && keep[p]) {
        const PeakPos q = peak_position(a, p);
        const double ZZon = record_ZZon(a, p, q);
        const double apk = 1.0 / ZZon;
        const double D = growth_D(a.atab, a.Dtab, a.ntab, apk);
        const float Df = (float)D;
        const hnt_results& R = a.res;
        float* o = rows + (slot[p] - s0) * a.nf;
        o[0] = (float)q.x;
        o[1] = (float)q.y;
        o[2] = (float)q.z;
        for (int L = 0; L < 3; ++L) o[3 + L] = (float)R.Sbar[p * 3 + L] * Df;
        o[6] = (float)((double)(float)R.RTHL[p] * a.pp.alatt);
        float coef = 0.0f;
        if (a.pp.ilpt >= 2) {
            // omega_m_a (Cosmology.jl:50); coef = -3/7 Om_a^(-1/143) D^2 (MultiResolution.jl:1188)
            const double a3 = apk * apk * apk;
            const double Om_a = (a.pp.Om / a3) / (a.pp.Om / a3 + a.pp.OL);
            coef = (float)(-3.0 / 7.0 * pow(Om_a, -1.0 / 143) * (D * D));
        }
        for (int L = 0; L < 3; ++L)
            o[7 + L] = a.pp.ilpt >= 2 ? (float)R.Sbar2[p * 3 + L] * coef : 0.0f;
        o[10] = (float)R.Fbarx[p];
        if (a.nf == 33) {
            const double* st = R.strain + p * 9;  // symmetric 3x3
            o[11] = (float)R.e_v[p];
            o[12] = (float)R.p_v[p];
            o[13] = (float)st[0];  // 11
            o[14] = (float)st[4];  // 22
            o[15] = (float)st[8];  // 33
            o[16] = (float)st[5];  // 23
            o[17] = (float)st[2];  // 13
            o[18] = (float)st[1];  // 12
            o[19] = (float)R.d2F[p];
            o[20] = (float)R.zvir_half[p];
            for (int L = 0; L < 3; ++L) o[21 + L] = (float)R.gradpk[p * 3 + L];
            for (int L = 0; L < 3; ++L) o[24 + L] = (float)R.gradpkf[p * 3 + L];
            o[27] = (float)a.Rf[p];
            o[28] = a.FcRf[p];
            o[29] = a.d2Rf ? a.d2Rf[p] : 0.0f;
            for (int L = 0; L < 3; ++L) o[30 + L] = (float)R.gradpkrf[p * 3 + L];
        }
    }
    __syncthreads();
    // records s0 .. s0 + nkeep - 1, those below capacity
    const int64_t nwrite = max((int64_t)0, min(s0 + nkeep, capacity) - s0) * a.nf;
    float* dst = out + s0 * a.nf;
    for (int64_t e = threadIdx.x; e < nwrite; e += REC_BLOCK) dst[e] = rows[e];
}


struct FinArgs {
    const float* rec;
    float* out;
    int64_t N;
    int nf;
    const double* atab;
    const double* ftab;
    int ntab;
    const double* chiz;
    const double* chid;
    int nchi;
    double chi_zmax;
    double Om, OL, ox, oy, oz, a_snap;
    int ievol;
};

// finalize_eulerian (Merger.jl:123-148) per halo, in the reference's operation order
// One block finalizes FIN_T records: they are loaded into LDS as one contiguous run
// (coalesced), each thread rewrites its record there, and the run is stored back.  A block
// reads all of its records before it writes any, and blocks own disjoint runs, so out may
// equal rec.  The record stride (11 or 33 floats) is odd: no LDS bank conflicts.
constexpr int FIN_T = 256;

__global__ __launch_bounds__(FIN_T) void fin_kernel(FinArgs a) {
    __shared__ float tile[FIN_T * 33];
    const int64_t base = (int64_t)blockIdx.x * FIN_T;
    const int cnt = (int)(a.N - base < FIN_T ? a.N - base : FIN_T);
    const int tot = cnt * a.nf;
    const float* src = a.rec + base * a.nf;
    for (int q = threadIdx.x; q < tot; q += FIN_T) tile[q] = src[q];
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
        float* h = tile + threadIdx.x * a.nf;
        const double d1x = h[3], d1y = h[4], d1z = h[5];
        const double d2x = h[7], d2y = h[8], d2z = h[9];
        const double ex = (double)h[0] + d1x + d2x, ey = (double)h[1] + d1y + d2y,
                     ez = (double)h[2] + d1z + d2z;
        double av = a.a_snap;
        if (a.ievol == 1) {
            const double rx = ex - a.ox, ry = ey - a.oy, rz = ez - a.oz;
            const double r = sqrt(rx * rx + ry * ry + rz * rz);
            av = r > 0.0 ? 1.0 / (1.0 + chi_to_z(a.chiz, a.chid, a.nchi, a.chi_zmax, r)) : 1.0;
        }
        const double E = sqrt(a.Om * pow_m3(av) + a.OL);  // Julia's a^-3
        const double f = growth_f(a.atab, a.ftab, a.ntab, av);
        const double vfac = av * 100.0 * E * f;
        h[0] = (float)ex;
        h[1] = (float)ey;
        h[2] = (float)ez;
        h[3] = (float)(vfac * (d1x + 2.0 * d2x));
        h[4] = (float)(vfac * (d1y + 2.0 * d2y));
        h[5] = (float)(vfac * (d1z + 2.0 * d2z));
        h[7] = 0.0f;
        h[8] = 0.0f;
        h[9] = 0.0f;
    }
    __syncthreads();
    float* dst = a.out + base * a.nf;
    for (int q = threadIdx.x; q < tot; q += FIN_T) dst[q] = tile[q];
}

}  // namespace

extern "C" hnt_status hnt_pack_records(hnt_ctx* ctx, const hnt_batch* batch,
                                       const hnt_results* results, const float* FcRf,
                                       const float* d2Rf, const hnt_cosmo_tables* tabs,
                                       const hnt_pack_params* pp, float* records,
                                       int64_t capacity, int64_t* count) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(batch && results && tabs && pp && count, "hnt_pack_records: NULL argument");
    *count = 0;
    const int64_t np = batch->npeaks;
    HNT_REQUIRE(np >= 0 && capacity >= 0, "hnt_pack_records: negative size");
    if (np == 0) return HNT_OK;
    HNT_REQUIRE(batch->pi && batch->pj && batch->pk && batch->Rf && FcRf,
                "hnt_pack_records: NULL peak array");
    const hnt_results& R = *results;
    HNT_REQUIRE(R.RTHL && R.Fbarx && R.Sbar && R.Sbar2, "hnt_pack_records: NULL result array");
    const int nf = pp->ioutshear >= 1 ? 33 : 11;
    if (nf == 33)
        HNT_REQUIRE(R.e_v && R.p_v && R.strain && R.d2F && R.zvir_half && R.gradpk &&
                        R.gradpkf && R.gradpkrf,
                    "hnt_pack_records: NULL result array (extended records)");
    HNT_REQUIRE(capacity == 0 || records, "hnt_pack_records: NULL records");
    HNT_REQUIRE(pp->n >= 1 && pp->alatt > 0.0, "hnt_pack_records: bad tile geometry");
    const int ntab = tabs->ntab;
    HNT_REQUIRE(ntab >= 4 && tabs->atab && tabs->D_atab, "hnt_pack_records: growth table of "
                "%d points (>= 4 needed)", ntab);
    for (int q = 1; q < ntab; ++q)
        HNT_REQUIRE(tabs->atab[q] > tabs->atab[q - 1], "hnt_pack_records: growth table not "
                    "ascending at %d", q);
    // The reference's Dlinear_ab indexes before its table for a in (atab[2], atab[3]) (1-based);
    // the scale factors used here must stay above that.
    const double a_floor = tabs->atab[2];
    if (pp->ievol) {
        HNT_REQUIRE(tabs->nchi >= 2 && tabs->chi2z_z && tabs->chi2z_chi,
                    "hnt_pack_records: lightcone mode needs the chi -> z table");
        HNT_REQUIRE(pp->z_max > -1.0 && 1.0 / (1.0 + pp->z_max) > a_floor,
                    "hnt_pack_records: z_max %g beyond the growth table", pp->z_max);
    } else {
        HNT_REQUIRE(pp->ZZon > 0.0 && 1.0 / pp->ZZon > a_floor,
                    "hnt_pack_records: ZZon %g beyond the growth table", pp->ZZon);
    }
    hipStream_t st = ctx->stream;

    RecArgs A;
    memset(&A, 0, sizeof(A));
    A.pi = batch->pi;
    A.pj = batch->pj;
    A.pk = batch->pk;
    A.Rf = batch->Rf;
    A.res = R;
    A.FcRf = FcRf;
    A.d2Rf = d2Rf;
    A.ntab = ntab;
    A.nchi = pp->ievol ? tabs->nchi : 0;
    A.chi_zmax = tabs->chi2z_zmax;
    A.pp = *pp;
    A.npeaks = np;
    A.nf = nf;
    const size_t tab_doubles = (size_t)ntab * 2 + (size_t)A.nchi * 2;
    HNT_TRY(reserve(ctx, ctx->rec_tabs, tab_doubles * 8));
    double* dt = (double*)ctx->rec_tabs.p;
    HNT_HIP(hipMemcpyAsync(dt, tabs->atab, (size_t)ntab * 8, hipMemcpyHostToDevice, st));
    HNT_HIP(hipMemcpyAsync(dt + ntab, tabs->D_atab, (size_t)ntab * 8, hipMemcpyHostToDevice, st));
    A.atab = dt;
    A.Dtab = dt + ntab;
    if (A.nchi) {
        HNT_HIP(hipMemcpyAsync(dt + 2 * ntab, tabs->chi2z_z, (size_t)A.nchi * 8,
                               hipMemcpyHostToDevice, st));
        HNT_HIP(hipMemcpyAsync(dt + 2 * ntab + A.nchi, tabs->chi2z_chi, (size_t)A.nchi * 8,
                               hipMemcpyHostToDevice, st));
        A.chiz = dt + 2 * ntab;
        A.chid = dt + 2 * ntab + A.nchi;
    }
    // flags, their exclusive scan (slot of each record), the records
    HNT_TRY(reserve(ctx, ctx->rec_keep, (size_t)np * 4));
    HNT_TRY(reserve(ctx, ctx->rec_slot, (size_t)(np + 1) * 8));
    size_t tmp = 0;
    HNT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (const int32_t*)ctx->rec_keep.p,
                                             (int64_t*)ctx->rec_slot.p, (int64_t)np, st));
    HNT_TRY(reserve(ctx, ctx->cub_tmp, tmp > 0 ? tmp : 16));
    const unsigned blocks = (unsigned)((np + REC_BLOCK - 1) / REC_BLOCK);
    hipLaunchKernelGGL(rec_flag_kernel, dim3(blocks), dim3(REC_BLOCK), 0, st, A,
                       (int32_t*)ctx->rec_keep.p);
    HNT_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp.p, tmp, (const int32_t*)ctx->rec_keep.p,
                                             (int64_t*)ctx->rec_slot.p, (int64_t)np, st));
    hipLaunchKernelGGL(rec_write_kernel, dim3(blocks), dim3(REC_BLOCK), 0, st, A,
                       (const int32_t*)ctx->rec_keep.p, (const int64_t*)ctx->rec_slot.p, records,
                       capacity);
    int64_t last_slot = 0;
    int32_t last_keep = 0;
    HNT_HIP(hipMemcpyAsync(&last_slot, (int64_t*)ctx->rec_slot.p + (np - 1), 8,
                           hipMemcpyDeviceToHost, st));
    HNT_HIP(hipMemcpyAsync(&last_keep, (int32_t*)ctx->rec_keep.p + (np - 1), 4,
                           hipMemcpyDeviceToHost, st));
    HNT_TRY(finish_call(ctx, "hnt_pack_records"));
    const int64_t total = last_slot + last_keep;
    *count = total;
    if (total > capacity) {
        set_error("hnt_pack_records: %lld records, capacity %lld", (long long)total,
                  (long long)capacity);
        return HNT_ERR_CAPACITY;
    }
    return HNT_OK;
}

extern "C" hnt_status hnt_finalize_eulerian(hnt_ctx* ctx, const float* records, int64_t nhalo,
                                            int32_t nfields, const hnt_cosmo_tables* tabs,
                                            double Om, double OL, const double* obs,
                                            int32_t ievol, double z_out, float* out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(tabs && obs && (nhalo == 0 || (records && out)),
                "hnt_finalize_eulerian: NULL argument");
    HNT_REQUIRE(nfields == 11 || nfields == 33, "hnt_finalize_eulerian: records of %d floats "
                "(11 or 33)", nfields);
    // a launch holds fewer than 2^32 work items
    HNT_REQUIRE(nhalo >= 0 && nhalo <= ((int64_t)1 << 32) - FIN_T,
                "hnt_finalize_eulerian: %lld halos", (long long)nhalo);
    HNT_REQUIRE(ievol == 0 || ievol == 1, "hnt_finalize_eulerian: ievol %d", ievol);
    HNT_REQUIRE(tabs->ntab >= 4 && tabs->atab && tabs->dlnD_atab,
                "hnt_finalize_eulerian: the growth tables (atab, dlnD_atab, >= 4 points) are "
                "needed");
    // the smallest scale factor reached: 1/(1 + z_max) of the chi -> z table, or of z_out;
    // above the table's third node (the reference's lookup is defined there)
    if (ievol == 1) {
        HNT_REQUIRE(tabs->nchi >= 2 && tabs->chi2z_z && tabs->chi2z_chi,
                    "hnt_finalize_eulerian: ievol = 1 needs the chi -> z table");
        HNT_REQUIRE(1.0 / (1.0 + tabs->chi2z_zmax) > tabs->atab[2],
                    "hnt_finalize_eulerian: z up to %g is beyond the growth table",
                    tabs->chi2z_zmax);
    } else {
        HNT_REQUIRE(z_out > -1.0 && 1.0 / (1.0 + z_out) > tabs->atab[2],
                    "hnt_finalize_eulerian: z_out %g is beyond the growth table", z_out);
    }
    if (nhalo == 0) return HNT_OK;
    hipStream_t st = ctx->stream;
    const int nt = tabs->ntab, nc = ievol == 1 ? tabs->nchi : 0;
    std::vector<double> host((size_t)2 * nt + 2 * (size_t)nc);
    std::copy(tabs->atab, tabs->atab + nt, host.begin());
    std::copy(tabs->dlnD_atab, tabs->dlnD_atab + nt, host.begin() + nt);
    if (nc) {
        std::copy(tabs->chi2z_z, tabs->chi2z_z + nc, host.begin() + 2 * nt);
        std::copy(tabs->chi2z_chi, tabs->chi2z_chi + nc, host.begin() + 2 * nt + nc);
    }
    HNT_TRY(reserve(ctx, ctx->rec_tabs, sizeof(double) * host.size()));
    HNT_HIP(hipMemcpyAsync(ctx->rec_tabs.p, host.data(), sizeof(double) * host.size(),
                           hipMemcpyHostToDevice, st));
    const double* d = (const double*)ctx->rec_tabs.p;
    FinArgs a;
    a.rec = records;
    a.out = out;
    a.N = nhalo;
    a.nf = nfields;
    a.atab = d;
    a.ftab = d + nt;
    a.ntab = nt;
    a.chiz = nc ? d + 2 * nt : nullptr;
    a.chid = nc ? d + 2 * nt + nc : nullptr;
    a.nchi = nc;
    a.chi_zmax = tabs->chi2z_zmax;
    a.Om = Om;
    a.OL = OL;
    a.ox = obs[0];
    a.oy = obs[1];
    a.oz = obs[2];
    a.a_snap = 1.0 / (1.0 + z_out);
    a.ievol = ievol;
    hipLaunchKernelGGL(fin_kernel, dim3((unsigned)((nhalo + FIN_T - 1) / FIN_T)), dim3(FIN_T),
                       0, st, a);
    // the host tables are read by the copy above before this returns
    return finish_call(ctx, "hnt_finalize_eulerian");
}

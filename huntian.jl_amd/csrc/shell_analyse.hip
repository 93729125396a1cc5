// shell_analyse.hip -- radial-shell homogeneous-ellipsoid collapse of every peak candidate.
//
// Replaces _shell_gather_full_kernel! + _post_process_kernel! and their launchers
// analyse_peaks_gpu_cuda_multirf / analyse_peak_gpu_cuda (ext/CUDAExt.jl:752-924,
// 1213-1912, 1914-2344).  Semantics follow the FP64 CPU path: RadialShell.analyse_peak
// (src/HaloFinder/RadialShell.jl:281-865, variant A) and ShellAnalysisGPU.analyse_peak_gpu
// (src/ShellAnalysisGPU.jl:81-535, variants B and C), with the same operation order, so
// the results equal the FP64 CPU restatement bit for bit except where log10 / pow differ
// by an ulp between the device and host math libraries.
//
// Design (MI355X):
//  * repack_kernel turns the caller's eight column-major fields into one 32-byte record
//    per cell {delta, delta^T, psi1 xyz, psi2 xyz}: a cell visit is two 16-byte loads from
//    one cache line instead of eight scattered 4-byte loads.  delta^T[i,j,k] = delta[k,i,j]
//    (the transposed read of the reference, RadialShell.jl:441) goes through an LDS-tiled
//    transpose so both the reads and the writes stay coalesced.
//  * Peaks are sorted by (ir2min, Morton code of the cell): a wavefront holds 64 peaks of
//    one filter scale that sit close together, so their shell balls overlap in L2.
//  * One lane per peak.  All 64 lanes of a wavefront walk the same static shell table in
//    lock step (the offset table is read with scalar loads), each gathering its own peak's
//    cells: no cross-lane reductions, and each lane sums its shell in the reference's cell
//    order, in FP64 with the Float32 products the reference forms.  Per-shell profiles go
//    to a per-lane slot of scratch memory (L2-resident while the lane post-processes).
//  * A lane stops gathering once its profile can no longer change any result (first
//    shell past ir2min with Fbar < fcrit, plus 3 cells, and past the filter-scale
//    gradient window); the wave stops when its last lane stops.
//  * Persistent grid: one wavefront per workgroup, as many as are resident; eight work
//    queues, one per XCD (workgroup i runs on XCD i % 8), each owning a contiguous eighth
//    of the sorted peak list, so neighbouring peaks share an XCD's L2.  A wave that
//    drains its queue moves on to the next queue; every wave exits once all eight are
//    empty.
#include <hipcub/hipcub.hpp>

#include <mutex>

#include "hnt_features.h"
#include "hnt_internal.h"

namespace {

using namespace hnt;

// Doubles per shell in a profile slot: 22 used, padded to two whole 128-B lines (written
// in full, so that no line of the slot is left half written for the L2 to evict: 24 doubles,
// 1.5 lines per shell, took the 1024^3 shell kernel 6 ms longer).
constexpr int PW = 32;
// First line, what the radial logic reads at every shell:
// [0..2] G/rad  [3..5] Gf/rad  [6..11] SR+SR^T over rad: 11 12 13 22 23 33
// [12] Fbar  [13] running lapd sum  [14] running lapd count  [15] zero
// Second line, read only at the shell below the walk's start m0 (Sbar, Sbar2), which lies
// below the crossing shell:
// [16] cumulative n  [17..19] cumulative S  [20..22] cumulative S2  [23..31] zero
constexpr int K_F = 12, K_LAP = 13, K_LAPN = 14, K_N = 16, K_S = 17, K_S2 = 20;
constexpr int K_LINE2 = 16;

struct SAArgs {
    const TileDev* tiles;
    const int32_t* slot;   // [ngroups * 64]: peak index within its tile, -1 = empty lane
    const int32_t* gtile;  // [ngroups]: tile of each group
    int64_t ngroups;
    const int4* cells;
    const int32_t* shell_start;
    const int32_t* shell_count;
    const int32_t* shell_r2;
    const double* rad;  // 1-based
    const double* kw;   // kernel_strain weight table
    const int32_t* kidx;
    int NS, rmax;
    CTabView ct;
    const double* akk;
    double alatt, wRnor, aRnor;
    int nbuff, variant, early;
    double* prof;
    int64_t slot_doubles;
    const int32_t* stop_hint;  // exact stop shell per global peak id (crossing pass) or null
    const float* pred;         // bracket network outputs per global peak id (learned order)
    double rmax2rs;            // > 0: each peak's shells limited to r^2 <= (rmax2rs Rf/alatt)^2
    const int8_t* skip;        // per global peak id: 1 = no collapse without analysis, or null
    int32_t* paint;            // per global peak id: r^2 of a halo's mask ball (-1: none), set
                               // for paint_mask_kernel; null: post_process marks the mask
    unsigned long long* qctr;
    int64_t qbeg[8], qend[8];
};

// ------------------------------------------------------------------ repack

__global__ __launch_bounds__(256) void repack_kernel(const float* __restrict__ d,
                                                     const float* __restrict__ ex,
                                                     const float* __restrict__ ey,
                                                     const float* __restrict__ ez,
                                                     const float* __restrict__ e2x,
                                                     const float* __restrict__ e2y,
                                                     const float* __restrict__ e2z, int n,
                                                     float4* __restrict__ recs) {
    __shared__ float tile[64][65];
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * 64, k0 = blockIdx.y * 64, j = blockIdx.z;
    const int64_t nn = n, n2 = (int64_t)n * n;
    // tile[ii][kk] = delta[k0+kk, i0+ii, j]  (delta^T at (i0+ii, j, k0+kk))
    for (int r = 0; r < 16; ++r) {
        int ii = (t >> 6) + 4 * r, kk = t & 63;
        if (i0 + ii < n && k0 + kk < n) tile[ii][kk] = d[(k0 + kk) + (i0 + ii) * nn + j * n2];
    }
    __syncthreads();
    for (int r = 0; r < 16; ++r) {
        int ii = t & 63, kk = (t >> 6) + 4 * r;
        int i = i0 + ii, k = k0 + kk;
        if (i < n && k < n) {
            int64_t c = i + j * nn + k * n2;
            float4 a = make_float4(d[c], tile[ii][kk], ex[c], ey[c]);
            float4 b = make_float4(ez[c], e2x ? e2x[c] : 0.0f, e2y ? e2y[c] : 0.0f,
                                   e2z ? e2z[c] : 0.0f);
            recs[2 * c] = a;
            recs[2 * c + 1] = b;
        }
    }
}

// ------------------------------------------------------------------ sort keys

__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}

// Sort key of a tile's peaks, value = global peak id base + q.  Bits, high to low: tile (14),
// Morton code of the peak's brick of 2^L cells a side (39 - 3L), primary (11), Morton code
// of the cell within its brick (3L).  Consecutive peaks therefore share a brick (and the
// waves resident on one XCD a neighbourhood of bricks, whose records stay in its L2), and
// within a brick they are ordered by the primary: ir2min for the first order, the stop
// shell for the second, so that a wavefront's lanes have similar gather depths.
__global__ void tile_key_kernel(const int32_t* __restrict__ pi, const int32_t* __restrict__ pj,
                                const int32_t* __restrict__ pk, const int32_t* __restrict__ prim,
                                int64_t np, uint64_t tile, int64_t base, int L,
                                uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const uint64_t mort = spread3((uint32_t)(pi[q] - 1) & 0x1fff) |
                          (spread3((uint32_t)(pj[q] - 1) & 0x1fff) << 1) |
                          (spread3((uint32_t)(pk[q] - 1) & 0x1fff) << 2);
    const uint64_t g = (uint64_t)min(max(prim[q], 0), 2047);
    const int lb = 3 * L;
    const uint64_t local = mort & ((1ull << lb) - 1ull);
    const uint64_t brick = mort >> lb;
    keys[base + q] = (tile << 50) | (brick << (lb + 11)) | (g << lb) | local;
    vals[base + q] = (int32_t)(base + q);
}

// Sorted position q -> lane slot: tile t's peaks occupy sorted positions [base_t, base_t +
// n_t) and slots [pad_t, pad_t + n_t); slots up to the next multiple of 64 stay -1.
__global__ void slot_kernel(const uint64_t* __restrict__ keys, const int32_t* __restrict__ vals,
                            int64_t ntot, const int64_t* __restrict__ tile_base,
                            const int64_t* __restrict__ tile_pad, int32_t* __restrict__ slot) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ntot) return;
    const int t = (int)(keys[q] >> 50);
    slot[tile_pad[t] + (q - tile_base[t])] = (int32_t)(vals[q] - tile_base[t]);
}

// Tile of every group of 64 slots.
__global__ void group_tile_kernel(const int64_t* __restrict__ tile_pad, int ntiles,
                                  int64_t ngroups, int32_t* __restrict__ gtile) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    int lo = 0, hi = ntiles - 1;  // last tile with tile_pad <= 64 g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_pad[mid] <= g * 64) lo = mid;
        else hi = mid - 1;
    }
    gtile[g] = lo;
}

__global__ void fill_i32(int32_t* p, int64_t n, int32_t v) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

__global__ void fill_i8(int8_t* p, int64_t n, int8_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Flags peaks outside the grid or with ZZon <= -1 (bit 0 / bit 1 of *bad).
__global__ void validate_kernel(const int32_t* __restrict__ pi, const int32_t* __restrict__ pj,
                                const int32_t* __restrict__ pk, const double* __restrict__ ZZon,
                                int64_t np, int n, unsigned int* bad) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    unsigned int b = 0;
    if (pi[q] < 1 || pi[q] > n || pj[q] < 1 || pj[q] > n || pk[q] < 1 || pk[q] > n) b |= 1u;
    if (ZZon && !(ZZon[q] > -1.0)) b |= 2u;
    if (b) atomicOr(bad, b);
}

__global__ void fill_u64(unsigned long long* p, int n, unsigned long long v) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

// ------------------------------------------------------------------ device helpers

struct KS {
    double E11, E12, E13, E22, E23, E33, g0, g1, g2, f0, f1, f2;
};

// kernel_strain (RadialShell.jl:225-274) over the lane's profile slot P.
// The shells with u0^2 < 4 form the contiguous range [kidx[3mp], kidx[3mp+1]]; their
// weights hRinteg(u0, u12) / (u12 - u0^2)^2 come precomputed from the table kw.
__device__ KS kstrain(const SAArgs& a, const double* __restrict__ P, int mp, int mlow,
                      int mupp) {
    const double* __restrict__ rad = a.rad;
    const double* __restrict__ akk = a.akk;
    const double wRnor = a.wRnor, aRnor = a.aRnor;
    KS r = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (mp > 1) {
        const double rmp = rad[mp];
        const int lo0 = a.kidx[3 * mp], hi0 = a.kidx[3 * mp + 1];
        const double* __restrict__ W = a.kw + (a.kidx[3 * mp + 2] - lo0);
        const int lo = mlow > lo0 ? mlow : lo0, hi = mupp < hi0 ? mupp : hi0;
        for (int m1 = lo; m1 <= hi; ++m1) {
            const double wt = W[m1];
            const double* k = P + (int64_t)m1 * PW;
            const double hw = 0.5 * wt;
            r.g0 += wt * k[0];
            r.g1 += wt * k[1];
            r.g2 += wt * k[2];
            r.f0 += wt * k[3];
            r.f1 += wt * k[4];
            r.f2 += wt * k[5];
            r.E11 -= hw * k[6];
            r.E12 -= hw * k[7];
            r.E13 -= hw * k[8];
            r.E22 -= hw * k[9];
            r.E23 -= hw * k[10];
            r.E33 -= hw * k[11];
        }
        const double sc = wRnor / rmp;
        r.E11 *= sc; r.E12 *= sc; r.E13 *= sc; r.E22 *= sc; r.E23 *= sc; r.E33 *= sc;
        r.g0 *= sc; r.g1 *= sc; r.g2 *= sc; r.f0 *= sc; r.f1 *= sc; r.f2 *= sc;
    } else {
        for (int m1 = mlow; m1 <= mupp; ++m1) {
            const double rm1 = rad[m1];
            const double con = 100.0 * rm1 * rm1;
            const double fl = floor(con);
            if (!(fl >= 0.0 && fl < 1100.0)) continue;
            const int icon = (int)fl;
            const double diff = con - (double)icon;
            const double aww = akk[icon] + diff * (akk[icon + 1] - akk[icon]);
            const double wt = 0.5 * aww / rm1;
            const double* k = P + (int64_t)m1 * PW;
            const double hw = 0.5 * wt;
            r.g0 += wt * k[0];
            r.g1 += wt * k[1];
            r.g2 += wt * k[2];
            r.f0 += wt * k[3];
            r.f1 += wt * k[4];
            r.f2 += wt * k[5];
            r.E11 -= hw * k[6];
            r.E12 -= hw * k[7];
            r.E13 -= hw * k[8];
            r.E22 -= hw * k[9];
            r.E23 -= hw * k[10];
            r.E33 -= hw * k[11];
        }
        r.E11 *= aRnor; r.E12 *= aRnor; r.E13 *= aRnor; r.E22 *= aRnor; r.E23 *= aRnor;
        r.E33 *= aRnor;
        r.g0 *= aRnor; r.g1 *= aRnor; r.g2 *= aRnor; r.f0 *= aRnor; r.f1 *= aRnor;
        r.f2 *= aRnor;
    }
    return r;
}

__device__ __forceinline__ void eig_ks(const KS& k, double lam[3]) {
    eig3_sym(k.E11, k.E12, k.E13, k.E22, k.E23, k.E33, lam);
}

// first index s in [lo, hi] with r2[s-1] > x (shells are 1-based; r2 is 0-based)
__device__ __forceinline__ int first_r2_gt(const int32_t* r2, int lo, int hi, int x) {
    int a = lo, b = hi + 1;
    while (a < b) {
        int mid = (a + b) >> 1;
        if (r2[mid - 1] > x) b = mid;
        else a = mid + 1;
    }
    return a;
}

// shell whose r^2 is closest to rf2, first one on ties, within s in [lo, hi]
__device__ __forceinline__ int closest_shell(const int32_t* r2, int lo, int hi, double rf2,
                                             double* dist) {
    int a = lo, b = hi + 1;  // first s with (double)r2 >= rf2
    while (a < b) {
        int mid = (a + b) >> 1;
        if ((double)r2[mid - 1] >= rf2) b = mid;
        else a = mid + 1;
    }
    int best = -1;
    double bd = 0.0;
    if (a - 1 >= lo) {
        best = a - 1;
        bd = fabs((double)r2[a - 2] - rf2);
    }
    if (a <= hi) {
        double d = fabs((double)r2[a - 1] - rf2);
        if (best < 0 || d < bd) {
            best = a;
            bd = d;
        }
    }
    *dist = bd;
    return best;
}

// Filter-scale gradient shell mrf and its window (RadialShell.jl:450-482 for A,
// ShellAnalysisGPU.jl:233-266 for B/C): the first shell whose r^2 is nearest to
// (Rf/alatt)^2, if within 10, recorded two shells below (A; one below for r^2 = 1) or one
// below (B, C).
// The last shell a peak's analysis uses: NS - 1 (the last shell is never finalised), or with
// rmax2rs > 0 the number of leading shells of 1 .. NS - 1 with r^2 <= (rmax2rs Rf / alatt)^2
// (ShellAnalysisGPU.jl:89-97, ext/CUDAExt.jl:2176-2185).
__device__ __forceinline__ int peak_last_shell(const SAArgs& a, double Rf) {
    if (!(a.rmax2rs > 0.0)) return a.NS - 1;
    double lim = a.rmax2rs * Rf / a.alatt;
    lim = lim * lim;
    int lo = 1, hi = a.NS;  // first shell s in [1, NS - 1] with r^2 > lim, else NS
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (!((double)a.shell_r2[mid - 1] <= lim)) hi = mid;
        else lo = mid + 1;
    }
    return lo - 1 < a.NS - 1 ? lo - 1 : a.NS - 1;
}

__device__ void rf_window(const SAArgs& a, double Rf, int& mrf, int& mlow_rf, int& mupp_rf,
                          int m_last = -1) {
    const double* __restrict__ rad = a.rad;
    const int NS = a.NS;
    const int m = m_last > 0 ? m_last : NS - 1;
    mrf = 1;
    const double rf = Rf / a.alatt;
    const double rf2 = rf * rf;
    double dist;
    if (a.variant == HNT_VARIANT_A) {
        int s = closest_shell(a.shell_r2, 2, NS, rf2, &dist);
        if (s >= 2 && dist < 10.0) mrf = (s == 2) ? 1 : s - 2;
    } else {
        int s = closest_shell(a.shell_r2, 2, m, rf2, &dist);
        if (s >= 2 && dist < 10.0) mrf = s - 1;
    }
    double rlow_rf = rad[mrf] - 2.0;
    if (rlow_rf < 0.0) rlow_rf = 0.0;
    mlow_rf = first_gt(rad, 1, mrf - 1, rlow_rf);
    if (mlow_rf > mrf - 1) mlow_rf = mrf;
    mupp_rf = first_ge(rad, mrf + 1, m, rad[mrf] + 2.0) - 1;
}

// Per-shell sums of Float32 terms.  SumF64: FP64 accumulation, the reference CPU
// arithmetic.  SumKahan: FP32 accumulation with Kahan compensation (the running error c is
// subtracted from each new term), read out as s - c in FP64.
struct SumF64 {
    double s;
    __device__ __forceinline__ void zero() { s = 0.0; }
    __device__ __forceinline__ void add(float v) { s += (double)v; }
    // add v if `on`, branch-free: a sum that starts at +0.0 is never -0.0, so adding +0.0
    // leaves it bit for bit as it was
    __device__ __forceinline__ void add_if(bool on, float v) { s += on ? (double)v : 0.0; }
    __device__ __forceinline__ double val() const { return s; }
};
struct SumKahan {
    float s, c;
    __device__ __forceinline__ void zero() {
        s = 0.0f;
        c = 0.0f;
    }
    __device__ __forceinline__ void add(float v) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ void add_if(bool on, float v) {
        const float y = v - c;
        const float t = s + y;
        const float cn = (t - s) - y;
        c = on ? cn : c;
        s = on ? t : s;
    }
    __device__ __forceinline__ double val() const { return (double)s - (double)c; }
};

template <class S>
struct Acc {
    S F, S0, S1, S2, T0, T1, T2, G0, G1, G2, H0, H1, H2;
    S R00, R01, R02, R10, R11, R12, R20, R21, R22;
    double lap;  // running over all shells so far, FP64 in both modes
    int64_t lapn, nin;
    __device__ __forceinline__ void zero() {
        F.zero(); S0.zero(); S1.zero(); S2.zero(); T0.zero(); T1.zero(); T2.zero();
        G0.zero(); G1.zero(); G2.zero(); H0.zero(); H1.zero(); H2.zero();
        R00.zero(); R01.zero(); R02.zero(); R10.zero(); R11.zero(); R12.zero();
        R20.zero(); R21.zero(); R22.zero();
    }
};

template <class S>
__device__ __forceinline__ void acc_cell(Acc<S>& A, const float4 v0, const float4 v1, float fx,
                                         float fy, float fz, bool eta2) {
    A.F.add(v0.x);
    A.S0.add(v0.z);
    A.S1.add(v0.w);
    A.S2.add(v1.x);
    if (eta2) {
        A.T0.add(v1.y);
        A.T1.add(v1.z);
        A.T2.add(v1.w);
    }
    // Float32 products (Float32 * Int32 is Float32 in the reference)
    A.G0.add(v0.x * fx);
    A.G1.add(v0.x * fy);
    A.G2.add(v0.x * fz);
    A.H0.add(v0.y * fx);
    A.H1.add(v0.y * fy);
    A.H2.add(v0.y * fz);
    A.R00.add(v0.z * fx);
    A.R01.add(v0.z * fy);
    A.R02.add(v0.z * fz);
    A.R10.add(v0.w * fx);
    A.R11.add(v0.w * fy);
    A.R12.add(v0.w * fz);
    A.R20.add(v1.x * fx);
    A.R21.add(v1.x * fy);
    A.R22.add(v1.x * fz);
}

// Gather one shell for one lane.  CHECK = the wave holds a peak whose ball leaves the grid.
template <bool CHECK, class S>
__device__ __forceinline__ void gather_shell(const SAArgs& a, const TileDev& T, int st, int nc, int ci, int cj,
                                             int ck, int64_t c0, Acc<S>& A, bool& oobA) {
    const bool eta2 = T.has_eta2;
    const bool lapd = T.has_lapd;
    const bool varA = a.variant == HNT_VARIANT_A;
    if (!CHECK) {
        int c = 0;
        for (; c + 4 <= nc; c += 4) {
            int4 o[4];
            float4 v0[4], v1[4];
            float lv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                o[u] = a.cells[st + c + u];
                const int64_t off = c0 + o[u].x + o[u].y * T.s1 + o[u].z * T.s2;
                v0[u] = T.recs[2 * off];
                v1[u] = T.recs[2 * off + 1];
                lv[u] = lapd ? T.lapd[off] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc_cell(A, v0[u], v1[u], (float)o[u].x, (float)o[u].y, (float)o[u].z, eta2);
                if (lapd) {
                    A.lap += (double)lv[u];
                    A.lapn += 1;
                }
            }
        }
        for (; c < nc; ++c) {
            const int4 o = a.cells[st + c];
            const int64_t off = c0 + o.x + o.y * T.s1 + o.z * T.s2;
            const float4 v0 = T.recs[2 * off], v1 = T.recs[2 * off + 1];
            acc_cell(A, v0, v1, (float)o.x, (float)o.y, (float)o.z, eta2);
            if (lapd) {
                A.lap += (double)T.lapd[off];
                A.lapn += 1;
            }
        }
        A.nin += nc;
    } else {
        for (int c = 0; c < nc; ++c) {
            const int4 o = a.cells[st + c];
            const int i1 = ci + o.x, i2 = cj + o.y, i3 = ck + o.z;
            if (i1 < 1 || i1 > T.n || i2 < 1 || i2 > T.n || i3 < 1 || i3 > T.n) {
                oobA = true;
                continue;
            }
            const int64_t off = c0 + o.x + o.y * T.s1 + o.z * T.s2;
            const float4 v0 = T.recs[2 * off], v1 = T.recs[2 * off + 1];
            acc_cell(A, v0, v1, (float)o.x, (float)o.y, (float)o.z, eta2);
            A.nin += 1;
            if (lapd && (!varA || !oobA)) {
                A.lap += (double)T.lapd[off];
                A.lapn += 1;
            }
        }
    }
}

__device__ void write_no_collapse(const TileDev& T, int p) {
    const hnt_results& r = T.res;
    r.RTHL[p] = -1.0;
    r.Fbarx[p] = 0.0;
    r.e_v[p] = 0.0;
    r.p_v[p] = 0.0;
    r.Srb[p] = 0.0;
    r.d2F[p] = 0.0;
    r.zvir_half[p] = -1.0;
    for (int q = 0; q < 9; ++q) r.strain[(int64_t)p * 9 + q] = 0.0;
    for (int L = 0; L < 3; ++L) {
        const int64_t o = (int64_t)p * 3 + L;
        r.eigs[o] = 0.0;
        r.Sbar[o] = 0.0;
        r.Sbar2[o] = 0.0;
        r.gradpk[o] = 0.0;
        r.gradpkf[o] = 0.0;
        r.gradpkrf[o] = 0.0;
    }
}

// The radial logic of one peak from its profile slot P (RadialShell.jl:449-865 for A,
// ShellAnalysisGPU.jl:231-535 for B; C = B + zvir_half of A).
// walk_lo: the walk's first phase runs from the crossing down to this shell (the learned
// bracket's lo - 2); a walk that passes it without collapsing goes on to shell 1, the full
// walk.  *lo_true (if not NULL) receives the shell holding RTHL, 0 without a collapse.
__device__ void post_process(const SAArgs& a, const TileDev& T, int p, const double* __restrict__ P, int ci, int cj,
                             int ck, double ZZon, double fcrit, int ir2min, int mrf,
                             int mlow_rf, int mupp_rf, int walk_lo = 1, int* lo_true = nullptr,
                             int m_last = -1) {
    if (lo_true) *lo_true = 0;
    const double* __restrict__ rad = a.rad;
    const int NS = a.NS;
    const int m = m_last > 0 ? m_last : NS - 1;
    const int variant = a.variant;
    auto Fb = [&](int s) -> double { return P[(int64_t)s * PW + K_F]; };

    const KS krf = kstrain(a, P, mrf, mlow_rf, mupp_rf);

    // m0: first shell with r^2 > ir2min, else the last finalised shell
    int m0 = first_r2_gt(a.shell_r2, 2, m, ir2min);
    if (m0 > m) m0 = m;

    int mupp;
    if (Fb(m0) >= fcrit) {
        int found = 0;
        for (int mm = m0; mm <= m; ++mm)
            if (Fb(mm) < fcrit) {
                m0 = mm;
                found = 1;
                break;
            }
        if (!found) {
            mupp = m;
        } else {
            mupp = first_ge(rad, m0 + 1, m, rad[m0] + 2.0) - 1;
        }
    } else {
        int mstart = m0 - 1 > 1 ? m0 - 1 : 1;
        int found = 0;
        for (int mp = mstart; mp >= 1; --mp) {
            if (Fb(mp) >= fcrit) {
                found = 1;
                break;
            }
            m0 = mp;
        }
        if (!found) {
            write_no_collapse(T, p);
            return;
        }
        mupp = first_ge(rad, m0 + 1, m, rad[m0] + 2.0) - 1;
    }

    double rlow = rad[m0] - 2.0;
    if (rlow < 0.0) rlow = 0.0;
    int mlow = first_gt(rad, 1, m0 - 1, rlow);
    if (mlow > m0 - 1) mlow = m0;

    const KS k0 = kstrain(a, P, m0, mlow, mupp);
    double lam[3], Fr, ev, pv;
    eig_ks(k0, lam);
    ev_from_lam(lam, &Fr, &ev, &pv);
    double zvir1 = -1.0, zvir1p = -1.0;
    double Frhoh = Fr;
    KS last = k0, prev = k0;  // Ebar_last/gradpk_last and strain_mat/gradpk_prev
    if (m0 == 1) {
        write_no_collapse(T, p);
        return;
    }
    bool collapsed = false;
    // phase 1: mp = m0 - 1 .. walk_lo; phase 2 (the fallback), only if phase 1 found no
    // collapse: walk_lo - 1 .. 1 -- one loop, the same steps in the same order
    (void)walk_lo;
    for (int mp = m0 - 1; mp >= 1; --mp) {
        const double rupp_mp = rad[mp] + 2.0;
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
        const KS kk = kstrain(a, P, mp, mlow, mupp);
        double e_mp, p_mp, Fr_mp;
        eig_ks(kk, lam);
        ev_from_lam(lam, &Fr_mp, &e_mp, &p_mp);
        const double Frho_mp = Fb(mp);
        double zv = -1.0;
        if (Frho_mp > 0.0) {
            const double poe = e_mp < 1e-5 ? 0.0 : p_mp / e_mp;
            zv = interp3(a.ct, log10(Frho_mp), e_mp, poe);
        }
        if (zv >= ZZon) {
            m0 = mp + 1;
            zvir1p = zv;
            last = kk;
            Frhoh = Fr_mp;
            collapsed = true;
            break;
        }
        zvir1 = zv;
        prev = kk;
        Frhoh = Fr_mp;
        last = kk;
    }
    if (!collapsed) {
        write_no_collapse(T, p);
        return;
    }

    const double dZvir = zvir1p - zvir1;
    double RTHL;
    if (zvir1 > 0.0 && dZvir != 0.0) {
        const double RTHL3 =
            cube(rad[m0 - 1]) + (cube(rad[m0]) - cube(rad[m0 - 1])) * (zvir1p - ZZon) / dZvir;
        RTHL = pow(RTHL3, 1.0 / 3.0);
    } else {
        RTHL = rad[m0 - 1];
    }
    if (RTHL <= 0.0) {
        write_no_collapse(T, p);
        return;
    }
    if (lo_true) *lo_true = RTHL == rad[m0 - 1] ? m0 - 1 : m0;
    const double RTHL3 = cube(RTHL);
    const double RTHL5 = RTHL3 * RTHL * RTHL;
    const double rad3p = cube(rad[m0 - 1]);
    const double rad3 = cube(rad[m0]);
    const double drad3 = rad3 - rad3p;
    const double dFb = Fb(m0) - Fb(m0 - 1);
    double Fbarx;
    double S11, S12, S13, S22, S23, S33, gk0, gk1, gk2, gf0, gf1, gf2;
    if (zvir1 > 0.0 && drad3 != 0.0) {
        const double frac = (RTHL3 - rad3p) / drad3;
        Fbarx = Fb(m0 - 1) + frac * dFb;
        S11 = last.E11 + frac * (prev.E11 - last.E11);
        S12 = last.E12 + frac * (prev.E12 - last.E12);
        S13 = last.E13 + frac * (prev.E13 - last.E13);
        S22 = last.E22 + frac * (prev.E22 - last.E22);
        S23 = last.E23 + frac * (prev.E23 - last.E23);
        S33 = last.E33 + frac * (prev.E33 - last.E33);
        gk0 = last.g0 + frac * (prev.g0 - last.g0);
        gk1 = last.g1 + frac * (prev.g1 - last.g1);
        gk2 = last.g2 + frac * (prev.g2 - last.g2);
        gf0 = last.f0 + frac * (prev.f0 - last.f0);
        gf1 = last.f1 + frac * (prev.f1 - last.f1);
        gf2 = last.f2 + frac * (prev.f2 - last.f2);
    } else {
        Fbarx = Fb(m0 - 1);
        S11 = last.E11; S12 = last.E12; S13 = last.E13;
        S22 = last.E22; S23 = last.E23; S33 = last.E33;
        gk0 = prev.g0; gk1 = prev.g1; gk2 = prev.g2;
        gf0 = prev.f0; gf1 = prev.f1; gf2 = prev.f2;
    }
    // normalize_strain (RadialShell.jl:207-210) and the final eigenvalues
    const double tr = S11 + S22 + S33;
    if (tr != 0.0) {
        const double sc = Fbarx / tr;
        S11 *= sc; S12 *= sc; S13 *= sc; S22 *= sc; S23 *= sc; S33 *= sc;
    }
    double lamf[3], Frhoc, e_f, p_f;
    eig3_sym(S11, S12, S13, S22, S23, S33, lamf);
    ev_from_lam(lamf, &Frhoc, &e_f, &p_f);
    double F11 = last.E11, F12 = last.E12, F13 = last.E13, F22 = last.E22, F23 = last.E23,
           F33 = last.E33;
    if (variant == HNT_VARIANT_A || Frhoh != 0.0) {
        const double sc2 = Fbarx / Frhoh;
        F11 *= sc2; F12 *= sc2; F13 *= sc2; F22 *= sc2; F23 *= sc2; F33 *= sc2;
    }

    // Sbar, Sbar2: running sums over shells 1..m0-1 kept in the slot
    const double* X = P + (int64_t)(m0 - 1) * PW;
    double Sb0 = X[K_S], Sb1 = X[K_S + 1], Sb2v = X[K_S + 2];
    double Tb0 = X[K_S2], Tb1 = X[K_S2 + 1], Tb2 = X[K_S2 + 2];
    const double nS = X[K_N];
    if (nS > 0.0) {
        Sb0 /= nS; Sb1 /= nS; Sb2v /= nS;
        Tb0 /= nS; Tb1 /= nS; Tb2 /= nS;
    }

    double Srb = 0.0, rad5p = 0.0;
    if (m0 > 2)
        for (int mp = 2; mp <= m0 - 1; ++mp) {
            const double rad5 = pow5(rad[mp]);
            Srb += 0.5 * (Fb(mp - 1) + Fb(mp)) * (rad5 - rad5p);
            rad5p = rad5;
        }
    if (zvir1 > 0.0 && dZvir != 0.0) Srb += 0.5 * (Fb(m0 - 1) + Fbarx) * (RTHL5 - rad5p);
    if (Fbarx * RTHL5 > 0.0) Srb /= (Fbarx * RTHL5);

    // last shell within RTHL: d2F and the mask
    const int sR = first_gt(rad, 1, NS, RTHL) - 1;
    double d2F = 0.0;
    if (T.has_lapd) {
        const double* Y = P + (int64_t)sR * PW;
        if (Y[K_LAPN] > 0.0) d2F = Y[K_LAP] / Y[K_LAPN];
    }
    if (T.mask && a.paint) {
        // shells 1 .. sR are the cells with d^2 <= r^2 of shell sR: paint_mask_kernel marks them
        a.paint[T.base + p] = a.shell_r2[sR - 1];
    } else if (T.mask) {
        const int nb = a.nbuff;
        for (int s = 1; s <= sR; ++s) {
            const int st = a.shell_start[s - 1], nc = a.shell_count[s - 1];
            for (int c = 0; c < nc; ++c) {
                const int4 o = a.cells[st + c];
                const int i1 = ci + o.x, i2 = cj + o.y, i3 = ck + o.z;
                if (i1 > nb && i1 <= T.n - nb && i2 > nb && i2 <= T.n - nb && i3 > nb &&
                    i3 <= T.n - nb && i1 >= 1 && i2 >= 1 && i3 >= 1)
                    T.mask[(int64_t)(i1 - 1) + (i2 - 1) * T.s1 + (i3 - 1) * T.s2] = 1;
            }
        }
    }

    // formation redshift (RadialShell.jl:801-859); variant B returns 0
    double zvir_half = -1.0;
    if (variant == HNT_VARIANT_B) {
        zvir_half = 0.0;
    } else if (RTHL / 1.0 >= 3.0) {
        for (int jj = 1; jj <= 10; ++jj) {
            const double rcur = RTHL * pow((double)(jj - 1) / 9.0 * 0.5 + 0.5, 1.0 / 3.0);
            int mupp_rc = 1;
            if (m - 1 >= 2) {
                int f = first_gt(rad, 2, m - 1, rcur + 2.0);
                mupp_rc = f <= m - 1 ? f : m - 1;
            }
            double rlow_rc = rcur - 2.0;
            if (rlow_rc < 0.0) rlow_rc = 0.0;
            int mlow_rc = 1;
            {
                int f = first_gt(rad, 1, m - 1, rlow_rc);
                if (f <= m - 1) mlow_rc = f;
            }
            int m0_rc = 1;
            if (m - 1 >= 1) {
                int f = first_gt(rad, 1, m - 1, rcur);
                m0_rc = f <= m - 1 ? f : m - 1;
            }
            const KS kr = kstrain(a, P, m0_rc, mlow_rc, mupp_rc);
            double Frr, e_rc, p_rc;
            eig_ks(kr, lam);
            ev_from_lam(lam, &Frr, &e_rc, &p_rc);
            const double Frhoc_rc = Fb(m0_rc);
            const double poe = e_rc < 1e-5 ? 0.0 : p_rc / e_rc;
            double zr = -1.0;
            if (Frhoc_rc > 0.0) zr = interp3(a.ct, log10(Frhoc_rc), e_rc, poe);
            const double zs = zr - 1.0;
            if (zs > zvir_half) zvir_half = zs;
        }
    }

    const hnt_results& R = T.res;
    R.RTHL[p] = RTHL;
    R.Fbarx[p] = Fbarx;
    R.e_v[p] = e_f;
    R.p_v[p] = p_f;
    R.Srb[p] = Srb;
    R.d2F[p] = d2F;
    R.zvir_half[p] = zvir_half;
    double* sm = R.strain + (int64_t)p * 9;
    sm[0] = F11; sm[1] = F12; sm[2] = F13;
    sm[3] = F12; sm[4] = F22; sm[5] = F23;
    sm[6] = F13; sm[7] = F23; sm[8] = F33;
    const int64_t o3 = (int64_t)p * 3;
    R.eigs[o3] = lamf[0]; R.eigs[o3 + 1] = lamf[1]; R.eigs[o3 + 2] = lamf[2];
    R.Sbar[o3] = Sb0; R.Sbar[o3 + 1] = Sb1; R.Sbar[o3 + 2] = Sb2v;
    R.Sbar2[o3] = Tb0; R.Sbar2[o3 + 1] = Tb1; R.Sbar2[o3 + 2] = Tb2;
    R.gradpk[o3] = gk0; R.gradpk[o3 + 1] = gk1; R.gradpk[o3 + 2] = gk2;
    R.gradpkf[o3] = gf0; R.gradpkf[o3 + 1] = gf1; R.gradpkf[o3 + 2] = gf2;
    R.gradpkrf[o3] = krf.g0; R.gradpkrf[o3 + 1] = krf.g1; R.gradpkrf[o3 + 2] = krf.g2;
}

// A shell index from a network output log(1 + s), rounded, clamped to [1, m].
__device__ __forceinline__ int bracket_shell(float y, int m) {
    float s = expf(y) - 1.0f;
    if (!(s >= 1.0f)) s = 1.0f;  // NaN too
    if (s > (float)m) s = (float)m;
    return (int)(s + 0.5f);
}

template <class S>
__device__ void analyse_group(const SAArgs& a, int64_t grp, double* __restrict__ P) {
    const int lane = threadIdx.x;
    const TileDev& T = a.tiles[a.gtile[grp]];
    const int p = a.slot[grp * 64 + lane];
    const bool valid = p >= 0;
    int ci = 1, cj = 1, ck = 1, ir2min = 0;
    double ZZon = 0.0, fcrit = 0.0, Rf = 0.0;
    if (valid) {
        ci = T.pi[p];
        cj = T.pj[p];
        ck = T.pk[p];
        ir2min = T.ir2min[p];
        Rf = T.Rf[p];
        ZZon = T.ZZon ? T.ZZon[p] : T.ZZon_s;
        fcrit = T.fcrit ? T.fcrit[p] : T.fcrit_s;
    }
    const double* __restrict__ rad = a.rad;
    const int NS = a.NS;
    const int m = NS - 1;
    // this lane's last shell (fewer with rmax2rs > 0), and peaks that get no analysis
    const int m_lane = valid ? peak_last_shell(a, Rf) : m;
    bool none = valid && (m_lane < 2 || (a.skip && a.skip[T.base + p]));
    if (none) {
        write_no_collapse(T, p);
        if (T.res.nshells_used) T.res.nshells_used[p] = 0;
        if (T.res.crossing_shell) T.res.crossing_shell[p] = 0;
        if (T.res.bracket) T.res.bracket[p] = 0;
    }

    int mrf = 1, mlow_rf = 1, mupp_rf = 1;
    if (valid && !none) rf_window(a, Rf, mrf, mlow_rf, mupp_rf, m_lane);

    // ---- gather ----
    const int64_t c0 = (int64_t)(ci - 1) + (cj - 1) * T.s1 + (ck - 1) * T.s2;
    const int R = a.rmax;
    const bool inside = ci - R >= 1 && ci + R <= T.n && cj - R >= 1 && cj + R <= T.n &&
                        ck - R >= 1 && ck + R <= T.n;
    const bool wave_inside = __ballot(!valid || inside) == ~0ull;

    double Fprev = 0.0, dFbarp = 0.0, cumN = 0.0, cS0 = 0.0, cS1 = 0.0, cS2 = 0.0, cT0 = 0.0,
           cT1 = 0.0, cT2 = 0.0, lap = 0.0;
    int64_t lapn = 0;
    if (valid) {
        const float4 v0 = T.recs[2 * c0], v1 = T.recs[2 * c0 + 1];
        Fprev = (double)v0.x;
        dFbarp = Fprev;
        cumN = 1.0;
        cS0 = 0.0 + (double)v0.z;
        cS1 = 0.0 + (double)v0.w;
        cS2 = 0.0 + (double)v1.x;
        if (T.has_eta2) {
            cT0 = 0.0 + (double)v1.y;
            cT1 = 0.0 + (double)v1.z;
            cT2 = 0.0 + (double)v1.w;
        }
        if (T.has_lapd && a.variant != HNT_VARIANT_A) {
            lap = 0.0 + (double)T.lapd[c0];
            lapn = 1;
        }
        double* X = P + PW;
        for (int q = 0; q < 12; ++q) X[q] = 0.0;
        X[K_F] = Fprev;
        X[K_LAP] = lap;
        X[K_LAPN] = (double)lapn;
        X[K_LINE2 - 1] = 0.0;
        X[K_N] = cumN;
        X[K_S] = cS0; X[K_S + 1] = cS1; X[K_S + 2] = cS2;
        X[K_S2] = cT0; X[K_S2 + 1] = cT1; X[K_S2 + 2] = cT2;
        for (int q = K_S2 + 3; q < PW; ++q) X[q] = 0.0;
    }

    bool active = valid && !none;
    bool past = false, crossed = false, stop_set = false, oobA = false;
    double rstop = 0.0;
    int stop = m_lane, scross = 0;
    for (int s = 2; s <= m; ++s) {
        if (s > m_lane) active = false;
        if (__ballot(active) == 0ull) break;
        const int st = a.shell_start[s - 1], nc = a.shell_count[s - 1];
        const double radS = rad[s], radP = rad[s - 1];
        if (active) {
            Acc<S> A;
            A.zero();
            A.lap = lap; A.lapn = lapn; A.nin = 0;
            if (wave_inside) gather_shell<false, S>(a, T, st, nc, ci, cj, ck, c0, A, oobA);
            else gather_shell<true, S>(a, T, st, nc, ci, cj, ck, c0, A, oobA);
            lap = A.lap;
            lapn = A.lapn;
            const int64_t nin = A.nin;
            const double dFbar = nin > 0 ? A.F.val() / (double)nin : 0.0;
            const double rad3p = cube(radP), rad3 = cube(radS);
            const double Fb = (rad3p * Fprev + 0.5 * (dFbarp + dFbar) * (rad3 - rad3p)) / rad3;
            Fprev = Fb;
            dFbarp = dFbar;
            if (nin > 0) {
                cumN += (double)nin;
                cS0 += A.S0.val(); cS1 += A.S1.val(); cS2 += A.S2.val();
                cT0 += A.T0.val(); cT1 += A.T1.val(); cT2 += A.T2.val();
            }
            const double r00 = A.R00.val() / radS, r01 = A.R01.val() / radS,
                         r02 = A.R02.val() / radS;
            const double r10 = A.R10.val() / radS, r11 = A.R11.val() / radS,
                         r12 = A.R12.val() / radS;
            const double r20 = A.R20.val() / radS, r21 = A.R21.val() / radS,
                         r22 = A.R22.val() / radS;
            double* X = P + (int64_t)s * PW;
            X[0] = A.G0.val() / radS; X[1] = A.G1.val() / radS; X[2] = A.G2.val() / radS;
            X[3] = A.H0.val() / radS; X[4] = A.H1.val() / radS; X[5] = A.H2.val() / radS;
            X[6] = r00 + r00; X[7] = r01 + r10; X[8] = r02 + r20;
            X[9] = r11 + r11; X[10] = r12 + r21; X[11] = r22 + r22;
            X[K_F] = Fb;
            X[K_LAP] = lap;
            X[K_LAPN] = (double)lapn;
            X[K_LINE2 - 1] = 0.0;
            // the second line only up to the crossing shell: the radial logic reads it at
            // m0 - 1 < crossing shell, and whole lines left out cost no partial writes
            if (!crossed) {
                X[K_N] = cumN;
                X[K_S] = cS0; X[K_S + 1] = cS1; X[K_S + 2] = cS2;
                X[K_S2] = cT0; X[K_S2 + 1] = cT1; X[K_S2 + 2] = cT2;
#pragma unroll
                for (int q = K_S2 + 3; q < PW; ++q) X[q] = 0.0;
            }
            // early exit: with s_c the first shell past ir2min where Fbar < fcrit, the radial
            // logic reads Fbar up to s_c + 1 and kernel_strain inputs of shells with
            // rad < rad[s_c + 1] + 2 (and the filter-scale window up to mupp_rf)
            if (!past && a.shell_r2[s - 1] > ir2min) past = true;
            if (past && !crossed && Fb < fcrit) {
                crossed = true;
                scross = s;
                rstop = rad[s + 1] + 2.0;
            }
            if (!stop_set && crossed && rad[s + 1] >= rstop && s >= mupp_rf) {
                stop = s;
                stop_set = true;
                if (a.early) active = false;
            }
        }
    }
    if (!valid || none) return;
    if (T.res.nshells_used) T.res.nshells_used[p] = stop;
    if (T.res.crossing_shell) T.res.crossing_shell[p] = scross;
    if (!a.pred) {
        post_process(a, T, p, P, ci, cj, ck, ZZon, fcrit, ir2min, mrf, mlow_rf, mupp_rf, 1,
                     nullptr, m_lane);
        if (T.res.bracket) T.res.bracket[p] = 0;
        return;
    }
    // learned bracket [lo, hi]: the gather verified when it ended within hi + 2, the walk
    // when it collapsed at or above lo - 2
    const float* y = a.pred + (T.base + p) * HNT_MLP_OUTPUTS;
    const int lo = bracket_shell(y[0], m_lane), hi = bracket_shell(y[1], m_lane);
    const int walk_lo = lo - 2 > 1 ? lo - 2 : 1;
    int lo_true = 0;
    post_process(a, T, p, P, ci, cj, ck, ZZon, fcrit, ir2min, mrf, mlow_rf, mupp_rf, walk_lo,
                 &lo_true, m_lane);
    if (T.res.bracket)
        T.res.bracket[p] = (stop <= hi + 2 ? 1 : 0) | (lo_true >= walk_lo && lo_true > 0 ? 2 : 0) |
                           ((y[2] > 0.0f) == (lo_true > 0) ? 4 : 0);
}

// ------------------------------------------------------------------ ganged gather
// Sixteen lanes (one DPP row) per peak, four peaks per wavefront at a time.  Lane k of a
// gang sums cells k, k + 16, ... of each shell; the sixteen partial sums of each quantity
// are then added in lane order through LDS, one or two quantities per lane.  Compared with
// one lane per peak this keeps a quarter as many peaks in flight per XCD, so the cell
// records they read stay in L2.  The sums are the same FP64 sums of the same Float32 terms
// in another order (results agree with the oracle to ~1e-13 instead of bit for bit).
// A wavefront still owns a group of 64 peaks: it gathers them four at a time (16 rounds,
// in schedule order, so the four have similar depths), then post-processes all 64 with
// one lane per peak.
constexpr int GANG = 16;
constexpr int QP = 26;  // LDS row of one lane's partial sums (25 used)
// row layout: F, S0..2, T0..2, G0..2, H0..2, R00 R01 R02 R10 R11 R12 R20 R21 R22, nin, lap, lapn
constexpr int Q_F = 0, Q_S = 1, Q_T = 4, Q_G = 7, Q_H = 10, Q_R = 13, Q_NIN = 22, Q_LAP = 23,
              Q_LAPN = 24;

template <class S, bool CHECK>
__device__ void gang_gather(const SAArgs& a, const TileDev& T, int p, double* __restrict__ P,
                            double* __restrict__ lds) {
    const int lane = threadIdx.x, gang = lane >> 4, k = lane & (GANG - 1), row0 = gang * GANG;
    const bool valid = p >= 0;
    int ci = 1, cj = 1, ck = 1, ir2min = 0;
    double fcrit = 0.0, Rf = 0.0;
    if (valid) {
        ci = T.pi[p];
        cj = T.pj[p];
        ck = T.pk[p];
        ir2min = T.ir2min[p];
        Rf = T.Rf[p];
        fcrit = T.fcrit ? T.fcrit[p] : T.fcrit_s;
    }
    const double* __restrict__ rad = a.rad;
    const int m = a.NS - 1;
    int mrf = 1, mlow_rf = 1, mupp_rf = 1;
    if (valid) rf_window(a, Rf, mrf, mlow_rf, mupp_rf);
    const int64_t c0 = (int64_t)(ci - 1) + (cj - 1) * T.s1 + (ck - 1) * T.s2;
    const bool varA = a.variant == HNT_VARIANT_A;
    const bool eta2 = T.has_eta2, haslap = T.has_lapd;

    // gang-uniform: Fprev, dFbarp (every lane of the gang computes them identically);
    // lane-held running values: k = 0 cumulative n, 1..3 cumulative S, 4..6 cumulative S2,
    // 10 running lapd sum, 11 running lapd count
    double Fprev = 0.0, dFbarp = 0.0, run = 0.0;
    if (valid) {
        const float4 v0 = T.recs[2 * c0], v1 = T.recs[2 * c0 + 1];
        Fprev = (double)v0.x;
        dFbarp = Fprev;
        const float e3[3] = {v0.z, v0.w, v1.x}, t3[3] = {v1.y, v1.z, v1.w};
        if (k == 0) run = 1.0;
        else if (k <= 3) run = 0.0 + (double)e3[k - 1];
        else if (k <= 6) run = eta2 ? 0.0 + (double)t3[k - 4] : 0.0;
        else if (k == 10) run = (haslap && !varA) ? 0.0 + (double)T.lapd[c0] : 0.0;
        else if (k == 11) run = (haslap && !varA) ? 1.0 : 0.0;
        double* X = P + PW;
        if (k < 12) X[k] = 0.0;
        if (k == 0) {
            X[K_F] = Fprev;
            X[K_N] = run;
        } else if (k <= 3) {
            X[K_S + k - 1] = run;
        } else if (k <= 6) {
            X[K_S2 + k - 4] = run;
        } else if (k == 10) {
            X[K_LAP] = run;
        } else if (k == 11) {
            X[K_LAPN] = run;
        }
    }

    bool active = valid, past = false, crossed = false, stop_set = false, oobA = false;
    double rstop = 0.0;
    int stop = m, scross = 0;
    double* __restrict__ row = lds + lane * QP;
    for (int s = 2; s <= m; ++s) {
        if (__ballot(active) == 0ull) break;
        const int st = a.shell_start[s - 1], nc = a.shell_count[s - 1];
        const double radS = rad[s], radP = rad[s - 1];
        // variant A stops adding lapd at the first out-of-bounds cell in cell order
        int cutA = nc;
        if (CHECK && varA && haslap) {
            int first = nc;
            if (active)
                for (int c = k; c < nc; c += GANG) {
                    const int4 o = a.cells[st + c];
                    const int i1 = ci + o.x, i2 = cj + o.y, i3 = ck + o.z;
                    if (i1 < 1 || i1 > T.n || i2 < 1 || i2 > T.n || i3 < 1 || i3 > T.n) {
                        first = c;
                        break;
                    }
                }
#pragma unroll
            for (int w = 1; w < GANG; w <<= 1) first = min(first, __shfl_xor(first, w));
            cutA = oobA ? 0 : first;
        }
        Acc<S> A;
        A.zero();
        A.lap = 0.0;
        A.lapn = 0;
        A.nin = 0;
        if (active) {
            if (!CHECK) {
                int c = k;
                for (; c + GANG < nc; c += 2 * GANG) {
                    const int4 o0 = a.cells[st + c], o1 = a.cells[st + c + GANG];
                    const int64_t f0 = c0 + o0.x + o0.y * T.s1 + o0.z * T.s2;
                    const int64_t f1 = c0 + o1.x + o1.y * T.s1 + o1.z * T.s2;
                    const float4 a0 = T.recs[2 * f0], b0 = T.recs[2 * f0 + 1];
                    const float4 a1 = T.recs[2 * f1], b1 = T.recs[2 * f1 + 1];
                    const float l0 = haslap ? T.lapd[f0] : 0.0f;
                    const float l1 = haslap ? T.lapd[f1] : 0.0f;
                    acc_cell(A, a0, b0, (float)o0.x, (float)o0.y, (float)o0.z, eta2);
                    acc_cell(A, a1, b1, (float)o1.x, (float)o1.y, (float)o1.z, eta2);
                    A.lap += (double)l0;
                    A.lap += (double)l1;
                }
                if (c < nc) {
                    const int4 o0 = a.cells[st + c];
                    const int64_t f0 = c0 + o0.x + o0.y * T.s1 + o0.z * T.s2;
                    const float4 a0 = T.recs[2 * f0], b0 = T.recs[2 * f0 + 1];
                    acc_cell(A, a0, b0, (float)o0.x, (float)o0.y, (float)o0.z, eta2);
                    if (haslap) A.lap += (double)T.lapd[f0];
                }
                // cells of this lane: ceil((nc - k) / 16)
                A.nin = k < nc ? (nc - k + GANG - 1) / GANG : 0;
                A.lapn = haslap ? A.nin : 0;
            } else {
                for (int c = k; c < nc; c += GANG) {
                    const int4 o = a.cells[st + c];
                    const int i1 = ci + o.x, i2 = cj + o.y, i3 = ck + o.z;
                    if (i1 < 1 || i1 > T.n || i2 < 1 || i2 > T.n || i3 < 1 || i3 > T.n) continue;
                    const int64_t f = c0 + o.x + o.y * T.s1 + o.z * T.s2;
                    const float4 v0 = T.recs[2 * f], v1 = T.recs[2 * f + 1];
                    acc_cell(A, v0, v1, (float)o.x, (float)o.y, (float)o.z, eta2);
                    A.nin += 1;
                    if (haslap && (!varA || c < cutA)) {
                        A.lap += (double)T.lapd[f];
                        A.lapn += 1;
                    }
                }
            }
        }
        if (CHECK && varA && haslap && cutA < nc) oobA = true;
        // partial sums to LDS, then lane (gang, k) adds rows row0..row0+15 of its quantities
        row[Q_F] = A.F.val();
        row[Q_S] = A.S0.val(); row[Q_S + 1] = A.S1.val(); row[Q_S + 2] = A.S2.val();
        row[Q_T] = A.T0.val(); row[Q_T + 1] = A.T1.val(); row[Q_T + 2] = A.T2.val();
        row[Q_G] = A.G0.val(); row[Q_G + 1] = A.G1.val(); row[Q_G + 2] = A.G2.val();
        row[Q_H] = A.H0.val(); row[Q_H + 1] = A.H1.val(); row[Q_H + 2] = A.H2.val();
        row[Q_R] = A.R00.val(); row[Q_R + 1] = A.R01.val(); row[Q_R + 2] = A.R02.val();
        row[Q_R + 3] = A.R10.val(); row[Q_R + 4] = A.R11.val(); row[Q_R + 5] = A.R12.val();
        row[Q_R + 6] = A.R20.val(); row[Q_R + 7] = A.R21.val(); row[Q_R + 8] = A.R22.val();
        row[Q_NIN] = (double)A.nin;
        row[Q_LAP] = A.lap;
        row[Q_LAPN] = (double)A.lapn;
        __syncthreads();
        // quantity of lane k (and a second one for some lanes)
        int q1, q2 = -1;
        if (k <= 6) q1 = k;                                   // F, S, S2
        else if (k <= 9) { q1 = Q_G + (k - 7); q2 = Q_R + 4 * (k - 7); }  // G and R00/R11/R22
        else if (k == 10) { q1 = Q_H; q2 = Q_LAP; }
        else if (k == 11) { q1 = Q_H + 1; q2 = Q_LAPN; }
        else if (k == 12) { q1 = Q_H + 2; q2 = Q_NIN; }
        else if (k == 13) { q1 = Q_R + 1; q2 = Q_R + 3; }    // R01, R10
        else if (k == 14) { q1 = Q_R + 2; q2 = Q_R + 6; }    // R02, R20
        else { q1 = Q_R + 5; q2 = Q_R + 7; }                  // R12, R21
        double v1 = 0.0, v2 = 0.0;
#pragma unroll
        for (int r = 0; r < GANG; ++r) {
            const double* rr = lds + (row0 + r) * QP;
            v1 += rr[q1];
            if (q2 >= 0) v2 += rr[q2];
        }
        __syncthreads();
        // Fbar recurrence, identical in every lane of the gang
        const double Ftot = __shfl(v1, row0);
        const double nin = __shfl(v2, row0 + 12);
        const double dFbar = nin > 0.0 ? Ftot / nin : 0.0;
        const double rad3p = cube(radP), rad3 = cube(radS);
        const double Fb = (rad3p * Fprev + 0.5 * (dFbarp + dFbar) * (rad3 - rad3p)) / rad3;
        Fprev = Fb;
        dFbarp = dFbar;
        if (active) {
            double* X = P + (int64_t)s * PW;
            if (k == 0) {
                if (nin > 0.0) run += nin;
                X[K_F] = Fb;
                X[K_N] = run;
            } else if (k <= 3) {
                if (nin > 0.0) run += v1;
                X[K_S + k - 1] = run;
            } else if (k <= 6) {
                if (nin > 0.0) run += v1;
                X[K_S2 + k - 4] = run;
            } else if (k <= 9) {
                X[k - 7] = v1 / radS;
                const double rd = v2 / radS;
                X[k == 7 ? 6 : (k == 8 ? 9 : 11)] = rd + rd;
            } else if (k <= 12) {
                X[3 + (k - 10)] = v1 / radS;
                if (k == 10) {
                    run += v2;
                    X[K_LAP] = run;
                } else if (k == 11) {
                    run += v2;
                    X[K_LAPN] = run;
                }
            } else {
                const double ra = v1 / radS, rb = v2 / radS;
                X[k == 13 ? 7 : (k == 14 ? 8 : 10)] = ra + rb;
            }
            if (!past && a.shell_r2[s - 1] > ir2min) past = true;
            if (past && !crossed && Fb < fcrit) {
                crossed = true;
                scross = s;
                rstop = rad[s + 1] + 2.0;
            }
            if (!stop_set && crossed && rad[s + 1] >= rstop && s >= mupp_rf) {
                stop = s;
                stop_set = true;
                if (a.early) active = false;
            }
        }
    }
    if (valid && k == 0) {
        if (T.res.nshells_used) T.res.nshells_used[p] = stop;
        if (T.res.crossing_shell) T.res.crossing_shell[p] = scross;
    }
}

template <class S>
__global__ __launch_bounds__(64) void shell_gang_kernel(SAArgs a) {
    __shared__ double lds[64 * QP];
    double* Pbase = a.prof + (int64_t)blockIdx.x * 64 * a.slot_doubles;
    int q = blockIdx.x & 7;
    for (int tried = 0; tried < 8;) {
        unsigned long long g = 0;
        if (threadIdx.x == 0) g = atomicAdd(&a.qctr[q], 1ull);
        g = __shfl(g, 0);
        const int64_t grp = a.qbeg[q] + (int64_t)g;
        if (grp >= a.qend[q]) {
            ++tried;
            q = (q + 1) & 7;
            continue;
        }
        const TileDev& T = a.tiles[a.gtile[grp]];
        const int R = a.rmax;
        for (int r = 0; r < 64 / 4; ++r) {
            const int idx = r * 4 + (threadIdx.x >> 4);
            const int p = a.slot[grp * 64 + idx];
            if (__ballot(p >= 0) == 0ull) continue;
            bool inside = true;
            if (p >= 0) {
                const int ci = T.pi[p], cj = T.pj[p], ck = T.pk[p];
                inside = ci - R >= 1 && ci + R <= T.n && cj - R >= 1 && cj + R <= T.n &&
                         ck - R >= 1 && ck + R <= T.n;
            }
            double* P = Pbase + (int64_t)idx * a.slot_doubles;
            if (__ballot(!inside) == 0ull) gang_gather<S, false>(a, T, p, P, lds);
            else gang_gather<S, true>(a, T, p, P, lds);
        }
        // radial logic of the 64 peaks, one lane each
        const int lane = threadIdx.x;
        const int p = a.slot[grp * 64 + lane];
        if (p >= 0) {
            const int ci = T.pi[p], cj = T.pj[p], ck = T.pk[p];
            const double ZZon = T.ZZon ? T.ZZon[p] : T.ZZon_s;
            const double fcrit = T.fcrit ? T.fcrit[p] : T.fcrit_s;
            int mrf, mlow_rf, mupp_rf;
            rf_window(a, T.Rf[p], mrf, mlow_rf, mupp_rf);
            post_process(a, T, p, Pbase + (int64_t)lane * a.slot_doubles, ci, cj, ck, ZZon,
                         fcrit, T.ir2min[p], mrf, mlow_rf, mupp_rf);
        }
    }
}

// ------------------------------------------------------------------ wavefront gather
// One wavefront per peak.  The wavefront walks its peak's cell table in chunks of up to
// WCAP cells (ending at most WMAXSH shells later): first all 64 lanes load the chunk's cell
// records into LDS at once (the records of a peak's ball share cache lines, and lines
// requested together are fetched from HBM once), then 22 lanes each run one of the
// reference's shell sums over the chunk cell by cell in table order -- the same FP64 (or
// Kahan) additions of the same Float32 terms as the one-lane-per-peak gather, so the
// results are bit-identical to it -- and lane 22 the running lapd sum.  At every shell end
// the sums go to an LDS table; after the chunk, lane e of the wavefront turns each finished
// shell into profile entry e, and every lane evaluates the Fbar recurrence and the early-exit
// rule.  With the crossing pass's exact stop shell a chunk never reads past the stop; without
// it, at most one chunk is read beyond.  The 64 peaks of a group are gathered one after the
// other, then post-processed one lane each.
constexpr int WCAP = 256, WMAXSH = 32;
static_assert(WCAP % 8 == 0 && WMAXSH + 2 <= 64, "wave gather chunk shape");
// Lane e < 22 of the wavefront produces profile entry e in this numbering (Fbar, n, S, S2,
// lapd sum, lapd count after the twelve kernel_strain inputs); wave_pos gives its place in
// the slot.
constexpr int WQ_F = 12, WQ_N = 13, WQ_S = 14, WQ_S2 = 17, WQ_LAP = 20, WQ_LAPN = 21;
__device__ __forceinline__ int wave_pos(int e) {
    return e < WQ_F ? e : e == WQ_F ? K_F : e == WQ_N ? K_N : e < WQ_S2 ? K_S + e - WQ_S
         : e < WQ_LAP ? K_S2 + e - WQ_S2 : e == WQ_LAP ? K_LAP : K_LAPN;
}

// lane l's double, read by every lane (l uniform)
__device__ __forceinline__ double rl_double(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

struct WaveLds {
    float4 rec[2 * WCAP];      // cell records {delta, delta^T, psi1 xyz, psi2 xyz}
    float mult[3][WCAP];       // the cell's offsets as Float32 (x, y, z)
    float lap[WCAP];
    uint8_t in[WCAP];          // cell inside the grid
    double shs[WMAXSH][24];    // per finished shell: the 22 sums, running lap, lapn
    int shn[WMAXSH];           // cells inside the grid, per finished shell
};

// What the gather of one peak needs besides the tables, computed by the peak's own lane
// for all 64 peaks of a group at once and handed to the wavefront with readlane.
struct WavePeak {
    int ci, cj, ck, ir2min, mupp_rf, s_lim;
    double fcrit;
    float4 v0c, v1c;  // the peak cell's record
    float lapc;
};

__device__ __forceinline__ float rl_float(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ float4 rl_float4(float4 v, int l) {
    return make_float4(rl_float(v.x, l), rl_float(v.y, l), rl_float(v.z, l), rl_float(v.w, l));
}

// Registers of one chunk in flight: lane l holds cells cA + l + 64u.
struct WaveFetch {
    int4 o[WCAP / 64];
    float4 r0[WCAP / 64], r1[WCAP / 64];
    float lp[WCAP / 64];
    bool ok[WCAP / 64];
};

template <bool CHECK>
__device__ __forceinline__ void wave_fetch(const SAArgs& a, const TileDev& T, const WavePeak& K,
                                           int64_t c0, int cA, int cB, WaveFetch& F) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int u = 0; u < WCAP / 64; ++u) {
        const int c = cA + lane + 64 * u;
        F.o[u] = c < cB ? a.cells[c] : make_int4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < WCAP / 64; ++u) {
        const int c = cA + lane + 64 * u;
        const int4 o = F.o[u];
        bool ok = c < cB;
        if (CHECK) {
            const int i1 = K.ci + o.x, i2 = K.cj + o.y, i3 = K.ck + o.z;
            ok = ok && i1 >= 1 && i1 <= T.n && i2 >= 1 && i2 <= T.n && i3 >= 1 && i3 <= T.n;
        }
        F.ok[u] = ok;
        if (ok) {
            const int64_t off = c0 + o.x + o.y * T.s1 + o.z * T.s2;
            F.r0[u] = T.recs[2 * off];
            F.r1[u] = T.recs[2 * off + 1];
            F.lp[u] = T.has_lapd ? T.lapd[off] : 0.0f;
        }
    }
}

template <class S, bool CHECK>
__device__ void wave_gather(const SAArgs& a, const TileDev& T, int p, const WavePeak& K,
                            double* __restrict__ P, WaveLds& L) {
    const int lane = threadIdx.x;
    const double* __restrict__ rad = a.rad;
    const int m = a.NS - 1;
    const int64_t c0 = (int64_t)(K.ci - 1) + (K.cj - 1) * T.s1 + (K.ck - 1) * T.s2;
    const bool varA = a.variant == HNT_VARIANT_A;
    const bool eta2 = T.has_eta2, haslap = T.has_lapd;
    const double fcrit = K.fcrit;

    // sum lanes: 0 F, 1-3 S, 4-6 S2 (psi2), 7-9 G, 10-12 H, 13-21 R00 R01 R02 R10 .. R22;
    // record float fidx, times the cell's offset component msel (0 = no product)
    int fidx = 0, msel = 0;
    if (lane <= 6) fidx = lane == 0 ? 0 : lane + 1;
    else if (lane <= 9) msel = lane - 6;
    else if (lane <= 12) { fidx = 1; msel = lane - 9; }
    else if (lane <= 21) { fidx = 2 + (lane - 13) / 3; msel = 1 + (lane - 13) % 3; }
    const bool summer = lane < 22 && (eta2 || lane < 4 || lane > 6);
    const bool lapper = lane == 22 && haslap;
    const float* __restrict__ recf = reinterpret_cast<const float*>(L.rec);
    const float* __restrict__ multf = L.mult[msel > 0 ? msel - 1 : 0];

    // shell 1: the peak cell
    const float4 v0c = K.v0c, v1c = K.v1c;
    double Fprev = (double)v0c.x, dFbarp = Fprev, cumN = 1.0;
    double run = 0.0;  // entry lanes 14-16 cumulative psi1, 17-19 cumulative psi2
    if (lane == 14) run = 0.0 + (double)v0c.z;
    else if (lane == 15) run = 0.0 + (double)v0c.w;
    else if (lane == 16) run = 0.0 + (double)v1c.x;
    else if (lane == 17) run = eta2 ? 0.0 + (double)v1c.y : 0.0;
    else if (lane == 18) run = eta2 ? 0.0 + (double)v1c.z : 0.0;
    else if (lane == 19) run = eta2 ? 0.0 + (double)v1c.w : 0.0;
    double lap = 0.0;  // lane 22
    int64_t lapn = 0;
    bool oobA = false;
    if (haslap && !varA) {
        lap = 0.0 + (double)K.lapc;
        lapn = 1;
    }
    if (lane < 22) {
        double x = 0.0;
        if (lane == WQ_F) x = Fprev;
        else if (lane == WQ_N) x = cumN;
        else if (lane >= WQ_S && lane < WQ_LAP) x = run;
        else if (lane == WQ_LAP) x = lap;
        else if (lane == WQ_LAPN) x = (double)lapn;
        P[PW + wave_pos(lane)] = x;
    }

    bool past = false, crossed = false, stop_set = false, done = false;
    double rstop = 0.0;
    int stop = m, scross = 0;
    int s_lim = K.s_lim;
    int s_cur = 2;  // first shell not yet finished (summed and turned into profile entries)
    // lane q: table end of shell s_cur + q
    auto shell_ends = [&](int s0) {
        const int sq = s0 + lane;
        return sq <= m ? a.shell_start[sq - 1] + a.shell_count[sq - 1] : 0x7fffffff;
    };
    int ends_v = shell_ends(s_cur);
    int cA = a.shell_start[1];
    int cB = min(cA + WCAP,
                 __builtin_amdgcn_readlane(ends_v, min(s_cur + WMAXSH - 1, s_lim) - s_cur));
    WaveFetch F;
    wave_fetch<CHECK>(a, T, K, c0, cA, cB, F);
    S acc;
    acc.zero();
    int nin_sh = 0;
    while (!done) {
        // ---- stage the chunk [cA, cB) ----
#pragma unroll
        for (int u = 0; u < WCAP / 64; ++u) {
            const int cl = lane + 64 * u;
            if (cA + cl < cB) {
                if (CHECK) L.in[cl] = F.ok[u];
                L.rec[2 * cl] = F.r0[u];
                L.rec[2 * cl + 1] = F.r1[u];
                L.lap[cl] = F.lp[u];
                L.mult[0][cl] = (float)F.o[u].x;
                L.mult[1][cl] = (float)F.o[u].y;
                L.mult[2][cl] = (float)F.o[u].z;
            }
        }
        __syncthreads();
        // ---- the next chunk follows from the table alone: fetch it during the sums ----
        const int j = __popcll(__ballot(ends_v <= cB));  // shells ending in this chunk
        const int s_next = s_cur + j;
        const bool more = s_next <= s_lim;
        int cB2 = cB;
        if (more) {
            cB2 = min(cB + WCAP, __builtin_amdgcn_readlane(
                                     ends_v, min(s_next + WMAXSH - 1, s_lim) - s_cur));
            wave_fetch<CHECK>(a, T, K, c0, cB, cB2, F);
        }
        // for the finishing pass, lane q: rad[s_cur - 1 + q] and the r^2 of shell s_cur + q
        const double rad_v = rad[min(s_cur - 1 + lane, m + 1)];
        const int r2_v = a.shell_r2[min(s_cur + lane, m) - 1];
        // ---- sums in table order ----
        int jj = 0;
        int e = __builtin_amdgcn_readlane(ends_v, 0);
        for (int cc = cA; cc < cB; cc += 8) {
            float t[8], lv[8];
            bool ok[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int cl = cc - cA + u;  // < WCAP: WCAP is a multiple of 8
                ok[u] = CHECK ? L.in[cl] != 0 : true;
                const float v = recf[8 * cl + fidx];
                const float prod = v * multf[cl];
                t[u] = msel ? prod : v;
                lv[u] = L.lap[cl];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (cc + u < cB) {
                    acc.add_if(summer && ok[u], t[u]);
                    // lane 22: running lapd sum; variant A stops at the first cell outside
                    oobA = oobA || !ok[u];
                    const bool addlap = lapper && ok[u] && (!varA || !oobA);
                    lap += addlap ? (double)lv[u] : 0.0;
                    lapn += addlap ? 1 : 0;
                    nin_sh += ok[u] ? 1 : 0;
                    if (cc + u + 1 == e) {
                        if (lane < 22) L.shs[jj][lane] = acc.val();
                        if (lane == 22) {
                            L.shs[jj][22] = lap;
                            L.shs[jj][23] = (double)lapn;
                        }
                        if (lane == 0) L.shn[jj] = nin_sh;
                        ++jj;
                        acc.zero();
                        nin_sh = 0;
                        e = __builtin_amdgcn_readlane(ends_v, jj);
                    }
                }
            }
        }
        __syncthreads();
        // ---- finish the shells that ended in this chunk ----
        for (int q = 0; q < j && !done; ++q) {
            const int s = s_cur + q;
            const double radP = rl_double(rad_v, q), radS = rl_double(rad_v, q + 1),
                         radN = rl_double(rad_v, q + 2);
            const int r2s = __builtin_amdgcn_readlane(r2_v, q);
            const double* sh = L.shs[q];
            const int nin = L.shn[q];
            const double dFbar = nin > 0 ? sh[0] / (double)nin : 0.0;
            const double rad3p = cube(radP), rad3 = cube(radS);
            const double Fb = (rad3p * Fprev + 0.5 * (dFbarp + dFbar) * (rad3 - rad3p)) / rad3;
            Fprev = Fb;
            dFbarp = dFbar;
            if (nin > 0) cumN += (double)nin;
            if (lane < 22) {
                double x;
                if (lane < 3) {
                    x = sh[7 + lane] / radS;
                } else if (lane < 6) {
                    x = sh[10 + lane - 3] / radS;
                } else if (lane < 12) {
                    // SR + SR^T: r00+r00, r01+r10, r02+r20, r11+r11, r12+r21, r22+r22
                    const int ia = lane == 6 ? 13 : lane == 7 ? 14 : lane == 8 ? 15
                                 : lane == 9 ? 17 : lane == 10 ? 18 : 21;
                    const int ib = lane == 6 ? 13 : lane == 7 ? 16 : lane == 8 ? 19
                                 : lane == 9 ? 17 : lane == 10 ? 20 : 21;
                    x = sh[ia] / radS + sh[ib] / radS;
                } else if (lane == WQ_F) {
                    x = Fb;
                } else if (lane == WQ_N) {
                    x = cumN;
                } else if (lane < WQ_LAP) {
                    if (nin > 0) run += sh[lane < WQ_S2 ? 1 + lane - WQ_S : 4 + lane - WQ_S2];
                    x = run;
                } else {
                    x = sh[lane == WQ_LAP ? 22 : 23];
                }
                P[(int64_t)s * PW + wave_pos(lane)] = x;
            }
            if (!past && r2s > K.ir2min) past = true;
            if (past && !crossed && Fb < fcrit) {
                crossed = true;
                scross = s;
                rstop = radN + 2.0;
            }
            if (!stop_set && crossed && radN >= rstop && s >= K.mupp_rf) {
                stop = s;
                stop_set = true;
                if (a.early) done = true;
            }
        }
        s_cur = s_next;
        if (s_cur > m) done = true;
        if (done) break;
        ends_v = shell_ends(s_cur);
        if (!more) {
            // a stop hint that fell short only bounded the chunks: go on to the last shell
            s_lim = m;
            cB2 = min(cB + WCAP,
                      __builtin_amdgcn_readlane(ends_v, min(s_cur + WMAXSH - 1, s_lim) - s_cur));
            wave_fetch<CHECK>(a, T, K, c0, cB, cB2, F);
        }
        cA = cB;
        cB = cB2;
        __syncthreads();
    }
    if (lane == 0) {
        if (T.res.nshells_used) T.res.nshells_used[p] = stop;
        if (T.res.crossing_shell) T.res.crossing_shell[p] = scross;
    }
}

template <class S>
__global__ __launch_bounds__(64) void shell_wave_kernel(SAArgs a) {
    __shared__ WaveLds L;
    double* Pbase = a.prof + (int64_t)blockIdx.x * 64 * a.slot_doubles;
    const int lane = threadIdx.x;
    int q = blockIdx.x & 7;
    for (int tried = 0; tried < 8;) {
        unsigned long long g = 0;
        if (lane == 0) g = atomicAdd(&a.qctr[q], 1ull);
        g = __shfl(g, 0);
        const int64_t grp = a.qbeg[q] + (int64_t)g;
        if (grp >= a.qend[q]) {
            ++tried;
            q = (q + 1) & 7;
            continue;
        }
        const TileDev& T = a.tiles[a.gtile[grp]];
        const int R = a.rmax, m = a.NS - 1;
        // every lane prepares its own peak
        const int pl = a.slot[grp * 64 + lane];
        WavePeak mine{};
        int mrf = 1, mlow_rf = 1, inside = 1;
        double ZZon = 0.0;
        if (pl >= 0) {
            mine.ci = T.pi[pl];
            mine.cj = T.pj[pl];
            mine.ck = T.pk[pl];
            mine.ir2min = T.ir2min[pl];
            mine.fcrit = T.fcrit ? T.fcrit[pl] : T.fcrit_s;
            ZZon = T.ZZon ? T.ZZon[pl] : T.ZZon_s;
            rf_window(a, T.Rf[pl], mrf, mlow_rf, mine.mupp_rf);
            mine.s_lim = m;
            if (a.early && a.stop_hint) mine.s_lim = min(max(a.stop_hint[T.base + pl], 2), m);
            const int64_t c0 = (int64_t)(mine.ci - 1) + (mine.cj - 1) * T.s1 + (mine.ck - 1) * T.s2;
            mine.v0c = T.recs[2 * c0];
            mine.v1c = T.recs[2 * c0 + 1];
            mine.lapc = T.has_lapd ? T.lapd[c0] : 0.0f;
            inside = mine.ci - R >= 1 && mine.ci + R <= T.n && mine.cj - R >= 1 &&
                     mine.cj + R <= T.n && mine.ck - R >= 1 && mine.ck + R <= T.n;
        }
        for (int idx = 0; idx < 64; ++idx) {
            const int p = __builtin_amdgcn_readlane(pl, idx);
            if (p < 0) continue;
            WavePeak K;
            K.ci = __builtin_amdgcn_readlane(mine.ci, idx);
            K.cj = __builtin_amdgcn_readlane(mine.cj, idx);
            K.ck = __builtin_amdgcn_readlane(mine.ck, idx);
            K.ir2min = __builtin_amdgcn_readlane(mine.ir2min, idx);
            K.mupp_rf = __builtin_amdgcn_readlane(mine.mupp_rf, idx);
            K.s_lim = __builtin_amdgcn_readlane(mine.s_lim, idx);
            K.fcrit = rl_double(mine.fcrit, idx);
            K.v0c = rl_float4(mine.v0c, idx);
            K.v1c = rl_float4(mine.v1c, idx);
            K.lapc = rl_float(mine.lapc, idx);
            double* P = Pbase + (int64_t)idx * a.slot_doubles;
            if (__builtin_amdgcn_readlane(inside, idx)) wave_gather<S, false>(a, T, p, K, P, L);
            else wave_gather<S, true>(a, T, p, K, P, L);
        }
        __syncthreads();  // profiles written by the whole wavefront, read per lane below
        if (pl >= 0)
            post_process(a, T, pl, Pbase + (int64_t)lane * a.slot_doubles, mine.ci, mine.cj,
                         mine.ck, ZZon, mine.fcrit, mine.ir2min, mrf, mlow_rf, mine.mupp_rf);
        __syncthreads();
    }
}

// ------------------------------------------------------------------ crossing pass
// Work of a peak = shells gathered, fixed by its crossing shell s_c (first shell past ir2min
// with Fbar < fcrit) and its filter-scale window.  This pass finds s_c from delta alone,
// with the main gather's arithmetic (same cell order, same FP64 sums and Fbar recurrence),
// so its stop shell is the one the main gather will reach; the main gather then runs over
// peaks ordered by that stop shell and a wavefront's 64 lanes finish together.
// Workgroup b takes group (b % 8) * G8 + b / 8: each XCD walks its own contiguous eighth of
// the (ir2min, Morton) order.
// Sum of delta over shell cells [st, st + nc) of the peak at (ci, cj, ck), whose cell is c0,
// read from the caller's field; cells outside the tile are skipped (CHECK) and nin counts
// the cells summed.  The reference's cell order, FP64 sum of the Float32 values.
template <bool CHECK>
__device__ __forceinline__ void cross_shell_hbm(const SAArgs& a, const TileDev& T, int st, int nc,
                                                int ci, int cj, int ck, int64_t c0, double& F,
                                                int64_t& nin) {
    const float* __restrict__ delta = T.delta;
    F = 0.0;
    nin = 0;
    if (!CHECK) {
        int c = 0;
        // eight loads in flight per step, then four (the 1024^3 pass, with the stage's sort:
        // four at a time 27.3 ms, eight 26.0, eight then four 25.7, sixteen 30.5)
        for (; c + 8 <= nc; c += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int4 o = a.cells[st + c + u];
                v[u] = delta[c0 + o.x + o.y * T.s1 + o.z * T.s2];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) F += (double)v[u];
        }
        for (; c + 4 <= nc; c += 4) {
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int4 o = a.cells[st + c + u];
                v[u] = delta[c0 + o.x + o.y * T.s1 + o.z * T.s2];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) F += (double)v[u];
        }
        for (; c < nc; ++c) {
            const int4 o = a.cells[st + c];
            F += (double)delta[c0 + o.x + o.y * T.s1 + o.z * T.s2];
        }
        nin = nc;
    } else {
        for (int c = 0; c < nc; ++c) {
            const int4 o = a.cells[st + c];
            const int i1 = ci + o.x, i2 = cj + o.y, i3 = ck + o.z;
            if (i1 < 1 || i1 > T.n || i2 < 1 || i2 > T.n || i3 < 1 || i3 > T.n) continue;
            F += (double)delta[c0 + o.x + o.y * T.s1 + o.z * T.s2];
            nin += 1;
        }
    }
}

// One step of the Fbar recurrence (RadialShell.jl:410-416) and the crossing test: shell s with
// delta sum F over nin cells; sets scross and clears active at the first shell past ir2min
// whose Fbar is below fcrit.
__device__ __forceinline__ void cross_step(const SAArgs& a, int s, double F, int64_t nin,
                                           int ir2min, double fcrit, double& Fprev,
                                           double& dFbarp, bool& past, bool& active,
                                           int& scross) {
    const double dFbar = nin > 0 ? F / (double)nin : 0.0;
    const double rad3p = cube(a.rad[s - 1]), rad3 = cube(a.rad[s]);
    const double Fb = (rad3p * Fprev + 0.5 * (dFbarp + dFbar) * (rad3 - rad3p)) / rad3;
    Fprev = Fb;
    dFbarp = dFbar;
    if (!past && a.shell_r2[s - 1] > ir2min) past = true;
    if (past && Fb < fcrit) {
        scross = s;
        active = false;
    }
}

// The gather extent of a peak whose crossing shell is scross (0: none within the hunt
// radius): the main gather stops at the first s >= max(s_c, mupp_rf) with
// rad[s+1] >= rad[s_c+1] + 2, else at the last shell.
__device__ int crossing_stop(const SAArgs& a, double Rf, int scross) {
    const int m = a.NS - 1;
    int stop = m;
    if (scross > 0) {
        int mrf, mlow_rf, mupp_rf;
        rf_window(a, Rf, mrf, mlow_rf, mupp_rf);
        const int s0 = scross > mupp_rf ? scross : mupp_rf;
        if (s0 <= m) {
            // first s in [s0, m] with rad[s + 1] >= rstop, i.e. first shell index t = s + 1
            // in [s0 + 1, m + 1] with rad[t] >= rstop
            const int t = first_ge(a.rad, s0 + 1, m + 1, a.rad[scross + 1] + 2.0);
            if (t <= m + 1) stop = t - 1;
        }
    }
    return stop;
}

// The crossing pass, one lane per peak of group g of the first order: delta only, the same
// arithmetic as the gather, so each peak's exact gather extent.  Workgroup b takes group
// (b % 8) * G8 + b / 8, G8 an eighth of the groups: each XCD walks its own contiguous eighth.
__global__ __launch_bounds__(64) void crossing_kernel(SAArgs a, int64_t G8,
                                                      int32_t* __restrict__ stop_out) {
    const int64_t g = (int64_t)(blockIdx.x & 7) * G8 + (blockIdx.x >> 3);
    if (g >= a.ngroups) return;
    const TileDev& T = a.tiles[a.gtile[g]];
    const int p = a.slot[g * 64 + threadIdx.x];
    const bool valid = p >= 0;
    int ci = 1, cj = 1, ck = 1, ir2min = 0;
    double fcrit = 0.0, Rf = 0.0;
    if (valid) {
        ci = T.pi[p];
        cj = T.pj[p];
        ck = T.pk[p];
        ir2min = T.ir2min[p];
        Rf = T.Rf[p];
        fcrit = T.fcrit ? T.fcrit[p] : T.fcrit_s;
    }
    const int m = a.NS - 1;
    const int64_t c0 = (int64_t)(ci - 1) + (cj - 1) * T.s1 + (ck - 1) * T.s2;
    const int R = a.rmax;
    const bool inside = ci - R >= 1 && ci + R <= T.n && cj - R >= 1 && cj + R <= T.n &&
                        ck - R >= 1 && ck + R <= T.n;
    const bool wave_inside = __ballot(!valid || inside) == ~0ull;

    double Fprev = valid ? (double)T.delta[c0] : 0.0;
    double dFbarp = Fprev;
    bool active = valid, past = false;
    int scross = 0;
    for (int s = 2; s <= m; ++s) {
        if (__ballot(active) == 0ull) break;
        const int st = a.shell_start[s - 1], nc = a.shell_count[s - 1];
        if (active) {
            double F;
            int64_t nin;
            if (wave_inside) cross_shell_hbm<false>(a, T, st, nc, ci, cj, ck, c0, F, nin);
            else cross_shell_hbm<true>(a, T, st, nc, ci, cj, ck, c0, F, nin);
            cross_step(a, s, F, nin, ir2min, fcrit, Fprev, dFbarp, past, active, scross);
        }
    }
    if (!valid) return;
    stop_out[T.base + p] = crossing_stop(a, Rf, scross);
}

// Bracket-network features of every peak in schedule order (same grouping and XCD mapping
// as the crossing pass), read from the cell records; feats[(base + p) * 18 + j].
__global__ __launch_bounds__(64) void features_sched_kernel(SAArgs a, int64_t G8,
                                                            float* __restrict__ feats) {
    const int64_t g = (int64_t)(blockIdx.x & 7) * G8 + (blockIdx.x >> 3);
    if (g >= a.ngroups) return;
    const TileDev& T = a.tiles[a.gtile[g]];
    const int p = a.slot[g * 64 + threadIdx.x];
    if (p < 0) return;
    const int ci = T.pi[p], cj = T.pj[p], ck = T.pk[p];
    const int64_t c0 = (int64_t)(ci - 1) + (cj - 1) * T.s1 + (ck - 1) * T.s2;
    const float* __restrict__ delta = T.delta;
    auto load = [&](int64_t off) -> float { return delta[off]; };
    peak_features(a.cells, a.shell_start, a.shell_count, ci, cj, ck, c0, T.n, T.s1, T.s2,
                  T.Rf[p] / a.alatt, load, feats + (T.base + p) * HNT_MLP_INPUTS);
}

// The bracket's hi, the predicted gather extent (network output 1 = log(1 + shells)),
// clamped to [1, m].
__global__ void pred_stop_kernel(const float* __restrict__ pred, int64_t np, int m,
                                 int32_t* __restrict__ stop) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    float s = expf(pred[q * HNT_MLP_OUTPUTS + 1]) - 1.0f;
    if (!(s >= 1.0f)) s = 1.0f;  // NaN too
    if (s > (float)m) s = (float)m;
    stop[q] = (int32_t)(s + 0.5f);
}



template <class S>
__global__ __launch_bounds__(64) void shell_analyse_kernel(SAArgs a) {
    double* P = a.prof + ((int64_t)blockIdx.x * 64 + threadIdx.x) * a.slot_doubles;
    int q = blockIdx.x & 7;  // workgroups are dispatched round-robin over the 8 XCDs
    for (int tried = 0; tried < 8;) {
        unsigned long long g = 0;
        if (threadIdx.x == 0) g = atomicAdd(&a.qctr[q], 1ull);
        g = __shfl(g, 0);
        const int64_t grp = a.qbeg[q] + (int64_t)g;
        if (grp >= a.qend[q]) {
            ++tried;
            q = (q + 1) & 7;
            continue;
        }
        analyse_group<S>(a, grp, P);
    }
}

}  // namespace

using namespace hnt;

namespace {

hnt_status check_tile(const hnt_fields* f, const hnt_batch* b, const hnt_results* r,
                      const hnt_shell_tables* tables, int t) {
    HNT_REQUIRE(f && b && r, "hnt_shell_analyse: NULL argument (tile %d)", t);
    const int n = f->n1;
    HNT_REQUIRE(f->n1 == f->n2 && f->n2 == f->n3 && n >= 1 && n <= 8192,
                "hnt_shell_analyse: the transposed read needs a cubic grid, got %d x %d x %d "
                "(tile %d)", f->n1, f->n2, f->n3, t);
    HNT_REQUIRE(f->delta && f->etax && f->etay && f->etaz,
                "hnt_shell_analyse: delta and eta fields are required (tile %d)", t);
    HNT_REQUIRE((f->eta2x == nullptr) == (f->eta2y == nullptr) &&
                    (f->eta2y == nullptr) == (f->eta2z == nullptr),
                "hnt_shell_analyse: eta2x/y/z must be all set or all NULL (tile %d)", t);
    HNT_REQUIRE(b->npeaks >= 0 && b->npeaks < (1LL << 31), "hnt_shell_analyse: npeaks");
    if (b->npeaks > 0) {
        HNT_REQUIRE(b->pi && b->pj && b->pk && b->Rf && b->ir2min,
                    "hnt_shell_analyse: NULL peak array (tile %d)", t);
        HNT_REQUIRE(r->RTHL && r->Fbarx && r->e_v && r->p_v && r->strain && r->eigs && r->Srb &&
                        r->Sbar && r->Sbar2 && r->gradpk && r->gradpkf && r->gradpkrf &&
                        r->d2F && r->zvir_half,
                    "hnt_shell_analyse: NULL result array (tile %d)", t);
        HNT_REQUIRE(b->ZZon != nullptr || b->ZZon_scalar > -1.0,
                    "hnt_shell_analyse: ZZon must be > -1");
        HNT_REQUIRE(tables->rmax < n,
                    "hnt_shell_analyse: shell radius %d not below the grid size %d (tile %d)",
                    tables->rmax, n, t);
    }
    return HNT_OK;
}

// Sorts the keys' low end_bit bits (the bits tile_key_kernel can set for the call's tiles).
void sort_pairs(SAWork& W, size_t tmp_bytes, int64_t ntot, int end_bit, hipStream_t st,
                hipError_t* e) {
    *e = hipcub::DeviceRadixSort::SortPairs(W.cub_tmp.p, tmp_bytes, (uint64_t*)W.keys_a.p,
                                            (uint64_t*)W.keys_b.p, (int32_t*)W.vals_a.p,
                                            (int32_t*)W.vals_b.p, (int)ntot, 0, end_bit, st);
}

// Variant B with rmax2rs > 0 skips a peak whose cell the mask holds beyond
// floor(Rf / alatt * rmax2rs) (ShellAnalysisGPU.jl:112-117).  The flags are taken from the
// mask as the call finds it, before any peak of the call marks its halo.
// The mask of the halos of group g (a wavefront per group of 64 schedule slots, workgroup b
// taking group (b % 8) * G8 + b / 8): for each halo in turn, with r2 the r^2 of its last shell
// within RTHL, every cell at d^2 <= r2 from the peak is set to 1 if it lies in the tile's
// interior (nbuff < index <= n - nbuff on every axis), as the reference marks the cells of
// shells 1 .. sR (RadialShell.jl:773-781, ShellAnalysisGPU.jl:512-527).  A lane takes one
// row (dj, dk) of the ball and writes its run of |di| <= sqrt(r2 - dj^2 - dk^2) bytes, so a
// store touches a row's one or two lines instead of one line per cell of 64 halos.
__device__ __forceinline__ int isqrt_floor(int v) {
    int h = (int)sqrtf((float)v);
    while (h * h > v) --h;
    while ((h + 1) * (h + 1) <= v) ++h;
    return h;
}

__global__ __launch_bounds__(64) void paint_mask_kernel(SAArgs a, int64_t G8) {
    const int64_t g = (int64_t)(blockIdx.x & 7) * G8 + (blockIdx.x >> 3);
    if (g >= a.ngroups) return;
    const TileDev& T = a.tiles[a.gtile[g]];
    if (!T.mask) return;
    const int lane = threadIdx.x;
    const int myp = a.slot[g * 64 + lane];
    const int myr2 = myp >= 0 ? a.paint[T.base + myp] : -1;
    // every lane's peak cell at once, read out per halo below
    const int mi = myr2 >= 0 ? T.pi[myp] : 0, mj = myr2 >= 0 ? T.pj[myp] : 0;
    const int mk = myr2 >= 0 ? T.pk[myp] : 0;
    const int n = T.n, nb = a.nbuff;
    for (int k = 0; k < 64; ++k) {
        const int r2 = __builtin_amdgcn_readlane(myr2, k);
        if (r2 < 0) continue;
        const int ci = __builtin_amdgcn_readlane(mi, k), cj = __builtin_amdgcn_readlane(mj, k);
        const int ck = __builtin_amdgcn_readlane(mk, k);
        const int r = isqrt_floor(r2), side = 2 * r + 1, nrows = side * side;
        for (int q0 = 0; q0 < nrows; q0 += 64) {
            const int q = q0 + lane;
            const int dj = q % side - r, dk = q / side - r;
            const int rem = r2 - dj * dj - dk * dk;
            const int j2 = cj + dj, k3 = ck + dk;
            if (q >= nrows || rem < 0 || j2 <= nb || j2 > n - nb || k3 <= nb || k3 > n - nb)
                continue;
            const int h = isqrt_floor(rem);
            const int lo = max(ci - h, nb + 1), hi = min(ci + h, n - nb);
            int8_t* row = T.mask + (int64_t)(j2 - 1) * T.s1 + (int64_t)(k3 - 1) * T.s2 - 1;
            for (int i = lo; i <= hi; ++i) row[i] = 1;
        }
    }
}

__global__ void premask_kernel(const int32_t* __restrict__ pi, const int32_t* __restrict__ pj,
                               const int32_t* __restrict__ pk, const double* __restrict__ Rf,
                               int64_t npeaks, int n, const int8_t* __restrict__ mask,
                               double rmax2rs, double alatt, int8_t* __restrict__ skip) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npeaks) return;
    const int64_t c = (int64_t)(pi[q] - 1) + (int64_t)(pj[q] - 1) * n +
                      (int64_t)(pk[q] - 1) * n * (int64_t)n;
    skip[q] = (double)mask[c] > floor(Rf[q] / alatt * rmax2rs) ? 1 : 0;
}

// What a replay of an analysis launches: the record repacks, the pre-mask flags and the
// main kernel.
struct RepackJob {
    const float *d, *ex, *ey, *ez, *e2x, *e2y, *e2z;
    int n;
    float4* recs;
};
struct PremaskJob {
    const int32_t *pi, *pj, *pk;
    const double* Rf;
    int64_t npeaks;
    int n;
    const int8_t* mask;
    int8_t* skip;
};
struct LaunchDesc {
    SAArgs A;
    int64_t nblocks = 0;
    int gang = 1;  // lanes per peak: 1, 16 or 64 (one wavefront)
    bool kahan = false;
    std::vector<RepackJob> repacks;
    std::vector<PremaskJob> premasks;
    int64_t ntot = 0;  // peaks of the call (the paint radii)
};

void launch_premasks(hipStream_t st, const LaunchDesc& d) {
    if (d.A.paint)
        hipLaunchKernelGGL(fill_i32, dim3((unsigned)((d.ntot + 255) / 256)), dim3(256), 0, st,
                           d.A.paint, d.ntot, -1);
    for (const PremaskJob& j : d.premasks)
        hipLaunchKernelGGL(premask_kernel, dim3((unsigned)((j.npeaks + 255) / 256)), dim3(256),
                           0, st, j.pi, j.pj, j.pk, j.Rf, j.npeaks, j.n, j.mask, d.A.rmax2rs,
                           d.A.alatt, j.skip);
}

void launch_repacks(hipStream_t st, const std::vector<RepackJob>& jobs) {
    for (const RepackJob& j : jobs) {
        dim3 g((j.n + 63) / 64, (j.n + 63) / 64, j.n);
        hipLaunchKernelGGL(repack_kernel, g, dim3(256), 0, st, j.d, j.ex, j.ey, j.ez, j.e2x,
                           j.e2y, j.e2z, j.n, j.recs);
    }
}

void launch_paint(hipStream_t st, const LaunchDesc& d) {
    if (!d.A.paint) return;
    const int64_t G8 = (d.A.ngroups + 7) / 8;
    hipLaunchKernelGGL(paint_mask_kernel, dim3((unsigned)(8 * G8)), dim3(64), 0, st, d.A, G8);
}

void launch_main(hipStream_t st, const LaunchDesc& d) {
    const dim3 grid((unsigned)d.nblocks), blk(64);
    if (d.gang == 64 && d.kahan)
        hipLaunchKernelGGL(shell_wave_kernel<SumKahan>, grid, blk, 0, st, d.A);
    else if (d.gang == 64)
        hipLaunchKernelGGL(shell_wave_kernel<SumF64>, grid, blk, 0, st, d.A);
    else if (d.gang == 16 && d.kahan)
        hipLaunchKernelGGL(shell_gang_kernel<SumKahan>, grid, blk, 0, st, d.A);
    else if (d.gang == 16)
        hipLaunchKernelGGL(shell_gang_kernel<SumF64>, grid, blk, 0, st, d.A);
    else if (d.kahan)
        hipLaunchKernelGGL(shell_analyse_kernel<SumKahan>, grid, blk, 0, st, d.A);
    else
        hipLaunchKernelGGL(shell_analyse_kernel<SumF64>, grid, blk, 0, st, d.A);
}

// All tiles of a call in one persistent launch: peaks of every tile share the eight queues;
// a wavefront's 64 lanes belong to one tile.  W holds the call's device buffers; desc (if
// not NULL) receives what a replay launches.
hnt_status analyse_tiles(hnt_ctx* ctx, SAWork& W, int ntiles, const hnt_fields* fields,
                         const hnt_batch* batches, const hnt_shell_tables* tables,
                         const hnt_ctab* ctab, const hnt_params* params, hnt_results* results,
                         int8_t* const* masks, LaunchDesc* desc) {
    HNT_REQUIRE(ntiles >= 1 && ntiles <= 8192, "hnt_shell_analyse: %d tiles", ntiles);
    HNT_REQUIRE(fields && batches && tables && ctab && params && results,
                "hnt_shell_analyse: NULL argument");
    HNT_REQUIRE(params->variant >= 0 && params->variant <= 2, "hnt_shell_analyse: bad variant");
    HNT_REQUIRE(params->order >= HNT_ORDER_CROSSING && params->order <= HNT_ORDER_GIVEN,
                "hnt_shell_analyse: bad order %d", params->order);
    HNT_REQUIRE(params->order != HNT_ORDER_GIVEN || params->extent,
                "hnt_shell_analyse: HNT_ORDER_GIVEN needs hnt_params.extent");
    HNT_REQUIRE(params->precision == HNT_PREC_FP64 || params->precision == HNT_PREC_FP32_KAHAN,
                "hnt_shell_analyse: bad precision %d", params->precision);
    HNT_REQUIRE(params->order != HNT_ORDER_LEARNED ||
                    (params->bracket && params->bracket->params && params->bracket->norm),
                "hnt_shell_analyse: HNT_ORDER_LEARNED needs the bracket network");
    HNT_REQUIRE(params->order != HNT_ORDER_LEARNED || tables->nshells > 12,
                "hnt_shell_analyse: HNT_ORDER_LEARNED needs more than 12 shells");
    HNT_REQUIRE(params->rmax2rs >= 0.0 && params->rmax2rs <= 1e300,
                "hnt_shell_analyse: rmax2rs must be finite and >= 0");
    const bool limited = params->rmax2rs > 0.0;
    if (limited && params->variant == HNT_VARIANT_A) {
        // RadialShell.analyse_peak limits by a cell count (RadialShell.jl), not by r^2
        set_error("hnt_shell_analyse: rmax2rs > 0 with variant A is not supported");
        return HNT_ERR_UNSUPPORTED;
    }
    HNT_REQUIRE(params->alatt > 0.0, "hnt_shell_analyse: alatt must be > 0");
    int64_t ntot = 0, ngroups = 0, ncell = 0;
    std::vector<int64_t> meta(2 * (size_t)ntiles);  // base, then padded slot offset
    for (int t = 0; t < ntiles; ++t) {
        HNT_TRY(check_tile(&fields[t], &batches[t], &results[t], tables, t));
        meta[t] = ntot;
        meta[ntiles + t] = ngroups * 64;
        ntot += batches[t].npeaks;
        ngroups += (batches[t].npeaks + 63) / 64;
        ncell += (int64_t)fields[t].n1 * fields[t].n1 * fields[t].n1;
    }
    HNT_REQUIRE(ntot < (1LL << 31), "hnt_shell_analyse: %lld peaks in one call",
                (long long)ntot);
    TablesDev& TT = W.own_tables ? W.tables : ctx->tables;
    CTabDev& CT = W.own_tables ? W.ctab : ctx->ctab;
    DevBuf& AKK = W.own_tables ? W.akk : ctx->akk;
    HNT_TRY(upload_tables_to(ctx, TT, tables, W.arena));
    HNT_TRY(upload_ctab_to(ctx, CT, ctab, W.arena));
    HNT_TRY(ensure_akk_to(ctx, AKK, W.arena));
    if (ntot == 0) return HNT_OK;
    hipStream_t st = ctx->stream;

    // ---- validate the peak cells and ZZon on the device before any gather ----
    stage_begin(ctx, HNT_STAGE_SA_PREP);
    HNT_TRY(reserve_in(ctx, W.flag, 64, W.arena));
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, st, (unsigned long long*)W.flag.p, 1,
                       0ull);
    for (int t = 0; t < ntiles; ++t) {
        const hnt_batch& b = batches[t];
        if (b.npeaks == 0) continue;
        hipLaunchKernelGGL(validate_kernel, dim3((unsigned)((b.npeaks + 255) / 256)), dim3(256),
                           0, st, b.pi, b.pj, b.pk, b.ZZon, b.npeaks, fields[t].n1,
                           (unsigned int*)W.flag.p);
    }
    {
        unsigned long long bad = 0;
        HNT_HIP(hipMemcpyAsync(&bad, W.flag.p, sizeof(bad), hipMemcpyDeviceToHost, st));
        HNT_HIP(hipStreamSynchronize(st));
        HNT_REQUIRE(!(bad & 1ull), "hnt_shell_analyse: a peak lies outside its grid");
        HNT_REQUIRE(!(bad & 2ull), "hnt_shell_analyse: per-peak ZZon must be > -1");
    }

    // ---- repack every tile's fields into 32-byte cell records ----
    HNT_TRY(reserve_in(ctx, W.recs, (size_t)ncell * 32, W.arena));
    std::vector<TileDev>& TD = W.tile_host;
    TD.assign((size_t)ntiles, TileDev{});
    std::vector<RepackJob> jobs;
    {
        int64_t off = 0;
        for (int t = 0; t < ntiles; ++t) {
            const hnt_fields& f = fields[t];
            const int n = f.n1;
            float4* recs = (float4*)W.recs.p + 2 * off;
            if (batches[t].npeaks > 0)
                jobs.push_back(RepackJob{f.delta, f.etax, f.etay, f.etaz, f.eta2x, f.eta2y,
                                         f.eta2z, n, recs});
            TileDev& T = TD[t];
            T.recs = recs;
            T.delta = f.delta;
            T.lapd = f.lapd;
            T.n = n;
            T.has_eta2 = f.eta2x != nullptr;
            T.has_lapd = f.lapd != nullptr;
            T.s1 = n;
            T.s2 = (int64_t)n * n;
            const hnt_batch& b = batches[t];
            T.pi = b.pi;
            T.pj = b.pj;
            T.pk = b.pk;
            T.ir2min = b.ir2min;
            T.Rf = b.Rf;
            T.ZZon = b.ZZon;
            T.fcrit = b.fcrit;
            T.ZZon_s = b.ZZon_scalar;
            T.fcrit_s = b.fcrit_scalar;
            T.mask = masks ? masks[t] : nullptr;
            T.res = results[t];
            T.npeaks = b.npeaks;
            T.base = meta[t];
            off += (int64_t)n * n * n;
        }
    }
    launch_repacks(st, jobs);
    HNT_TRY(reserve_in(ctx, W.tiles, TD.size() * sizeof(TileDev), W.arena));
    HNT_TRY(reserve_in(ctx, W.tilemeta, meta.size() * sizeof(int64_t), W.arena));
    W.meta_host = meta;
    HNT_HIP(hipMemcpyAsync(W.tiles.p, TD.data(), TD.size() * sizeof(TileDev),
                           hipMemcpyHostToDevice, st));
    HNT_HIP(hipMemcpyAsync(W.tilemeta.p, W.meta_host.data(), meta.size() * sizeof(int64_t),
                           hipMemcpyHostToDevice, st));
    const int64_t* tile_base = (const int64_t*)W.tilemeta.p;
    const int64_t* tile_pad = tile_base + ntiles;

    // ---- first order: (tile, ir2min, Morton) ----
    HNT_TRY(reserve_in(ctx, W.keys_a, (size_t)ntot * 8, W.arena));
    HNT_TRY(reserve_in(ctx, W.keys_b, (size_t)ntot * 8, W.arena));
    HNT_TRY(reserve_in(ctx, W.vals_a, (size_t)ntot * 4, W.arena));
    HNT_TRY(reserve_in(ctx, W.vals_b, (size_t)ntot * 4, W.arena));
    HNT_TRY(reserve_in(ctx, W.slots, (size_t)ngroups * 64 * 4, W.arena));
    HNT_TRY(reserve_in(ctx, W.gtiles, (size_t)ngroups * 4, W.arena));
    int32_t* slot = (int32_t*)W.slots.p;
    size_t tmp_bytes = 0;
    HNT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (uint64_t*)W.keys_a.p,
                                               (uint64_t*)W.keys_b.p, (int32_t*)W.vals_a.p,
                                               (int32_t*)W.vals_b.p, (int)ntot, 0, 64, st));
    HNT_TRY(reserve_in(ctx, W.cub_tmp, tmp_bytes > 0 ? tmp_bytes : 16, W.arena));
    // key bits in use: Morton codes of coordinates below 2^cb (3 cb bits), the 11-bit primary
    // above the 3 L bits of a brick's cells, and the tile index from bit 50 when the call has
    // several tiles
    int cb = 0, tb = 0;
    {
        int maxn = 1;
        for (int t = 0; t < ntiles; ++t) maxn = std::max(maxn, fields[t].n1);
        while ((1 << cb) < maxn) ++cb;
        while ((1LL << tb) < (int64_t)ntiles) ++tb;
    }
    auto order_by = [&](const int32_t* const* prim, int L) -> hnt_status {
        const int key_bits = ntiles > 1 ? std::min(64, 50 + tb)
                                        : std::min(50, std::max(3 * cb, 3 * L) + 11);
        for (int t = 0; t < ntiles; ++t) {
            const hnt_batch& b = batches[t];
            if (b.npeaks == 0) continue;
            hipLaunchKernelGGL(tile_key_kernel, dim3((unsigned)((b.npeaks + 255) / 256)),
                               dim3(256), 0, st, b.pi, b.pj, b.pk, prim[t], b.npeaks,
                               (uint64_t)t, meta[t], L, (uint64_t*)W.keys_a.p,
                               (int32_t*)W.vals_a.p);
        }
        hipError_t e;
        sort_pairs(W, tmp_bytes, ntot, key_bits, st, &e);
        HNT_HIP(e);
        hipLaunchKernelGGL(fill_i32, dim3((unsigned)((ngroups * 64 + 255) / 256)), dim3(256), 0,
                           st, slot, ngroups * 64, -1);
        hipLaunchKernelGGL(slot_kernel, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, st,
                           (const uint64_t*)W.keys_b.p, (const int32_t*)W.vals_b.p, ntot,
                           tile_base, tile_pad, slot);
        return HNT_OK;
    };
    std::vector<const int32_t*> prim((size_t)ntiles);
    for (int t = 0; t < ntiles; ++t) prim[t] = batches[t].ir2min;
    // first order: 16^3 bricks when it only feeds the crossing pass (its delta-only reads
    // favour small bricks), else the schedule's brick size
    // with rmax2rs > 0 every peak gathers all shells up to its limit (no early exit), as
    // the reference GPU path (ext/CUDAExt.jl:2235-2236): the schedule pass has nothing to order
    const bool early = params->early_exit && !limited;
    const bool two_orders = early && params->order != HNT_ORDER_PLAIN;
    HNT_TRY(order_by(prim.data(), two_orders ? 4 : ctx->brick_log2));
    hipLaunchKernelGGL(group_tile_kernel, dim3((unsigned)((ngroups + 255) / 256)), dim3(256), 0,
                       st, tile_pad, ntiles, ngroups, (int32_t*)W.gtiles.p);

    // ---- persistent grid and profile scratch ----
    SAArgs A;
    memset(&A, 0, sizeof(A));
    A.tiles = (const TileDev*)W.tiles.p;
    A.slot = slot;
    A.gtile = (const int32_t*)W.gtiles.p;
    A.ngroups = ngroups;
    A.cells = (const int4*)TT.cells.p;
    A.shell_start = (const int32_t*)TT.start.p;
    A.shell_count = (const int32_t*)TT.count.p;
    A.shell_r2 = (const int32_t*)TT.r2.p;
    A.rad = (const double*)TT.rad.p;
    A.kw = (const double*)TT.kw.p;
    A.kidx = (const int32_t*)TT.kw_idx.p;
    A.NS = tables->nshells;
    A.rmax = tables->rmax;
    A.ct = CTabView{(const float*)CT.table.p, ctab->nx, ctab->ny, ctab->nz, ctab->x1,
                    ctab->x2, ctab->y1, ctab->y2, ctab->z1, ctab->z2, ctab->out_val};
    A.akk = (const double*)AKK.p;
    A.alatt = params->alatt;
    // hlatt = 1 (RadialShell.jl:296-300)
    A.aRnor = 3.0 * 1.0 / params->alatt;
    A.wRnor = 3.0 * (1.0 / (4.0 * kPi)) / params->alatt;
    A.nbuff = params->nbuff;
    A.variant = params->variant;
    A.early = early ? 1 : 0;
    A.rmax2rs = params->rmax2rs;

    const bool kahan = params->precision == HNT_PREC_FP32_KAHAN;
    // the ganged kernels analyse every peak to the last shell: a shell limit takes the
    // lane-per-peak kernel
    const int gang = limited ? 1 : ctx->gang;
    std::vector<PremaskJob> premasks;
    if (limited && params->variant == HNT_VARIANT_B && masks) {
        for (int t = 0; t < ntiles; ++t)
            if (masks[t] && batches[t].npeaks > 0) premasks.push_back(PremaskJob{});
        if (!premasks.empty()) {
            HNT_TRY(reserve_in(ctx, W.skip, (size_t)ntot, W.arena));
            int8_t* skip = (int8_t*)W.skip.p;
            hipLaunchKernelGGL(fill_i8, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, st,
                               skip, ntot, (int8_t)0);
            premasks.clear();
            for (int t = 0; t < ntiles; ++t) {
                const hnt_batch& b = batches[t];
                if (!masks[t] || b.npeaks == 0) continue;
                premasks.push_back(PremaskJob{b.pi, b.pj, b.pk, b.Rf, b.npeaks, fields[t].n1,
                                              masks[t], skip + meta[t]});
            }
            A.skip = skip;
        }
    }
    // halos' masks are marked after the analysis, row by row (paint_mask_kernel)
    bool any_mask = false;
    for (int t = 0; t < ntiles; ++t) any_mask = any_mask || (masks && masks[t] && batches[t].npeaks);
    if (any_mask) {
        HNT_TRY(reserve_in(ctx, W.paint, (size_t)ntot * 4, W.arena));
        A.paint = (int32_t*)W.paint.p;
    }
    const void* kern = gang == 64 ? (kahan ? (const void*)shell_wave_kernel<SumKahan>
                                           : (const void*)shell_wave_kernel<SumF64>)
                     : gang == 16 ? (kahan ? (const void*)shell_gang_kernel<SumKahan>
                                           : (const void*)shell_gang_kernel<SumF64>)
                                  : (kahan ? (const void*)shell_analyse_kernel<SumKahan>
                                           : (const void*)shell_analyse_kernel<SumF64>);
    int per_cu = 0;
    HNT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, 0));
    if (per_cu < 1) per_cu = 1;
    if (ctx->waves_per_cu > 0 && ctx->waves_per_cu < per_cu) per_cu = ctx->waves_per_cu;
    A.slot_doubles = (int64_t)(A.NS + 1) * PW;
    const size_t slot_bytes = (size_t)A.slot_doubles * 8;
    HNT_TRY(reserve_in(ctx, W.queue, 8 * sizeof(unsigned long long), W.arena));
    A.qctr = (unsigned long long*)W.queue.p;
    const bool sched = A.early && params->order != HNT_ORDER_PLAIN;
    if (sched && W.arena) HNT_TRY(reserve_in(ctx, W.stops, (size_t)ntot * 4, W.arena));
    // profile scratch: 32 GiB at most, or what is left of a fixed region (fewer resident
    // wavefronts then, same results)
    size_t budget = (size_t)32 << 30;
    if (W.arena) {
        const size_t used = (W.arena->used + 255) & ~(size_t)255;
        budget = used < W.arena->cap ? W.arena->cap - used : 0;
    }
    int64_t nblocks = (int64_t)per_cu * ctx->num_cus;
    int64_t cap_blocks = (int64_t)(budget / (slot_bytes * 64));
    if (nblocks > cap_blocks) nblocks = cap_blocks;
    int64_t want = ((ngroups + 7) / 8) * 8;
    if (nblocks > want) nblocks = want;
    nblocks = (nblocks / 8) * 8;
    if (nblocks < 8) nblocks = 8;
    HNT_TRY(reserve_in(ctx, W.prof, (size_t)nblocks * 64 * slot_bytes, W.arena));
    A.prof = (double*)W.prof.p;
    for (int x = 0; x < 8; ++x) {
        A.qbeg[x] = ngroups * x / 8;
        A.qend[x] = ngroups * (x + 1) / 8;
    }
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, st, A.qctr, 8, 0ull);
    stage_end(ctx, HNT_STAGE_SA_PREP);

    // ---- schedule: order the peaks by their (exact or predicted) gather extent ----
    if (sched) {
        stage_begin(ctx, HNT_STAGE_SA_CROSS);
        if (!W.arena) HNT_TRY(reserve_in(ctx, W.stops, (size_t)ntot * 4, W.arena));
        int32_t* stops = (int32_t*)W.stops.p;
        if (params->order == HNT_ORDER_CROSSING) {
            const int64_t G8 = (ngroups + 7) / 8;
            hipLaunchKernelGGL(crossing_kernel, dim3((unsigned)(8 * G8)), dim3(64), 0, st, A, G8,
                               stops);
        } else if (params->order == HNT_ORDER_GIVEN) {
            // the caller's extents order the peaks; no pass of the library's own
            HNT_HIP(hipMemcpyAsync(stops, params->extent, (size_t)ntot * 4,
                                   hipMemcpyDeviceToDevice, st));
        } else {
            HNT_TRY(reserve_in(ctx, W.feats, (size_t)ntot * (HNT_MLP_INPUTS + HNT_MLP_OUTPUTS) * 4, W.arena));
            float* feats = (float*)W.feats.p;
            float* pred = feats + ntot * HNT_MLP_INPUTS;
            const int64_t G8 = (ngroups + 7) / 8;
            hipLaunchKernelGGL(features_sched_kernel, dim3((unsigned)(8 * G8)), dim3(64), 0, st,
                               A, G8, feats);
            HNT_TRY(mlp_predict(ctx, params->bracket, feats, ntot, pred));
            hipLaunchKernelGGL(pred_stop_kernel, dim3((unsigned)((ntot + 255) / 256)), dim3(256),
                               0, st, pred, ntot, A.NS - 1, stops);
            A.pred = pred;  // the main kernel checks each peak against its bracket
        }
        for (int t = 0; t < ntiles; ++t) prim[t] = stops + meta[t];
        if (params->order == HNT_ORDER_CROSSING) {
            A.stop_hint = stops;
            for (int t = 0; t < ntiles; ++t)
                if (results[t].schedule_extent && batches[t].npeaks > 0)
                    HNT_HIP(hipMemcpyAsync(results[t].schedule_extent, stops + meta[t],
                                           (size_t)batches[t].npeaks * 4,
                                           hipMemcpyDeviceToDevice, st));
        }
        HNT_TRY(order_by(prim.data(), ctx->brick_log2));
        stage_end(ctx, HNT_STAGE_SA_CROSS);
    }

    stage_begin(ctx, HNT_STAGE_SA_KERNEL);
    LaunchDesc local;
    LaunchDesc& D = desc ? *desc : local;
    D.A = A;
    D.nblocks = nblocks;
    D.gang = gang;
    D.kahan = kahan;
    D.repacks = std::move(jobs);
    D.premasks = std::move(premasks);
    D.ntot = ntot;
    launch_premasks(st, D);
    launch_main(st, D);
    stage_end(ctx, HNT_STAGE_SA_KERNEL);
    stage_begin(ctx, HNT_STAGE_SA_MASK);
    launch_paint(st, D);
    stage_end(ctx, HNT_STAGE_SA_MASK);
    return HNT_OK;
}

}  // namespace

extern "C" hnt_status hnt_shell_analyse(hnt_ctx* ctx, const hnt_fields* fields,
                                        const hnt_batch* batch, const hnt_shell_tables* tables,
                                        const hnt_ctab* ctab, const hnt_params* params,
                                        hnt_results* results, int8_t* mask) {
    HNT_TRY(begin_call(ctx));
    HNT_TRY(analyse_tiles(ctx, ctx->sa, 1, fields, batch, tables, ctab, params, results, &mask,
                          nullptr));
    return finish_call(ctx, "hnt_shell_analyse");
}

extern "C" hnt_status hnt_shell_analyse_tiles(hnt_ctx* ctx, int32_t ntiles,
                                              const hnt_fields* fields,
                                              const hnt_batch* batches,
                                              const hnt_shell_tables* tables,
                                              const hnt_ctab* ctab, const hnt_params* params,
                                              hnt_results* results, int8_t* const* masks) {
    HNT_TRY(begin_call(ctx));
    HNT_TRY(analyse_tiles(ctx, ctx->sa, ntiles, fields, batches, tables, ctab, params, results,
                          masks, nullptr));
    return finish_call(ctx, "hnt_shell_analyse_tiles");
}

// ------------------------------------------------------------------ plans (hipGraph replay)

struct hnt_plan {
    hnt_ctx* ctx = nullptr;
    SAWork W;
    LaunchDesc D;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void plan_free(hnt_plan* P) {
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->cap) (void)hipStreamDestroy(P->cap);
    release_work(P->W);
    delete P;
}

static hnt_status plan_build(hnt_plan* P, int32_t ntiles, const hnt_fields* fields,
                             const hnt_batch* batches, const hnt_shell_tables* tables,
                             const hnt_ctab* ctab, const hnt_params* params,
                             hnt_results* results, int8_t* const* masks) {
    hnt_ctx* ctx = P->ctx;
    HNT_TRY(analyse_tiles(ctx, P->W, ntiles, fields, batches, tables, ctab, params, results,
                          masks, &P->D));
    HNT_HIP(hipStreamSynchronize(ctx->stream));
    if (P->D.nblocks == 0) return HNT_OK;  // no peaks: launching the plan does nothing
    HNT_HIP(hipStreamCreateWithFlags(&P->cap, hipStreamNonBlocking));
    HNT_HIP(hipStreamBeginCapture(P->cap, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, P->cap, P->D.A.qctr, 8, 0ull);
    launch_repacks(P->cap, P->D.repacks);
    launch_premasks(P->cap, P->D);
    launch_main(P->cap, P->D);
    launch_paint(P->cap, P->D);
    HNT_HIP(hipStreamEndCapture(P->cap, &P->graph));
    // the graph must hold kernels only (no memset or copy nodes)
    size_t nn = 0;
    HNT_HIP(hipGraphGetNodes(P->graph, nullptr, &nn));
    std::vector<hipGraphNode_t> nodes(nn);
    HNT_HIP(hipGraphGetNodes(P->graph, nodes.data(), &nn));
    for (hipGraphNode_t n : nodes) {
        hipGraphNodeType ty;
        HNT_HIP(hipGraphNodeGetType(n, &ty));
        HNT_REQUIRE(ty == hipGraphNodeTypeKernel, "hnt_plan_create: graph holds a non-kernel node");
    }
    HNT_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    return HNT_OK;
}

extern "C" hnt_status hnt_plan_create(hnt_ctx* ctx, int32_t ntiles, const hnt_fields* fields,
                                      const hnt_batch* batches, const hnt_shell_tables* tables,
                                      const hnt_ctab* ctab, const hnt_params* params,
                                      hnt_results* results, int8_t* const* masks,
                                      hnt_plan** out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(out, "hnt_plan_create: NULL output");
    *out = nullptr;
    hnt_plan* P = new hnt_plan;
    P->ctx = ctx;
    P->W.own_tables = true;  // replays read the plan's copies of the tables
    const hnt_status st = plan_build(P, ntiles, fields, batches, tables, ctab, params, results,
                                     masks);
    if (st != HNT_OK) {
        plan_free(P);
        return st;
    }
    *out = P;
    return finish_call(ctx, "hnt_plan_create");
}

extern "C" hnt_status hnt_plan_launch(hnt_plan* plan) {
    HNT_REQUIRE(plan, "hnt_plan_launch: NULL plan");
    HNT_TRY(begin_call(plan->ctx));
    if (plan->exec) HNT_HIP(hipGraphLaunch(plan->exec, plan->ctx->stream));
    return finish_call(plan->ctx, "hnt_plan_launch");
}

extern "C" hnt_status hnt_plan_destroy(hnt_plan* plan) {
    clear_error();
    if (!plan) return HNT_OK;
    (void)hipSetDevice(plan->ctx->device);
    (void)hipStreamSynchronize(plan->ctx->stream);
    plan_free(plan);
    return HNT_OK;
}

// ------------------------------------------------------------------ fixed-scratch entry point
// hnt_shell_analyse_noalloc runs hnt_shell_analyse with every device buffer -- cell records,
// sort keys, lane slots, tile descriptor, queues, stop shells, profile slots, the shell,
// collapse and SPH kernel tables -- carved from `scratch`, a static device array each GPU
// holds from the library's load to the end of the process.  The host only ever uses the
// array's device address from hipGetSymbolAddress (per device, looked up by
// noalloc_scratch_init from hnt_ctx_create); the array's name on the host is not a device
// address.

namespace {

constexpr size_t kNoallocScratchBytes = (size_t)256 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoallocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

}  // namespace

namespace hnt {

// With `device` current: the address of that device's copy of `scratch`, once per device.
hnt_status noalloc_scratch_init(int device) {
    HNT_REQUIRE(device >= 0 && device < kMaxDevices, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return HNT_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        g_scratch_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        set_error("hipGetSymbolAddress(scratch) on device %d: %s", device,
                  hipGetErrorString(g_scratch_err[device]));
        return HNT_ERR_DEVICE;
    }
    g_scratch_ptr[device] = p;
    return HNT_OK;
}

}  // namespace hnt

extern "C" hnt_status hnt_shell_analyse_noalloc(hnt_ctx* ctx, const hnt_fields* fields,
                                                const hnt_batch* batch,
                                                const hnt_shell_tables* tables,
                                                const hnt_ctab* ctab, const hnt_params* params,
                                                hnt_results* results, int8_t* mask) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(params, "hnt_shell_analyse_noalloc: NULL params");
    if (params->order == HNT_ORDER_LEARNED) {
        set_error("hnt_shell_analyse_noalloc: the learned schedule needs the network's "
                  "scratch; use HNT_ORDER_CROSSING or HNT_ORDER_PLAIN");
        return HNT_ERR_UNSUPPORTED;
    }
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch_ptr[ctx->device];
    }
    if (!base) {
        // the context was created before the lookup succeeded: try once more, device current
        HNT_TRY(noalloc_scratch_init(ctx->device));
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = g_scratch_ptr[ctx->device];
    }
    Arena ar;
    ar.base = (char*)base;
    ar.cap = kNoallocScratchBytes;
    SAWork W;
    W.arena = &ar;
    W.own_tables = true;
    const hnt_status st =
        analyse_tiles(ctx, W, 1, fields, batch, tables, ctab, params, results, &mask, nullptr);
    if (st == HNT_ERR_CAPACITY) {
        char msg[1024];
        hnt_last_error(msg, sizeof(msg));
        set_error("hnt_shell_analyse_noalloc: the call needs more than the library's fixed "
                  "%zu MiB of scratch (%s); use hnt_shell_analyse", kNoallocScratchBytes >> 20,
                  msg);
        return st;
    }
    HNT_TRY(st);
    return finish_call(ctx, "hnt_shell_analyse_noalloc");
}

// peak_find.hip -- multi-filter 26-neighbour peak finding on smoothed density fields.
//
// Replaces the per-filter loop of peak_find_tile_gpu (ext/CUDAExt.jl:2985-3012, kernel
// _find_peaks_kernel! at :2876-2936) and of the CPU drivers (src/MultiResolution.jl:1046-1066
// calling PeakFind.find_peaks, src/HaloFinder/PeakFind.jl:15-78).
//
// Rule restated: the mask is read and written only at the peak cell itself, so processing
// the filters in order with a running mask is the same as "a cell is a peak of the first
// filter (in the given order) where it passes the mask/threshold/local-max test".  That
// makes every cell independent, so one pass decides all filters.
//
// Kernels
//  pf_mark    one workgroup per 64 x 8 x 16 brick of the inner cube.  For each filter the
//             workgroup sweeps the brick along k with a ring of three LDS planes (66 x 10
//             floats each: brick + 1-cell halo), so every field value is read from HBM about
//             once (x1.13 for the k halo, halo rows/columns mostly hit L2).  The claim (first
//             filter index, 255 = none) of each cell is kept in registers until the end.
//  pf_count   per 4096-cell chunk of the claim array (inner-cube scan order k, j, i): peaks
//             per filter.
//  (hipcub exclusive scan over the filter-major count table)
//  pf_scatter writes the peaks of each chunk at their scan offsets: output order is
//             filter, then k, j, i -- the order of the reference CPU loop.
#include <hipcub/hipcub.hpp>

#include "hnt_internal.h"

namespace {

constexpr int TI = 64, TJ = 8, TK = 16;
constexpr int HI = TI + 2, HJ = TJ + 2;
constexpr int CHUNK = 4096;
constexpr uint8_t NONE = 255, BLOCKED = 254;

__global__ __launch_bounds__(256) void pf_mark(const float* const* __restrict__ fields,
                                               const float* __restrict__ fcrit, int nfilt,
                                               int n, int nbuff, const int8_t* __restrict__ mask,
                                               uint8_t* __restrict__ claim) {
    __shared__ float pl[3][HJ][HI];
    const int m = n - 2 * nbuff;
    const int ti = threadIdx.x & 63, tj = threadIdx.x >> 6;  // tj in 0..3, rows tj and tj+4
    const int i0 = blockIdx.x * TI, j0 = blockIdx.y * TJ, k0 = blockIdx.z * TK;
    const int64_t s1 = n, s2 = (int64_t)n * n;

    uint8_t cl[2][TK];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            int ii = i0 + ti, jj = j0 + tj + 4 * c, kq = k0 + kk;
            uint8_t v = BLOCKED;
            if (ii < m && jj < m && kq < m) {
                int64_t a = (int64_t)(ii + nbuff) + (jj + nbuff) * s1 + (kq + nbuff) * s2;
                v = (mask[a] == 1) ? BLOCKED : NONE;
            }
            cl[c][kk] = v;
        }
    }

    for (int f = 0; f < nfilt; ++f) {
        const float* __restrict__ F = fields[f];
        const float fc = fcrit[f];
#pragma unroll
        for (int step = 0; step < TK + 2; ++step) {
            const int p = k0 - 1 + step;  // inner-cube plane index, -1 .. k0+TK
            const int ka = min(p + nbuff, n - 1);
            const int slot = step % 3;
            for (int e = threadIdx.x; e < HJ * HI; e += 256) {
                int jh = e / HI, ih = e - jh * HI;
                int ia = min(i0 - 1 + ih + nbuff, n - 1);
                int ja = min(j0 - 1 + jh + nbuff, n - 1);
                pl[slot][jh][ih] = F[(int64_t)ia + ja * s1 + ka * s2];
            }
            __syncthreads();
            if (step >= 2) {
                const int kk = step - 2;
                const int sb = (step - 2) % 3, sc = (step - 1) % 3, sa = step % 3;
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    if (cl[c][kk] == NONE) {
                        const int jh = tj + 4 * c + 1, ih = ti + 1;
                        const float ff = pl[sc][jh][ih];
                        // the reference skips `ff < fcrit`; a NaN is therefore not skipped
                        if (!(ff < fc)) {
                            bool ismax = true;
#pragma unroll
                            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                                for (int dx = -1; dx <= 1; ++dx) {
                                    ismax &= !(ff < pl[sb][jh + dy][ih + dx]);
                                    ismax &= !(ff < pl[sc][jh + dy][ih + dx]);
                                    ismax &= !(ff < pl[sa][jh + dy][ih + dx]);
                                }
                            if (ismax) cl[c][kk] = (uint8_t)f;
                        }
                    }
                }
            }
            __syncthreads();
        }
    }

    const int64_t mm = m, mm2 = (int64_t)m * m;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            int ii = i0 + ti, jj = j0 + tj + 4 * c, kq = k0 + kk;
            if (ii < m && jj < m && kq < m)
                claim[ii + jj * mm + kq * mm2] = (cl[c][kk] == BLOCKED) ? NONE : cl[c][kk];
        }
    }
}

__global__ __launch_bounds__(256) void pf_count(const uint8_t* __restrict__ claim, int64_t ncell,
                                                int nfilt, int64_t nchunks,
                                                unsigned long long* __restrict__ counts) {
    __shared__ unsigned int c[256];
    c[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * CHUNK;
    for (int e = threadIdx.x; e < CHUNK; e += 256) {
        int64_t idx = base + e;
        if (idx < ncell) {
            uint8_t v = claim[idx];
            if (v != NONE) atomicAdd(&c[v], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < nfilt) counts[(int64_t)threadIdx.x * nchunks + blockIdx.x] = c[threadIdx.x];
}

__global__ __launch_bounds__(256) void pf_scatter(
    const uint8_t* __restrict__ claim, int64_t ncell, int64_t nchunks,
    const unsigned long long* __restrict__ offsets, int n, int nbuff,
    const float* const* __restrict__ fields, int8_t* __restrict__ mask, int32_t* __restrict__ oi,
    int32_t* __restrict__ oj, int32_t* __restrict__ ok, int32_t* __restrict__ of,
    float* __restrict__ od, int64_t capacity) {
    __shared__ int s_scan[256];
    __shared__ int s_cell[CHUNK];
    __shared__ uint8_t s_f[CHUNK];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * CHUNK;
    // each thread owns 16 consecutive cells of the chunk
    int mine = 0;
    uint8_t v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        int64_t idx = base + t * 16 + q;
        v[q] = idx < ncell ? claim[idx] : NONE;
        mine += v[q] != NONE;
    }
    // block exclusive scan of the per-thread counts (Hillis-Steele in LDS)
    s_scan[t] = mine;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        int add = t >= d ? s_scan[t - d] : 0;
        __syncthreads();
        s_scan[t] += add;
        __syncthreads();
    }
    const int total = s_scan[255];
    int pos = s_scan[t] - mine;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        if (v[q] != NONE) {
            s_cell[pos] = t * 16 + q;
            s_f[pos] = v[q];
            ++pos;
        }
    }
    __syncthreads();
    const int m = n - 2 * nbuff;
    const int64_t mm = m, mm2 = (int64_t)m * m;
    const int64_t s1 = n, s2 = (int64_t)n * n;
    for (int e = t; e < total; e += 256) {
        const uint8_t f = s_f[e];
        int rank = 0;
        for (int q = 0; q < e; ++q) rank += (s_f[q] == f);
        const unsigned long long dst = offsets[(int64_t)f * nchunks + blockIdx.x] + rank;
        if (dst < (unsigned long long)capacity) {
            const int64_t idx = base + s_cell[e];
            const int kq = (int)(idx / mm2);
            const int64_t rem = idx - kq * mm2;
            const int jj = (int)(rem / mm);
            const int ii = (int)(rem - jj * mm);
            const int64_t a = (int64_t)(ii + nbuff) + (jj + nbuff) * s1 + (kq + nbuff) * s2;
            oi[dst] = ii + nbuff + 1;
            oj[dst] = jj + nbuff + 1;
            ok[dst] = kq + nbuff + 1;
            of[dst] = f;
            od[dst] = fields[f][a];
            mask[a] = 1;
        }
    }
}

}  // namespace

using namespace hnt;

extern "C" hnt_status hnt_peak_find(hnt_ctx* ctx, const float* const* smoothed,
                                    const float* fcrit, int32_t nfilt, int32_t n, int32_t nbuff,
                                    int8_t* mask, hnt_peaks* out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(smoothed && fcrit && mask && out, "hnt_peak_find: NULL argument");
    HNT_REQUIRE(nfilt >= 1 && nfilt <= 254, "hnt_peak_find: nfilt %d not in 1..254", nfilt);
    HNT_REQUIRE(n >= 3 && n <= 4096, "hnt_peak_find: n %d not in 3..4096", n);
    HNT_REQUIRE(nbuff >= 1, "hnt_peak_find: nbuff must be >= 1 (the 26-neighbour test reads "
                            "one cell beyond the inner cube), got %d", nbuff);
    HNT_REQUIRE(out->capacity >= 0, "hnt_peak_find: negative capacity");
    HNT_REQUIRE(out->capacity == 0 || (out->i && out->j && out->k && out->filt && out->delta),
                "hnt_peak_find: NULL output array");
    for (int f = 0; f < nfilt; ++f)
        HNT_REQUIRE(smoothed[f] != nullptr, "hnt_peak_find: field %d is NULL", f);
    out->count = 0;
    out->total = 0;
    const int m = n - 2 * nbuff;
    if (m <= 0) return HNT_OK;
    hipStream_t st = ctx->stream;
    const int64_t ncell = (int64_t)m * m * m;
    const int64_t nchunks = (ncell + CHUNK - 1) / CHUNK;
    const int64_t ntab = nchunks * nfilt;

    HNT_TRY(reserve(ctx, ctx->fieldptrs, sizeof(float*) * (size_t)nfilt));
    HNT_TRY(reserve(ctx, ctx->fcrit, sizeof(float) * (size_t)nfilt));
    HNT_TRY(reserve(ctx, ctx->claim, (size_t)ncell));
    HNT_TRY(reserve(ctx, ctx->counts, sizeof(unsigned long long) * (size_t)ntab));
    HNT_TRY(reserve(ctx, ctx->offsets, sizeof(unsigned long long) * (size_t)ntab));
    size_t tmp_bytes = 0;
    HNT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes,
                                             (unsigned long long*)ctx->counts.p,
                                             (unsigned long long*)ctx->offsets.p, (int)ntab, st));
    HNT_TRY(reserve(ctx, ctx->cub_tmp, tmp_bytes > 0 ? tmp_bytes : 16));
    HNT_TRY(reserve(ctx, ctx->small, 64));

    HNT_HIP(hipMemcpyAsync(ctx->fieldptrs.p, smoothed, sizeof(float*) * (size_t)nfilt,
                           hipMemcpyHostToDevice, st));
    HNT_HIP(hipMemcpyAsync(ctx->fcrit.p, fcrit, sizeof(float) * (size_t)nfilt,
                           hipMemcpyHostToDevice, st));
    const float* const* dfields = (const float* const*)ctx->fieldptrs.p;

    dim3 g1((m + TI - 1) / TI, (m + TJ - 1) / TJ, (m + TK - 1) / TK);
    stage_begin(ctx, HNT_STAGE_PF_MARK);
    hipLaunchKernelGGL(pf_mark, g1, dim3(256), 0, st, dfields, (const float*)ctx->fcrit.p, nfilt,
                       n, nbuff, (const int8_t*)mask, (uint8_t*)ctx->claim.p);
    stage_end(ctx, HNT_STAGE_PF_MARK);
    stage_begin(ctx, HNT_STAGE_PF_COMPACT);
    hipLaunchKernelGGL(pf_count, dim3((unsigned)nchunks), dim3(256), 0, st,
                       (const uint8_t*)ctx->claim.p, ncell, nfilt, nchunks,
                       (unsigned long long*)ctx->counts.p);
    HNT_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp.p, tmp_bytes,
                                             (unsigned long long*)ctx->counts.p,
                                             (unsigned long long*)ctx->offsets.p, (int)ntab, st));
    hipLaunchKernelGGL(pf_scatter, dim3((unsigned)nchunks), dim3(256), 0, st,
                       (const uint8_t*)ctx->claim.p, ncell, nchunks,
                       (const unsigned long long*)ctx->offsets.p, n, nbuff, dfields, mask, out->i,
                       out->j, out->k, out->filt, out->delta, out->capacity);
    stage_end(ctx, HNT_STAGE_PF_COMPACT);
    unsigned long long tail[2] = {0, 0};
    HNT_HIP(hipMemcpyAsync(&tail[0], (unsigned long long*)ctx->offsets.p + (ntab - 1),
                           sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HNT_HIP(hipMemcpyAsync(&tail[1], (unsigned long long*)ctx->counts.p + (ntab - 1),
                           sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HNT_TRY(finish_call(ctx, "hnt_peak_find"));
    const int64_t total = (int64_t)(tail[0] + tail[1]);
    out->total = total;
    out->count = total < out->capacity ? total : out->capacity;
    if (total > out->capacity) {
        set_error("hnt_peak_find: %lld peaks found, capacity %lld", (long long)total,
                  (long long)out->capacity);
        return HNT_ERR_CAPACITY;
    }
    return HNT_OK;
}

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
//  pf_mark    a wavefront owns 62 i x 8 j x 32 k cells of the inner cube; a workgroup is
//             4 x 2 wavefronts along i and j.  For each filter a wavefront streams the 10
//             rows (its 8 + a 1-cell halo) of each plane with one coalesced 256-byte load
//             per row, one plane ahead of the arithmetic; the plane is reduced to 3x3 maxima
//             (DPP lane shifts along i, then three rows), and the k direction keeps the last
//             planes' maxima in registers (no LDS).  One barrier per plane keeps the
//             workgroup's wavefronts on the same plane, so the cache lines they share (halo
//             columns and rows) come from HBM once: 1.45x the field bytes at 1024^3, against
//             2.1x for free-running wavefronts.  A peak is written to the claim array
//             (filter index; 255 = none) by the one lane owning the cell; a per-lane bitset
//             marks cells masked or already taken by an earlier filter.
//  pf_count   per 4096-cell chunk of the claim array (inner-cube scan order k, j, i): peaks
//             per filter.
//  (hipcub exclusive scan over the filter-major count table)
//  pf_scatter writes the peaks of each chunk at their scan offsets: output order is
//             filter, then k, j, i -- the order of the reference CPU loop.
#include <hipcub/hipcub.hpp>
#include <hipfft/hipfft.h>

#include "hnt_internal.h"

namespace {

// pf_mark geometry: a wavefront owns OI = 62 consecutive i (its 64 lanes load i-1 .. i+62),
// TJ rows of j and TK planes of k; a workgroup is WI x WJ wavefronts along i and j.
constexpr int OI = 62, TJ = 8, TK = 32, WI = 4, WJ = 2;
constexpr int PF_THREADS = 64 * WI * WJ;
constexpr int CHUNK = 4096;
constexpr uint8_t NONE = 255, BLOCKED = 254;  // claim of a cell: filter index, or these
// Word 3 of a raw (stride 0) buffer resource on CDNA: 32-bit data format, no swizzle.
constexpr int BUF_DWORD3 = 0x00020000;
static_assert(TJ * 4 == 32 && TK % 4 == 0, "claim bits: one 32-bit word per 4 planes of TJ=8");

// Left and right neighbour of a lane's value within the wavefront (lanes 0 and 63 receive
// values that are never used: they are halo columns).
// DPP wave_shr:1 / wave_shl:1 (whole-wavefront shifts by one lane, a VALU operand modifier).
__device__ __forceinline__ float from_lane_below(float v) {
    const int x = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(x, x, 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_lane_above(float v) {
    const int x = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(x, x, 0x130, 0xF, 0xF, false));
}

// The 26-neighbour test of the reference, !(ff < x) for every neighbour x and for fcrit, is
// the same as !(ff < max(neighbours)) with IEEE maxNum (fmaxf): a NaN neighbour drops out of
// the max exactly as !(ff < NaN) drops out of the conjunction, and an all-NaN max passes.
// So each plane is reduced once to its 3x3 maxima (row max through lane shuffles, then over
// three rows) and the k direction keeps the last three planes' maxima in registers.
// One launch runs filters fbase .. fbase + nfilt - 1 (fields[f] is filter fbase + f).  The
// cells not to claim come from the mask (the caller cleared the claim array), or, in the
// TILE instance with from_claim = 1, from the bitset `takenw` that carries the taken bits
// from one launch to the next (one 32-bit word per lane column, 8 rows and 4 planes, as the
// registers hold them: masked cells, and peaks of earlier launches); a hit there reads
// nothing, so no wavefront waits on a load in the claim path.  A claimed cell also records
// its field value in cellval (and, with celld2, the value of lapfields[f]), indexed like
// the claim array.
template <bool TILE>
__global__ __launch_bounds__(PF_THREADS) void pf_mark(const float* const* __restrict__ fields,
                                               const float* __restrict__ fcrit, int nfilt,
                                               int fbase, int n, int nbuff, int fside,
                                               int fshift,
                                               const int8_t* __restrict__ mask, int from_claim,
                                               uint8_t* __restrict__ claim,
                                               float* __restrict__ cellval,
                                               const float* const* __restrict__ lapfields,
                                               float* __restrict__ celld2,
                                               uint32_t* __restrict__ takenw) {
    const int m = n - 2 * nbuff;
    // the wave index is uniform across the wavefront: say so, so row offsets stay scalar
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i0 = (blockIdx.x * WI + w % WI) * OI;
    const int j0 = (blockIdx.y * WJ + w / WI) * TJ;
    const int k0 = blockIdx.z * TK;
    // A wavefront past the inner cube still streams (clamped) planes and meets every barrier;
    // all its cells are outside and so marked taken below.
    const int io = i0 - 1 + lane;              // inner-cube i of this lane's column, >= -1
    const bool out_i = lane >= 1 && lane <= OI && io < m;
    // The fields are fside^3 arrays whose index 0 is tile index fshift along each axis (the
    // whole tile: fside = n, fshift = 0; the inner cube and its one-cell halo: fside = m + 2,
    // fshift = nbuff - 1).  The reads stay in tile indices nbuff - 1 .. nbuff + m.
    const int fb = nbuff - fshift;
    // column offset inside a plane; io is clamped to m (a halo column: values unused)
    const int col_bytes = (min(io, m) + fb) * 4;
    const int64_t s2 = (int64_t)n * n, fs2 = (int64_t)fside * fside;
    const int row_bytes = 4 * fside, plane_bytes = (int)(4 * fs2);  // fside <= 4096
    const int64_t mm = m, mm2 = (int64_t)m * m;

    // Cells already taken (masked, or a peak of an earlier filter): bit 8*(kk%4)+jr of
    // word kk/4.  The words rotate by one per group of 4 planes, so word 0 is always current.
    uint32_t taken[TK / 4];
    // word of bitset `takenw` for 4-plane group q of this lane's column (TILE, from_claim)
    const int NJ = (int)gridDim.y * WJ, NKG = (int)gridDim.z * (TK / 4);
    auto word = [&](int q) -> int64_t {
        return (int64_t)io + (int64_t)m * ((j0 / TJ) + (int64_t)NJ * (k0 / 4 + q));
    };
#pragma unroll
    for (int q = 0; q < TK / 4; ++q) {
        if (TILE && from_claim) {
            taken[q] = out_i ? takenw[word(q)] : 0xFFFFFFFFu;
            continue;
        }
        uint32_t bits = 0;
#pragma unroll 4
        for (int e = 0; e < 32; ++e) {
            const int kk = 4 * q + (e >> 3), jr = e & 7;
            const int jj = j0 + jr, kq = k0 + kk;
            bool t = true;  // cells outside the inner cube count as taken
            if (out_i && jj < m && kq < m)
                t = (TILE && from_claim)
                        ? false  // the claim array is checked at a hit instead
                        : mask[(int64_t)(io + nbuff) + (int64_t)(jj + nbuff) * n +
                               (kq + nbuff) * s2] == 1;
            bits |= (uint32_t)t << e;
        }
        taken[q] = bits;
    }

    for (int f = 0; f < nfilt; ++f) {
        const float* __restrict__ F = fields[f];
        const float fc = fcrit[f];
        // Rows j0-1 .. j0+TJ of inner-cube plane kp (clamped to m: beyond it nothing is
        // tested), as buffer loads off a per-plane resource: the lane's column offset is the
        // one vector register, the row offset is scalar.
        auto load = [&](int kp, float (&v)[TJ + 2]) {
            const float* P = F + (int64_t)(min(kp, m) + fb) * fs2;
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc((void*)P, 0, plane_bytes, BUF_DWORD3);
#pragma unroll
            for (int r = 0; r < TJ + 2; ++r)
                v[r] = __builtin_bit_cast(
                    float, __builtin_amdgcn_raw_buffer_load_b32(
                               rs, col_bytes, (min(j0 - 1 + r, m) + fb) * row_bytes, 0));
        };
        auto reduce = [&](const float (&v)[TJ + 2], float (&mx)[TJ]) {
            float h[TJ + 2];
#pragma unroll
            for (int r = 0; r < TJ + 2; ++r)
                h[r] = fmaxf(v[r], fmaxf(from_lane_below(v[r]), from_lane_above(v[r])));
#pragma unroll
            for (int r = 0; r < TJ; ++r) mx[r] = fmaxf(h[r], fmaxf(h[r + 1], h[r + 2]));
        };
        // State entering the test of plane kk: mx = 3x3 maxima of plane kk, cc = its cell
        // values, and bit r of out = cell (r, kk) already fails against fcrit or against
        // planes kk-1 and kk.  !(ff < max(a, b, c)) splits exactly into !(ff < max(a, b))
        // and !(ff < c) under maxNum, so plane kk+1 completes the test.
        float va[TJ + 2], mx[TJ], cc[TJ];
        uint32_t out = 0;
        {
            float mprev[TJ];
            load(k0 - 1, va);
            reduce(va, mprev);
            load(k0, va);
            reduce(va, mx);
#pragma unroll
            for (int r = 0; r < TJ; ++r) {
                cc[r] = va[r + 1];
                // the reference skips `ff < fcrit`; a NaN is therefore not skipped
                out |= (uint32_t)((cc[r] < fc) || (cc[r] < fmaxf(mprev[r], mx[r]))) << r;
            }
        }
        load(k0 + 1, va);  // plane kk+1, in flight
#pragma unroll 1
        for (int g = 0; g < TK / 4; ++g) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int kk = 4 * g + u;
                float cur[TJ + 2], mn[TJ];
#pragma unroll
                for (int r = 0; r < TJ + 2; ++r) cur[r] = va[r];
                load(k0 + kk + 2, va);  // next plane in flight while this one is reduced
                // Keep the workgroup's wavefronts on the same plane: the cache lines they
                // share (halo rows and columns) are then fetched from HBM once.
                __syncthreads();
                reduce(cur, mn);
                uint32_t hit = 0, nout = 0;
                float cv[TJ];  // plane kk's cell values, the claimed cells' FcRf
#pragma unroll
                for (int r = 0; r < TJ; ++r) {
                    cv[r] = cc[r];
                    hit |= (uint32_t)!(cc[r] < mn[r]) << (8 * u + r);
                    const float c = cur[r + 1];
                    nout |= (uint32_t)((c < fc) || (c < fmaxf(mx[r], mn[r]))) << r;
                    cc[r] = c;
                    mx[r] = mn[r];
                }
                hit &= ~(taken[0] | (out << (8 * u)));
                out = nout;
                if (hit & (0xFFu << (8 * u))) {
#pragma unroll
                    for (int r = 0; r < TJ; ++r)
                        if (hit & (1u << (8 * u + r))) {
                            const int64_t ci = io + (j0 + r) * mm + (k0 + kk) * mm2;
                            claim[ci] = (uint8_t)(fbase + f);
                            if (TILE && cellval) {
                                cellval[ci] = cv[r];
                                if (celld2) {
                                    const int64_t a = (int64_t)(io + fb) +
                                                      (int64_t)(j0 + r + fb) * fside +
                                                      (k0 + kk + fb) * fs2;
                                    celld2[ci] = lapfields[f][a];
                                }
                            }
                        }
                }
                taken[0] |= hit;
            }
            const uint32_t t0 = taken[0];
#pragma unroll
            for (int q = 0; q + 1 < TK / 4; ++q) taken[q] = taken[q + 1];
            taken[TK / 4 - 1] = t0;
        }
    }
    // the words have rotated through a whole cycle: taken[q] is group q again
    if (TILE && from_claim && out_i) {
#pragma unroll
        for (int q = 0; q < TK / 4; ++q) takenw[word(q)] = taken[q];
    }
    (void)NKG;
}

// Bitset of pf_mark's TILE launches before the first filter: for every lane column io of
// the inner cube, 8-row block jb and 4-plane group kg (the grid's, some past the cube), bit
// 8 kk + jr = cell (io, 8 jb + jr, 4 kg + kk) is outside the inner cube or masked.
__global__ __launch_bounds__(256) void pf_taken_init(const int8_t* __restrict__ mask, int n,
                                                     int nbuff, int NJ, int NKG,
                                                     uint32_t* __restrict__ takenw) {
    const int m = n - 2 * nbuff;
    const int64_t nw = (int64_t)m * NJ * NKG, s2 = (int64_t)n * n;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw;
         w += (int64_t)gridDim.x * blockDim.x) {
        const int io = (int)(w % m);
        const int64_t r = w / m;
        const int jb = (int)(r % NJ), kg = (int)(r / NJ);
        uint32_t bits = 0;
        for (int e = 0; e < 32; ++e) {
            const int jj = 8 * jb + (e & 7), kq = 4 * kg + (e >> 3);
            const bool t = jj >= m || kq >= m ||
                           mask[(int64_t)(io + nbuff) + (int64_t)(jj + nbuff) * n +
                                (kq + nbuff) * s2] == 1;
            bits |= (uint32_t)t << e;
        }
        takenw[w] = bits;
    }
}

__global__ __launch_bounds__(256) void pf_count(const uint8_t* __restrict__ claim, int64_t ncell,
                                                int nfilt, int64_t nchunks,
                                                unsigned long long* __restrict__ counts) {
    __shared__ unsigned int c[256];
    c[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * CHUNK;
    if (base + CHUNK <= ncell) {
        // a whole chunk: 16 consecutive claim bytes per thread in one load
        const uint4 w = *reinterpret_cast<const uint4*>(claim + base + threadIdx.x * 16);
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t v = (ws[q >> 2] >> (8 * (q & 3))) & 0xffu;
            if (v < BLOCKED) atomicAdd(&c[v], 1u);
        }
    } else {
        for (int e = threadIdx.x; e < CHUNK; e += 256) {
            int64_t idx = base + e;
            if (idx < ncell) {
                uint8_t v = claim[idx];
                if (v < BLOCKED) atomicAdd(&c[v], 1u);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < nfilt) counts[(int64_t)threadIdx.x * nchunks + blockIdx.x] = c[threadIdx.x];
}

__global__ __launch_bounds__(256) void pf_scatter(
    const uint8_t* __restrict__ claim, int64_t ncell, int64_t nchunks,
    const unsigned long long* __restrict__ offsets, int n, int nbuff,
    const float* const* __restrict__ fields, const float* __restrict__ cellval,
    const float* __restrict__ celld2, int8_t* __restrict__ mask, int32_t* __restrict__ oi,
    int32_t* __restrict__ oj, int32_t* __restrict__ ok, int32_t* __restrict__ of,
    float* __restrict__ od, float* __restrict__ od2, int64_t capacity) {
    __shared__ int s_scan[256];
    __shared__ int s_cell[CHUNK];
    __shared__ uint8_t s_f[CHUNK];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * CHUNK;
    // each thread owns 16 consecutive cells of the chunk
    int mine = 0;
    uint8_t v[16];
    if (base + CHUNK <= ncell) {
        const uint4 w = *reinterpret_cast<const uint4*>(claim + base + t * 16);
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = (uint8_t)(ws[q >> 2] >> (8 * (q & 3)));
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            int64_t idx = base + t * 16 + q;
            v[q] = idx < ncell ? claim[idx] : NONE;
        }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) mine += v[q] < BLOCKED;
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
        if (v[q] < BLOCKED) {
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
            od[dst] = cellval ? cellval[idx] : fields[f][a];
            if (od2) od2[dst] = celld2 ? celld2[idx] : 0.0f;
            mask[a] = 1;
        }
    }
}

// Claim array of the tile path before its first filter: BLOCKED where the mask is 1.
// Grid (ceil(m / 256), m, m): workgroup (x, jj, kq) sets claim cells ii = 256 x + t of row
// (jj, kq) of the inner cube.
__global__ __launch_bounds__(256) void pf_claim_init(const int8_t* __restrict__ mask, int n,
                                                     int nbuff, uint8_t* __restrict__ claim) {
    const int m = n - 2 * nbuff;
    const int ii = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (ii >= m) return;
    const int jj = (int)blockIdx.y, kq = (int)blockIdx.z;
    const int64_t c = ii + (int64_t)m * (jj + (int64_t)m * kq);
    const int64_t a = (int64_t)(ii + nbuff) + (int64_t)(jj + nbuff) * n +
                      (int64_t)(kq + nbuff) * n * n;
    claim[c] = mask[a] == 1 ? BLOCKED : NONE;
}

// Window product of _smooth_window_kernel! (ext/CUDAExt.jl:2831-2871) on the rfft spectrum
// [k][j][0..n/2] of a column-major tile: kx from the halved index, ky and kz signed; W in FP64
// (top-hat 3(sin x - x cos x)/x^3, Gaussian exp(-x^2/8), or k^2 times it), rounded to Float32
// and applied to both parts; k = 0 is left unchanged.  k^2 = kx^2 + ky^2 + kz^2 takes the
// same FP64 value for +-ky and +-kz (negation is exact), so one thread evaluates W once for
// (kx, |ky|, |kz|) and applies it to the up to four elements with those magnitudes.  The
// inverse transform's 1/n^3 is applied here too (for a power-of-two n it is exact, so it
// equals scaling afterwards).
// The Float32 window of one filter at k^2 > 0 (FP64, then rounded).
__device__ __forceinline__ float window32(double k2, double Rf, int wsmooth) {
    const double k = sqrt(k2), kR = k * Rf;
    double w;
    if (wsmooth == 0) {
        w = exp(-kR * kR * 0.125);
    } else if (wsmooth == 3) {
        w = exp(-kR * kR * 0.125) * k2;
    } else {
        double sn, cs;
        sincos(kR, &sn, &cs);
        w = 3.0 * (sn - kR * cs) / (kR * kR * kR);
    }
    return (float)w;
}

// The window of one filter as a table over (kx, |ky|, |kz|), nk = n/2 + 1 entries a side,
// with k from those magnitudes as fb_window computes it, and 1 at k = 0 (where
// fb_window leaves the spectrum as it is: v * 1 = v).  The library's inverse transform reads
// it as it first reads the spectrum (FFT_OP_WINDOW), so the windowed spectrum is never
// written: 0.54 GB of table instead of 8.6 GB of spectrum traffic per filter at 1024^3.
// The table is laid out as the z pass reads it, |kz| fastest: entry az + nk (ix + nk ay).
// One workgroup per (kx, |ky|) row of the table: blockIdx.x + nk blockIdx.y is
// ix + nk ay for ix < nk, ay < nk (no 64-bit division per element).
__global__ __launch_bounds__(256) void fb_wtable(int n, double dk, double Rf, int wsmooth,
                                                 float* __restrict__ w) {
    const int nk = n / 2 + 1;
    const int ix = (int)blockIdx.x, ay = (int)blockIdx.y;
    const double kx = (double)ix * dk, ky = (double)ay * dk;
    float* row = w + (int64_t)nk * (ix + (int64_t)nk * ay);
    for (int az = threadIdx.x; az < nk; az += 256) {
        const double kz = (double)az * dk;
        const double k2 = kx * kx + ky * ky + kz * kz;
        row[az] = k2 != 0.0 ? window32(k2, Rf, wsmooth) : 1.0f;
    }
}

__global__ __launch_bounds__(256) void fb_window(const float2* __restrict__ spec,
                                                 float2* __restrict__ out, int n, double dk,
                                                 double Rf, int wsmooth, float inv_n3) {
    const int nk = n / 2 + 1, nyq = n / 2;
    const int64_t total = (int64_t)nk * (nyq + 1) * (nyq + 1);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int ix = (int)(e % nk);
        const int64_t rem = e / nk;
        const int ay = (int)(rem % (nyq + 1)), az = (int)(rem / (nyq + 1));  // |ky|, |kz| / dk
        const double kx = (double)ix * dk, ky = (double)ay * dk, kz = (double)az * dk;
        const double k2 = kx * kx + ky * ky + kz * kz;
        float w32 = 1.0f;
        bool apply = false;
        if (k2 != 0.0) {
            w32 = window32(k2, Rf, wsmooth);
            apply = true;
        }
        // rows iy = ay and n - ay (one row when ay is 0, or n/2 for an even n), likewise iz
        const bool even = (n & 1) == 0;
        const int ys[2] = {ay, n - ay}, zs[2] = {az, n - az};
        const int ny = (ay == 0 || (even && ay == nyq)) ? 1 : 2;
        const int nz = (az == 0 || (even && az == nyq)) ? 1 : 2;
        for (int b = 0; b < nz; ++b)
            for (int a = 0; a < ny; ++a) {
                const int64_t q = ix + (int64_t)nk * (ys[a] + (int64_t)n * zs[b]);
                float2 v = spec[q];
                if (apply) {
                    v.x *= w32;
                    v.y *= w32;
                }
                v.x *= inv_n3;
                v.y *= inv_n3;
                out[q] = v;
            }
    }
}

}  // namespace

using namespace hnt;

namespace hnt {
void release_fft(hnt_ctx* ctx) {
    if (ctx->fft_r2c) (void)hipfftDestroy((hipfftHandle)ctx->fft_r2c);
    if (ctx->fft_c2r) (void)hipfftDestroy((hipfftHandle)ctx->fft_c2r);
    ctx->fft_r2c = ctx->fft_c2r = nullptr;
    ctx->fft_n = 0;
}

void release_field_plans(hnt_ctx* ctx) {
    for (hnt_ctx::FftPlans& p : ctx->fs_plans) {
        if (p.r2c) (void)hipfftDestroy((hipfftHandle)p.r2c);
        if (p.c2r) (void)hipfftDestroy((hipfftHandle)p.c2r);
    }
    ctx->fs_plans.clear();
}
}  // namespace hnt

#define HNT_FFT(expr)                                                                   \
    do {                                                                                \
        const hipfftResult r_ = (expr);                                                 \
        if (r_ != HIPFFT_SUCCESS) {                                                     \
            hnt::set_error("%s: hipFFT status %d (%s:%d)", #expr, (int)r_, __FILE__,    \
                           __LINE__);                                                   \
            return HNT_ERR_DEVICE;                                                      \
        }                                                                               \
    } while (0)

namespace {

// Buffers of the claim array and the filter-major count table for an inner cube of ncell
// cells and nfilt filters.
hnt_status reserve_claims(hnt_ctx* ctx, int64_t ncell, int nfilt, size_t& tmp_bytes) {
    const int64_t nchunks = (ncell + CHUNK - 1) / CHUNK;
    const int64_t ntab = nchunks * nfilt;
    HNT_TRY(reserve(ctx, ctx->claim, (size_t)ncell));
    HNT_TRY(reserve(ctx, ctx->counts, sizeof(unsigned long long) * (size_t)ntab));
    HNT_TRY(reserve(ctx, ctx->offsets, sizeof(unsigned long long) * (size_t)ntab));
    tmp_bytes = 0;
    HNT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes,
                                             (unsigned long long*)ctx->counts.p,
                                             (unsigned long long*)ctx->offsets.p, (int)ntab,
                                             ctx->stream));
    HNT_TRY(reserve(ctx, ctx->cub_tmp, tmp_bytes > 0 ? tmp_bytes : 16));
    return HNT_OK;
}

// Count, scan and scatter the claimed cells into `out` (filter, then k, j, i), mask them,
// wait for the stream and fill out->count / out->total.
hnt_status compact_claims(hnt_ctx* ctx, const char* what, int n, int nbuff, int nfilt,
                          size_t tmp_bytes, const float* const* dfields, const float* cellval,
                          const float* celld2, int8_t* mask, hnt_peaks* out, float* od2) {
    hipStream_t st = ctx->stream;
    const int m = n - 2 * nbuff;
    const int64_t ncell = (int64_t)m * m * m;
    const int64_t nchunks = (ncell + CHUNK - 1) / CHUNK;
    const int64_t ntab = nchunks * nfilt;
    stage_begin(ctx, HNT_STAGE_PF_COMPACT);
    hipLaunchKernelGGL(pf_count, dim3((unsigned)nchunks), dim3(256), 0, st,
                       (const uint8_t*)ctx->claim.p, ncell, nfilt, nchunks,
                       (unsigned long long*)ctx->counts.p);
    HNT_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp.p, tmp_bytes,
                                             (unsigned long long*)ctx->counts.p,
                                             (unsigned long long*)ctx->offsets.p, (int)ntab, st));
    hipLaunchKernelGGL(pf_scatter, dim3((unsigned)nchunks), dim3(256), 0, st,
                       (const uint8_t*)ctx->claim.p, ncell, nchunks,
                       (const unsigned long long*)ctx->offsets.p, n, nbuff, dfields, cellval,
                       celld2, mask, out->i, out->j, out->k, out->filt, out->delta, od2,
                       out->capacity);
    stage_end(ctx, HNT_STAGE_PF_COMPACT);
    unsigned long long tail[2] = {0, 0};
    HNT_HIP(hipMemcpyAsync(&tail[0], (unsigned long long*)ctx->offsets.p + (ntab - 1),
                           sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HNT_HIP(hipMemcpyAsync(&tail[1], (unsigned long long*)ctx->counts.p + (ntab - 1),
                           sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HNT_TRY(finish_call(ctx, what));
    const int64_t total = (int64_t)(tail[0] + tail[1]);
    out->total = total;
    out->count = total < out->capacity ? total : out->capacity;
    if (total > out->capacity) {
        set_error("%s: %lld peaks found, capacity %lld", what, (long long)total,
                  (long long)out->capacity);
        return HNT_ERR_CAPACITY;
    }
    return HNT_OK;
}

dim3 mark_grid(int m) {
    return dim3((m + OI * WI - 1) / (OI * WI), (m + TJ * WJ - 1) / (TJ * WJ), (m + TK - 1) / TK);
}

}  // namespace

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

    size_t tmp_bytes = 0;
    HNT_TRY(reserve_claims(ctx, ncell, nfilt, tmp_bytes));
    HNT_TRY(reserve(ctx, ctx->fieldptrs, sizeof(float*) * (size_t)nfilt));
    HNT_TRY(reserve(ctx, ctx->fcrit, sizeof(float) * (size_t)nfilt));
    HNT_HIP(hipMemcpyAsync(ctx->fieldptrs.p, smoothed, sizeof(float*) * (size_t)nfilt,
                           hipMemcpyHostToDevice, st));
    HNT_HIP(hipMemcpyAsync(ctx->fcrit.p, fcrit, sizeof(float) * (size_t)nfilt,
                           hipMemcpyHostToDevice, st));
    const float* const* dfields = (const float* const*)ctx->fieldptrs.p;

    stage_begin(ctx, HNT_STAGE_PF_MARK);
    HNT_HIP(hipMemsetAsync(ctx->claim.p, NONE, (size_t)ncell, st));
    hipLaunchKernelGGL(pf_mark<false>, mark_grid(m), dim3(PF_THREADS), 0, st, dfields,
                       (const float*)ctx->fcrit.p, nfilt, 0, n, nbuff, n, 0, (const int8_t*)mask, 0,
                       (uint8_t*)ctx->claim.p, nullptr, nullptr, nullptr, nullptr);
    stage_end(ctx, HNT_STAGE_PF_MARK);
    return compact_claims(ctx, "hnt_peak_find", n, nbuff, nfilt, tmp_bytes, dfields, nullptr,
                          nullptr, mask, out, nullptr);
}

extern "C" hnt_status hnt_peak_find_tile(hnt_ctx* ctx, const float* delta, int32_t n,
                                         int32_t nbuff, const hnt_filter_bank* bank,
                                         int8_t* mask, hnt_peaks* out, float* d2Rf,
                                         float* const* smoothed_out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(delta && bank && mask && out, "hnt_peak_find_tile: NULL argument");
    const int nfilt = bank->nfilt;
    HNT_REQUIRE(nfilt >= 1 && nfilt <= 254, "hnt_peak_find_tile: nfilt %d not in 1..254",
                nfilt);
    HNT_REQUIRE(bank->Rf && bank->fcrit, "hnt_peak_find_tile: NULL filter radii / thresholds");
    HNT_REQUIRE(bank->wsmooth == 0 || bank->wsmooth == 1 || bank->wsmooth == 3,
                "hnt_peak_find_tile: wsmooth %d not 0, 1 or 3", bank->wsmooth);
    HNT_REQUIRE(bank->alatt > 0.0, "hnt_peak_find_tile: alatt must be > 0");
    HNT_REQUIRE(n >= 3 && n <= 2048, "hnt_peak_find_tile: n %d not in 3..2048", n);
    HNT_REQUIRE(nbuff >= 1, "hnt_peak_find_tile: nbuff must be >= 1, got %d", nbuff);
    HNT_REQUIRE(out->capacity >= 0, "hnt_peak_find_tile: negative capacity");
    HNT_REQUIRE(out->capacity == 0 || (out->i && out->j && out->k && out->filt && out->delta),
                "hnt_peak_find_tile: NULL output array");
    out->count = 0;
    out->total = 0;
    const int m = n - 2 * nbuff;
    if (m <= 0) return HNT_OK;
    hipStream_t st = ctx->stream;
    const int64_t ncell = (int64_t)m * m * m;
    // spectrum buffers: the library's padded rows (hipFFT's n/2 + 1 fit in them)
    const int64_t nreal = (int64_t)n * n * n, ncplx = fft_plane(n) * n;
    const bool lap = d2Rf != nullptr;

    size_t tmp_bytes = 0;
    HNT_TRY(reserve_claims(ctx, ncell, nfilt, tmp_bytes));
    HNT_TRY(reserve(ctx, ctx->fieldptrs, sizeof(float*) * 2));
    HNT_TRY(reserve(ctx, ctx->fcrit, sizeof(float) * (size_t)nfilt));
    HNT_TRY(reserve(ctx, ctx->fb_spec, sizeof(float2) * (size_t)ncplx));
    HNT_TRY(reserve(ctx, ctx->fb_work, sizeof(float2) * (size_t)ncplx));
    HNT_TRY(reserve(ctx, ctx->fb_field, sizeof(float) * (size_t)nreal));
    HNT_TRY(reserve(ctx, ctx->fb_val, sizeof(float) * (size_t)ncell));
    if (lap) {
        HNT_TRY(reserve(ctx, ctx->fb_lap, sizeof(float) * (size_t)nreal));
        HNT_TRY(reserve(ctx, ctx->fb_d2, sizeof(float) * (size_t)ncell));
    }
    const int64_t nwtab = (int64_t)(n / 2 + 1) * (n / 2 + 1) * (n / 2 + 1);
    // transforms: the library's own for power-of-two n (tuning fft = 0), else hipFFT, whose
    // 32-bit 3-D plans are only valid below 2^31 elements
    const bool own = ctx->fft_mode == 0 && fft_supported(n);
    HNT_REQUIRE(own || nreal < (1LL << 31),
                "hnt_peak_find_tile: a %d^3 tile needs the library FFT (power-of-two n, tuning "
                "fft = 0)", n);
    if (own) HNT_TRY(reserve(ctx, ctx->fb_wtab, sizeof(float) * (size_t)nwtab));
    // taken bits carried from filter to filter (one word per lane column, 8 rows, 4 planes)
    const dim3 mg = mark_grid(m);
    const int NJ = (int)mg.y * WJ, NKG = (int)mg.z * (TK / 4);
    const int64_t ntw = (int64_t)m * NJ * NKG;
    HNT_TRY(reserve(ctx, ctx->fb_taken, sizeof(uint32_t) * (size_t)ntw));
    hipfftHandle r2c = nullptr, c2r = nullptr;
    if (!own) {
        if (ctx->fft_n != n) {
            release_fft(ctx);
            HNT_FFT(hipfftPlan3d(&r2c, n, n, n, HIPFFT_R2C));
            ctx->fft_r2c = r2c;
            HNT_FFT(hipfftPlan3d(&c2r, n, n, n, HIPFFT_C2R));
            ctx->fft_c2r = c2r;
            ctx->fft_n = n;
        }
        r2c = (hipfftHandle)ctx->fft_r2c;
        c2r = (hipfftHandle)ctx->fft_c2r;
        HNT_FFT(hipfftSetStream(r2c, st));
        HNT_FFT(hipfftSetStream(c2r, st));
    }
    float2* spec = (float2*)ctx->fb_spec.p;
    float2* work = (float2*)ctx->fb_work.p;
    float* field = (float*)ctx->fb_field.p;
    float* lapf = lap ? (float*)ctx->fb_lap.p : nullptr;
    const float* ptrs[2] = {field, lapf};
    HNT_HIP(hipMemcpyAsync(ctx->fieldptrs.p, ptrs, sizeof(ptrs), hipMemcpyHostToDevice, st));
    HNT_HIP(hipMemcpyAsync(ctx->fcrit.p, bank->fcrit, sizeof(float) * (size_t)nfilt,
                           hipMemcpyHostToDevice, st));
    const float* const* dptrs = (const float* const*)ctx->fieldptrs.p;

    // rfft of the tile (hipFFT from a copy: the caller's delta stays as it was)
    stage_begin(ctx, HNT_STAGE_PF_SMOOTH);
    if (own) {
        HNT_TRY(fft_forward(ctx, delta, n, n, spec));
    } else {
        HNT_HIP(hipMemcpyAsync(field, delta, sizeof(float) * (size_t)nreal,
                               hipMemcpyDeviceToDevice, st));
        HNT_FFT(hipfftExecR2C(r2c, (hipfftReal*)field, (hipfftComplex*)spec));
    }
    stage_end(ctx, HNT_STAGE_PF_SMOOTH);
    hipLaunchKernelGGL(pf_claim_init, dim3((unsigned)((m + 255) / 256), (unsigned)m, (unsigned)m),
                       dim3(256), 0, st, (const int8_t*)mask, n, nbuff, (uint8_t*)ctx->claim.p);
    hipLaunchKernelGGL(pf_taken_init, dim3(2048), dim3(256), 0, st, (const int8_t*)mask, n,
                       nbuff, NJ, NKG, (uint32_t*)ctx->fb_taken.p);
    const double dk = 2.0 * M_PI / ((double)n * bank->alatt);
    const float inv_n3 = (float)(1.0 / (double)nreal);
    const int64_t nwin = (int64_t)(n / 2 + 1) * (n / 2 + 1) * (n / 2 + 1);
    const unsigned wblocks = (unsigned)std::min<int64_t>((nwin + 255) / 256, 16384);
    for (int f = 0; f < nfilt; ++f) {
        // The peak test reads the inner cube and its one-cell halo, (m + 2)^3 cells from tile
        // index nbuff - 1: the library's inverse transform writes only those (the same values
        // as in the whole field) unless the caller wants the whole smoothed field.
        const bool whole = !own || (smoothed_out && smoothed_out[f]);
        const int fshift = whole ? 0 : nbuff - 1, fside = whole ? n : m + 2;
        // library path: the window table of this filter, then the inverse transform of the
        // spectrum times the table (and 1/n^3) into `dst`
        auto window_inverse = [&](int wsmooth, float* dst) -> hnt_status {
            float* wtab = (float*)ctx->fb_wtab.p;
            hipLaunchKernelGGL(fb_wtable, dim3((unsigned)(n / 2 + 1), (unsigned)(n / 2 + 1)),
                               dim3(256), 0, st, n, dk, bank->Rf[f], wsmooth, wtab);
            FftOp wop;
            wop.kind = FFT_OP_WINDOW;
            wop.L = n;
            wop.amp = wtab;
            wop.post = inv_n3;
            return fft_inverse(ctx, spec, work, n, fshift, fside, wop, 1.0f, 0, 0.0f, dst);
        };
        stage_begin(ctx, HNT_STAGE_PF_SMOOTH);
        if (own) {
            HNT_TRY(window_inverse(bank->wsmooth, field));
        } else {
            hipLaunchKernelGGL(fb_window, dim3(wblocks), dim3(256), 0, st,
                               (const float2*)ctx->fb_spec.p, (float2*)ctx->fb_work.p, n, dk,
                               bank->Rf[f], bank->wsmooth, inv_n3);
            HNT_FFT(hipfftExecC2R(c2r, (hipfftComplex*)work, (hipfftReal*)field));
        }
        if (smoothed_out && smoothed_out[f])
            HNT_HIP(hipMemcpyAsync(smoothed_out[f], field, sizeof(float) * (size_t)nreal,
                                   hipMemcpyDeviceToDevice, st));
        if (lap) {
            if (own) {
                HNT_TRY(window_inverse(3, lapf));
            } else {
                hipLaunchKernelGGL(fb_window, dim3(wblocks), dim3(256), 0, st,
                                   (const float2*)ctx->fb_spec.p, (float2*)ctx->fb_work.p, n,
                                   dk, bank->Rf[f], 3, inv_n3);
                HNT_FFT(hipfftExecC2R(c2r, (hipfftComplex*)work, (hipfftReal*)lapf));
            }
        }
        stage_end(ctx, HNT_STAGE_PF_SMOOTH);
        stage_begin(ctx, HNT_STAGE_PF_MARK);
        hipLaunchKernelGGL(pf_mark<true>, mg, dim3(PF_THREADS), 0, st, dptrs,
                           (const float*)ctx->fcrit.p + f, 1, f, n, nbuff, fside, fshift,
                           (const int8_t*)mask, 1, (uint8_t*)ctx->claim.p,
                           (float*)ctx->fb_val.p, lap ? dptrs + 1 : nullptr,
                           lap ? (float*)ctx->fb_d2.p : nullptr, (uint32_t*)ctx->fb_taken.p);
        stage_end(ctx, HNT_STAGE_PF_MARK);
    }
    return compact_claims(ctx, "hnt_peak_find_tile", n, nbuff, nfilt, tmp_bytes, dptrs,
                          (const float*)ctx->fb_val.p, lap ? (const float*)ctx->fb_d2.p : nullptr,
                          mask, out, d2Rf);
}

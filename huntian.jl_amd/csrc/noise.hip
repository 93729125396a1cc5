// noise.hip -- the tile driver's white noise on the device: the Threefry-2x64-20 Gaussian
// stream of the global N^3 grid, per tile (_generate_tile_noise, src/MultiResolution.jl:35-57),
// block-averaged onto the coarse grid (_downsample_noise, :64-87) and as the Hoffman-Ribak
// residual of an extended tile region (_generate_extended_residual, :190-227).
//
// Cell idx (0-based, column-major over the global grid) takes Gaussian (idx & 1) of pair
// idx >> 1 (RandomField._threefry_gaussian, src/InitialConditions/RandomField.jl:45-53):
// Threefry-2x64 with 20 rounds (Salmon et al., SC'11, the algorithm of Random123) of counter
// (pair, 0) under key (seed, 0), the two words turned into doubles in (0, 1] as
// ((w >> 11) + 1) 2^-53, then Box-Muller in FP64: sqrt(-2 log u1) (cos, sin)(2 pi u2).  Every
// cell depends only on (seed, idx), so threads compute their cells independently (the two
// cells of a pair together where they are neighbours).
#include "hnt_internal.h"

namespace {

using namespace hnt;

constexpr uint64_t kParity = 0x1BD11BDAA9FC1A22ull;  // Threefry key-schedule parity

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) {
    return (x << r) | (x >> (64 - r));
}

// Threefry-2x64, 20 rounds: rotations 16 42 12 31 16 32 24 21, a key injection after every
// fourth round (word 0 + ks[s mod 3], word 1 + ks[(s + 1) mod 3] + s).
__device__ inline void threefry2x64_20(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1,
                                       uint64_t& o0, uint64_t& o1) {
    const uint64_t ks[3] = {k0, k1, kParity ^ k0 ^ k1};
    constexpr int R[8] = {16, 42, 12, 31, 16, 32, 24, 21};
    uint64_t x0 = c0 + ks[0], x1 = c1 + ks[1];
#pragma unroll
    for (int r = 0; r < 20; ++r) {
        x0 += x1;
        x1 = rotl64(x1, R[r % 8]);
        x1 ^= x0;
        if (r % 4 == 3) {
            const int s = (r + 1) / 4;
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3] + (uint64_t)s;
        }
    }
    o0 = x0;
    o1 = x1;
}

// Both Gaussians of pair `pair` (cells 2 pair and 2 pair + 1), in FP64.
__device__ inline void gauss_pair(uint64_t seed, uint64_t pair, double& g1, double& g2) {
    uint64_t r0, r1;
    threefry2x64_20(pair, 0, seed, 0, r0, r1);
    const double u1 = ((double)(r0 >> 11) + 1.0) * 1.1102230246251565e-16;
    const double u2 = ((double)(r1 >> 11) + 1.0) * 1.1102230246251565e-16;
    const double r = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586 * u2;  // 2 pi (Float64) times u2
    g1 = r * cos(th);
    g2 = r * sin(th);
}

// The Gaussian of global cell idx0 (0-based linear index), in FP64.
__device__ inline double gauss_at(uint64_t seed, uint64_t idx0) {
    double g1, g2;
    gauss_pair(seed, idx0 >> 1, g1, g2);
    return (idx0 & 1) ? g2 : g1;
}

// mod1(x, N) - 1: the 0-based periodic index of 1-based x
__device__ __forceinline__ int64_t wrap0(int64_t x, int64_t N) {
    const int64_t m = (x - 1) % N;
    return m < 0 ? m + N : m;
}

// Cells (li, lj, lk) of a side^3 region whose first cell is global (i0, j0, k0) (1-based,
// wrapped): out = Float32(gauss), or with coarse, Float32(gauss) - coarse[block cell] * inv.
// For an even N the two cells of a pair are neighbours along x in one row (the parity of a
// cell's global index is that of its wrapped x index), so one thread takes a pair: the
// region's row is cut into pairs from the parity of its first cell, a lone cell at either
// end.  An odd N takes one cell per thread.
__device__ __forceinline__ void region_store(int64_t e, int64_t gi, int64_t gj, int64_t gk,
                                             double g, const float* __restrict__ coarse, int M,
                                             int64_t block, float inv, float* __restrict__ out) {
    const float fine = (float)g;
    if (!coarse) {
        out[e] = fine;
        return;
    }
    const int64_t ci = gi / block, cj = gj / block, ck = gk / block;
    const float mean = coarse[ci + (int64_t)M * (cj + (int64_t)M * ck)] * inv;
    out[e] = fine - mean;
}

__global__ __launch_bounds__(256) void region_noise_kernel(uint64_t seed, int64_t N,
                                                           int64_t i0, int64_t j0, int64_t k0,
                                                           int side, const float* __restrict__ coarse,
                                                           int M, int64_t block, float inv,
                                                           float* __restrict__ out) {
    const bool paired = (N & 1) == 0;
    const int64_t gi0 = wrap0(i0, N);                  // global x of the row's first cell
    const int shift = paired ? (int)(gi0 & 1) : 0;     // 1: the first cell ends a pair
    const int per_row = paired ? (side + shift + 1) / 2 : side;
    const int64_t rows = (int64_t)side * side, total = rows * per_row;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(e % per_row);
        const int64_t row = e / per_row;
        const int lj = (int)(row % side), lk = (int)(row / side);
        const int64_t gj = wrap0(j0 + lj, N), gk = wrap0(k0 + lk, N);
        const int64_t obase = (int64_t)side * row;
        if (!paired) {
            const int64_t gi = wrap0(i0 + t, N);
            region_store(obase + t, gi, gj, gk,
                         gauss_at(seed, (uint64_t)(gi + gj * N + gk * N * N)), coarse, M,
                         block, inv, out);
            continue;
        }
        const int la = 2 * t - shift;  // the pair's even-index cell, maybe outside the row
        const int64_t gi = wrap0(i0 + (la >= 0 ? la : la + 1), N) - (la >= 0 ? 0 : 1);
        const uint64_t idx0 = (uint64_t)(gi + gj * N + gk * N * N);  // even
        double g1, g2;
        gauss_pair(seed, idx0 >> 1, g1, g2);
        if (la >= 0) region_store(obase + la, gi, gj, gk, g1, coarse, M, block, inv, out);
        if (la + 1 < side)
            region_store(obase + la + 1, gi + 1, gj, gk, g2, coarse, M, block, inv, out);
    }
}

// Coarse cell (I, J, K): the FP64 sum of its block^3 fine Gaussians in the reference's loop
// order (dk, dj, di, di fastest), divided by sqrt(block^3), rounded to Float32.
__global__ __launch_bounds__(256) void downsample_kernel(uint64_t seed, int64_t N, int M,
                                                         int64_t block, double norm,
                                                         float* __restrict__ coarse) {
    const int64_t total = (int64_t)M * M * M;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t I = e % M, rem = e / M;
        const int64_t J = rem % M, K = rem / M;
        double val = 0.0;
        for (int64_t dk = 0; dk < block; ++dk)
            for (int64_t dj = 0; dj < block; ++dj) {
                const int64_t row = (I * block) + (J * block + dj) * N + (K * block + dk) * N * N;
                if ((block & 1) == 0 && (N & 1) == 0) {
                    // an even block in an even grid starts every row on a pair: one pair
                    // per two cells, added in cell order
                    for (int64_t di = 0; di < block; di += 2) {
                        double g1, g2;
                        gauss_pair(seed, (uint64_t)(row + di) >> 1, g1, g2);
                        val += g1;
                        val += g2;
                    }
                } else {
                    for (int64_t di = 0; di < block; ++di)
                        val += gauss_at(seed, (uint64_t)(row + di));
                }
            }
        coarse[e] = (float)(val / norm);
    }
}

unsigned grid_for(int64_t total) {
    int64_t b = (total + 255) / 256;
    if (b > (1 << 20)) b = 1 << 20;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

extern "C" hnt_status hnt_tile_noise(hnt_ctx* ctx, int32_t it, int32_t jt, int32_t kt,
                                     int32_t nsub, int32_t nmesh, int64_t N, uint64_t seed,
                                     float* out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(out, "hnt_tile_noise: NULL output");
    HNT_REQUIRE(N >= 1 && nsub >= 1 && nmesh >= 1 && nmesh <= 16384 && it >= 1 && jt >= 1 &&
                    kt >= 1,
                "hnt_tile_noise: bad sizes (N %lld, nsub %d, nmesh %d, tile %d %d %d)",
                (long long)N, nsub, nmesh, it, jt, kt);
    const int64_t nbuff = (nmesh - nsub) / 2;
    const int64_t i0 = (int64_t)(it - 1) * nsub + 1 - nbuff;
    const int64_t j0 = (int64_t)(jt - 1) * nsub + 1 - nbuff;
    const int64_t k0 = (int64_t)(kt - 1) * nsub + 1 - nbuff;
    const int64_t total = (int64_t)nmesh * nmesh * nmesh;
    hipLaunchKernelGGL(region_noise_kernel, dim3(grid_for(total)), dim3(256), 0, ctx->stream,
                       seed, N, i0, j0, k0, nmesh, nullptr, 1, (int64_t)1, 0.0f, out);
    HNT_HIP(hipGetLastError());
    return finish_call(ctx, "hnt_tile_noise");
}

extern "C" hnt_status hnt_downsample_noise(hnt_ctx* ctx, int64_t N, int32_t M, uint64_t seed,
                                           float* coarse) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(coarse, "hnt_downsample_noise: NULL output");
    HNT_REQUIRE(M >= 1 && N >= M && N % M == 0,
                "hnt_downsample_noise: N = %lld must be a multiple of M = %d", (long long)N, M);
    const int64_t block = N / M;
    // sqrt(Float64(block)^3): block^3 as x*x*x, exact for any block below 2^17
    const double b = (double)block;
    const double norm = sqrt(b * b * b);
    hipLaunchKernelGGL(downsample_kernel, dim3(grid_for((int64_t)M * M * M)), dim3(256), 0,
                       ctx->stream, seed, N, M, block, norm, coarse);
    HNT_HIP(hipGetLastError());
    return finish_call(ctx, "hnt_downsample_noise");
}

extern "C" hnt_status hnt_extended_residual(hnt_ctx* ctx, int32_t it, int32_t jt, int32_t kt,
                                            int32_t nsub, int32_t nmesh, int64_t N,
                                            uint64_t seed, const float* coarse, int32_t M,
                                            int32_t nshell, float* out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(out && coarse, "hnt_extended_residual: NULL argument");
    HNT_REQUIRE(M >= 1 && N >= M && N % M == 0,
                "hnt_extended_residual: N = %lld must be a multiple of M = %d", (long long)N, M);
    HNT_REQUIRE(nsub >= 1 && nmesh >= 1 && nshell >= 0 && nmesh + 2 * nshell <= 16384 &&
                    it >= 1 && jt >= 1 && kt >= 1,
                "hnt_extended_residual: bad sizes (nsub %d, nmesh %d, nshell %d, tile %d %d %d)",
                nsub, nmesh, nshell, it, jt, kt);
    const int64_t block = N / M;
    const int64_t nbuff = (nmesh - nsub) / 2;
    const int side = nmesh + 2 * nshell;
    const int64_t i0 = (int64_t)(it - 1) * nsub + 1 - nbuff - nshell;
    const int64_t j0 = (int64_t)(jt - 1) * nsub + 1 - nbuff - nshell;
    const int64_t k0 = (int64_t)(kt - 1) * nsub + 1 - nbuff - nshell;
    // Float32(1 / sqrt(Float64(block)^3)): the block mean from the coarse block sum
    const double b = (double)block;
    const float inv = (float)(1.0 / sqrt(b * b * b));
    const int64_t total = (int64_t)side * side * side;
    hipLaunchKernelGGL(region_noise_kernel, dim3(grid_for(total)), dim3(256), 0, ctx->stream,
                       seed, N, i0, j0, k0, side, coarse, M, block, inv, out);
    HNT_HIP(hipGetLastError());
    return finish_call(ctx, "hnt_extended_residual");
}

// fields.hip -- tile field synthesis: isolated (zero-padded) convolutions with the tabulated
// sqrt P(k) and the LPT transfers, periodic 2LPT and Laplacian, isolated Poisson displacement
// and short-range potential, and the tri-cubic coarse -> tile interpolation.
//
// Replaces the CUDA methods of ext/CUDAExt.jl:2346-2825 and 3062-3176 (stubs
// src/Huntian.jl:283-385), called by the tile driver src/MultiResolution.jl:881-1010.
//
// Every transform is one hipFFT R2C or C2R over the whole tile (plans cached per size).  The
// k-space steps are single passes that read the shared spectrum and write the per-output
// copy, so one forward transform serves all outputs without a separate copy; the crop, the
// sign flip and the 2LPT source update are fused into one pass each over the real field.
// k-space arithmetic follows the reference kernels operation for operation (Float64
// frequencies and coefficients, Float32 spectrum); the tri-cubic interpolation is Float32 in
// the reference's order, so it matches the reference bit for bit.
#include <hipfft/hipfft.h>

#include <cmath>

#include "hnt_internal.h"

using namespace hnt;

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

constexpr int FT = 256;  // threads per workgroup

unsigned grid_for(int64_t n) {
    return (unsigned)std::min<int64_t>((n + FT - 1) / FT, 1 << 20);
}

// _interp_pk_table (ext/CUDAExt.jl:2351-2364), Float32
__device__ __forceinline__ float interp_pk(const float* __restrict__ tab, float log_k,
                                           float log_k_min, float d_log_k, int ntab) {
    const float fpos = (log_k - log_k_min) / d_log_k;
    const int i0 = (int)floorf(fpos) + 1;  // 1-based
    if (i0 < 1) return tab[0];
    if (i0 >= ntab) return tab[ntab - 1];
    const float frac = fpos - (float)(i0 - 1);
    return tab[i0 - 1] + frac * (tab[i0] - tab[i0 - 1]);
}

__device__ __forceinline__ double pick(int d, double kx, double ky, double kz) {
    return d == 1 ? kx : (d == 2 ? ky : kz);
}

// Rows: every kernel below walks whole rows along x, one wavefront per row (the row index is
// a 32-bit (iy, iz) pair, so no 64-bit division per element), lanes over x.
constexpr int WAVES = FT / 64;

#define FOR_ROWS(nrows)                                                                     \
    for (int row = (int)blockIdx.x * WAVES + (int)(threadIdx.x >> 6); row < (nrows);        \
         row += (int)gridDim.x * WAVES)

unsigned grid_rows(int64_t nrows) {
    return (unsigned)std::min<int64_t>((nrows + WAVES - 1) / WAVES, 1 << 20);
}

// _transfer_kernel! (ext/CUDAExt.jl:2367-2432): work = transfer(base), over the half spectrum
// (nk = n2/2 + 1, x fastest).  ky, kz of index >= n2/2 are negative (fftfreq).  The rows
// (iy, iz), (-iy, iz), (iy, -iz), (-iy, -iz) share k^2 bit for bit (the negated frequency is
// the exact negation), so one wavefront takes the group and looks sqrt P up once per ix.
__global__ void fs_transfer(const float2* __restrict__ base, float2* __restrict__ work, int n2,
                            double dk, const float* __restrict__ tab, int ntab, float log_k_min,
                            float d_log_k, int id, int d1, int d2) {
    const int nk = n2 / 2 + 1, half = n2 / 2;
    const int lane = threadIdx.x & 63;
    FOR_ROWS((half + 1) * (half + 1)) {
        const int ay = row % (half + 1), az = row / (half + 1);
        // the group's rows: index and signed frequency; index half is -half and unpaired
        int ry[2], rz[2];
        double sy[2], sz[2];
        int ny = 1, nz = 1;
        ry[0] = ay;
        sy[0] = ay < half ? (double)ay * dk : (double)(ay - n2) * dk;
        if (ay != 0 && ay != half) {
            ry[1] = n2 - ay;
            sy[1] = (double)(-ay) * dk;
            ny = 2;
        }
        rz[0] = az;
        sz[0] = az < half ? (double)az * dk : (double)(az - n2) * dk;
        if (az != 0 && az != half) {
            rz[1] = n2 - az;
            sz[1] = (double)(-az) * dk;
            nz = 2;
        }
        const double kyz = sy[0] * sy[0];
        const double kzz = sz[0] * sz[0];
        for (int ix = lane; ix < nk; ix += 64) {
            const double kx = (double)ix * dk;
            const double k2 = kx * kx + kyz + kzz;
            const double sqrt_pk =
                k2 == 0.0 ? 0.0
                          : (double)interp_pk(tab, (float)(log(k2) * 0.5), log_k_min, d_log_k,
                                              ntab);
            for (int qz = 0; qz < nz; ++qz)
                for (int qy = 0; qy < ny; ++qy) {
                    const int64_t t = ((int64_t)ry[qy] + (int64_t)n2 * rz[qz]) * nk + ix;
                    if (k2 == 0.0) {
                        work[t] = make_float2(0.0f, 0.0f);
                        continue;
                    }
                    const double ky = sy[qy], kz = sz[qz];
                    const float2 v = base[t];
                    float2 o;
                    if (id == 0) {
                        o = make_float2((float)((double)v.x * sqrt_pk),
                                        (float)((double)v.y * sqrt_pk));
                    } else if (id == 1 || id == 2) {
                        const double coeff = sqrt_pk * pick(d1, kx, ky, kz) / k2;
                        const double a = v.x, b = v.y;
                        o = id == 1 ? make_float2((float)(-b * coeff), (float)(a * coeff))
                                    : make_float2((float)(b * coeff), (float)(-a * coeff));
                    } else {
                        float c;
                        if (id == 3)
                            c = (float)(-sqrt_pk * pick(d1, kx, ky, kz) * pick(d2, kx, ky, kz) /
                                        k2);
                        else if (id == 4)
                            c = (float)(sqrt_pk * k2);
                        else
                            c = (float)(-sqrt_pk / k2);
                        o = make_float2(v.x * c, v.y * c);
                    }
                    work[t] = o;
                }
        }
    }
}

// _hermitize_planes_kernel! (ext/CUDAExt.jl:2439-2473) on the kx = 0 and kx = n2/2 planes:
// the lower cell of each conjugate pair (iy, iz) ~ (-iy, -iz) writes both.
__global__ void fs_hermitize(float2* __restrict__ a, int n2) {
    const int nk = n2 / 2 + 1;
    const int64_t total = (int64_t)n2 * n2;
    for (int64_t t = (int64_t)blockIdx.x * FT + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * FT) {
        const int iy = (int)(t % n2), iz = (int)(t / n2);
        const int iyc = iy == 0 ? 0 : n2 - iy, izc = iz == 0 ? 0 : n2 - iz;
        const bool self = iy == iyc && iz == izc;
        const bool lower = iy < iyc || (iy == iyc && iz < izc);
        if (!(lower || self)) continue;
        for (int p = 0; p < 2; ++p) {
            const int ix = p == 0 ? 0 : nk - 1;
            float2* me = a + ix + (int64_t)nk * (iy + (int64_t)n2 * iz);
            if (self) {
                me->y = 0.0f;
            } else {
                float2* co = a + ix + (int64_t)nk * (iyc + (int64_t)n2 * izc);
                const float2 vm = *me, vc = *co;
                const float2 avg = make_float2((vm.x + vc.x) * 0.5f, (vm.y - vc.y) * 0.5f);
                *me = avg;
                *co = make_float2(avg.x, -avg.y);
            }
        }
    }
}

// _periodic_kernel! (ext/CUDAExt.jl:2629-2677): dst = kernel(src) on an n^3 half spectrum;
// k = 0 zeroed, and the Nyquist planes too except for phi_ij (id 3).  ky, kz of index n/2
// count as +n/2 (iy <= nyq + 1, 1-based).  id 2: -i k_d1/k^2, 3: -k_d1 k_d2/k^2, 4: k^2.
__global__ void fs_periodic(const float2* __restrict__ src, float2* __restrict__ dst, int n,
                            double dk, int id, int d1, int d2) {
    const int nk = n / 2 + 1, nyq = n / 2;
    const int lane = threadIdx.x & 63;
    FOR_ROWS(n * n) {
        const int iy = row % n, iz = row / n;
        const double ky = (double)(iy <= nyq ? iy : iy - n) * dk;
        const double kz = (double)(iz <= nyq ? iz : iz - n) * dk;
        const double kyz = ky * ky;
        const bool nyq_row = iy == nyq || iz == nyq;
        const int64_t r0 = (int64_t)row * nk;
        for (int ix = lane; ix < nk; ix += 64) {
            const int64_t t = r0 + ix;
            const double kx = (double)ix * dk;
            const double k2 = kx * kx + kyz + kz * kz;
            if (k2 == 0.0 || (id != 3 && (ix == nk - 1 || nyq_row))) {
                dst[t] = make_float2(0.0f, 0.0f);
                continue;
            }
            const float2 v = src[t];
            float2 o;
            if (id == 2) {
                const double coeff = pick(d1, kx, ky, kz) / k2;
                const double a = v.x, b = v.y;
                o = make_float2((float)(b * coeff), (float)(-a * coeff));
            } else {
                const float c = id == 3
                                    ? (float)(-pick(d1, kx, ky, kz) * pick(d2, kx, ky, kz) / k2)
                                    : (float)k2;
                o = make_float2(v.x * c, v.y * c);
            }
            dst[t] = o;
        }
    }
}

// _short_pot_kernel (ext/CUDAExt.jl:2722), in place on a (2n)^3 half spectrum; the index
// n2/2 counts as +n2/2 (only k^2 enters).
__global__ void fs_short_pot(float2* __restrict__ a, int n2, double dk, double rs2) {
    const int nk = n2 / 2 + 1, half = n2 / 2;
    const int lane = threadIdx.x & 63;
    FOR_ROWS(n2 * n2) {
        const int iy = row % n2, iz = row / n2;
        const double ky = (double)(iy <= half ? iy : iy - n2) * dk;
        const double kz = (double)(iz <= half ? iz : iz - n2) * dk;
        const double kyz = ky * ky;
        const int64_t r0 = (int64_t)row * nk;
        for (int ix = lane; ix < nk; ix += 64) {
            const double kx = (double)ix * dk;
            const double k2 = kx * kx + kyz + kz * kz;
            const float c = (float)(k2 == 0.0 ? -rs2 : -(1.0 - exp(-k2 * rs2)) / k2);
            const float2 v = a[r0 + ix];
            a[r0 + ix] = make_float2(v.x * c, v.y * c);
        }
    }
}

// dst (n2^3) = src (nin^3) in the low corner, zeros elsewhere
__global__ void fs_pad(const float* __restrict__ src, int nin, int n2, float* __restrict__ dst) {
    const int lane = threadIdx.x & 63;
    FOR_ROWS(n2 * n2) {
        const int iy = row % n2, iz = row / n2;
        float* d = dst + (int64_t)row * n2;
        if (iy < nin && iz < nin) {
            const float* sr = src + (int64_t)nin * (iy + (int64_t)nin * iz);
            for (int ix = lane; ix < n2; ix += 64) d[ix] = ix < nin ? sr[ix] : 0.0f;
        } else {
            for (int ix = lane; ix < n2; ix += 64) d[ix] = 0.0f;
        }
    }
}

// out (n^3) = scale * big[s + i, s + j, s + k] of an m^3 field (scale: the inverse
// transform's 1/m^3, negated where the reference negates the result); in place when m == n
// and s == 0 (each element is read and written by one lane)
__global__ void fs_crop(const float* big, int m, int n, int s, float scale, float* out) {
    const int lane = threadIdx.x & 63;
    FOR_ROWS(n * n) {
        const int j = row % n, k = row / n;
        const float* b = big + (int64_t)m * ((s + j) + (int64_t)m * (s + k)) + s;
        float* o = out + (int64_t)row * n;
        for (int i = lane; i < n; i += 64) o[i] = b[i] * scale;
    }
}

// src2 = delta^2 * 0.5 (Float32, compute_2lpt_gpu's broadcast)
__global__ void fs_src2_init(const float* __restrict__ d, int64_t n3, float* __restrict__ s2) {
    for (int64_t t = (int64_t)blockIdx.x * FT + threadIdx.x; t < n3;
         t += (int64_t)gridDim.x * FT) {
        const float v = d[t];
        s2[t] = (v * v) * 0.5f;
    }
}

// src2 -= phi^2 * coef, phi = the unnormalised inverse transform times scale (1/n^3)
__global__ void fs_src2_sub(const float* __restrict__ phi, int64_t n3, float scale, float coef,
                            float* __restrict__ s2) {
    for (int64_t t = (int64_t)blockIdx.x * FT + threadIdx.x; t < n3;
         t += (int64_t)gridDim.x * FT) {
        const float v = phi[t] * scale;
        s2[t] = s2[t] - (v * v) * coef;
    }
}

// mod1(i, M) of the reference (_mod1_i32, ext/CUDAExt.jl:3074-3081)
__device__ __forceinline__ int mod1(int i, int M) {
    int r = (i - 1) % M;
    if (r < 0) r += M;
    return r + 1;
}

// _catmull_rom_f32 (ext/CUDAExt.jl:3065-3072)
__device__ __forceinline__ void catmull_rom(float t, float w[4]) {
    const float t2 = t * t, t3 = t2 * t;
    w[0] = -0.5f * t3 + t2 - 0.5f * t;
    w[1] = 1.5f * t3 - 2.5f * t2 + 1.0f;
    w[2] = -1.5f * t3 + 2.0f * t2 + 0.5f * t;
    w[3] = 0.5f * t3 - 0.5f * t2;
}

// Catmull-Rom weights and wrapped coarse indices (0-based) of global fine cell g (1-based)
__device__ __forceinline__ void stencil(int g, int block, int M, float w[4], int idx[4]) {
    const double f = ((double)g - 0.5) / (double)block + 0.5;
    const int i = (int)floor(f);
    catmull_rom((float)(f - (double)i), w);
    for (int q = 0; q < 4; ++q) idx[q] = mod1(i - 1 + q, M) - 1;
}

// The stencils of the tile's cells along each axis (axis a, 0-based cell l: global cell
// mod1(start_a + l, N)), computed once per call: Catmull-Rom weights w[a * nmesh + l] and
// coarse indices idx[a * nmesh + l].  Every row of the tile shares the x stencils, every
// plane the y ones.
__global__ void fs_tricubic_axes(int M, int N, int block, int i0, int j0, int k0, int nmesh,
                                 float4* __restrict__ w, int4* __restrict__ idx) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= 3 * nmesh) return;
    const int axis = t / nmesh, l = t - axis * nmesh;
    const int s = axis == 0 ? i0 : (axis == 1 ? j0 : k0);
    float ww[4];
    int ii[4];
    stencil(mod1(s + l, N), block, M, ww, ii);
    w[t] = make_float4(ww[0], ww[1], ww[2], ww[3]);
    idx[t] = make_int4(ii[0], ii[1], ii[2], ii[3]);
}

// _tricubic_kernel! (ext/CUDAExt.jl:3083-3140): one wavefront per tile row (lj, lk), whose y
// and z weights and coarse rows are the same for every lane.  Each lane takes a run of
// consecutive cells of the row: cells of one coarse cell (N / M of them) share the x indices,
// so the lane keeps the 4 x 4 x 4 coarse values in registers and reloads them only when the
// x indices change.  The sum runs over (z, y, x) with ((wx wy) wz) f as the reference does.
__global__ __launch_bounds__(FT) void fs_tricubic(const float* __restrict__ coarse, int M,
                                                  int nmesh, const float4* __restrict__ sw,
                                                  const int4* __restrict__ si,
                                                  float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int run = (nmesh + 63) >> 6;
    const int l0 = min(lane * run, nmesh), l1 = min(l0 + run, nmesh);
    FOR_ROWS(nmesh * nmesh) {
        const int lj = row % nmesh, lk = row / nmesh;
        const float4 wy4 = sw[nmesh + lj], wz4 = sw[2 * nmesh + lk];
        const int4 ys4 = si[nmesh + lj], zs4 = si[2 * nmesh + lk];
        const float wy[4] = {wy4.x, wy4.y, wy4.z, wy4.w}, wz[4] = {wz4.x, wz4.y, wz4.z, wz4.w};
        const int ys[4] = {ys4.x, ys4.y, ys4.z, ys4.w}, zs[4] = {zs4.x, zs4.y, zs4.z, zs4.w};
        float* o = out + (int64_t)row * nmesh;
        float cv[4][4][4];
        int4 have = make_int4(-1, -1, -1, -1);
        for (int li = l0; li < l1; ++li) {
            const float4 wx4 = sw[li];
            const int4 xs4 = si[li];
            if (xs4.x != have.x || xs4.y != have.y || xs4.z != have.z || xs4.w != have.w) {
                const int xs[4] = {xs4.x, xs4.y, xs4.z, xs4.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float* cz = coarse + (int64_t)M * M * zs[c];
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const float* cy = cz + (int64_t)M * ys[b];
#pragma unroll
                        for (int a = 0; a < 4; ++a) cv[c][b][a] = cy[xs[a]];
                    }
                }
                have = xs4;
            }
            const float wx[4] = {wx4.x, wx4.y, wx4.z, wx4.w};
            float val = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int b = 0; b < 4; ++b)
#pragma unroll
                    for (int a = 0; a < 4; ++a) val += wx[a] * wy[b] * wz[c] * cv[c][b][a];
            o[li] = val;
        }
    }
}

// fs_tricubic for a power-of-two block (fine cells per coarse cell) of at most TRI_MAXB.
// Then a cell's x stencil depends only on its position in its coarse cell, which is that of
// tile cell li mod block: (g - 1/2) / block + 1/2 is exact in double, so cells li and
// li mod block have the same weights bit for bit.  Each row forms the 64 products
// ((wx wy) wz) of each of the block positions once, into the wavefront's LDS table, and a
// cell takes 64 multiplies and adds instead of 208 operations: the same products summed in
// the same order, so the same values bit for bit.
constexpr int TRI_MAXB = 32;

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(FT) void fs_tricubic_tab(const float* __restrict__ coarse, int M,
                                                      int nmesh, int block,
                                                      const float4* __restrict__ sw,
                                                      const int4* __restrict__ si,
                                                      float* __restrict__ out) {
    __shared__ float4 wtab[WAVES][TRI_MAXB * 16];  // [position][4 z + y], x in the float4
    const int lane = threadIdx.x & 63;
    float4* W = wtab[threadIdx.x >> 6];
    const int run = (nmesh + 63) >> 6;
    const int l0 = min(lane * run, nmesh), l1 = min(l0 + run, nmesh);
    FOR_ROWS(nmesh * nmesh) {
        const int lj = row % nmesh, lk = row / nmesh;
        const float4 wy4 = sw[nmesh + lj], wz4 = sw[2 * nmesh + lk];
        const int4 ys4 = si[nmesh + lj], zs4 = si[2 * nmesh + lk];
        const float wy[4] = {wy4.x, wy4.y, wy4.z, wy4.w}, wz[4] = {wz4.x, wz4.y, wz4.z, wz4.w};
        const int ys[4] = {ys4.x, ys4.y, ys4.z, ys4.w}, zs[4] = {zs4.x, zs4.y, zs4.z, zs4.w};
        wave_lds_fence();  // the previous row's table reads are done
        for (int e = lane; e < block * 16; e += 64) {
            const int q = e >> 4, c = (e >> 2) & 3, b = e & 3;
            const float4 wx = sw[q];
            const float y = wy[b], z = wz[c];
            W[e] = make_float4(wx.x * y * z, wx.y * y * z, wx.z * y * z, wx.w * y * z);
        }
        wave_lds_fence();
        float* o = out + (int64_t)row * nmesh;
        float cv[4][4][4];
        int4 have = make_int4(-1, -1, -1, -1);
        for (int li = l0; li < l1; ++li) {
            const int4 xs4 = si[li];
            if (xs4.x != have.x || xs4.y != have.y || xs4.z != have.z || xs4.w != have.w) {
                const int xs[4] = {xs4.x, xs4.y, xs4.z, xs4.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float* cz = coarse + (int64_t)M * M * zs[c];
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const float* cy = cz + (int64_t)M * ys[b];
#pragma unroll
                        for (int a = 0; a < 4; ++a) cv[c][b][a] = cy[xs[a]];
                    }
                }
                have = xs4;
            }
            const float4* Wq = W + 16 * (li & (block - 1));
            float val = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const float4 w = Wq[4 * c + b];
                    val += w.x * cv[c][b][0];
                    val += w.y * cv[c][b][1];
                    val += w.z * cv[c][b][2];
                    val += w.w * cv[c][b][3];
                }
            o[li] = val;
        }
    }
}

// hipFFT R2C / C2R plans of an n^3 real field (cached per size on the context).  The 32-bit
// plan call gives wrong transforms from 2^31 elements on (2048^3: an impulse response loses
// its axis symmetry at 1e-3, a smooth field is off by half its norm), so larger sizes are
// refused; the library's own FFTs cover the power-of-two sizes up to 2048.
hnt_status plans_for(hnt_ctx* ctx, int n, hipfftHandle& r2c, hipfftHandle& c2r) {
    if ((int64_t)n * n * n >= (1LL << 31)) {
        set_error("hipFFT path: %d^3 real transforms are not supported (2^31 elements or more); "
                  "power-of-two sizes up to 2048 use the library's FFTs (tuning fft = 0)", n);
        return HNT_ERR_UNSUPPORTED;
    }
    for (hnt_ctx::FftPlans& p : ctx->fs_plans)
        if (p.n == n) {
            r2c = (hipfftHandle)p.r2c;
            c2r = (hipfftHandle)p.c2r;
            HNT_FFT(hipfftSetStream(r2c, ctx->stream));
            HNT_FFT(hipfftSetStream(c2r, ctx->stream));
            return HNT_OK;
        }
    // keep at most two sizes (the padded and the periodic tile)
    if (ctx->fs_plans.size() >= 2) release_field_plans(ctx);
    r2c = nullptr;
    c2r = nullptr;
    HNT_FFT(hipfftPlan3d(&r2c, n, n, n, HIPFFT_R2C));
    const hipfftResult rc = hipfftPlan3d(&c2r, n, n, n, HIPFFT_C2R);
    if (rc != HIPFFT_SUCCESS) {
        (void)hipfftDestroy(r2c);
        set_error("hipfftPlan3d C2R %d^3: hipFFT status %d", n, (int)rc);
        return HNT_ERR_DEVICE;
    }
    ctx->fs_plans.push_back({n, (void*)r2c, (void*)c2r});
    HNT_FFT(hipfftSetStream(r2c, ctx->stream));
    HNT_FFT(hipfftSetStream(c2r, ctx->stream));
    return HNT_OK;
}

int64_t cube(int n) { return (int64_t)n * n * n; }
// elements of a half spectrum buffer: the library's padded rows (hipFFT's n/2 + 1 fit in them)
int64_t half_spectrum(int n) { return fft_plane(n) * n; }

// Zero-pad a device n^3 field to (2n)^3 in fs_real and transform it into fs_base.
hnt_status padded_rfft(hnt_ctx* ctx, const float* f, int nin, int n2, hipfftHandle r2c) {
    hipStream_t st = ctx->stream;
    HNT_TRY(reserve(ctx, ctx->fs_real, sizeof(float) * (size_t)cube(n2)));
    HNT_TRY(reserve(ctx, ctx->fs_base, sizeof(float2) * (size_t)half_spectrum(n2)));
    hipLaunchKernelGGL(fs_pad, dim3(grid_rows((int64_t)n2 * n2)), dim3(FT), 0, st, f, nin, n2,
                       (float*)ctx->fs_real.p);
    HNT_FFT(hipfftExecR2C(r2c, (hipfftReal*)ctx->fs_real.p, (hipfftComplex*)ctx->fs_base.p));
    return HNT_OK;
}

}  // namespace

extern "C" hnt_status hnt_isolated_convolve(hnt_ctx* ctx, const float* noise, int32_t n_input,
                                            int32_t n, double boxsize_local,
                                            const hnt_pk_table* table,
                                            const hnt_transfer* kernels, int32_t nkernels,
                                            int32_t nshell, float* const* out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(noise && table && table->sqrt_pk && kernels && out,
                "hnt_isolated_convolve: NULL argument");
    HNT_REQUIRE(n >= 1 && nshell >= 0 && n_input == n + 2 * nshell && n_input <= 2048,
                "hnt_isolated_convolve: n_input %d must be n + 2 nshell (n %d, nshell %d), "
                "at most 2048", n_input, n, nshell);
    HNT_REQUIRE(boxsize_local > 0.0 && table->n >= 2 && table->d_log_k > 0.0f,
                "hnt_isolated_convolve: boxsize %g, table of %d points, step %g", boxsize_local,
                table->n, (double)table->d_log_k);
    HNT_REQUIRE(nkernels >= 1, "hnt_isolated_convolve: no transfer");
    for (int q = 0; q < nkernels; ++q) {
        const hnt_transfer& k = kernels[q];
        HNT_REQUIRE(out[q] && k.id >= 0 && k.id <= 5 &&
                        ((k.id != 1 && k.id != 2 && k.id != 3) || (k.dim1 >= 1 && k.dim1 <= 3)) &&
                        (k.id != 3 || (k.dim2 >= 1 && k.dim2 <= 3)),
                    "hnt_isolated_convolve: transfer %d: id %d dims %d %d", q, k.id, k.dim1,
                    k.dim2);
    }
    hipStream_t st = ctx->stream;
    const int n2 = 2 * n_input;
    const double dx = boxsize_local / n;
    const double dk = 2.0 * M_PI / (n2 * dx);
    HNT_TRY(reserve(ctx, ctx->fs_tab, sizeof(float) * (size_t)table->n));
    HNT_HIP(hipMemcpyAsync(ctx->fs_tab.p, table->sqrt_pk, sizeof(float) * (size_t)table->n,
                           hipMemcpyHostToDevice, st));
    if (ctx->fft_mode == 0 && fft_supported(n2)) {
        // library FFTs: pruned forward transform, one sqrt P table, pruned inverse per output
        const int64_t nhalf = n2 / 2 + 1;
        HNT_TRY(reserve(ctx, ctx->fs_base, sizeof(float2) * (size_t)half_spectrum(n2)));
        HNT_TRY(reserve(ctx, ctx->fs_work, sizeof(float2) * (size_t)half_spectrum(n2)));
        HNT_TRY(reserve(ctx, ctx->fs_aux, sizeof(float) * (size_t)(nhalf * nhalf * nhalf)));
        float2* base = (float2*)ctx->fs_base.p;
        HNT_TRY(fft_forward(ctx, noise, n_input, n2, base));
        HNT_TRY(fft_amp_table(ctx, n2, dk, (const float*)ctx->fs_tab.p, table->n,
                              table->log_k_min, table->d_log_k, (float*)ctx->fs_aux.p));
        const float scale = (float)(1.0 / (double)cube(n2));
        for (int q = 0; q < nkernels; ++q) {
            FftOp op;
            op.kind = FFT_OP_TRANSFER;
            op.L = n2;
            op.dk = dk;
            op.amp = (const float*)ctx->fs_aux.p;
            op.id = kernels[q].id;
            op.d1 = kernels[q].dim1;
            op.d2 = kernels[q].dim2;
            HNT_TRY(fft_inverse(ctx, base, (float2*)ctx->fs_work.p, n2, nshell, n, op, scale, 0,
                                0.0f, out[q]));
        }
        return finish_call(ctx, "hnt_isolated_convolve");
    }
    hipfftHandle r2c, c2r;
    HNT_TRY(plans_for(ctx, n2, r2c, c2r));
    HNT_TRY(padded_rfft(ctx, noise, n_input, n2, r2c));
    HNT_TRY(reserve(ctx, ctx->fs_work, sizeof(float2) * (size_t)half_spectrum(n2)));
    for (int q = 0; q < nkernels; ++q) {
        const hnt_transfer& k = kernels[q];
        hipLaunchKernelGGL(fs_transfer, dim3(grid_rows((int64_t)(n2 / 2 + 1) * (n2 / 2 + 1))),
                           dim3(FT), 0, st,
                           (const float2*)ctx->fs_base.p, (float2*)ctx->fs_work.p, n2, dk,
                           (const float*)ctx->fs_tab.p, table->n, table->log_k_min,
                           table->d_log_k, k.id, k.dim1, k.dim2);
        hipLaunchKernelGGL(fs_hermitize, dim3(grid_for((int64_t)n2 * n2)), dim3(FT), 0, st,
                           (float2*)ctx->fs_work.p, n2);
        HNT_FFT(hipfftExecC2R(c2r, (hipfftComplex*)ctx->fs_work.p, (hipfftReal*)ctx->fs_real.p));
        hipLaunchKernelGGL(fs_crop, dim3(grid_rows((int64_t)n * n)), dim3(FT), 0, st,
                           (const float*)ctx->fs_real.p, n2, n, nshell,
                           (float)(1.0 / (double)cube(n2)), out[q]);
    }
    return finish_call(ctx, "hnt_isolated_convolve");
}

extern "C" hnt_status hnt_compute_2lpt(hnt_ctx* ctx, const float* delta, int32_t n,
                                       double boxsize_local, float* const* psi2) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(delta && psi2 && psi2[0] && psi2[1] && psi2[2], "hnt_compute_2lpt: NULL argument");
    HNT_REQUIRE(n >= 2 && n <= 4096 && boxsize_local > 0.0, "hnt_compute_2lpt: n %d, boxsize %g",
                n, boxsize_local);
    hipStream_t st = ctx->stream;
    const int64_t n3 = cube(n), nh = half_spectrum(n);
    const double dk = 2.0 * M_PI / boxsize_local;
    const float inv3 = (float)(1.0 / (double)n3);
    if (ctx->fft_mode == 0 && fft_supported(n)) {
        HNT_TRY(reserve(ctx, ctx->fs_aux, sizeof(float) * (size_t)n3));
        HNT_TRY(reserve(ctx, ctx->fs_base, sizeof(float2) * (size_t)nh));
        HNT_TRY(reserve(ctx, ctx->fs_work, sizeof(float2) * (size_t)nh));
        float* s2 = (float*)ctx->fs_aux.p;
        float2* base = (float2*)ctx->fs_base.p;
        float2* work = (float2*)ctx->fs_work.p;
        HNT_TRY(fft_forward(ctx, delta, n, n, base));
        hipLaunchKernelGGL(fs_src2_init, dim3(grid_for(n3)), dim3(FT), 0, st, delta, n3, s2);
        const int prs[6][2] = {{1, 1}, {2, 2}, {3, 3}, {1, 2}, {1, 3}, {2, 3}};
        FftOp op;
        op.kind = FFT_OP_PERIODIC;
        op.L = n;
        op.dk = dk;
        for (int q = 0; q < 6; ++q) {
            op.id = 3;
            op.d1 = prs[q][0];
            op.d2 = prs[q][1];
            HNT_TRY(fft_inverse(ctx, base, work, n, 0, n, op, inv3, 1, q < 3 ? 0.5f : 1.0f, s2));
        }
        HNT_TRY(fft_forward(ctx, s2, n, n, base));
        for (int d = 0; d < 3; ++d) {
            op.id = 2;
            op.d1 = d + 1;
            op.d2 = 0;
            HNT_TRY(fft_inverse(ctx, base, work, n, 0, n, op, inv3, 0, 0.0f, psi2[d]));
        }
        return finish_call(ctx, "hnt_compute_2lpt");
    }
    hipfftHandle r2c, c2r;
    HNT_TRY(plans_for(ctx, n, r2c, c2r));
    HNT_TRY(reserve(ctx, ctx->fs_real, sizeof(float) * (size_t)n3));
    HNT_TRY(reserve(ctx, ctx->fs_aux, sizeof(float) * (size_t)n3));
    HNT_TRY(reserve(ctx, ctx->fs_base, sizeof(float2) * (size_t)nh));
    HNT_TRY(reserve(ctx, ctx->fs_work, sizeof(float2) * (size_t)nh));
    float* real = (float*)ctx->fs_real.p;
    float* src2 = (float*)ctx->fs_aux.p;
    float2* base = (float2*)ctx->fs_base.p;
    float2* work = (float2*)ctx->fs_work.p;
    const unsigned g3 = grid_for(n3), gh = grid_rows((int64_t)n * n), gc = gh;
    // the transform may use its input as scratch: delta is copied first
    HNT_HIP(hipMemcpyAsync(real, delta, sizeof(float) * (size_t)n3, hipMemcpyDeviceToDevice, st));
    HNT_FFT(hipfftExecR2C(r2c, (hipfftReal*)real, (hipfftComplex*)base));
    hipLaunchKernelGGL(fs_src2_init, dim3(g3), dim3(FT), 0, st, delta, n3, src2);
    const float inv = (float)(1.0 / (double)n3);
    // src2 = delta^2/2 - sum_i phi_ii^2/2 - sum_{i<j} phi_ij^2 (compute_2lpt_gpu's order)
    const int pairs[6][2] = {{1, 1}, {2, 2}, {3, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int q = 0; q < 6; ++q) {
        hipLaunchKernelGGL(fs_periodic, dim3(gh), dim3(FT), 0, st, (const float2*)base, work, n,
                           dk, 3, pairs[q][0], pairs[q][1]);
        HNT_FFT(hipfftExecC2R(c2r, (hipfftComplex*)work, (hipfftReal*)real));
        hipLaunchKernelGGL(fs_src2_sub, dim3(g3), dim3(FT), 0, st, (const float*)real, n3, inv,
                           q < 3 ? 0.5f : 1.0f, src2);
    }
    HNT_FFT(hipfftExecR2C(r2c, (hipfftReal*)src2, (hipfftComplex*)base));
    for (int d = 0; d < 3; ++d) {
        hipLaunchKernelGGL(fs_periodic, dim3(gh), dim3(FT), 0, st, (const float2*)base, work, n,
                           dk, 2, d + 1, 0);
        HNT_FFT(hipfftExecC2R(c2r, (hipfftComplex*)work, (hipfftReal*)psi2[d]));
        hipLaunchKernelGGL(fs_crop, dim3(gc), dim3(FT), 0, st, (const float*)psi2[d], n, n, 0,
                           inv, psi2[d]);
    }
    return finish_call(ctx, "hnt_compute_2lpt");
}

extern "C" hnt_status hnt_compute_laplacian(hnt_ctx* ctx, const float* delta, int32_t n,
                                            double boxsize_local, float* lapd) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(delta && lapd, "hnt_compute_laplacian: NULL argument");
    HNT_REQUIRE(n >= 2 && n <= 4096 && boxsize_local > 0.0,
                "hnt_compute_laplacian: n %d, boxsize %g", n, boxsize_local);
    hipStream_t st = ctx->stream;
    const int64_t n3 = cube(n), nh = half_spectrum(n);
    if (ctx->fft_mode == 0 && fft_supported(n)) {
        HNT_TRY(reserve(ctx, ctx->fs_base, sizeof(float2) * (size_t)nh));
        HNT_TRY(reserve(ctx, ctx->fs_work, sizeof(float2) * (size_t)nh));
        HNT_TRY(fft_forward(ctx, delta, n, n, (float2*)ctx->fs_base.p));
        FftOp op;
        op.kind = FFT_OP_PERIODIC;
        op.L = n;
        op.dk = 2.0 * M_PI / boxsize_local;
        op.id = 4;
        HNT_TRY(fft_inverse(ctx, (const float2*)ctx->fs_base.p, (float2*)ctx->fs_work.p, n, 0, n,
                            op, (float)(1.0 / (double)n3), 0, 0.0f, lapd));
        return finish_call(ctx, "hnt_compute_laplacian");
    }
    hipfftHandle r2c, c2r;
    HNT_TRY(plans_for(ctx, n, r2c, c2r));
    HNT_TRY(reserve(ctx, ctx->fs_real, sizeof(float) * (size_t)n3));
    HNT_TRY(reserve(ctx, ctx->fs_base, sizeof(float2) * (size_t)nh));
    HNT_TRY(reserve(ctx, ctx->fs_work, sizeof(float2) * (size_t)nh));
    float* real = (float*)ctx->fs_real.p;
    float2* base = (float2*)ctx->fs_base.p;
    float2* work = (float2*)ctx->fs_work.p;
    HNT_HIP(hipMemcpyAsync(real, delta, sizeof(float) * (size_t)n3, hipMemcpyDeviceToDevice, st));
    HNT_FFT(hipfftExecR2C(r2c, (hipfftReal*)real, (hipfftComplex*)base));
    hipLaunchKernelGGL(fs_periodic, dim3(grid_rows((int64_t)n * n)), dim3(FT), 0, st, (const float2*)base,
                       work, n, 2.0 * M_PI / boxsize_local, 4, 0, 0);
    HNT_FFT(hipfftExecC2R(c2r, (hipfftComplex*)work, (hipfftReal*)lapd));
    hipLaunchKernelGGL(fs_crop, dim3(grid_rows((int64_t)n * n)), dim3(FT), 0, st, (const float*)lapd, n, n, 0,
                       (float)(1.0 / (double)n3), lapd);
    return finish_call(ctx, "hnt_compute_laplacian");
}

extern "C" hnt_status hnt_isolated_poisson_psi(hnt_ctx* ctx, const float* dS, int32_t n,
                                               double boxsize_local, float* const* psi) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(dS && psi && psi[0] && psi[1] && psi[2], "hnt_isolated_poisson_psi: NULL argument");
    HNT_REQUIRE(n >= 1 && n <= 2048 && boxsize_local > 0.0,
                "hnt_isolated_poisson_psi: n %d, boxsize %g", n, boxsize_local);
    hipStream_t st = ctx->stream;
    const int n2 = 2 * n;
    const double dk = 2.0 * M_PI / (2.0 * boxsize_local);
    if (ctx->fft_mode == 0 && fft_supported(n2)) {
        HNT_TRY(reserve(ctx, ctx->fs_base, sizeof(float2) * (size_t)half_spectrum(n2)));
        HNT_TRY(reserve(ctx, ctx->fs_work, sizeof(float2) * (size_t)half_spectrum(n2)));
        HNT_TRY(fft_forward(ctx, dS, n, n2, (float2*)ctx->fs_base.p));
        FftOp op;
        op.kind = FFT_OP_PERIODIC;
        op.L = n2;
        op.dk = dk;
        op.id = 2;
        for (int d = 0; d < 3; ++d) {
            op.d1 = d + 1;
            // the 1LPT displacement is the negative of kernel 2's output
            HNT_TRY(fft_inverse(ctx, (const float2*)ctx->fs_base.p, (float2*)ctx->fs_work.p, n2,
                                0, n, op, (float)(-1.0 / (double)cube(n2)), 0, 0.0f, psi[d]));
        }
        return finish_call(ctx, "hnt_isolated_poisson_psi");
    }
    hipfftHandle r2c, c2r;
    HNT_TRY(plans_for(ctx, n2, r2c, c2r));
    HNT_TRY(padded_rfft(ctx, dS, n, n2, r2c));
    HNT_TRY(reserve(ctx, ctx->fs_work, sizeof(float2) * (size_t)half_spectrum(n2)));
    for (int d = 0; d < 3; ++d) {
        hipLaunchKernelGGL(fs_periodic, dim3(grid_rows((int64_t)n2 * n2)), dim3(FT), 0, st,
                           (const float2*)ctx->fs_base.p, (float2*)ctx->fs_work.p, n2, dk, 2,
                           d + 1, 0);
        HNT_FFT(hipfftExecC2R(c2r, (hipfftComplex*)ctx->fs_work.p, (hipfftReal*)ctx->fs_real.p));
        // the 1LPT displacement is the negative of kernel 2's output
        hipLaunchKernelGGL(fs_crop, dim3(grid_rows((int64_t)n * n)), dim3(FT), 0, st,
                           (const float*)ctx->fs_real.p, n2, n, 0,
                           (float)(-1.0 / (double)cube(n2)), psi[d]);
    }
    return finish_call(ctx, "hnt_isolated_poisson_psi");
}

extern "C" hnt_status hnt_isolated_short_pot(hnt_ctx* ctx, const float* delta, int32_t n,
                                             double boxsize_local, double rs, float* out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(delta && out, "hnt_isolated_short_pot: NULL argument");
    HNT_REQUIRE(n >= 1 && n <= 2048 && boxsize_local > 0.0 && rs >= 0.0,
                "hnt_isolated_short_pot: n %d, boxsize %g, rs %g", n, boxsize_local, rs);
    hipStream_t st = ctx->stream;
    const int n2 = 2 * n;
    if (ctx->fft_mode == 0 && fft_supported(n2)) {
        HNT_TRY(reserve(ctx, ctx->fs_base, sizeof(float2) * (size_t)half_spectrum(n2)));
        HNT_TRY(reserve(ctx, ctx->fs_work, sizeof(float2) * (size_t)half_spectrum(n2)));
        HNT_TRY(fft_forward(ctx, delta, n, n2, (float2*)ctx->fs_base.p));
        FftOp op;
        op.kind = FFT_OP_SHORTPOT;
        op.L = n2;
        op.dk = 2.0 * M_PI / (2.0 * boxsize_local);
        op.rs2 = rs * rs;
        HNT_TRY(fft_inverse(ctx, (const float2*)ctx->fs_base.p, (float2*)ctx->fs_work.p, n2, 0,
                            n, op, (float)(1.0 / (double)cube(n2)), 0, 0.0f, out));
        return finish_call(ctx, "hnt_isolated_short_pot");
    }
    hipfftHandle r2c, c2r;
    HNT_TRY(plans_for(ctx, n2, r2c, c2r));
    HNT_TRY(padded_rfft(ctx, delta, n, n2, r2c));
    hipLaunchKernelGGL(fs_short_pot, dim3(grid_rows((int64_t)n2 * n2)), dim3(FT), 0, st,
                       (float2*)ctx->fs_base.p, n2, 2.0 * M_PI / (2.0 * boxsize_local), rs * rs);
    HNT_FFT(hipfftExecC2R(c2r, (hipfftComplex*)ctx->fs_base.p, (hipfftReal*)ctx->fs_real.p));
    hipLaunchKernelGGL(fs_crop, dim3(grid_rows((int64_t)n * n)), dim3(FT), 0, st,
                       (const float*)ctx->fs_real.p, n2, n, 0, (float)(1.0 / (double)cube(n2)),
                       out);
    return finish_call(ctx, "hnt_isolated_short_pot");
}

extern "C" hnt_status hnt_interpolate_to_tile(hnt_ctx* ctx, const float* coarse, int32_t M,
                                              int32_t N, int32_t it, int32_t jt, int32_t kt,
                                              int32_t nsub, int32_t nmesh, float* tile) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(coarse && tile, "hnt_interpolate_to_tile: NULL argument");
    HNT_REQUIRE(M >= 1 && N >= M && N % M == 0 && nmesh >= 1 && nsub >= 1 && nmesh >= nsub &&
                    nmesh <= 4096 && N < (1 << 30),
                "hnt_interpolate_to_tile: M %d, N %d, nsub %d, nmesh %d", M, N, nsub, nmesh);
    HNT_REQUIRE(it >= 1 && jt >= 1 && kt >= 1, "hnt_interpolate_to_tile: tile (%d, %d, %d)", it,
                jt, kt);
    const int block = N / M;
    const int nbuff = (nmesh - nsub) / 2;
    const int64_t i0 = (int64_t)(it - 1) * nsub + 1 - nbuff;
    const int64_t j0 = (int64_t)(jt - 1) * nsub + 1 - nbuff;
    const int64_t k0 = (int64_t)(kt - 1) * nsub + 1 - nbuff;
    HNT_REQUIRE(std::max(i0, std::max(j0, k0)) + nmesh < (1LL << 31),
                "hnt_interpolate_to_tile: tile start beyond Int32");
    // the three axes' stencils: 3 nmesh x (float4 weights + int4 indices)
    HNT_TRY(reserve(ctx, ctx->fs_tab, (size_t)3 * nmesh * 32));
    float4* sw = (float4*)ctx->fs_tab.p;
    int4* si = (int4*)(sw + 3 * (size_t)nmesh);
    hipLaunchKernelGGL(fs_tricubic_axes, dim3((unsigned)((3 * nmesh + 255) / 256)), dim3(256), 0,
                       ctx->stream, M, N, block, (int)i0, (int)j0, (int)k0, nmesh, sw, si);
    if ((block & (block - 1)) == 0 && block <= TRI_MAXB && nmesh >= block && ctx->tricubic_tab)
        hipLaunchKernelGGL(fs_tricubic_tab, dim3(grid_rows((int64_t)nmesh * nmesh)), dim3(FT), 0,
                           ctx->stream, coarse, M, nmesh, block, (const float4*)sw,
                           (const int4*)si, tile);
    else
        hipLaunchKernelGGL(fs_tricubic, dim3(grid_rows((int64_t)nmesh * nmesh)), dim3(FT), 0,
                           ctx->stream, coarse, M, nmesh, (const float4*)sw, (const int4*)si,
                           tile);
    return finish_call(ctx, "hnt_interpolate_to_tile");
}

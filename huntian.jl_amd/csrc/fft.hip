// fft.hip -- power-of-two 3-D real FFTs for the field-synthesis entries.
//
// A forward real-to-complex transform of a w^3 field zero-padded to L^3, and an inverse
// complex-to-real transform whose output is cropped to n^3 cells starting at s, with an
// element-wise operator (transfer, periodic kernel, short-range kernel) applied as the
// spectrum is first read.  Layouts are hipFFT's for an L^3 real field with x fastest:
// half spectrum (L/2 + 1) x L x L complex.  Both transforms are unnormalised.
//
// Each axis is one pass over HBM.  Along y and z a workgroup loads C contiguous x-columns of
// length L (C x 8 B segments, c fastest across lanes) into LDS, runs a Stockham radix-8/4 FFT
// on every column, and writes back only the requested range; along x a workgroup takes G rows,
// packs the even/odd reals of a row into an L/2-point complex FFT and untwists it.  The zero
// padding is pruned (the forward x pass reads only the w^2 non-empty rows, y only the w
// non-empty planes) and so is the crop (the inverse z pass writes n of L planes, y works on
// those n planes only and writes n rows each, x transforms the n^2 kept rows).  The inverse x
// pass uses only the real parts of the kx = 0 and kx = L/2 coefficients, which is the same as
// making those two planes Hermitian (averaging conjugate pairs) before the transform.
#include <cmath>

#include "hnt_internal.h"

namespace hnt {

namespace {

// Workgroup shapes hold at most 8 points per thread (<= 64 VGPRs: 8 waves per SIMD, some
// workgroups loading or storing while others transform): 512 threads and 35 KB of LDS (four
// workgroups per CU) for columns up to 512 and for all rows; 1024 threads and 68 KB (two
// workgroups per CU) for 1024-point columns (eight per workgroup, 64-B pieces of each line
// instead of the small shape's 32 B: 7 % faster per transform) and 2048-point columns (four).
// Sixteen 1024-point columns (whole 128-B lines) need 16 points per thread and 139 KB of
// LDS, i.e. 4 waves per SIMD, and measured slower (9 % at 1024, 10 % at 2048).
constexpr int SMALL_T = 512, SMALL_LDS = 4368, SMALL_DATA = 4096;  // 4 padded columns of 1024
constexpr int BIG_T = 1024, BIG_LDS = 8720, BIG_DATA = 8192;       // 8 of 1024, 4 of 2048


int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

// LDS slot of element i of a column: one spare slot after every 16, so that the strided
// writes of the first Stockham stages (stride R complex) spread over the banks
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int padded_len(int L) { return L + (L >> 4) + 1; }
// Column stride of the column passes: the padded length raised to LS = 16 / C (mod 16), so
// that the C columns a 16-lane group touches when it loads or stores one row of each (c
// fastest across lanes) fall in distinct 8-B bank slots.
inline int cols_ls(int L, int C) {
    const int base = padded_len(L);
    const int target = (16 / C) & 15;
    return base + ((target - base) % 16 + 16) % 16;
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a * w, or a * conj(w) for the inverse direction
template <bool INV>
__device__ __forceinline__ float2 ctw(float2 a, float2 w) {
    return INV ? make_float2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y)
               : make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}
// -i a (forward) or +i a (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
    const float2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3);
    const float2 a3 = rot<INV>(csub(v1, v3));
    v0 = cadd(a0, a2);
    v2 = csub(a0, a2);
    v1 = cadd(a1, a3);
    v3 = csub(a1, a3);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&v)[R]) {
    if constexpr (R == 2) {
        const float2 t = v[0];
        v[0] = cadd(t, v[1]);
        v[1] = csub(t, v[1]);
    } else if constexpr (R == 4) {
        dft4<INV>(v[0], v[1], v[2], v[3]);
    } else {
        float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4<INV>(e0, e1, e2, e3);
        dft4<INV>(o0, o1, o2, o3);
        const float h = 0.70710678118654752f;
        // o1 *= w8, o2 *= w8^2, o3 *= w8^3 with w8 = exp(-+ i pi / 4)
        o1 = INV ? make_float2((o1.x - o1.y) * h, (o1.x + o1.y) * h)
                 : make_float2((o1.x + o1.y) * h, (o1.y - o1.x) * h);
        o2 = rot<INV>(o2);
        o3 = INV ? make_float2(-(o3.x + o3.y) * h, (o3.x - o3.y) * h)
                 : make_float2((o3.y - o3.x) * h, -(o3.x + o3.y) * h);
        v[0] = cadd(e0, o0);
        v[4] = csub(e0, o0);
        v[1] = cadd(e1, o1);
        v[5] = csub(e1, o1);
        v[2] = cadd(e2, o2);
        v[6] = csub(e2, o2);
        v[3] = cadd(e3, o3);
        v[7] = csub(e3, o3);
    }
}

// One Stockham stage of radix R on `ncol` columns of length L held in LDS (column c at
// buf + c * LS): butterfly j of a column reads x[j + r L/R], twiddles by W^(r k L/(Ns R)),
// k = j mod Ns, and writes x[(j/Ns) Ns R + k + r Ns].  All reads precede all writes.
// Row-major LDS layout of the direct-load column passes (C = 8 or 4 columns): element i of
// column c at (i + i / (128 / C)) * C + c, i.e. rows of C complex (8 C bytes) in blocks of
// 128 / C rows (1 KB, what one wavefront's 16-byte direct loads fill) with one spare row
// after each block.
template <int C>
__device__ __forceinline__ int rm_slot(int c, int i) {
    constexpr int LGC = C == 8 ? 3 : 2, LGB = 7 - LGC;
    return ((i + (i >> LGB)) << LGC) + c;
}

template <int R, bool INV, int NT, int DATA, int RMC = 0>
__device__ __forceinline__ void stage(float2* buf, int LS, int ncol, int lgL, int lgNs,
                                      const float2* __restrict__ W) {
    constexpr int LGR = R == 8 ? 3 : (R == 4 ? 2 : 1);
    constexpr int MAXB = (DATA + R * NT - 1) / (R * NT);  // butterflies per thread
    const int lgnb = lgL - LGR;
    const int nb = 1 << lgnb;
    const int total = nb * ncol;
    const int Ns = 1 << lgNs;
    const int tw_shift = lgL - lgNs - LGR;
    // column-major (RMC = 0): a thread's butterflies run down one column; row-major (RMC
    // columns): the column index is the fastest, so RMC lanes read or write one row
    constexpr bool RM = RMC != 0;
    constexpr int LGRM = RMC == 8 ? 3 : 2;
    auto at = [&](int c, int i) -> float2& {
        if constexpr (RM) return buf[rm_slot<RMC>(c, i)];
        else return buf[c * LS + pad(i)];
    };
    float2 v[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int g = (int)threadIdx.x + q * NT;
        if (g < total) {
            const int c = RM ? (g & (RMC - 1)) : g >> lgnb, j = RM ? (g >> LGRM) : g & (nb - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = at(c, j + r * nb);
            const int k = j & (Ns - 1);
            if (k != 0) {
                // W^k from the table, the higher powers by products (at most four roundings
                // more than the table's): one table read per butterfly instead of R - 1, whose
                // latency each stage waited for (computing W^k with sincospi was slower)
                const float2 w1 = W[k << tw_shift];
                if constexpr (R == 2) {
                    v[q][1] = ctw<INV>(v[q][1], w1);
                } else {
                    const float2 w2 = ctw<false>(w1, w1);
                    v[q][1] = ctw<INV>(v[q][1], w1);
                    v[q][2] = ctw<INV>(v[q][2], w2);
                    const float2 w3 = ctw<false>(w1, w2);
                    v[q][3] = ctw<INV>(v[q][3], w3);
                    if constexpr (R == 8) {
                        const float2 w4 = ctw<false>(w2, w2);
                        v[q][4] = ctw<INV>(v[q][4], w4);
                        v[q][5] = ctw<INV>(v[q][5], ctw<false>(w1, w4));
                        v[q][6] = ctw<INV>(v[q][6], ctw<false>(w2, w4));
                        v[q][7] = ctw<INV>(v[q][7], ctw<false>(w3, w4));
                    }
                }
            }
            dft<R, INV>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int g = (int)threadIdx.x + q * NT;
        if (g < total) {
            const int c = RM ? (g & (RMC - 1)) : g >> lgnb, j = RM ? (g >> LGRM) : g & (nb - 1);
            const int k = j & (Ns - 1);
            const int d0 = ((j >> lgNs) << (lgNs + LGR)) + k;
#pragma unroll
            for (int r = 0; r < R; ++r) at(c, d0 + r * Ns) = v[q][r];
        }
    }
    __syncthreads();
}

// FFT of length 2^lgL of every column, in place, natural order in and out (from stage
// lgNs0 on: the caller has run the stages before it).
template <bool INV, int NT, int DATA, int RMC = 0>
__device__ __forceinline__ void stockham(float2* buf, int LS, int ncol, int lgL,
                                         const float2* __restrict__ W, int lgNs0 = 0) {
    int lgNs = lgNs0;
    while (lgNs < lgL) {
        const int rem = lgL - lgNs;
        if (rem == 1) {
            stage<2, INV, NT, DATA, RMC>(buf, LS, ncol, lgL, lgNs, W);
            lgNs += 1;
        } else if (rem == 2 || rem == 4) {
            stage<4, INV, NT, DATA, RMC>(buf, LS, ncol, lgL, lgNs, W);
            lgNs += 2;
        } else {
            stage<8, INV, NT, DATA, RMC>(buf, LS, ncol, lgL, lgNs, W);
            lgNs += 3;
        }
    }
}

// ---------------------------------------------------------------- spectrum operators

__device__ __forceinline__ double pick(int d, double kx, double ky, double kz) {
    return d == 1 ? kx : (d == 2 ? ky : kz);
}

// Operator on element (ix, iy, iz) of the half spectrum of an L^3 field, as the reference
// kernels compute it (_transfer_kernel!, _periodic_kernel!, _short_pot_kernel).
// Entry of element (ix, iy, iz) in a (L/2 + 1)^3 table over (ix, |iy|, |iz|), |iz| fastest:
// the z pass, which reads it, walks iz down a column, so a wavefront's entries are runs of
// contiguous floats (with ix fastest every entry was a 4-byte piece of its own line, 4 MB
// from the next: 12 of the 34 ms of a 2048-point inverse z pass with the transfer).
__device__ __forceinline__ int64_t table_index(int L, int ix, int iy, int iz) {
    const int half = L / 2;
    const int ay = iy <= half ? iy : L - iy, az = iz <= half ? iz : L - iz;
    return az + (int64_t)(half + 1) * (ix + (int64_t)(half + 1) * ay);
}

// `tab` is op.amp[table_index(...)] for WINDOW and TRANSFER (loaded by the caller together
// with the spectrum element), unused otherwise.
__device__ __forceinline__ float2 apply_op(const FftOp& op, float2 v, int ix, int iy, int iz,
                                           float tab) {
    const int L = op.L, half = L / 2;
    if (op.kind == FFT_OP_WINDOW) {
        const float w = tab;
        v.x *= w;
        v.y *= w;
        return make_float2(v.x * op.post, v.y * op.post);
    }
    if (op.kind == FFT_OP_TRANSFER) {
        const double kx = (double)ix * op.dk;
        const double ky = (double)(iy < half ? iy : iy - L) * op.dk;
        const double kz = (double)(iz < half ? iz : iz - L) * op.dk;
        const double k2 = kx * kx + ky * ky + kz * kz;
        if (k2 == 0.0) return make_float2(0.0f, 0.0f);
        const double sqrt_pk = (double)tab;
        if (op.id == 0)
            return make_float2((float)((double)v.x * sqrt_pk), (float)((double)v.y * sqrt_pk));
        if (op.id == 1 || op.id == 2) {
            const double coeff = sqrt_pk * pick(op.d1, kx, ky, kz) / k2;
            const double a = v.x, b = v.y;
            return op.id == 1 ? make_float2((float)(-b * coeff), (float)(a * coeff))
                              : make_float2((float)(b * coeff), (float)(-a * coeff));
        }
        float c;
        if (op.id == 3)
            c = (float)(-sqrt_pk * pick(op.d1, kx, ky, kz) * pick(op.d2, kx, ky, kz) / k2);
        else if (op.id == 4)
            c = (float)(sqrt_pk * k2);
        else
            c = (float)(-sqrt_pk / k2);
        return make_float2(v.x * c, v.y * c);
    }
    if (op.kind == FFT_OP_PERIODIC) {
        const double kx = (double)ix * op.dk;
        const double ky = (double)(iy <= half ? iy : iy - L) * op.dk;
        const double kz = (double)(iz <= half ? iz : iz - L) * op.dk;
        const double k2 = kx * kx + ky * ky + kz * kz;
        if (k2 == 0.0 || (op.id != 3 && (ix == half || iy == half || iz == half)))
            return make_float2(0.0f, 0.0f);
        if (op.id == 2) {
            const double coeff = pick(op.d1, kx, ky, kz) / k2;
            const double a = v.x, b = v.y;
            return make_float2((float)(b * coeff), (float)(-a * coeff));
        }
        const float c = op.id == 3
                            ? (float)(-pick(op.d1, kx, ky, kz) * pick(op.d2, kx, ky, kz) / k2)
                            : (float)k2;
        return make_float2(v.x * c, v.y * c);
    }
    if (op.kind == FFT_OP_SHORTPOT) {
        const double kx = (double)ix * op.dk;
        const double ky = (double)(iy <= half ? iy : iy - L) * op.dk;
        const double kz = (double)(iz <= half ? iz : iz - L) * op.dk;
        const double k2 = kx * kx + ky * ky + kz * kz;
        const float c = (float)(k2 == 0.0 ? -op.rs2 : -(1.0 - exp(-k2 * op.rs2)) / k2);
        return make_float2(v.x * c, v.y * c);
    }
    return v;
}

// ---------------------------------------------------------------- kernels

// Transform along a strided axis: element (x, a, b) at x + a strideA + b strideB, x over the
// nx contiguous columns (C per workgroup), a the transformed axis (length 2^lgL), b =
// b0 + blockIdx.y.  Reads a in [r0, r1) (zero elsewhere), applies `op` when set (zpass:
// a = iz, b = iy), writes a in [w0, w1).  In place when in == out: a workgroup reads all of
// its columns before it writes.
// Workgroups are dealt to the 8 XCDs in turn; tile t = xcd * per + (id / 8) keeps the tiles
// that share 128-B lines (adjacent x) on one XCD and its L2, one after another.
template <bool INV, int NT, int NLDS, int DATA, int WAVES>
__global__ __launch_bounds__(NT, WAVES) void fft_cols(
    const float2* in, float2* out, int lgL, int64_t strideA, int64_t strideB, int nx, int lgC,
    int gx, int ntiles, int LS, int b0, int r0, int r1, int w0, int w1, FftOp op,
    const float2* __restrict__ W) {
    __shared__ float2 buf[NLDS];
    const int per = (ntiles + 7) >> 3;
    const int tile = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (tile >= ntiles) return;
    const int L = 1 << lgL, C = 1 << lgC;
    const int x0 = (tile % gx) * C;
    const int b = b0 + tile / gx;
    const int64_t boff = (int64_t)b * strideB;
    // two elements per step, both loads issued before either is used
    for (int e = threadIdx.x; e < C * L; e += 2 * NT) {
        const int e1 = e + NT;
        const int c = e & (C - 1), a = e >> lgC, x = x0 + c;
        const int c1 = e1 & (C - 1), a1 = e1 >> lgC, x1 = x0 + c1;
        const bool has1 = e1 < C * L;
        const bool in0 = x < nx && a >= r0 && a < r1;
        const bool in1 = has1 && x1 < nx && a1 >= r0 && a1 < r1;
        float2 v = make_float2(0.0f, 0.0f), v1 = make_float2(0.0f, 0.0f);
        if (in0) v = in[x + a * strideA + boff];
        if (in1) v1 = in[x1 + a1 * strideA + boff];
        // the operator's table entries (window, sqrt P) loaded together with the spectrum
        float t = 0.0f, t1 = 0.0f;
        if (op.kind == FFT_OP_WINDOW || op.kind == FFT_OP_TRANSFER) {
            if (in0) t = op.amp[table_index(L, x, b, a)];
            if (in1) t1 = op.amp[table_index(L, x1, b, a1)];
        }
        if (op.kind != FFT_OP_NONE) {
            if (in0) v = apply_op(op, v, x, b, a, t);
            if (in1) v1 = apply_op(op, v1, x1, b, a1, t1);
        }
        buf[c * LS + pad(a)] = v;
        if (has1) buf[c1 * LS + pad(a1)] = v1;
    }
    __syncthreads();
    stockham<INV, NT, DATA>(buf, LS, C, lgL, W);
    for (int e = threadIdx.x; e < C * L; e += NT) {
        const int c = e & (C - 1), a = e >> lgC, x = x0 + c;
        if (x < nx && a >= w0 && a < w1) out[x + a * strideA + boff] = buf[c * LS + pad(a)];
    }
}

// fft_cols for 1024-point columns, 8 per workgroup, and 2048-point columns, 4 per
// workgroup, reading rows [0, r1) with r1 a multiple of 128 / C (the rest are zero): the
// spectrum goes from HBM straight into LDS (16-byte direct loads, no registers held), so a
// workgroup has all of its 64 KB in flight at once instead of two elements per thread.  Each wavefront load fills one 1 KB
// block of the row-major layout (rm_slot<C>): lane l brings columns 2 (l % (C/2)) and
// 2 (l % (C/2)) + 1 of row (128/C) b + l / (C/2).  The operator's table entries are loaded
// into registers meanwhile and the operator is applied in LDS, element by element as the
// register path does, so the values are the same bit for bit.
template <int L, int C>
struct DirectShape {
    static constexpr int LGL = L == 2048 ? 11 : 10, LGC = C == 8 ? 3 : 2;
    static constexpr int ROWS = 128 / C;        // rows per 1 KB block
    static constexpr int NB = L / ROWS;         // blocks
    static constexpr int LDS = NB * (ROWS + 1) * C;  // float2, a spare row per block
    static_assert((L == 1024 && C == 8) || (L == 2048 && C == 4), "direct column shapes");
};
// Twiddles the Stockham stages read: index (k << (lgL - lgNs - lgR)) with k < Ns is below
// L / R of the last stage, i.e. below 256 (1024 = 8 8 4 4) and 512 (2048 = 8 8 8 4).
constexpr int DIRECT_TW = 512;
template <bool INV, int NT, int WAVES, int L, int C>
__global__ __launch_bounds__(NT, WAVES) void fft_cols_direct(
    const float2* in, float2* out, int64_t strideA, int64_t strideB, int nx, int gx,
    int ntiles, int b0, int r1, int w0, int w1, FftOp op, const float2* __restrict__ W) {
    using S = DirectShape<L, C>;
    constexpr int lgL = S::LGL, LGC = S::LGC, NB = S::NB, ROWS = S::ROWS;
    constexpr int EPT = C * L / NT;  // elements per thread
    __shared__ float2 buf[S::LDS];
    __shared__ float2 tw[DIRECT_TW];  // the twiddles, read by every butterfly of the stages
    const int per = (ntiles + 7) >> 3;
    const int tile = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (tile >= ntiles) return;
    const int x0 = (tile % gx) * C;
    const int b = b0 + tile / gx;
    const int64_t boff = (int64_t)b * strideB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the twiddles (4 KB) into LDS with the first four wavefronts' first direct loads: the
    // stages then read them from LDS instead of waiting on L2 at every stage
    if (wave < DIRECT_TW / 128)
        __builtin_amdgcn_global_load_lds(
            W + 128 * wave + 2 * lane, (__attribute__((address_space(3))) void*)(tw + 128 * wave),
            16, 0, 0);
    // rows of the block, two columns per lane; columns past nx (the last tile) read the
    // row's pitch padding, whose values are never stored
    // (the blocks from r1 on are zero: the padding of a forward transform)
    const int lrow = lane >> (LGC - 1), lcol = 2 * (lane & (C / 2 - 1));
    const int nread = r1 / ROWS;
    for (int blk = wave; blk < NB; blk += NT / 64) {
        float2* dst = buf + (ROWS + 1) * C * blk;
        if (blk < nread) {
            const int a = ROWS * blk + lrow;
            const float2* g = in + (x0 + lcol) + (int64_t)a * strideA + boff;
            __builtin_amdgcn_global_load_lds(
                g, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
        } else {
            dst[2 * lane] = make_float2(0.0f, 0.0f);
            dst[2 * lane + 1] = make_float2(0.0f, 0.0f);
        }
    }
    float t[EPT];
    if (op.kind == FFT_OP_WINDOW || op.kind == FFT_OP_TRANSFER) {
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = (int)threadIdx.x + q * NT;
            const int c = e & (C - 1), a = e >> LGC, x = x0 + c;
            t[q] = x < nx ? op.amp[table_index(L, x, b, a)] : 0.0f;
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (op.kind != FFT_OP_NONE) {
        // the first radix-8 stage (span 1, no twiddles) with the operator applied to its
        // inputs as they are read: thread (c, j) takes elements j + (L/8) r of column c, the
        // very elements whose table entries it loaded (t[r])
        static_assert(EPT == 8 && NT == 1024, "one radix-8 butterfly per thread");
        const int c = (int)threadIdx.x & (C - 1), j = (int)threadIdx.x >> LGC, x = x0 + c;
        float2 v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int a = j + (L / 8) * r;
            v[r] = buf[rm_slot<C>(c, a)];
            if (x < nx && a < r1) v[r] = apply_op(op, v[r], x, b, a, t[r]);
        }
        dft<8, INV>(v);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; ++r) buf[rm_slot<C>(c, 8 * j + r)] = v[r];
        __syncthreads();
        stockham<INV, NT, C * L, C>(buf, 0, C, lgL, tw, 3);
    } else {
        stockham<INV, NT, C * L, C>(buf, 0, C, lgL, tw);
    }
    for (int e = threadIdx.x; e < C * L; e += NT) {
        const int c = e & (C - 1), a = e >> LGC, x = x0 + c;
        if (x < nx && a >= w0 && a < w1) out[x + a * strideA + boff] = buf[rm_slot<C>(c, a)];
    }
}

// Forward x pass: rows (iy, iz) of a w^3 real field (iy, iz < w), zero-extended to L = 2M,
// into rows of the (M + 1) x L x L half spectrum.  z[m] = x[2m] + i x[2m + 1], Z = FFT_M(z),
// X[k] = E[k] + W_L^k O[k] with E = (Z[k] + conj Z[M-k]) / 2, O = (Z[k] - conj Z[M-k]) / 2i.
__global__ __launch_bounds__(SMALL_T, 8) void fft_rows_r2c(
    const float* __restrict__ f, int w, int lgM, int lgG, float2* __restrict__ spec,
    const float2* __restrict__ WM, const float2* __restrict__ WL) {
    __shared__ float2 buf[SMALL_LDS];
    const int M = 1 << lgM, G = 1 << lgG, LS = padded_len(M), L = 2 * M, nk = M + 1;
    const int row0 = (int)blockIdx.x * G;
    const int nrows = w * w;
    // two elements per step, both loads issued before either is used
    for (int e0 = threadIdx.x; e0 < G * M; e0 += 2 * SMALL_T) {
        float2 v[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = e0 + u * SMALL_T;
            const int g = e >> lgM, m = e & (M - 1), row = row0 + g;
            v[u] = make_float2(0.0f, 0.0f);
            if (e < G * M && row < nrows) {
                const float* src = f + (int64_t)row * w;
                const int x = 2 * m;
                v[u] = make_float2(x < w ? src[x] : 0.0f, x + 1 < w ? src[x + 1] : 0.0f);
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = e0 + u * SMALL_T;
            if (e < G * M) buf[(e >> lgM) * LS + pad(e & (M - 1))] = v[u];
        }
    }
    __syncthreads();
    stockham<false, SMALL_T, SMALL_DATA>(buf, LS, G, lgM, WM);
    for (int e = threadIdx.x; e < G * nk; e += SMALL_T) {
        const int g = e / nk, k = e - g * nk, row = row0 + g;
        if (row >= nrows) continue;
        const int iy = row % w, iz = row / w;
        const float2 zk = buf[g * LS + pad(k & (M - 1))];
        const float2 zm = buf[g * LS + pad((M - k) & (M - 1))];
        const float2 ev = make_float2((zk.x + zm.x) * 0.5f, (zk.y - zm.y) * 0.5f);
        // (zk - conj zm) / 2i
        const float2 od = make_float2((zk.y + zm.y) * 0.5f, -(zk.x - zm.x) * 0.5f);
        const float2 t = ctw<false>(od, WL[k]);
        spec[k + (int64_t)fft_pitch(L) * iy + fft_plane(L) * iz] = cadd(ev, t);
    }
}

// Inverse x pass: rows (s + jy, s + jz), jy, jz < n, of the half spectrum to real rows, of
// which x in [s, s + n) is kept: out[(jy + n jz) n + x - s] = scale * x (mode 0), or
// s2 -= (scale * x)^2 * coef (mode 1).  X[0] and X[M] enter through their real parts.
// Z[k] = (X[k] + conj X[M-k]) + i (X[k] - conj X[M-k]) W_L^-k, z = IFFT_M(Z), x[2m] = Re z[m],
// x[2m + 1] = Im z[m].
__global__ __launch_bounds__(SMALL_T, 8) void fft_rows_c2r(
    const float2* __restrict__ spec, int lgM, int lgG, int s, int n, float scale, int mode,
    float coef, float* __restrict__ out, const float2* __restrict__ WM,
    const float2* __restrict__ WL) {
    __shared__ float2 buf[SMALL_LDS];
    const int M = 1 << lgM, G = 1 << lgG, LS = padded_len(M), L = 2 * M, nk = M + 1;
    const int row0 = (int)blockIdx.x * G;
    const int nrows = n * n;
    // two elements per step, both loads issued before either is used
    for (int e0 = threadIdx.x; e0 < G * nk; e0 += 2 * SMALL_T) {
        float2 v[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = e0 + u * SMALL_T;
            const int g = e / nk, k = e - g * nk, row = row0 + g;
            v[u] = make_float2(0.0f, 0.0f);
            if (e < G * nk && row < nrows) {
                const int iy = s + row % n, iz = s + row / n;
                v[u] = spec[k + (int64_t)fft_pitch(L) * iy + fft_plane(L) * iz];
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = e0 + u * SMALL_T;
            if (e < G * nk) {
                const int g = e / nk;
                buf[g * LS + pad(e - g * nk)] = v[u];
            }
        }
    }
    __syncthreads();
    // untwist into Z (registers first: Z[k] needs X[M - k] of the same row)
    constexpr int MAXB = SMALL_DATA / SMALL_T;
    float2 z[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int e = (int)threadIdx.x + q * SMALL_T;
        if (e < G * M) {
            const int g = e >> lgM, k = e & (M - 1);
            float2 a = buf[g * LS + pad(k)];
            float2 b = buf[g * LS + pad(M - k)];
            if (k == 0) {
                a.y = 0.0f;
                b.y = 0.0f;
            }
            const float2 sum = make_float2(a.x + b.x, a.y - b.y);   // a + conj b
            const float2 dif = make_float2(a.x - b.x, a.y + b.y);   // a - conj b
            const float2 t = ctw<true>(dif, WL[k]);                 // dif W_L^-k
            z[q] = make_float2(sum.x - t.y, sum.y + t.x);           // sum + i t
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int e = (int)threadIdx.x + q * SMALL_T;
        if (e < G * M) buf[(e >> lgM) * LS + pad(e & (M - 1))] = z[q];
    }
    __syncthreads();
    stockham<true, SMALL_T, SMALL_DATA>(buf, LS, G, lgM, WM);
    for (int e = threadIdx.x; e < G * n; e += SMALL_T) {
        const int g = e / n, xo = e - g * n, row = row0 + g;
        if (row >= nrows) continue;
        const int xx = s + xo;
        const float2 zz = buf[g * LS + pad(xx >> 1)];
        const float v = ((xx & 1) ? zz.y : zz.x) * scale;
        const int64_t o = (int64_t)row * n + xo;
        if (mode == 0) out[o] = v;
        else out[o] = out[o] - (v * v) * coef;
    }
}

// fft_rows_c2r for L = 1024 (M = 512, 8 rows per workgroup) and L = 2048 (M = 1024, 4
// rows) with the rows loaded straight into LDS: each row's first M coefficients as M / 128
// 1 KB direct loads into an unpadded staging layout (row g at g P, as in the spectrum), the
// Nyquist coefficient X[M] by an ordinary load; the untwist reads the staging layout and
// writes the padded FFT layout over it after a barrier.  Same arithmetic as fft_rows_c2r,
// so the same values bit for bit.
template <int LGM>
__global__ __launch_bounds__(SMALL_T, 8) void fft_rows_c2r_direct(
    const float2* __restrict__ spec, int s, int n, float scale, int mode, float coef,
    float* __restrict__ out, const float2* __restrict__ WM, const float2* __restrict__ WL) {
    constexpr int lgM = LGM, M = 1 << LGM, G = SMALL_DATA / M, L = 2 * M, P = fft_pitch(L);
    constexpr int Q = M / 128, LGQ = LGM - 7;  // 1 KB chunks per row
    static_assert((LGM == 9 || LGM == 10) && G * padded_len(M) <= SMALL_LDS, "row shapes");
    static_assert(G * P <= SMALL_LDS, "staging rows fit the FFT buffer");
    __shared__ float2 buf[SMALL_LDS];
    const int LS = padded_len(M);
    const int row0 = (int)blockIdx.x * G;
    const int nrows = n * n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;  // 8 wavefronts
    // 32 chunks (G rows x Q) over the 8 wavefronts
    for (int u = wave; u < Q * G; u += SMALL_T / 64) {
        const int g = u >> LGQ, q = u & (Q - 1);
        int row = row0 + g;
        if (row >= nrows) row = nrows - 1;  // a row past the end: loads a kept row, unused
        const int iy = s + row % n, iz = s + row / n;
        const float2* src = spec + (int64_t)P * iy + fft_plane(L) * iz + 128 * q + 2 * lane;
        __builtin_amdgcn_global_load_lds(
            src, (__attribute__((address_space(3))) void*)(buf + g * P + 128 * q), 16, 0, 0);
    }
    float2 nyq = make_float2(0.0f, 0.0f);
    if (threadIdx.x < G && row0 + (int)threadIdx.x < nrows) {
        const int row = row0 + threadIdx.x;
        const int iy = s + row % n, iz = s + row / n;
        nyq = spec[M + (int64_t)P * iy + fft_plane(L) * iz];
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x < G) buf[threadIdx.x * P + M] = nyq;
    __syncthreads();
    constexpr int MAXB = SMALL_DATA / SMALL_T;
    float2 z[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int e = (int)threadIdx.x + q * SMALL_T;
        const int g = e >> lgM, k = e & (M - 1);
        float2 a = buf[g * P + k];
        float2 b = buf[g * P + (M - k)];
        if (k == 0) {
            a.y = 0.0f;
            b.y = 0.0f;
        }
        const float2 sum = make_float2(a.x + b.x, a.y - b.y);   // a + conj b
        const float2 dif = make_float2(a.x - b.x, a.y + b.y);   // a - conj b
        const float2 t = ctw<true>(dif, WL[k]);                 // dif W_L^-k
        z[q] = make_float2(sum.x - t.y, sum.y + t.x);           // sum + i t
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int e = (int)threadIdx.x + q * SMALL_T;
        buf[(e >> lgM) * LS + pad(e & (M - 1))] = z[q];
    }
    __syncthreads();
    stockham<true, SMALL_T, SMALL_DATA>(buf, LS, G, lgM, WM);
    for (int e = threadIdx.x; e < G * n; e += SMALL_T) {
        const int g = e / n, xo = e - g * n, row = row0 + g;
        if (row >= nrows) continue;
        const int xx = s + xo;
        const float2 zz = buf[g * LS + pad(xx >> 1)];
        const float v = ((xx & 1) ? zz.y : zz.x) * scale;
        const int64_t o = (int64_t)row * n + xo;
        if (mode == 0) out[o] = v;
        else out[o] = out[o] - (v * v) * coef;
    }
}

// sqrt P per (ix, |iy|, |iz|) of the transfer: the (L/2 + 1)^3 table the inverse z pass reads
// (rows of equal k^2 share it; k^2 as _transfer_kernel! computes it), laid out as
// table_index: a wavefront per (ix, |iy|) row, lanes over |iz|.
__global__ void fft_amp(float* __restrict__ amp, int L, double dk, const float* __restrict__ tab,
                        int ntab, float log_k_min, float d_log_k) {
    const int half = L / 2, nh = half + 1;
    const int lane = threadIdx.x & 63;
    for (int row = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6); row < nh * nh;
         row += (int)gridDim.x * 4) {
        const int ix = row % nh, ay = row / nh;
        const double kx = (double)ix * dk, ky = (double)ay * dk;
        const double kyz = ky * ky;
        for (int az = lane; az < nh; az += 64) {
            const double kz = (double)az * dk;
            const double kzz = kz * kz;
            const double k2 = kx * kx + kyz + kzz;
            float v = 0.0f;
            if (k2 != 0.0) {
                const float lk = (float)(log(k2) * 0.5);
                const float fpos = (lk - log_k_min) / d_log_k;
                const int i0 = (int)floorf(fpos) + 1;
                if (i0 < 1) v = tab[0];
                else if (i0 >= ntab) v = tab[ntab - 1];
                else {
                    const float frac = fpos - (float)(i0 - 1);
                    v = tab[i0 - 1] + frac * (tab[i0] - tab[i0 - 1]);
                }
            }
            amp[az + (int64_t)nh * row] = v;
        }
    }
}

// grid of the XCD-dealt column tiles: 8 * ceil(tiles / 8) workgroups
unsigned cols_grid(unsigned gx, int nb) {
    const int64_t t = (int64_t)gx * nb;
    return (unsigned)(((t + 7) / 8) * 8);
}

bool cols_big(int L) { return L >= 1024; }



int cols_lgC(int L) {
    // as many padded columns as the shape's LDS holds, at most 8 (64-B segments)
    const int cap = cols_big(L) ? BIG_LDS : SMALL_LDS;
    int lg = 3;
    while (lg > 0 && (1 << lg) * cols_ls(L, 1 << lg) > cap) --lg;
    return lg;
}

int rows_lgG(int M) {
    int lg = 4;
    while (lg > 0 && (1 << lg) * padded_len(M) > SMALL_LDS) --lg;
    return lg;
}

template <bool INV>
void launch_cols(hipStream_t st, int L, unsigned gx, int nb, const float2* in, float2* out,
                 int64_t strideA, int64_t strideB, int nx, int b0, int r0, int r1, int w0,
                 int w1, const FftOp& op, const float2* W, bool direct = false) {
    const int lgC = cols_lgC(L);
    const int LS = cols_ls(L, 1 << lgC);
    if (direct && r0 == 0 && r1 <= L && r1 % 16 == 0 && L == 1024 && lgC == 3) {
        hipLaunchKernelGGL((fft_cols_direct<INV, BIG_T, 8, 1024, 8>), dim3(cols_grid(gx, nb)),
                           dim3(BIG_T), 0, st, in, out, strideA, strideB, nx, (int)gx,
                           (int)gx * nb, b0, r1, w0, w1, op, W);
        return;
    }
    if (direct && r0 == 0 && r1 <= L && r1 % 32 == 0 && L == 2048 && lgC == 2) {
        hipLaunchKernelGGL((fft_cols_direct<INV, BIG_T, 8, 2048, 4>), dim3(cols_grid(gx, nb)),
                           dim3(BIG_T), 0, st, in, out, strideA, strideB, nx, (int)gx,
                           (int)gx * nb, b0, r1, w0, w1, op, W);
        return;
    }
    if (cols_big(L))
        hipLaunchKernelGGL((fft_cols<INV, BIG_T, BIG_LDS, BIG_DATA, 8>), dim3(cols_grid(gx, nb)), dim3(BIG_T),
                           0, st, in, out, ilog2(L), strideA, strideB, nx, lgC, (int)gx,
                           (int)gx * nb, LS, b0, r0, r1, w0, w1, op, W);
    else
        hipLaunchKernelGGL((fft_cols<INV, SMALL_T, SMALL_LDS, SMALL_DATA, 8>), dim3(cols_grid(gx, nb)),
                           dim3(SMALL_T), 0, st, in, out, ilog2(L), strideA, strideB, nx, lgC,
                           (int)gx, (int)gx * nb, LS, b0, r0, r1, w0, w1, op, W);
}

}  // namespace

bool fft_supported(int L) { return L >= 8 && L <= 2048 && (L & (L - 1)) == 0; }

// Forward twiddles W[k] = exp(-2 pi i k / L), k < L, computed in double, kept per L.
hnt_status fft_twiddles(hnt_ctx* ctx, int L, const float2** out) {
    for (hnt_ctx::FftTw& t : ctx->fft_tw)
        if (t.L == L && t.buf.p) {
            *out = (const float2*)t.buf.p;
            return HNT_OK;
        }
    std::vector<float2> h(L);
    for (int k = 0; k < L; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)L;
        h[k] = make_float2((float)cos(a), (float)sin(a));
    }
    // filled in a local buffer and cached only once it holds the table, so a failed
    // allocation or copy leaves no entry behind
    DevBuf b;
    HNT_TRY(reserve(ctx, b, sizeof(float2) * (size_t)L));
    hipError_t e = hipMemcpyAsync(b.p, h.data(), sizeof(float2) * (size_t)L,
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        release(b);
        set_error("fft_twiddles: %s", hipGetErrorString(e));
        return HNT_ERR_DEVICE;
    }
    ctx->fft_tw.push_back({L, b});
    *out = (const float2*)b.p;
    return HNT_OK;
}

hnt_status fft_forward(hnt_ctx* ctx, const float* f, int w, int L, float2* spec) {
    hipStream_t st = ctx->stream;
    const int M = L / 2, nk = M + 1;
    const float2 *WL, *WM;
    HNT_TRY(fft_twiddles(ctx, L, &WL));
    HNT_TRY(fft_twiddles(ctx, M, &WM));
    const int lgG = rows_lgG(M);
    const int64_t rows = (int64_t)w * w;
    hipLaunchKernelGGL(fft_rows_r2c, dim3((unsigned)((rows + (1 << lgG) - 1) >> lgG)),
                       dim3(SMALL_T), 0, st, f, w, ilog2(M), lgG, spec, WM, WL);
    FftOp none;
    const int lgC = cols_lgC(L);
    const unsigned gx = (unsigned)((nk + (1 << lgC) - 1) >> lgC);
    // y: planes iz < w, rows iy < w are filled
    const int64_t P = fft_pitch(L), PL = fft_plane(L);
    launch_cols<false>(st, L, gx, w, spec, spec, P, PL, nk, 0, 0, w, 0,
                       L, none, WL, ctx->fft_direct != 0);
    // z: every (ix, iy) column, planes iz < w are filled
    launch_cols<false>(st, L, gx, L, spec, spec, PL, P, nk, 0, 0, w, 0,
                       L, none, WL, ctx->fft_direct != 0);
    return HNT_OK;
}

hnt_status fft_inverse(hnt_ctx* ctx, const float2* spec, float2* work, int L, int s, int n,
                       const FftOp& op, float scale, int mode, float coef, float* out) {
    hipStream_t st = ctx->stream;
    const int M = L / 2, nk = M + 1;
    const float2 *WL, *WM;
    HNT_TRY(fft_twiddles(ctx, L, &WL));
    HNT_TRY(fft_twiddles(ctx, M, &WM));
    const int lgC = cols_lgC(L);
    const unsigned gx = (unsigned)((nk + (1 << lgC) - 1) >> lgC);
    FftOp none;
    // z: every (ix, iy) column, the operator applied as read; planes iz in [s, s + n) kept
    const int64_t P = fft_pitch(L), PL = fft_plane(L);
    launch_cols<true>(st, L, gx, L, spec, work, PL, P, nk, 0, 0, L, s,
                      s + n, op, WL, ctx->fft_direct != 0);
    // y: planes iz in [s, s + n), rows iy in [s, s + n) kept
    launch_cols<true>(st, L, gx, n, work, work, P, PL, nk, s, 0, L, s,
                      s + n, none, WL, ctx->fft_direct != 0);
    const int lgG = rows_lgG(M);
    const int64_t rows = (int64_t)n * n;
    if (ctx->fft_direct && L == 1024 && lgG == 3) {
        hipLaunchKernelGGL(fft_rows_c2r_direct<9>, dim3((unsigned)((rows + 7) >> 3)),
                           dim3(SMALL_T), 0, st, (const float2*)work, s, n, scale, mode, coef,
                           out, WM, WL);
        return HNT_OK;
    }
    if (ctx->fft_direct && L == 2048 && lgG == 2) {
        hipLaunchKernelGGL(fft_rows_c2r_direct<10>, dim3((unsigned)((rows + 3) >> 2)),
                           dim3(SMALL_T), 0, st, (const float2*)work, s, n, scale, mode, coef,
                           out, WM, WL);
        return HNT_OK;
    }
    hipLaunchKernelGGL(fft_rows_c2r, dim3((unsigned)((rows + (1 << lgG) - 1) >> lgG)),
                       dim3(SMALL_T), 0, st, (const float2*)work, ilog2(M), lgG, s, n, scale,
                       mode, coef, out, WM, WL);
    return HNT_OK;
}

hnt_status fft_amp_table(hnt_ctx* ctx, int L, double dk, const float* tab, int ntab,
                         float log_k_min, float d_log_k, float* amp) {
    const int64_t nh = L / 2 + 1;
    hipLaunchKernelGGL(fft_amp, dim3((unsigned)std::min<int64_t>((nh * nh + 3) / 4, 1 << 20)),
                       dim3(256), 0, ctx->stream, amp, L, dk, tab, ntab, log_k_min, d_log_k);
    return HNT_OK;
}

}  // namespace hnt

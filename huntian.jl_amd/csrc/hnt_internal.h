// hnt_internal.h -- context, device-buffer cache and error plumbing of libhuntian_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/huntian_abi.h"
#include "hnt_math.h"

namespace hnt {

void set_error(const char* fmt, ...);
void clear_error();

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Device copy of the shell tables, with the host copy used to detect a change.
struct TablesDev {
    int32_t rmax = -1, nshells = 0;
    int64_t ncells = 0;
    std::vector<int32_t> host;  // di, dj, dk, start, count, r2 concatenated
    DevBuf cells;               // int4 {di, dj, dk, 0} per cell
    DevBuf start, count, r2;    // int32 [nshells]
    DevBuf rad;                 // double [nshells + 2], rad[s] = sqrt(r2 of shell s), 1-based
    // kernel_strain weights hRinteg(u0, u12) / (u12 - u0^2)^2 of shell m1 around shell mp,
    // for the shells with u0^2 < 4 (u0 = rad[mp] - rad[m1]): m1 in [kw_lo[mp], kw_hi[mp]],
    // weight at kw[kw_off[mp] + m1 - kw_lo[mp]]; mp = 2 .. nshells
    DevBuf kw;                  // double
    DevBuf kw_idx;              // int32 [3 * (nshells + 2)]: lo, hi, off of each mp
};

// One tile of a call, in device memory.  A wavefront's peaks all belong to one tile, so a
// wave reads its tile's descriptor with scalar loads.
struct TileDev {
    const float4* recs;   // 32-byte cell records of this tile
    const float* delta;   // the caller's delta (crossing pass)
    const float* lapd;
    int n, has_eta2, has_lapd;
    int64_t s1, s2;
    const int32_t* pi;
    const int32_t* pj;
    const int32_t* pk;
    const int32_t* ir2min;
    const double* Rf;
    const double* ZZon;
    const double* fcrit;
    double ZZon_s, fcrit_s;
    int8_t* mask;
    hnt_results res;
    int64_t npeaks;
    int64_t base;  // global id of the tile's first peak
};


// scratch of hnt_merge_catalog: sort keys and permutations, ranks, the cell hash,
// containment edges, decision states, overlap pairs, volume deficits, new radii
struct MergeBufs {
    DevBuf a, b, c, d, rank, cidx, cstart, cnt, off, e1, e2, e3, e4, state, instart, dv, pk,
        pv, newr, pos, prank, ext;
};

// scratch of the bracket MLP (bf16 weights and activations, gradients)
struct MlpScratch {
    DevBuf wt, x0, z, h, y, dy, dyf, dz, gw, part, gradv;
    int64_t Bp = 0;
};

struct CTabDev {
    int32_t nx = 0, ny = 0, nz = 0;
    std::vector<float> host;
    DevBuf table;
};

// Bump allocator over a fixed device region: buffers carved from it are never freed, the
// region is reused from its start by the next call that uses it.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
};

// Device buffers of one shell analysis (the context holds one for direct calls, each plan
// its own): cell records, sort keys, lane slots, tile descriptors, profile slots, queues.
// With own_tables the shell tables, collapse table and SPH kernel table the kernels read
// are copies held here (a plan's replays must not depend on what the context uploads later);
// otherwise they are the context's.  With an arena every buffer is carved from it.
struct SAWork {
    DevBuf recs, keys_a, keys_b, vals_a, vals_b, cub_tmp, slots, gtiles, tiles, tilemeta, stops,
        feats, prof, queue, flag, skip, paint;
    std::vector<TileDev> tile_host;
    std::vector<int64_t> meta_host;
    bool own_tables = false;
    TablesDev tables;
    CTabDev ctab;
    DevBuf akk;
    Arena* arena = nullptr;
};

}  // namespace hnt

struct hnt_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    // buffers reused across calls (grow only)
    hnt::DevBuf claim, counts, offsets, cub_tmp, fieldptrs, fcrit, akk, small;
    hnt::SAWork sa;  // shell analysis buffers of direct calls
    hnt::TablesDev tables;
    hnt::CTabDev ctab;
    hnt::MlpScratch mlp;
    hnt::DevBuf tiles, tilemeta, slots, gtiles;  // multi-tile shell analysis
    // tile peak finding: spectrum, windowed spectrum, smoothed fields, per-cell peak values
    hnt::DevBuf fb_spec, fb_work, fb_field, fb_lap, fb_val, fb_d2, fb_wtab, fb_taken;
    hnt::DevBuf rec_tabs, rec_keep, rec_slot;  // halo records: tables, flags, slots
    hnt::MergeBufs mg;                         // catalog merge
    // tile field synthesis: hipFFT plan pairs by size {n, r2c, c2r}, real and two complex
    // scratch arrays, the uploaded sqrt P(k) table
    struct FftPlans { int n; void* r2c; void* c2r; };
    std::vector<FftPlans> fs_plans;
    hnt::DevBuf fs_real, fs_base, fs_work, fs_aux, fs_tab;
    // the library's own power-of-two FFTs: twiddle tables by length; fft_mode 0 uses them
    // where the size allows, 1 always uses hipFFT
    struct FftTw { int L; hnt::DevBuf buf; };
    std::vector<FftTw> fft_tw;
    int fft_mode = 0;
    int fft_direct = 1;    // 1024- and 2048-point passes load straight into LDS
    int tricubic_tab = 1;  // interpolation weight products tabulated per row
    void* fft_r2c = nullptr;  // hipfftHandle of an n^3 real-to-complex plan
    void* fft_c2r = nullptr;
    int fft_n = 0;
    std::vector<hnt::TileDev> tile_host;
    int brick_log2 = 6;  // peak-order bricks of 2^brick_log2 cells a side (0..13)
    int waves_per_cu = 0;  // cap on resident shell-analysis waves per CU (0 = occupancy)
    int gang = 1;          // lanes per peak in the shell gather: 1 (exact order) or 16
    std::vector<int64_t> meta_host;
    // stage timers (hnt_ctx_set_profiling)
    bool profiling = false;
    std::vector<hipEvent_t> ev[HNT_NSTAGES];  // begin/end pairs of the current call
    int ev_n[HNT_NSTAGES] = {};                // pairs recorded in the current call
    bool ev_open[HNT_NSTAGES] = {};
    double stage_ms[HNT_NSTAGES] = {};
    int64_t stage_calls[HNT_NSTAGES] = {};
};

namespace hnt {

hnt_status reserve(hnt_ctx* ctx, DevBuf& b, size_t bytes);
void release(DevBuf& b);
void release_work(SAWork& w);
void release_fft(hnt_ctx* ctx);  // FFT plans of hnt_peak_find_tile
void release_field_plans(hnt_ctx* ctx);  // FFT plans of the field-synthesis entries

// Power-of-two 3-D real FFTs (fft.hip).  Operators applied to the spectrum as the inverse
// transform first reads it (the reference's _transfer_kernel!, _periodic_kernel!,
// _short_pot_kernel).
enum {
    FFT_OP_NONE = 0,
    FFT_OP_TRANSFER = 1,
    FFT_OP_PERIODIC = 2,
    FFT_OP_SHORTPOT = 3,
    FFT_OP_WINDOW = 4  // v * amp[(ix, |iy|, |iz|)], then * post (the filter bank's window)
};
struct FftOp {
    int kind = FFT_OP_NONE;
    int L = 0;                 // transform length
    double dk = 0.0;
    const float* amp = nullptr;  // per (ix, |iy|, |iz|), (L/2 + 1)^3 with |iz| fastest: sqrt P, or the window
    int id = 0, d1 = 0, d2 = 0;
    double rs2 = 0.0;
    float post = 1.0f;
};
bool fft_supported(int L);
// Row pitch (complex elements) of the library's half spectra: L/2 + 1 rounded up to 16, so
// that every row starts on a 128-B line and the column passes' 64-B pieces pair up in lines.
// A spectrum's rows are fft_pitch(L) elements apart (columns from L/2 + 1 on are never
// used) and its planes fft_plane(L): fft_pitch(L) x L and one 128-B line more, because the
// unpadded plane stride is a multiple of 32 KB for L >= 512 and put every row of a z column
// in the same L2 set and channel (measured: 1.5x the bytes of the 2048-point z pass, and
// half its speed); a spectrum takes fft_plane(L) x L elements.
__host__ __device__ constexpr int fft_pitch(int L) { return ((L / 2 + 1) + 15) & ~15; }
__host__ __device__ constexpr int64_t fft_plane(int L) { return (int64_t)fft_pitch(L) * L + 16; }
hnt_status fft_twiddles(hnt_ctx* ctx, int L, const float2** out);
// spec ((L/2+1) x L x L) = R2C of the w^3 field f zero-padded to L^3
hnt_status fft_forward(hnt_ctx* ctx, const float* f, int w, int L, float2* spec);
// out (n^3) = scale * C2R(op(spec)) cropped to [s, s + n)^3 (mode 0), or out -= (that)^2 * coef
// (mode 1); work: a spectrum-sized scratch
hnt_status fft_inverse(hnt_ctx* ctx, const float2* spec, float2* work, int L, int s, int n,
                       const FftOp& op, float scale, int mode, float coef, float* out);
hnt_status fft_amp_table(hnt_ctx* ctx, int L, double dk, const float* tab, int ntab,
                         float log_k_min, float d_log_k, float* amp);
// reserve(), or a carve of `bytes` from `ar` when it is not NULL (HNT_ERR_CAPACITY if the
// arena is full)
hnt_status reserve_in(hnt_ctx* ctx, DevBuf& b, size_t bytes, Arena* ar);
// Uploads into the context's copies, or into T / D / akk (carved from ar if not NULL)
hnt_status upload_tables(hnt_ctx* ctx, const hnt_shell_tables* t);
hnt_status upload_ctab(hnt_ctx* ctx, const hnt_ctab* c);
hnt_status ensure_akk(hnt_ctx* ctx);
hnt_status upload_tables_to(hnt_ctx* ctx, TablesDev& T, const hnt_shell_tables* t, Arena* ar);
hnt_status upload_ctab_to(hnt_ctx* ctx, CTabDev& D, const hnt_ctab* c, Arena* ar);
hnt_status ensure_akk_to(hnt_ctx* ctx, DevBuf& akk, Arena* ar);
// The fixed scratch of hnt_shell_analyse_noalloc: its device address on ctx's device,
// looked up once per device (shell_analyse.hip)
hnt_status noalloc_scratch_init(int device);
hnt_status begin_call(hnt_ctx* ctx);
hnt_status finish_call(hnt_ctx* ctx, const char* what);
hnt_status mlp_predict(hnt_ctx* ctx, const hnt_mlp* net, const float* x, int64_t B, float* out);
hnt_status bracket_features(hnt_ctx* ctx, const hnt_fields* fields, const hnt_batch* batch,
                            double alatt, float* out);
// bracket a stage with events when profiling is on (no-ops otherwise)
void stage_begin(hnt_ctx* ctx, int stage);
void stage_end(hnt_ctx* ctx, int stage);

}  // namespace hnt

#define HNT_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            hnt::set_error("%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__,       \
                           __LINE__);                                                      \
            return HNT_ERR_DEVICE;                                                         \
        }                                                                                  \
    } while (0)

#define HNT_TRY(expr)                       \
    do {                                    \
        hnt_status s_ = (expr);             \
        if (s_ != HNT_OK) return s_;        \
    } while (0)

#define HNT_REQUIRE(cond, ...)              \
    do {                                    \
        if (!(cond)) {                      \
            hnt::set_error(__VA_ARGS__);    \
            return HNT_ERR_INVALID;         \
        }                                   \
    } while (0)

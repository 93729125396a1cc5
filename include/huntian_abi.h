/*
 * huntian_abi.h -- C ABI of libhuntian_hip.so, the gfx950 (MI355X) implementation of
 * Huntian.jl's halo-finding hot path: multi-filter peak finding over smoothed density
 * fields, then the radial-shell homogeneous-ellipsoid collapse of every candidate.
 *
 * The ABI replaces the CUDA.jl package-extension methods that Huntian's drivers look up
 * with getglobal (src/MultiResolution.jl:492,559,1037).  Each entry point below names the
 * reference interface it replaces.  A Julia package extension binds these with ccall.
 *
 * Conventions
 *  - Plain C types only.  No exceptions cross the boundary: every call returns an
 *    hnt_status and, on failure, stores a message readable with hnt_last_error().
 *  - The caller owns every buffer.  Arrays documented as "device" are HIP device
 *    pointers (hipMalloc'd or equivalent); arrays documented as "host" are host memory.
 *  - 3-D fields keep the reference (Julia) layout: column-major Float32, i fastest,
 *    flat index i + (j-1)*n1 + (k-1)*n1*n2 with 1-based i, j, k.
 *  - Peak coordinates are 1-based, as in PeakCandidate (src/HaloFinder/PeakFind.jl:3-13).
 *  - One context per device; a context is not safe for concurrent calls from several
 *    threads (the reference binds one Julia task per device, src/Huntian.jl:414-424).
 *  - All work of a call is enqueued on the context's stream; the call returns after that
 *    stream has completed it (the reference's host copies are synchronous as well).
 */
#ifndef HUNTIAN_ABI_H
#define HUNTIAN_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HNT_ABI_VERSION 3

typedef enum {
    HNT_OK = 0,
    HNT_ERR_INVALID = 1,     /* bad argument (shape, NULL pointer, out-of-range index) */
    HNT_ERR_DEVICE = 2,      /* a HIP runtime call failed */
    HNT_ERR_CAPACITY = 3,    /* output capacity too small; hnt_peaks.total says how much is needed */
    HNT_ERR_UNSUPPORTED = 4, /* option not implemented by this library */
    HNT_ERR_ALLOC = 5        /* device or host allocation failed */
} hnt_status;

/* Result variants of the collapse step.  The reference holds three forms of it:
 *  A: RadialShell.analyse_peak, the FP64 CPU production path (src/HaloFinder/RadialShell.jl:281-865)
 *  B: ShellAnalysisGPU.analyse_peak_gpu, FP64 CPU form of the GPU algorithm
 *     (src/ShellAnalysisGPU.jl:81-535); zvir_half is returned as 0
 *  C: the CUDA extension's semantics (ext/CUDAExt.jl:1213-1912): B plus the formation
 *     redshift zvir_half of variant A.
 * All three are computed here in FP64 with the reference CPU path's operation order. */
enum { HNT_VARIANT_A = 0, HNT_VARIANT_B = 1, HNT_VARIANT_C = 2 };

typedef struct hnt_ctx hnt_ctx;

/* Context on HIP device `device` (0-based).  Replaces set_cuda_device! (src/Huntian.jl:414-424). */
hnt_status hnt_ctx_create(int device, hnt_ctx** out);
hnt_status hnt_ctx_destroy(hnt_ctx* ctx);
/* Free the device scratch the context keeps between calls (sized by the largest call so far:
 * cell records, spectra, sort and hash buffers) and its FFT plans; uploaded tables and
 * networks stay, and plans keep their own buffers.  No counterpart in the reference, which
 * frees its temporaries after every call (CUDA.unsafe_free!). */
hnt_status hnt_ctx_release_scratch(hnt_ctx* ctx);
/* Run subsequent calls on the caller's hipStream_t, used as given (NULL is the device's
 * default stream).  A new context uses a stream of its own until this is called. */
hnt_status hnt_ctx_set_stream(hnt_ctx* ctx, void* hip_stream);
/* Copy the calling thread's last error message into buf (NUL-terminated); returns its length. */
size_t hnt_last_error(char* buf, size_t len);
int hnt_abi_version(void);

/* Per-stage device timers, the counterpart of the reference's profile=true stage timers
 * ("08_peak_find", "09_shell_analysis", src/MultiResolution.jl:862-875,1068,1149), split
 * per kernel.  hnt_ctx_set_profiling(ctx, 1) zeroes the accumulators and brackets every
 * stage of later calls with HIP events on the context's stream;
 * hnt_ctx_get_profile fills ms[s] (summed device time) and calls[s] for s < nstages. */
enum {
    HNT_STAGE_PF_MARK = 0,      /* peak finding: filter sweep */
    HNT_STAGE_PF_COMPACT = 1,   /* peak finding: count, scan, scatter */
    HNT_STAGE_SA_PREP = 2,      /* shell analysis: validation, record repack, peak sort */
    HNT_STAGE_SA_KERNEL = 3,    /* shell analysis: gather and radial logic */
    HNT_STAGE_SA_CROSS = 4,     /* shell analysis: crossing pass and work-ordered peak sort */
    HNT_STAGE_PF_SMOOTH = 5,    /* tile peak finding: FFTs and window products */
    HNT_STAGE_SA_MASK = 6,      /* shell analysis: marking the halos' cells in the mask */
    HNT_NSTAGES = 7
};
hnt_status hnt_ctx_set_profiling(hnt_ctx* ctx, int on);
hnt_status hnt_ctx_get_profile(hnt_ctx* ctx, double* ms, int64_t* calls, int32_t nstages);

/* Launch-shape knobs of a context (no counterpart in the reference; set by the autotuner,
 * huntian_hip/autotune.py).  They change speed, never results.
 *  BRICK_LOG2:   shell-analysis peak order groups peaks in bricks of 2^v cells a side
 *                (0..13; 13 = no bricks), default 6
 *  WAVES_PER_CU: cap on resident shell-analysis wavefronts per CU (0 = as many as fit)
 *  GANG:         lanes per peak in the shell gather.  1: each lane sums its peak's cells in
 *                the reference's order (results equal the FP64 CPU path bit for bit);
 *                16: sixteen lanes share a peak and their partial sums are added in lane
 *                order (the same FP64 sums in another order, ~1e-13 relative);
 *                64: a wavefront loads a peak's cells together into LDS and 22 of its
 *                lanes each run one sum in the reference's order (bit for bit, like 1)
 *  FFT:          transforms of the field-synthesis entries.  0 (default): the library's own
 *                FFTs for power-of-two sizes 8..2048 (hipFFT otherwise); 1: hipFFT always.
 *                Both are Float32 transforms; results differ by their rounding.
 *  FFT_DIRECT:   1 (default): the library's 1024- and 2048-point column passes (reading
 *                whole 1 KB blocks of rows, the rest zero) and its 1024- and 2048-point
 *                complex-to-real row passes load straight into LDS (no registers held, a
 *                workgroup's 64 KB in flight at once); 0: through registers, two elements
 *                per thread at a time.  Speed only: the results are the same bit for bit.
 *  TRICUBIC_TAB: 1 (default): hnt_interpolate_to_tile with a power-of-two block of at most
 *                32 fine cells per coarse cell forms each row's weight products once per
 *                position in the coarse cell (a table in LDS); 0: per cell.  Speed only:
 *                the same products in the same order, the same results bit for bit. */
enum { HNT_TUNE_BRICK_LOG2 = 0, HNT_TUNE_WAVES_PER_CU = 1, HNT_TUNE_GANG = 2, HNT_TUNE_FFT = 3,
       HNT_TUNE_FFT_DIRECT = 4, HNT_TUNE_TRICUBIC_TAB = 5 };
hnt_status hnt_ctx_set_tuning(hnt_ctx* ctx, int32_t what, int32_t value);

/* ---------------------------------------------------------------- shells and tables */

/* Lattice offsets within radius rmax, sorted by r^2 with a stable sort, grouped into shells
 * of equal r^2 (RadialShell.precompute_shells, src/HaloFinder/RadialShell.jl:151-166;
 * ShellAnalysisGPU.build_shell_tables, src/ShellAnalysisGPU.jl:44-71).  All arrays host. */
typedef struct {
    int32_t rmax;
    int32_t nshells;           /* distinct r^2 values, shell 1 (index 0 here) is r^2 = 0 */
    int64_t ncells;
    const int32_t* di;         /* [ncells] */
    const int32_t* dj;
    const int32_t* dk;
    const int32_t* shell_start;  /* [nshells], 0-based offset into di/dj/dk */
    const int32_t* shell_count;  /* [nshells] */
    const int32_t* shell_r2;     /* [nshells] */
} hnt_shell_tables;

/* Number of cells and shells for rmax (to size the arrays of hnt_build_shell_tables). */
hnt_status hnt_shell_table_size(int32_t rmax, int64_t* ncells, int32_t* nshells);
/* Fill caller arrays (host) with the shell tables of radius rmax. */
hnt_status hnt_build_shell_tables(int32_t rmax, int32_t* di, int32_t* dj, int32_t* dk,
                                  int32_t* shell_start, int32_t* shell_count, int32_t* shell_r2);

/* HomelTab collapse table: value 1 + z_vir on a regular grid of (log10 F, e, p/e).
 * Layout of read_homeltab (src/EllipsoidalCollapse/CollapseTable.jl:96-115): column-major
 * Float32 table, axis bounds widened to Float64.  Queries outside the box return out_val. */
typedef struct {
    int32_t nx, ny, nz;
    double x1, x2, y1, y2, z1, z2;
    double out_val;            /* -1 in the reference (CollapseTable.jl:132) */
    const float* table;        /* host, [nx*ny*nz] */
} hnt_ctab;

/* fsc_of_z: critical linear overdensity for spherical collapse at redshift z, by 50-step
 * bisection of the table (src/HaloFinder/RadialShell.jl:181-195).  Host computation. */
double hnt_fsc_of_z(const hnt_ctab* ct, double z);

/* ---------------------------------------------------------------- peak finding */

typedef struct {
    int64_t capacity;   /* length of the arrays below */
    int32_t* i;         /* device [capacity], 1-based grid indices */
    int32_t* j;
    int32_t* k;
    int32_t* filt;      /* device [capacity], 0-based index into the filter list passed */
    float* delta;       /* device [capacity], smoothed delta at the peak (FcRf) */
    int64_t count;      /* out: peaks written = min(total, capacity) */
    int64_t total;      /* out: peak cells found, i.e. the capacity an uncapped call needs.
                           (The reference leaves a peak beyond max_peaks unmasked, so a later
                           filter can count its cell again; its counter is not returned.) */
} hnt_peaks;

/*
 * hnt_peak_find -- replaces the peak-finding loop of peak_find_tile_gpu
 * (ext/CUDAExt.jl:2938-3060, stub src/Huntian.jl:387-412) and the CPU loop
 * `for filter: find_peaks(...)` (src/MultiResolution.jl:1046-1066,
 * src/HaloFinder/PeakFind.jl:15-78) on already-smoothed fields.
 *
 * smoothed: host array of nfilt device pointers, each an n^3 Float32 field, in the
 *           order the filters are processed (the reference sorts largest radius first).
 * fcrit:    host [nfilt] Float32 peak thresholds.
 * mask:     device n^3 Int8, read and updated: every peak sets its own cell to 1.
 * A cell of the inner cube (nbuff+1 : n-nbuff)^3 is a peak of filter f when the mask is
 * not 1 there before the call, delta_f >= fcrit[f], no 26-neighbour of delta_f is larger
 * (ties accepted), and no earlier filter took it.  Output order is the reference CPU
 * order: filter, then k, j, i.  If total > capacity, the first `capacity` peaks in that
 * order are written (and masked) and HNT_ERR_CAPACITY is returned.
 */
hnt_status hnt_peak_find(hnt_ctx* ctx, const float* const* smoothed, const float* fcrit,
                         int32_t nfilt, int32_t n, int32_t nbuff, int8_t* mask, hnt_peaks* out);

/* Filter bank of hnt_peak_find_tile (the `filters`, `fcrits`, `wsmooth` arguments of
 * peak_find_tile_gpu, src/Huntian.jl:387-412). */
typedef struct {
    int32_t nfilt;          /* 1..254 */
    const double* Rf;       /* host [nfilt] smoothing radii, Mpc/h, in processing order */
    const float* fcrit;     /* host [nfilt] peak thresholds */
    int32_t wsmooth;        /* window: 0 Gaussian exp(-(kR/2)^2/2), 1 top-hat
                               3(sin x - x cos x)/x^3, 3 k^2 times the Gaussian */
    double alatt;           /* cell size, Mpc/h: the tile is periodic with box n * alatt */
} hnt_filter_bank;

/*
 * hnt_peak_find_tile -- replaces peak_find_tile_gpu (ext/CUDAExt.jl:2938-3060, stub
 * src/Huntian.jl:387-412): filter-bank smoothing of the tile followed by the peak test of
 * hnt_peak_find, on the device.  delta (device n^3 Float32, column-major) is transformed once
 * (real-to-complex, Float32, as the reference's cuFFT rfft); then for each filter in order
 * the spectrum times W(kR) (W in FP64, rounded to Float32, k = 0 left unchanged,
 * _smooth_window_kernel! ext/CUDAExt.jl:2831-2871) is transformed back (normalised by 1/n^3)
 * and its peaks are claimed at once, so one smoothed field is held at a time.  Peaks, order,
 * mask update and capacity rule are those of hnt_peak_find on the same smoothed fields.
 * d2Rf:         device [out->capacity] or NULL.  If given, each filter's field is also
 *               smoothed with window 3 (k^2 Gaussian) and sampled at that filter's peaks
 *               (the lapd_s of the reference, its d2Rf output).
 * smoothed_out: host array of nfilt device pointers (n^3 Float32 each; an entry may be
 *               NULL), or NULL: copies of the smoothed fields, for checking.
 */
hnt_status hnt_peak_find_tile(hnt_ctx* ctx, const float* delta, int32_t n, int32_t nbuff,
                              const hnt_filter_bank* bank, int8_t* mask, hnt_peaks* out,
                              float* d2Rf, float* const* smoothed_out);

/* ---------------------------------------------------------------- shell analysis */

typedef struct {
    int32_t n1, n2, n3;        /* the transposed read delta[k,i,j] needs n1 == n2 == n3 */
    const float* delta;        /* device fields, n1*n2*n3 Float32 each */
    const float* etax;         /* 1LPT displacement psi_1 */
    const float* etay;
    const float* etaz;
    const float* eta2x;        /* 2LPT displacement psi_2, or NULL (read as zeros) */
    const float* eta2y;
    const float* eta2z;
    const float* lapd;         /* Laplacian of delta for d2F, or NULL (d2F = 0) */
} hnt_fields;

typedef struct {
    int64_t npeaks;
    const int32_t* pi;         /* device [npeaks], 1-based peak cell */
    const int32_t* pj;
    const int32_t* pk;
    const double* Rf;          /* device [npeaks], smoothing radius that found the peak (Mpc/h) */
    const int32_t* ir2min;     /* device [npeaks], first-crossing search floor in r^2 */
    const double* ZZon;        /* device [npeaks] 1 + z_collapse, or NULL: use ZZon_scalar */
    const double* fcrit;       /* device [npeaks] crossing threshold, or NULL: fcrit_scalar */
    double ZZon_scalar;
    double fcrit_scalar;
} hnt_batch;

typedef struct {
    double alatt;              /* lattice spacing (Mpc/h) */
    int32_t nbuff;             /* mask is only written on (nbuff+1 : n-nbuff)^3 */
    double rmax2rs;            /* 0: every shell of the tables.  > 0 (variants B and C; A is
                                  HNT_ERR_UNSUPPORTED): a peak uses only the leading shells with
                                  r^2 <= (rmax2rs Rf / alatt)^2, none if fewer than 2
                                  (ShellAnalysisGPU.jl:90-99, ext/CUDAExt.jl:2176-2185); its
                                  gather takes all of them (early_exit and order are ignored,
                                  ext/CUDAExt.jl:2235-2236); variant B with a mask skips a peak
                                  whose cell the mask holds beyond floor(Rf / alatt rmax2rs)
                                  as the call finds it (ShellAnalysisGPU.jl:112-117) */
    int32_t variant;           /* HNT_VARIANT_A / _B / _C */
    int32_t early_exit;        /* 1: stop gathering once no later shell can change a result */
    int32_t order;             /* wavefront schedule of the peaks, HNT_ORDER_* (0: crossing pass);
                                  it changes the speed, never a result */
    int32_t precision;         /* per-shell sums: HNT_PREC_FP64 (0, the reference CPU arithmetic)
                                  or HNT_PREC_FP32_KAHAN */
    const struct hnt_mlp* bracket;  /* network for HNT_ORDER_LEARNED, else unused */
    const int32_t* extent;     /* HNT_ORDER_GIVEN: device [total peaks of the call, tiles in
                                  order] gather extent (last shell) per peak, e.g. the
                                  nshells_used of an earlier call; else unused */
} hnt_params;

/* Per-shell sums of the gather.  FP64: FP64 sums of the Float32 terms, as the reference CPU
 * path.  FP32_KAHAN: FP32 sums with Kahan compensation; the radial logic stays FP64. */
enum { HNT_PREC_FP64 = 0, HNT_PREC_FP32_KAHAN = 1 };

/* Peak schedules of hnt_shell_analyse (early_exit = 1).  None changes a result.
 *  CROSSING: a delta-only pass finds each peak's crossing shell, hence its exact gather
 *            extent; wavefronts hold peaks of equal extent.
 *  PLAIN:    (ir2min, Morton code) order, no extra pass.
 *  LEARNED:  the bracket network reads 18 features of the first 12 shells and proposes
 *            an RTHL bracket [lo, hi] (lo: the shell holding RTHL, hi: the last shell the
 *            exact search reads, i.e. the gather extent) and a collapse logit; peaks are
 *            ordered by hi.  The exact search then runs inside the bracket: every lane
 *            gathers until the exact stop rule holds, and the walk takes the shells from
 *            the crossing down to max(lo - 2, 1) first; a gather that runs past hi + 2 or
 *            a walk that passes lo - 2 without collapsing fails verification and goes on as
 *            the full scan (hnt_results.bracket records which).
 *  GIVEN:    peaks ordered by the extents in hnt_params.extent (no extra pass): with the
 *            exact extents of an earlier call, the best order any bracket can give. */
enum { HNT_ORDER_CROSSING = 0, HNT_ORDER_PLAIN = 1, HNT_ORDER_LEARNED = 2, HNT_ORDER_GIVEN = 3 };

/* Per-peak results, device arrays of Float64 (caller-allocated, npeaks entries each;
 * vector fields hold 3 consecutive values per peak, strain 9 in column-major order).
 * No-collapse peaks get RTHL = -1, zvir_half = -1 and zeros elsewhere
 * (RadialShell.no_collapse, src/HaloFinder/RadialShell.jl:86-90). */
typedef struct {
    double* RTHL;              /* Lagrangian radius in lattice units */
    double* Fbarx;             /* mean overdensity within RTHL */
    double* e_v;
    double* p_v;
    double* strain;            /* [9*npeaks] */
    double* eigs;              /* [3*npeaks] eigenvalues of the trace-matched strain */
    double* Srb;
    double* Sbar;              /* [3*npeaks] */
    double* Sbar2;             /* [3*npeaks] */
    double* gradpk;            /* [3*npeaks] */
    double* gradpkf;           /* [3*npeaks] */
    double* gradpkrf;          /* [3*npeaks] */
    double* d2F;
    double* zvir_half;
    int32_t* nshells_used;     /* optional (NULL): shells gathered per peak, incl. the centre */
    int32_t* crossing_shell;   /* optional (NULL): first shell (1-based) past ir2min whose Fbar
                                  is below fcrit, 0 if there is none within the hunt radius */
    int32_t* bracket;          /* optional (NULL), with one lane per peak (HNT_TUNE_GANG 1):
                                  HNT_ORDER_LEARNED: bit 0 the gather ended within hi + 2,
                                  bit 1 the peak collapsed and the shell holding RTHL is
                                  >= lo - 2 (both set: the bracket verified), bit 2 the
                                  collapse logit predicted the collapse flag; other orders: 0 */
    int32_t* schedule_extent;  /* optional (NULL), HNT_ORDER_CROSSING: each peak's gather extent
                                  as the crossing pass finds it, the schedule's sort key (equal
                                  to nshells_used with FP64 sums); written by the call that
                                  orders the peaks (a plan: when it is built); other orders:
                                  left as it is */
} hnt_results;

/*
 * hnt_shell_analyse -- replaces analyse_peaks_gpu_cuda_multirf (ext/CUDAExt.jl:2085-2344,
 * stub src/Huntian.jl:259-281) and analyse_peak_gpu_cuda (ext/CUDAExt.jl:1914-2078), i.e.
 * the radial-shell collapse of every peak in `batch`: shell gather of delta, psi_1,
 * psi_2, the transposed delta and lapd; Fbar crossing of fcrit; inward walk with
 * kernel_strain, eigenvalues and the HomelTab lookup until zvir >= ZZon; RTHL and the
 * record fields.  mask (device n^3 Int8 or NULL) gets 1 on non-buffer cells within RTHL
 * of every collapsed peak.
 */
hnt_status hnt_shell_analyse(hnt_ctx* ctx, const hnt_fields* fields, const hnt_batch* batch,
                             const hnt_shell_tables* tables, const hnt_ctab* ctab,
                             const hnt_params* params, hnt_results* results, int8_t* mask);

/*
 * hnt_shell_analyse_noalloc -- no counterpart in the reference.  hnt_shell_analyse for
 * callers that cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc): same arguments, same results bit for bit, but it
 * calls no hipMalloc / hipFree and keeps nothing in the context.  All of its device scratch
 * (cell records 32 B per cell, sort keys and lane slots ~30 B per peak, the shell, collapse
 * and SPH kernel tables, profile slots 192 B per shell per resident lane) lives in one
 * fixed 256 MiB device array that each GPU the library runs on holds until the process
 * ends; fewer wavefronts stay resident when the profile slots do not fit beside the rest.
 * A call that needs more (e.g. a 128^3 tile with hunt radius 15 fits, a 256^3 tile does
 * not: its records alone are 512 MiB) returns HNT_ERR_CAPACITY with the sizes in
 * hnt_last_error.  Every call on one GPU shares that array: calls that overlap on a GPU
 * (several host threads, or several contexts' streams) race and must be serialised by the
 * caller.  The learned schedule (HNT_ORDER_LEARNED) is not available here
 * (HNT_ERR_UNSUPPORTED).
 */
hnt_status hnt_shell_analyse_noalloc(hnt_ctx* ctx, const hnt_fields* fields,
                                     const hnt_batch* batch, const hnt_shell_tables* tables,
                                     const hnt_ctab* ctab, const hnt_params* params,
                                     hnt_results* results, int8_t* mask);

/*
 * hnt_shell_analyse_tiles -- hnt_shell_analyse over many tiles in one persistent launch (no
 * counterpart in the reference, whose drivers call analyse_peaks_gpu_cuda_multirf once per
 * tile, src/MultiResolution.jl:532-644).  fields, batches, results: arrays of ntiles
 * structs; masks: host array of ntiles device pointers (entries may be NULL) or NULL.  Tiles
 * may differ in size; peaks of all tiles share the work queues, so tiles of uneven
 * candidate count and shell depth keep every CU busy.  Results equal those of one
 * hnt_shell_analyse call per tile.
 */
hnt_status hnt_shell_analyse_tiles(hnt_ctx* ctx, int32_t ntiles, const hnt_fields* fields,
                                   const hnt_batch* batches, const hnt_shell_tables* tables,
                                   const hnt_ctab* ctab, const hnt_params* params,
                                   hnt_results* results, int8_t* const* masks);

/*
 * Plans: hnt_shell_analyse_tiles recorded once and replayed as one hipGraph (no counterpart
 * in the reference).  hnt_plan_create runs the analysis eagerly (results are written) and
 * records its device work -- the record repacks, variant B's pre-mask flags (rmax2rs > 0)
 * and the persistent analysis kernel -- in a graph of kernel nodes only.  hnt_plan_launch replays it on the context's stream: it
 * re-reads the fields, peak arrays, ZZon and fcrit at the same addresses (contents may
 * change between launches, sizes and pointers may not) and rewrites results and masks.  The
 * peak schedule of the creating call is kept, so a replay over changed fields is exact but
 * may be slower.  A plan owns its device buffers; it must not outlive its context.
 */
typedef struct hnt_plan hnt_plan;
hnt_status hnt_plan_create(hnt_ctx* ctx, int32_t ntiles, const hnt_fields* fields,
                           const hnt_batch* batches, const hnt_shell_tables* tables,
                           const hnt_ctab* ctab, const hnt_params* params, hnt_results* results,
                           int8_t* const* masks, hnt_plan** out);
hnt_status hnt_plan_launch(hnt_plan* plan);
hnt_status hnt_plan_destroy(hnt_plan* plan);

/* ---------------------------------------------------------------- halo records */

/* Cosmology tables of hnt_pack_records: host arrays the caller builds once, as the reference
 * driver does (src/MultiResolution.jl:721,743).
 *  growth: Dlinear_tables (src/Cosmology/Cosmology.jl:161-284): ntab ascending scale factors
 *          atab and D(a)/a in D_atab, normalised to D(1) = 1;
 *  chi2z:  build_chi_to_z (:343-350): nchi redshifts (uniform from 0 to chi2z_zmax) and their
 *          comoving distances, Mpc/h; needed in lightcone mode only (nchi = 0 otherwise);
 *  dlnD_atab: the growth rate dlnD/dlna at atab (Dlinear_tables' third column); needed by
 *          hnt_finalize_eulerian only (NULL otherwise). */
typedef struct {
    int32_t ntab;
    const double* atab;
    const double* D_atab;
    int32_t nchi;
    const double* chi2z_z;
    const double* chi2z_chi;
    double chi2z_zmax;
    const double* dlnD_atab;
} hnt_cosmo_tables;

/* What the record loop of the tile driver knows (src/MultiResolution.jl:1159-1219). */
typedef struct {
    double alatt;           /* cell size, Mpc/h */
    int32_t n;              /* tile side; peak position x = xbx + alatt (i - (n + 1)/2) */
    double xbx, ybx, zbx;   /* tile centre, Mpc/h */
    double ZZon;            /* 1 + z of the output (ievol == 0) */
    int32_t ievol;          /* 1: lightcone, each peak's z from its distance to obs */
    double obs[3];          /* observer, Mpc/h */
    double z_max;           /* lightcone: a peak with z > z_max gives no record */
    int32_t ilpt;           /* >= 2: 2LPT terms vx2, vy2, vz2 (else 0) */
    double Om, OL;          /* Omega_m, Omega_Lambda of omega_m_a */
    int32_t ioutshear;      /* 0: 11-float HaloRecord, >= 1: 33-float ExtHaloRecord */
} hnt_pack_params;

/*
 * hnt_pack_records -- replaces the record loop of the tile driver after the shell analysis
 * (src/MultiResolution.jl:1159-1219) in the HaloRecord / ExtHaloRecord layout of
 * src/IO/Catalog.jl:4-46, on the device.  For every peak p of `batch` (pi, pj, pk and Rf
 * are read) whose results->RTHL[p] > 0 (and, in lightcone mode, whose z <= z_max), in peak
 * order: D = Dlinear_ab(1/ZZon_p) (4-point Lagrange in the growth table), the record
 *   x y z, Sbar*D, RTHL*alatt, Sbar2*(-3/7 Omega_m(a)^(-1/143) D^2) (ilpt >= 2, else 0),
 *   Fbarx  [+ e_v p_v, strain 11 22 33 23 13 12, d2F, zvir_half, gradpk, gradpkf, Rf,
 *   FcRf, d2Rf, gradpkrf]
 * in Float32, from the Float32 values of the results as the reference's GPU results are.
 * records: device [capacity * (11 or 33)] Float32; *count: records of the tile.  FcRf:
 * device [npeaks] (the smoothed delta at each peak, hnt_peaks.delta); d2Rf: device [npeaks]
 * or NULL (zeros).  If *count > capacity the first `capacity` records are written and
 * HNT_ERR_CAPACITY is returned.
 */
hnt_status hnt_pack_records(hnt_ctx* ctx, const hnt_batch* batch, const hnt_results* results,
                            const float* FcRf, const float* d2Rf, const hnt_cosmo_tables* tabs,
                            const hnt_pack_params* pp, float* records, int64_t capacity,
                            int64_t* count);

/* hnt_finalize_eulerian -- replaces Merger.finalize_eulerian (src/Merger/Merger.jl:107-166):
 * a raw catalog (Lagrangian positions in x y z, 1LPT displacement in vx vy vz, 2LPT
 * displacement in vx2 vy2 vz2, Mpc/h) to Eulerian positions x + d1 + d2 and peculiar
 * velocities a 100 E(a) f(a) (d1 + 2 d2) km/s, E = sqrt(Om a^-3 + OL), f = Dlinear_ab(a)[2];
 * a from the Eulerian distance to obs through the chi -> z table (ievol = 1; a = 1 at the
 * observer) or 1 / (1 + z_out) (ievol = 0).  vx2 vy2 vz2 are set to 0, every other field is
 * kept.  records, out: device [nhalo * nfields] (11 or 33), out may equal records.
 * tabs needs atab and dlnD_atab (and the chi -> z table for ievol = 1). */
hnt_status hnt_finalize_eulerian(hnt_ctx* ctx, const float* records, int64_t nhalo,
                                 int32_t nfields, const hnt_cosmo_tables* tabs, double Om,
                                 double OL, const double* obs, int32_t ievol, double z_out,
                                 float* out);

/* ---------------------------------------------------------------- catalog merge */

/*
 * hnt_merge_catalog -- replaces Merger.merge_catalog (src/Merger/Merger.jl:23-89) with the
 * spatial hash, Lagrangian exclusion and volume reduction of src/Merger/Exclusion.jl:13-208,
 * on the device.  records: device [nhalo * nfields] Float32 (11 or 33 fields, Catalog.jl
 * layout; x, y, z are fields 0-2, RTHL field 6).  Halos are ranked by (-RTHL, x, y, z)
 * (catalog order breaks exact ties); a halo is excluded when a surviving halo of higher rank
 * contains its centre (distance < that halo's radius).  With volume_reduction, each
 * survivor loses the caps of its overlaps with the other survivors, summed in the
 * reference's order, r_new = cbrt(3 (4/3 pi r^3 - dV) / (4 pi)), and a halo left with no
 * volume is removed; fortran_ties counts equal-radius pairs twice, as the Fortran code.
 * out: device [nhalo * nfields], not overlapping records: the survivors in catalog order with
 * RTHL = r_new; *nout: their number.  Non-finite positions or radii, or a zero extent (one
 * halo without a positive radius), give HNT_ERR_INVALID.  Scratch kept by the context: about
 * 80 B per halo, 16 B per containment pair, 32 B per overlapping pair (with reduction) and
 * 4 B per hash cell (nc^3, nc = clamp(round(cbrt(nhalo)), 16, 1024)).
 */
hnt_status hnt_merge_catalog(hnt_ctx* ctx, const float* records, int64_t nhalo,
                             int32_t nfields, int32_t volume_reduction, int32_t fortran_ties,
                             float* out, int64_t* nout);

/* ---------------------------------------------------------------- tile field synthesis */
/* The GPU methods that build a tile's input fields (src/MultiResolution.jl:881-1010 calls
 * them through the stubs of src/Huntian.jl:283-385).  Fields are device Float32 arrays,
 * column-major [i,j,k] (i fastest), as the reference's CuArrays.  Transforms are the
 * library's own for power-of-two sizes up to 2048 and hipFFT otherwise (see HNT_TUNE_FFT);
 * inverse transforms are normalised (1/n^3), as AbstractFFTs' irfft. */

/* sqrt(P(k) dk^3 n2^3) tabulated at pk_table_n points uniform in ln k from ln(dk/2) to
 * ln(1.5 sqrt(3) (n2/2) dk) (host array; built by the caller as isolated_convolve_gpu does,
 * ext/CUDAExt.jl:2489-2497), with log_k_min = ln k of the first point and d_log_k the step,
 * both as Float32. */
typedef struct {
    const float* sqrt_pk;
    int32_t n;
    float log_k_min;
    float d_log_k;
} hnt_pk_table;

/* One output of an isolated convolution: id 0 delta (sqrt P), 1 1LPT (i k_d1/k^2 sqrt P),
 * 2 2LPT (-i k_d1/k^2 sqrt P), 3 phi_ij (-k_d1 k_d2/k^2 sqrt P), 4 k^2 sqrt P,
 * 5 -sqrt P/k^2; dims 1..3 (the reference's kernel_fn_id, dim1, dim2). */
typedef struct {
    int32_t id;
    int32_t dim1;
    int32_t dim2;
} hnt_transfer;

/* hnt_isolated_convolve -- replaces isolated_convolve_gpu and isolated_convolve_gpu_multi
 * (ext/CUDAExt.jl:2346-2622; stubs src/Huntian.jl:283-324).  noise: device n_input^3
 * (n_input = n + 2 nshell); the noise is zero-padded to n2 = 2 n_input, transformed once,
 * and for each of the nkernels transfers: multiplied by the table's sqrt P(k) and the
 * transfer (k = 0 set to 0), the kx = 0 and kx = Nyquist planes made Hermitian (conjugate
 * pairs averaged, self-conjugate modes made real), inverse transformed, and the cells
 * nshell+1 .. nshell+n of each axis written to out[q] (device n^3).  dk = 2 pi / (n2 dx),
 * dx = boxsize_local / n.  Scratch kept by the context: two half spectra ((n2/2+1) n2^2
 * complex) and a (n2/2+1)^3 float sqrt P table, ~73 GB at n_input = 1024. */
hnt_status hnt_isolated_convolve(hnt_ctx* ctx, const float* noise, int32_t n_input, int32_t n,
                                 double boxsize_local, const hnt_pk_table* table,
                                 const hnt_transfer* kernels, int32_t nkernels, int32_t nshell,
                                 float* const* out);

/* hnt_compute_2lpt -- replaces compute_2lpt_gpu (ext/CUDAExt.jl:2724-2791; stub
 * src/Huntian.jl:326-342): on the periodic n^3 tile, src2 = delta^2/2 - sum_i phi_ii^2/2 -
 * sum_{i<j} phi_ij^2 with phi_ij = irfft(-k_i k_j/k^2 rfft(delta)) (k = 0 zeroed), then
 * psi2[d] = irfft(-i k_d/k^2 rfft(src2)) with k = 0 and the Nyquist planes zeroed;
 * dk = 2 pi / boxsize_local.  delta is not modified; psi2: three device n^3 outputs. */
hnt_status hnt_compute_2lpt(hnt_ctx* ctx, const float* delta, int32_t n, double boxsize_local,
                            float* const* psi2);

/* hnt_compute_laplacian -- replaces compute_laplacian_gpu (ext/CUDAExt.jl:2793-2825; stub
 * src/Huntian.jl:344-353): lapd = irfft(k^2 rfft(delta)) with k = 0 and the Nyquist planes
 * zeroed (k^2, not -k^2, as the reference). */
hnt_status hnt_compute_laplacian(hnt_ctx* ctx, const float* delta, int32_t n,
                                 double boxsize_local, float* lapd);

/* hnt_isolated_poisson_psi -- replaces isolated_poisson_psi_gpu (ext/CUDAExt.jl:2679-2701;
 * stub src/Huntian.jl:355-363): psi[d] = the first n^3 cells of irfft(i k_d/k^2 rfft(dS
 * zero-padded to 2n)), k = 0 and the padded Nyquist planes zeroed, dk = 2 pi/(2 boxsize). */
hnt_status hnt_isolated_poisson_psi(hnt_ctx* ctx, const float* dS, int32_t n,
                                    double boxsize_local, float* const* psi);

/* hnt_isolated_short_pot -- replaces isolated_short_pot_gpu (ext/CUDAExt.jl:2703-2722; stub
 * src/Huntian.jl:365-373): the first n^3 cells of irfft(-(1 - exp(-k^2 rs^2))/k^2 rfft(delta
 * zero-padded to 2n)), -rs^2 at k = 0; rs in Mpc/h. */
hnt_status hnt_isolated_short_pot(hnt_ctx* ctx, const float* delta, int32_t n,
                                  double boxsize_local, double rs, float* out);

/* hnt_interpolate_to_tile -- replaces interpolate_to_tile_gpu (ext/CUDAExt.jl:3062-3176;
 * stub src/Huntian.jl:375-385): tri-cubic Catmull-Rom interpolation of the periodic M^3
 * coarse grid (device) onto tile (it, jt, kt)'s nmesh^3 cells, cell-centred on both grids,
 * in Float32 with the reference's operation order (block = N / M, nbuff = (nmesh-nsub)/2,
 * the tile starts at global cell (it-1) nsub + 1 - nbuff, wrapped mod N). */
hnt_status hnt_interpolate_to_tile(hnt_ctx* ctx, const float* coarse, int32_t M, int32_t N,
                                   int32_t it, int32_t jt, int32_t kt, int32_t nsub,
                                   int32_t nmesh, float* tile);

/* The tile driver's white noise (the Threefry-2x64-20 Gaussian stream of the global N^3
 * grid, RandomField._threefry_gaussian, src/InitialConditions/RandomField.jl:45-53: cell idx
 * (0-based, column-major) takes Gaussian idx & 1 of the Box-Muller pair of counter
 * (idx >> 1, 0) under key (seed, 0)).  Outputs are device Float32, column-major.
 *
 * hnt_tile_noise -- replaces _generate_tile_noise (src/MultiResolution.jl:35-57): the
 *   nmesh^3 cells of tile (it, jt, kt) (1-based), starting at global cell
 *   (it-1) nsub + 1 - (nmesh - nsub) / 2 on each axis, wrapped mod N.
 * hnt_downsample_noise -- replaces _downsample_noise (src/MultiResolution.jl:64-87): coarse
 *   [M^3], each the FP64 sum of its block^3 fine Gaussians (block = N / M, N a multiple of
 *   M) in the reference's loop order, divided by sqrt(block^3).
 * hnt_extended_residual -- replaces _generate_extended_residual (src/MultiResolution.jl:
 *   190-227): the (nmesh + 2 nshell)^3 cells around tile (it, jt, kt), each its Float32
 *   Gaussian less coarse[block cell] / sqrt(block^3) (Float32), coarse a device [M^3] array
 *   (the output of hnt_downsample_noise). */
hnt_status hnt_tile_noise(hnt_ctx* ctx, int32_t it, int32_t jt, int32_t kt, int32_t nsub,
                          int32_t nmesh, int64_t N, uint64_t seed, float* out);
hnt_status hnt_downsample_noise(hnt_ctx* ctx, int64_t N, int32_t M, uint64_t seed,
                                float* coarse);
hnt_status hnt_extended_residual(hnt_ctx* ctx, int32_t it, int32_t jt, int32_t kt,
                                 int32_t nsub, int32_t nmesh, int64_t N, uint64_t seed,
                                 const float* coarse, int32_t M, int32_t nshell, float* out);

/* ---------------------------------------------------------------- primitive entries */
/* Single-primitive entry points; the reference exposes them as test harnesses
 * (src/Huntian.jl:98-217). */

/* Full per-shell gather (shell_gather_full_gpu, ext/CUDAExt.jl:961-1026) for all shells of
 * every peak.  Device outputs, Float64, shell-major per peak:
 * F, lapd [nshells*np]; n int32 [nshells*np]; S, S2, G, Gf [3*nshells*np]; SR [9*nshells*np]
 * (component L fastest, then K for SR).  G, Gf, SR are not divided by the shell radius. */
hnt_status hnt_shell_gather_full(hnt_ctx* ctx, const hnt_fields* fields, int64_t npeaks,
                                 const int32_t* pi, const int32_t* pj, const int32_t* pk,
                                 const hnt_shell_tables* tables,
                                 double* F, int32_t* n, double* S, double* S2, double* G,
                                 double* Gf, double* SR, double* lapd);

/* Ascending eigenvalues of symmetric 3x3 matrices (eig3_symmetric_batch_gpu,
 * ext/CUDAExt.jl:1030-1058; CPU get_evals, RadialShell.jl:215-220).  mats: device
 * [9*n] column-major, upper triangle used; out: device [3*n]. */
hnt_status hnt_eig3_symmetric_batch(hnt_ctx* ctx, const double* mats, int64_t n, double* out);

/* Trilinear table lookup with out-of-box -> out_val (interp3_trilinear_gpu,
 * ext/CUDAExt.jl:1060-1110; CollapseTable.interpolate).  x, y, z, out: device [n]. */
hnt_status hnt_interp3_trilinear(hnt_ctx* ctx, const hnt_ctab* ctab, const double* x,
                                 const double* y, const double* z, int64_t n, double* out);

/* One kernel_strain evaluation (kernel_strain_gpu, ext/CUDAExt.jl:1113-1166;
 * RadialShell.kernel_strain, RadialShell.jl:225-274).  rad [nshells], G, Gf [3*nshells],
 * SR [9*nshells] are host arrays; mp, mlow, mupp are 1-based shell indices.
 * out (host, 12): Ebar 11,12,13,22,23,33, grad xyz, gradf xyz. */
hnt_status hnt_kernel_strain(hnt_ctx* ctx, const double* rad, const double* G,
                             const double* Gf, const double* SR, int32_t nshells, int32_t mp,
                             int32_t mlow, int32_t mupp, double wRnor, double aRnor,
                             double hlatt_1, double hlatt_2, double* out);

/* ---------------------------------------------------------------- bracket network */
/* No counterpart in the reference.  An MLP 18 -> 128 -> 128 -> 128 -> 3 with SiLU, run in
 * BF16 on the matrix cores with FP32 accumulation.  Inputs (hnt_bracket_features): mean
 * delta of shells 1..12, (e, p) of the delta quadrupole sum delta d_L d_K (d the cell
 * offset) over shells 2..6 and 2..12, mean delta over shells 1..12, Rf / alatt: delta only
 * (4 B per cell).  Outputs, the RTHL bracket and the
 * collapse logit: log(1 + lo), lo the shell holding RTHL (RTHL in (rad[lo - 1], rad[lo]],
 * 1 for no collapse); log(1 + hi), hi the gather extent in shells; the collapse logit.
 * Parameters: one FP32 vector, row-major W1 [18x128], b1, W2 [128x128], b2, W3 [128x128],
 * b3, W4 [128x3], b4. */
#define HNT_MLP_INPUTS 18
#define HNT_MLP_HIDDEN 128
#define HNT_MLP_OUTPUTS 3
#define HNT_MLP_NPARAM (18 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 3 + 3)

typedef struct hnt_mlp {
    float* params;             /* device [HNT_MLP_NPARAM] FP32 master weights */
    const float* norm;         /* device [2 * 18]: input mean, then input scale */
    float* adam_m;             /* device [HNT_MLP_NPARAM] Adam moments (training only) */
    float* adam_v;
    int64_t step;              /* Adam step count */
} hnt_mlp;

/* Features of every peak in batch (batch->Rf used, ZZon / fcrit ignored): out device
 * [npeaks * 18] FP32. */
hnt_status hnt_bracket_features(hnt_ctx* ctx, const hnt_fields* fields, const hnt_batch* batch,
                                const hnt_shell_tables* tables, double alatt, float* out);
/* out device [B * 3] = network(x), x device [B * 18]. */
hnt_status hnt_mlp_forward(hnt_ctx* ctx, const hnt_mlp* net, const float* x, int64_t B,
                           float* out);
/* Gradient of the mean loss over the batch, loss = (y0 - t0)^2 + (y1 - t1)^2 +
 * BCE-with-logits(y2, t2): grad device [HNT_MLP_NPARAM]; *loss (host, may be NULL). */
hnt_status hnt_mlp_gradient(hnt_ctx* ctx, const hnt_mlp* net, const float* x,
                            const float* target, int64_t B, float* grad, double* loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss; advances net->step. */
hnt_status hnt_mlp_train_step(hnt_ctx* ctx, hnt_mlp* net, const float* x, const float* target,
                              int64_t B, float lr, double* loss);

#ifdef __cplusplus
}
#endif

#endif /* HUNTIAN_ABI_H */

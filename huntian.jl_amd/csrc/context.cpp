// context.cpp -- host side of libhuntian_hip.so: contexts, error messages, buffer cache,
// shell-table and collapse-table uploads, and the host-only ABI functions.
#include <cmath>
#include <cstdio>
#include <mutex>

#include "hnt_internal.h"

namespace hnt {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = 0; }

hnt_status reserve(hnt_ctx* ctx, DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes && b.p) return HNT_OK;
    if (b.p) {
        HNT_HIP(hipStreamSynchronize(ctx->stream));
        HNT_HIP(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    if (bytes == 0) return HNT_OK;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        return HNT_ERR_ALLOC;
    }
    b.bytes = bytes;
    return HNT_OK;
}

void release(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

hnt_status reserve_in(hnt_ctx* ctx, DevBuf& b, size_t bytes, Arena* ar) {
    if (!ar) return reserve(ctx, b, bytes);
    const size_t off = (ar->used + 255) & ~(size_t)255;
    if (off > ar->cap || bytes > ar->cap - off) {
        set_error("fixed scratch of %zu bytes exhausted: %zu more bytes needed at offset %zu",
                  ar->cap, bytes, off);
        return HNT_ERR_CAPACITY;
    }
    b.p = ar->base + off;
    b.bytes = bytes;
    ar->used = off + bytes;
    return HNT_OK;
}

void release_work(SAWork& w) {
    if (w.arena) {  // carved from a fixed region: nothing to free
        w = SAWork();
        return;
    }
    DevBuf* bufs[] = {&w.recs, &w.keys_a, &w.keys_b, &w.vals_a, &w.vals_b, &w.cub_tmp, &w.slots,
                      &w.gtiles, &w.tiles, &w.tilemeta, &w.stops, &w.feats, &w.prof, &w.queue,
                      &w.flag, &w.skip, &w.paint, &w.tables.cells, &w.tables.start, &w.tables.count,
                      &w.tables.r2, &w.tables.rad, &w.tables.kw, &w.tables.kw_idx,
                      &w.ctab.table, &w.akk};
    for (DevBuf* b : bufs) release(*b);
    w.tables = TablesDev();
    w.ctab = CTabDev();
}

hnt_status begin_call(hnt_ctx* ctx) {
    clear_error();
    HNT_REQUIRE(ctx != nullptr, "NULL context");
    HNT_HIP(hipSetDevice(ctx->device));
    return HNT_OK;
}

// A stage may run several times in one call (the per-filter steps of the tile peak finder):
// each run records its own pair of events, and finish_call adds all of them up.
void stage_begin(hnt_ctx* ctx, int stage) {
    if (!ctx->profiling) return;
    std::vector<hipEvent_t>& ev = ctx->ev[stage];
    const size_t q = 2 * (size_t)ctx->ev_n[stage];
    while (ev.size() < q + 2) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return;
        ev.push_back(e);
    }
    ctx->ev_open[stage] = hipEventRecord(ev[q], ctx->stream) == hipSuccess;
}

void stage_end(hnt_ctx* ctx, int stage) {
    if (!ctx->profiling || !ctx->ev_open[stage]) return;
    ctx->ev_open[stage] = false;
    const size_t q = 2 * (size_t)ctx->ev_n[stage];
    if (hipEventRecord(ctx->ev[stage][q + 1], ctx->stream) == hipSuccess) ctx->ev_n[stage] += 1;
}

hnt_status finish_call(hnt_ctx* ctx, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: kernel launch failed: %s", what, hipGetErrorString(e));
        return HNT_ERR_DEVICE;
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return HNT_ERR_DEVICE;
    }
    if (ctx->profiling) {
        for (int s = 0; s < HNT_NSTAGES; ++s) {
            if (ctx->ev_n[s] == 0) continue;
            for (int q = 0; q < ctx->ev_n[s]; ++q) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, ctx->ev[s][2 * q], ctx->ev[s][2 * q + 1]) ==
                    hipSuccess)
                    ctx->stage_ms[s] += ms;
            }
            ctx->stage_calls[s] += 1;
            ctx->ev_n[s] = 0;
            ctx->ev_open[s] = false;
        }
    }
    return HNT_OK;
}

// Shell tables: validated on the host (the library does not trust the caller's grouping),
// then uploaded once and reused while the caller passes identical tables.
hnt_status upload_tables(hnt_ctx* ctx, const hnt_shell_tables* t) {
    return upload_tables_to(ctx, ctx->tables, t, nullptr);
}

hnt_status upload_tables_to(hnt_ctx* ctx, TablesDev& T, const hnt_shell_tables* t, Arena* ar) {
    HNT_REQUIRE(t && t->di && t->dj && t->dk && t->shell_start && t->shell_count && t->shell_r2,
                "shell tables: NULL array");
    HNT_REQUIRE(t->nshells >= 3 && t->ncells > 0 && t->ncells < (1LL << 31),
                "shell tables: need at least 3 shells (rmax >= 2), got %d", t->nshells);
    const int32_t ns = t->nshells;
    const int64_t nc = t->ncells;
    if (!ar && T.rmax == t->rmax && T.nshells == ns && T.ncells == nc &&
        T.host.size() == (size_t)(3 * nc + 3 * ns)) {
        const int32_t* H = T.host.data();
        const size_t b = sizeof(int32_t);
        if (!memcmp(H, t->di, nc * b) && !memcmp(H + nc, t->dj, nc * b) &&
            !memcmp(H + 2 * nc, t->dk, nc * b) && !memcmp(H + 3 * nc, t->shell_start, ns * b) &&
            !memcmp(H + 3 * nc + ns, t->shell_count, ns * b) &&
            !memcmp(H + 3 * nc + 2 * ns, t->shell_r2, ns * b))
            return HNT_OK;
    }
    std::vector<int32_t> h;
    h.reserve((size_t)(3 * nc + 3 * ns));
    h.insert(h.end(), t->di, t->di + nc);
    h.insert(h.end(), t->dj, t->dj + nc);
    h.insert(h.end(), t->dk, t->dk + nc);
    h.insert(h.end(), t->shell_start, t->shell_start + ns);
    h.insert(h.end(), t->shell_count, t->shell_count + ns);
    h.insert(h.end(), t->shell_r2, t->shell_r2 + ns);
    // consistency: contiguous shells covering all cells, increasing r^2, r^2 matching offsets
    int64_t expect = 0;
    int32_t prev = -1;
    for (int32_t s = 0; s < ns; ++s) {
        HNT_REQUIRE(t->shell_start[s] == expect && t->shell_count[s] > 0,
                    "shell tables: shell %d is not contiguous", s);
        HNT_REQUIRE(t->shell_r2[s] > prev, "shell tables: r^2 not increasing at shell %d", s);
        for (int32_t c = 0; c < t->shell_count[s]; ++c) {
            int64_t q = expect + c;
            int32_t di = t->di[q], dj = t->dj[q], dk = t->dk[q];
            HNT_REQUIRE(di * di + dj * dj + dk * dk == t->shell_r2[s],
                        "shell tables: cell %lld has the wrong r^2", (long long)q);
            HNT_REQUIRE(std::abs(di) <= t->rmax && std::abs(dj) <= t->rmax && std::abs(dk) <= t->rmax,
                        "shell tables: cell %lld beyond rmax", (long long)q);
        }
        expect += t->shell_count[s];
        prev = t->shell_r2[s];
    }
    HNT_REQUIRE(expect == nc, "shell tables: counts sum to %lld, ncells is %lld",
                (long long)expect, (long long)nc);
    HNT_REQUIRE(t->shell_r2[0] == 0, "shell tables: shell 1 must be the centre (r^2 = 0)");
    std::vector<int32_t> cells4((size_t)nc * 4);
    for (int64_t q = 0; q < nc; ++q) {
        cells4[q * 4 + 0] = t->di[q];
        cells4[q * 4 + 1] = t->dj[q];
        cells4[q * 4 + 2] = t->dk[q];
        cells4[q * 4 + 3] = 0;
    }
    std::vector<double> rad((size_t)ns + 2, 0.0);
    for (int32_t s = 1; s <= ns; ++s) rad[s] = std::sqrt((double)t->shell_r2[s - 1]);
    // kernel_strain weights (RadialShell.jl:236-244): the same IEEE operations the device
    // would do per term, done once per table
    std::vector<int32_t> kidx((size_t)3 * (ns + 2), 0);
    std::vector<double> kw;
    for (int32_t mp = 2; mp <= ns; ++mp) {
        const double rmp = rad[mp];
        int32_t lo = mp, hi = mp;
        while (lo > 1 && (rmp - rad[lo - 1]) * (rmp - rad[lo - 1]) < 4.0) --lo;
        while (hi < ns && (rmp - rad[hi + 1]) * (rmp - rad[hi + 1]) < 4.0) ++hi;
        kidx[3 * mp] = lo;
        kidx[3 * mp + 1] = hi;
        kidx[3 * mp + 2] = (int32_t)kw.size();
        for (int32_t m1 = lo; m1 <= hi; ++m1) {
            const double rm1 = rad[m1];
            const double u0 = rmp - rm1;
            const double u02 = u0 * u0;
            const double u12 = rmp * rmp + rm1 * rm1;
            const double dd = u12 - u02;
            kw.push_back(hRinteg(u0, u12) / (dd * dd));
        }
    }
    if (kw.empty()) kw.push_back(0.0);
    HNT_TRY(reserve_in(ctx, T.cells, cells4.size() * sizeof(int32_t), ar));
    HNT_TRY(reserve_in(ctx, T.start, (size_t)ns * sizeof(int32_t), ar));
    HNT_TRY(reserve_in(ctx, T.count, (size_t)ns * sizeof(int32_t), ar));
    HNT_TRY(reserve_in(ctx, T.r2, (size_t)ns * sizeof(int32_t), ar));
    HNT_TRY(reserve_in(ctx, T.rad, rad.size() * sizeof(double), ar));
    HNT_TRY(reserve_in(ctx, T.kw, kw.size() * sizeof(double), ar));
    HNT_TRY(reserve_in(ctx, T.kw_idx, kidx.size() * sizeof(int32_t), ar));
    HNT_HIP(hipMemcpyAsync(T.cells.p, cells4.data(), cells4.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, ctx->stream));
    HNT_HIP(hipMemcpyAsync(T.start.p, t->shell_start, ns * sizeof(int32_t), hipMemcpyHostToDevice,
                           ctx->stream));
    HNT_HIP(hipMemcpyAsync(T.count.p, t->shell_count, ns * sizeof(int32_t), hipMemcpyHostToDevice,
                           ctx->stream));
    HNT_HIP(hipMemcpyAsync(T.r2.p, t->shell_r2, ns * sizeof(int32_t), hipMemcpyHostToDevice,
                           ctx->stream));
    HNT_HIP(hipMemcpyAsync(T.rad.p, rad.data(), rad.size() * sizeof(double),
                           hipMemcpyHostToDevice, ctx->stream));
    HNT_HIP(hipMemcpyAsync(T.kw.p, kw.data(), kw.size() * sizeof(double), hipMemcpyHostToDevice,
                           ctx->stream));
    HNT_HIP(hipMemcpyAsync(T.kw_idx.p, kidx.data(), kidx.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, ctx->stream));
    HNT_HIP(hipStreamSynchronize(ctx->stream));
    T.rmax = t->rmax;
    T.nshells = ns;
    T.ncells = nc;
    T.host.swap(h);
    return HNT_OK;
}

hnt_status upload_ctab(hnt_ctx* ctx, const hnt_ctab* c) {
    return upload_ctab_to(ctx, ctx->ctab, c, nullptr);
}

hnt_status upload_ctab_to(hnt_ctx* ctx, CTabDev& D, const hnt_ctab* c, Arena* ar) {
    HNT_REQUIRE(c && c->table, "collapse table: NULL");
    HNT_REQUIRE(c->nx >= 1 && c->ny >= 1 && c->nz >= 1, "collapse table: bad shape");
    size_t n = (size_t)c->nx * c->ny * c->nz;
    if (!ar && D.nx == c->nx && D.ny == c->ny && D.nz == c->nz && D.host.size() == n &&
        std::memcmp(D.host.data(), c->table, n * sizeof(float)) == 0)
        return HNT_OK;
    HNT_TRY(reserve_in(ctx, D.table, n * sizeof(float), ar));
    HNT_HIP(hipMemcpyAsync(D.table.p, c->table, n * sizeof(float), hipMemcpyHostToDevice,
                           ctx->stream));
    HNT_HIP(hipStreamSynchronize(ctx->stream));
    D.nx = c->nx;
    D.ny = c->ny;
    D.nz = c->nz;
    D.host.assign(c->table, c->table + n);
    return HNT_OK;
}

// SPH kernel table of RadialShell._build_atab4 (RadialShell.jl:125-142), FP64.
hnt_status ensure_akk(hnt_ctx* ctx) { return ensure_akk_to(ctx, ctx->akk, nullptr); }

hnt_status ensure_akk_to(hnt_ctx* ctx, DevBuf& akk_buf, Arena* ar) {
    if (!ar && akk_buf.p) return HNT_OK;
    double akk[1101];
    const double ca = 3.0 / (2.0 * kPi);
    const double cb = 1.0 / (4.0 * kPi);
    const double two_thirds = 2.0 / 3.0;
    for (int i = 1; i <= 1100; ++i) {
        double u = (double)(i - 1) * 0.01;
        double ru = std::sqrt(u);
        double v;
        if (ru < 1.0) {
            v = ca * (two_thirds - u + 0.5 * u * ru);
        } else if (ru <= 2.0) {
            double t = 2.0 - ru;
            v = cb * (t * t * t);
        } else {
            v = 0.0;
        }
        akk[i - 1] = v;
    }
    akk[1100] = 0.0;
    HNT_TRY(reserve_in(ctx, akk_buf, sizeof(akk), ar));
    HNT_HIP(hipMemcpyAsync(akk_buf.p, akk, sizeof(akk), hipMemcpyHostToDevice, ctx->stream));
    HNT_HIP(hipStreamSynchronize(ctx->stream));
    return HNT_OK;
}

}  // namespace hnt

using namespace hnt;

extern "C" {

int hnt_abi_version(void) { return HNT_ABI_VERSION; }

size_t hnt_last_error(char* buf, size_t len) {
    size_t n = strlen(g_err);
    if (buf && len) {
        size_t c = n < len - 1 ? n : len - 1;
        memcpy(buf, g_err, c);
        buf[c] = 0;
    }
    return n;
}

hnt_status hnt_ctx_create(int device, hnt_ctx** out) {
    clear_error();
    HNT_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    HNT_HIP(hipGetDeviceCount(&ndev));
    HNT_REQUIRE(device >= 0 && device < ndev, "device %d out of range (%d devices)", device, ndev);
    HNT_HIP(hipSetDevice(device));
    hnt_ctx* c = new hnt_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_error("hipStreamCreate: %s", hipGetErrorString(e));
        return HNT_ERR_DEVICE;
    }
    c->stream = c->own_stream;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    if (c->num_cus <= 0) c->num_cus = 256;
    // the fixed scratch of hnt_shell_analyse_noalloc: its address on this device, once per
    // device; a failure here only disables that entry point (it reports the error)
    (void)noalloc_scratch_init(device);
    clear_error();
    *out = c;
    return HNT_OK;
}

// The per-call scratch: everything sized by a call's inputs (not the uploaded tables, the
// interpolation table or the network, which plans may reference).
static void release_scratch(hnt_ctx* ctx) {
    release_work(ctx->sa);
    for (DevBuf* b : {&ctx->claim, &ctx->counts, &ctx->offsets, &ctx->cub_tmp, &ctx->fb_spec,
                      &ctx->fb_work, &ctx->fb_field, &ctx->fb_lap, &ctx->fb_val, &ctx->fb_d2,
                      &ctx->fb_wtab, &ctx->fb_taken,
                      &ctx->rec_tabs, &ctx->rec_keep, &ctx->rec_slot, &ctx->fs_real,
                      &ctx->fs_base, &ctx->fs_work, &ctx->fs_aux, &ctx->fs_tab})
        release(*b);
    MergeBufs& m = ctx->mg;
    for (DevBuf* b : {&m.a, &m.b, &m.c, &m.d, &m.rank, &m.cidx, &m.cstart, &m.cnt, &m.off, &m.e1,
                      &m.e2, &m.e3, &m.e4, &m.state, &m.instart, &m.dv, &m.pk, &m.pv, &m.newr,
                      &m.pos, &m.prank, &m.ext})
        release(*b);
    release_fft(ctx);
    release_field_plans(ctx);
    for (hnt_ctx::FftTw& t : ctx->fft_tw) release(t.buf);
    ctx->fft_tw.clear();
}

hnt_status hnt_ctx_release_scratch(hnt_ctx* ctx) {
    HNT_TRY(begin_call(ctx));
    HNT_HIP(hipStreamSynchronize(ctx->stream));
    release_scratch(ctx);
    return HNT_OK;
}

hnt_status hnt_ctx_destroy(hnt_ctx* ctx) {
    clear_error();
    if (!ctx) return HNT_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_scratch(ctx);
    DevBuf* bufs[] = {&ctx->fieldptrs, &ctx->fcrit, &ctx->akk, &ctx->small,
                      &ctx->tables.cells, &ctx->tables.start, &ctx->tables.count,
                      &ctx->tables.r2, &ctx->tables.rad, &ctx->tables.kw,
                      &ctx->tables.kw_idx, &ctx->ctab.table, &ctx->mlp.wt, &ctx->mlp.x0,
                      &ctx->mlp.z, &ctx->mlp.h, &ctx->mlp.y, &ctx->mlp.dy, &ctx->mlp.dyf,
                      &ctx->mlp.dz, &ctx->mlp.gw, &ctx->mlp.part, &ctx->mlp.gradv};
    for (DevBuf* b : bufs) release(*b);
    for (std::vector<hipEvent_t>& v : ctx->ev)
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return HNT_OK;
}

hnt_status hnt_ctx_set_tuning(hnt_ctx* ctx, int32_t what, int32_t value) {
    clear_error();
    HNT_REQUIRE(ctx, "hnt_ctx_set_tuning: NULL context");
    if (what == HNT_TUNE_FFT) {
        HNT_REQUIRE(value == 0 || value == 1, "hnt_ctx_set_tuning: fft %d (0 or 1)", value);
        ctx->fft_mode = value;
        return HNT_OK;
    }
    if (what == HNT_TUNE_FFT_DIRECT) {
        HNT_REQUIRE(value == 0 || value == 1, "hnt_ctx_set_tuning: fft_direct %d (0 or 1)",
                    value);
        ctx->fft_direct = value;
        return HNT_OK;
    }
    if (what == HNT_TUNE_TRICUBIC_TAB) {
        HNT_REQUIRE(value == 0 || value == 1, "hnt_ctx_set_tuning: tricubic_tab %d (0 or 1)",
                    value);
        ctx->tricubic_tab = value;
        return HNT_OK;
    }
    if (what == HNT_TUNE_BRICK_LOG2) {
        HNT_REQUIRE(value >= 0 && value <= 13, "hnt_ctx_set_tuning: brick_log2 %d", value);
        ctx->brick_log2 = value;
    } else if (what == HNT_TUNE_GANG) {
        HNT_REQUIRE(value == 1 || value == 16 || value == 64,
                    "hnt_ctx_set_tuning: gang must be 1, 16 or 64");
        ctx->gang = value;
    } else if (what == HNT_TUNE_WAVES_PER_CU) {
        HNT_REQUIRE(value >= 0 && value <= 64, "hnt_ctx_set_tuning: waves_per_cu %d", value);
        ctx->waves_per_cu = value;
    } else {
        HNT_REQUIRE(false, "hnt_ctx_set_tuning: unknown knob %d", what);
    }
    return HNT_OK;
}

hnt_status hnt_ctx_set_profiling(hnt_ctx* ctx, int on) {
    HNT_TRY(begin_call(ctx));
    ctx->profiling = on != 0;
    for (int s = 0; s < HNT_NSTAGES; ++s) {
        ctx->stage_ms[s] = 0.0;
        ctx->stage_calls[s] = 0;
        ctx->ev_n[s] = 0;
        ctx->ev_open[s] = false;
    }
    return HNT_OK;
}

hnt_status hnt_ctx_get_profile(hnt_ctx* ctx, double* ms, int64_t* calls, int32_t nstages) {
    clear_error();
    HNT_REQUIRE(ctx && ms && calls, "hnt_ctx_get_profile: NULL argument");
    for (int s = 0; s < nstages; ++s) {
        ms[s] = s < HNT_NSTAGES ? ctx->stage_ms[s] : 0.0;
        calls[s] = s < HNT_NSTAGES ? ctx->stage_calls[s] : 0;
    }
    return HNT_OK;
}

hnt_status hnt_ctx_set_stream(hnt_ctx* ctx, void* hip_stream) {
    clear_error();
    HNT_REQUIRE(ctx != nullptr, "NULL context");
    ctx->stream = (hipStream_t)hip_stream;
    return HNT_OK;
}

hnt_status hnt_shell_table_size(int32_t rmax, int64_t* ncells, int32_t* nshells) {
    clear_error();
    HNT_REQUIRE(rmax >= 0 && rmax <= 1000, "rmax %d out of range", rmax);
    HNT_REQUIRE(ncells && nshells, "NULL output");
    int32_t r2max = rmax * rmax;
    std::vector<char> seen((size_t)r2max + 1, 0);
    int64_t c = 0;
    for (int32_t dk = -rmax; dk <= rmax; ++dk)
        for (int32_t dj = -rmax; dj <= rmax; ++dj)
            for (int32_t di = -rmax; di <= rmax; ++di) {
                int32_t r2 = di * di + dj * dj + dk * dk;
                if (r2 <= r2max) {
                    ++c;
                    seen[r2] = 1;
                }
            }
    int32_t s = 0;
    for (char v : seen) s += v;
    *ncells = c;
    *nshells = s;
    return HNT_OK;
}

// precompute_shells + build_shell_tables (RadialShell.jl:151-166, ShellAnalysisGPU.jl:44-71):
// generation order dk, dj, di then a stable grouping by r^2 (a counting sort is stable).
hnt_status hnt_build_shell_tables(int32_t rmax, int32_t* di, int32_t* dj, int32_t* dk,
                                  int32_t* shell_start, int32_t* shell_count, int32_t* shell_r2) {
    clear_error();
    HNT_REQUIRE(rmax >= 0 && rmax <= 1000, "rmax %d out of range", rmax);
    HNT_REQUIRE(di && dj && dk && shell_start && shell_count && shell_r2, "NULL output");
    int32_t r2max = rmax * rmax;
    std::vector<int64_t> pos((size_t)r2max + 2, 0);
    for (int32_t k = -rmax; k <= rmax; ++k)
        for (int32_t j = -rmax; j <= rmax; ++j)
            for (int32_t i = -rmax; i <= rmax; ++i) {
                int32_t r2 = i * i + j * j + k * k;
                if (r2 <= r2max) pos[r2 + 1]++;
            }
    int32_t ns = 0;
    for (int32_t v = 0; v <= r2max; ++v) {
        if (pos[v + 1] > 0) {
            shell_start[ns] = (int32_t)pos[v];
            shell_count[ns] = (int32_t)pos[v + 1];
            shell_r2[ns] = v;
            ++ns;
        }
        pos[v + 1] += pos[v];
    }
    for (int32_t k = -rmax; k <= rmax; ++k)
        for (int32_t j = -rmax; j <= rmax; ++j)
            for (int32_t i = -rmax; i <= rmax; ++i) {
                int32_t r2 = i * i + j * j + k * k;
                if (r2 <= r2max) {
                    int64_t q = pos[r2]++;
                    di[q] = i;
                    dj[q] = j;
                    dk[q] = k;
                }
            }
    return HNT_OK;
}

double hnt_fsc_of_z(const hnt_ctab* ct, double z) {
    clear_error();
    if (!ct || !ct->table) {
        set_error("hnt_fsc_of_z: NULL table");
        return NAN;
    }
    CTabView v{ct->table, ct->nx, ct->ny, ct->nz, ct->x1, ct->x2, ct->y1, ct->y2,
               ct->z1, ct->z2, ct->out_val};
    // 50-step bisection of the table at e = p/e = 0 (RadialShell.jl:181-195)
    double target = 1.0 + z;
    double lo = ct->x1, hi = ct->x2;
    for (int it = 0; it < 50; ++it) {
        double mid = 0.5 * (lo + hi);
        double zvir = interp3(v, mid, 0.0, 0.0);
        if (zvir > target) hi = mid;
        else lo = mid;
    }
    return std::pow(10.0, lo);
}

}  // extern "C"

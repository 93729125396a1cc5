// merge.hip -- Lagrangian exclusion and volume reduction of a halo catalog (hnt_merge_catalog).
//
// Replaces Merger.merge_catalog (src/Merger/Merger.jl:23-89) with Exclusion.build_hash,
// auto_hash_nc, lagrangian_exclusion!, sphere_overlap and volume_reduction!
// (src/Merger/Exclusion.jl:13-208).
//
// The reference walks the halos in (-r, x, y, z) order and lets every survivor kill the
// survivors of lower rank whose centres lie inside it.  A halo therefore survives exactly
// when no surviving halo of higher rank contains its centre: a greedy independent set with
// fixed priorities, whose result does not depend on the order the decisions are made in.
// Here each halo lists the lower-ranked halos it contains (edges), and rounds over all
// undecided halos apply "dead if a living container exists, alive once every container is
// dead" until every halo is decided -- the same survivors as the sequential pass.
//
// Volume reduction is a sum per halo whose order matters in the last bits.  The reference
// adds, for halo j, first the caps cut by larger overlapping halos (in their rank order, as
// its outer loop visits them), then its own caps against smaller halos in its cell-scan order
// (cells by dz, dy, dx over +-ceil(2 r / cell); within a cell the hash's linked list, i.e.
// descending catalog index).  Both orders are rebuilt with one sort: every overlapping pair
// writes the cap cut from the smaller halo keyed by (halo, rank of the larger) and the
// larger halo's own cap keyed by (halo, N + position of the smaller in a cell-sorted copy of
// the catalog), whose order is the reference's cell-scan order (same grid as the reference's
// hash: nc = clamp(round(cbrt(N)), 16, 1024) cells over the bounding box padded by 2 r_max);
// one thread per halo then adds its run.
//
// Layout: after the rank and cell sorts every pass works on positions p of the cell-sorted
// order, where halo idx[p] keeps (x, y, z, r) in one float4 and its rank beside it, so the
// neighbour scans read contiguous 16-B entries instead of 44- or 132-B records; states, edges
// and volume sums are indexed by position, and only the survivors' flags, new radii and the
// output go back to catalog order.
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "hnt_internal.h"

namespace {

using namespace hnt;

constexpr int MB = 256;  // threads per workgroup of the merge kernels

// The hash grid and the cell-sorted halos.
struct Grid {
    double ox, oy, oz, cell;     // origin and cell size
    int nc;
    int64_t N;
    const float4* pos;           // [N] (x, y, z, r) of halo idx[p]
    const int32_t* prank;        // [N] rank in (-r, x, y, z) order of halo idx[p]
    const int32_t* cell_start;   // [nc^3 + 1] first position of each cell
};

// _cell_idx (Exclusion.jl:70-76), 1-based
__device__ __forceinline__ int cell_of(double p, double o, double cell, int nc) {
    const double f = floor((p - o) / cell);
    int c = f < -1e9 ? -1000000000 : (f > 1e9 ? 1000000000 : (int)f);
    c += 1;
    return c < 1 ? 1 : (c > nc ? nc : c);
}

// Orderable key of a float (ascending with the float's order, -0 before +0).
__device__ __forceinline__ uint32_t fkey(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// first index in sorted a[0, n) whose value is >= v
template <class T>
__device__ __forceinline__ int64_t lower_bound(const T* __restrict__ a, int64_t n, T v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void merge_keys(const float* __restrict__ rec, int64_t N, int nf, int field,
                           int negate, const int32_t* __restrict__ perm,
                           uint32_t* __restrict__ keys) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    const float v = rec[(int64_t)perm[s] * nf + field];
    keys[s] = fkey(negate ? -v : v);
}

__global__ void merge_iota(int32_t* __restrict__ v, int64_t N, int descending) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < N) v[s] = (int32_t)(descending ? N - 1 - s : s);
}

__global__ void merge_rank(const int32_t* __restrict__ order, int64_t N, int32_t* __restrict__ rank) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < N) rank[order[s]] = (int32_t)s;
}

__global__ void merge_cell_keys(const float* __restrict__ rec, int64_t N, int nf, Grid g,
                                const int32_t* __restrict__ idx, uint32_t* __restrict__ keys) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    const float* h = rec + (int64_t)idx[s] * nf;
    const int ix = cell_of((double)h[0], g.ox, g.cell, g.nc);
    const int iy = cell_of((double)h[1], g.oy, g.cell, g.nc);
    const int iz = cell_of((double)h[2], g.oz, g.cell, g.nc);
    keys[s] = (uint32_t)((ix - 1) + (int64_t)g.nc * ((iy - 1) + (int64_t)g.nc * (iz - 1)));
}

// cell_start[c] = first position of cell c in the cell-sorted keys, c = 0 .. ncell
__global__ void merge_cell_start(const uint32_t* __restrict__ keys, int64_t N, int64_t ncell,
                                 int32_t* __restrict__ start) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    start[c] = c == ncell ? (int32_t)N : (int32_t)lower_bound(keys, N, (uint32_t)c);
}

// The compact cell-sorted copy: pos[p] = (x, y, z, r), prank[p] = rank of halo idx[p].
__global__ void merge_gather(const float* __restrict__ rec, int64_t N, int nf,
                             const int32_t* __restrict__ idx, const int32_t* __restrict__ rank,
                             float4* __restrict__ pos, int32_t* __restrict__ prank) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int64_t m = idx[p];
    const float* h = rec + m * nf;
    pos[p] = make_float4(h[0], h[1], h[2], h[6]);
    prank[p] = rank[m];
}

// Visit the positions of the halos in the cells within +-dcell of (xi, yi, zi)'s cell, in the
// reference's order: cells by dz, dy, dx; within a cell by descending catalog index.
template <class F>
__device__ __forceinline__ void for_neighbours(const Grid& g, double xi, double yi, double zi,
                                               int dcell, F&& f) {
    const int ci = cell_of(xi, g.ox, g.cell, g.nc), cj = cell_of(yi, g.oy, g.cell, g.nc),
              ck = cell_of(zi, g.oz, g.cell, g.nc);
    for (int dz = -dcell; dz <= dcell; ++dz) {
        const int jz = ck + dz;
        if (jz < 1 || jz > g.nc) continue;
        for (int dy = -dcell; dy <= dcell; ++dy) {
            const int jy = cj + dy;
            if (jy < 1 || jy > g.nc) continue;
            const int64_t row = (int64_t)g.nc * ((jy - 1) + (int64_t)g.nc * (jz - 1));
            for (int dx = -dcell; dx <= dcell; ++dx) {
                const int jx = ci + dx;
                if (jx < 1 || jx > g.nc) continue;
                const int64_t c = (jx - 1) + row;
                const int32_t e = g.cell_start[c + 1];
                for (int32_t q = g.cell_start[c]; q < e; ++q) f((int64_t)q);
            }
        }
    }
}

// Containment edges p -> q: rank q > rank p and |x_q - x_p| < r_p (lagrangian_exclusion!,
// Exclusion.jl:110-131).  write == 0 counts, write == 1 fills from off[p].
__global__ void merge_edges(Grid g, int write, int64_t* __restrict__ cnt,
                            const int64_t* __restrict__ off, int32_t* __restrict__ src,
                            int32_t* __restrict__ dst) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.N) return;
    const float4 me = g.pos[p];
    const double xi = me.x, yi = me.y, zi = me.z, ri = me.w;
    const int dcell = (int)ceil(ri / g.cell);
    const int32_t rkp = g.prank[p];
    int64_t n = 0, w = write ? off[p] : 0;
    for_neighbours(g, xi, yi, zi, dcell, [&](int64_t q) {
        if (g.prank[q] <= rkp) return;  // also q == p
        const float4 o = g.pos[q];
        const double dx = (double)o.x - xi, dy = (double)o.y - yi, dz = (double)o.z - zi;
        const double dist = sqrt(dx * dx + dy * dy + dz * dz);
        if (dist < ri) {
            if (write) {
                src[w] = (int32_t)p;
                dst[w] = (int32_t)q;
                ++w;
            } else {
                ++n;
            }
        }
    });
    if (!write) cnt[p] = n;
}

// One round of decisions: state 0 undecided, 1 alive, 2 dead.  in_start / in_src: containers
// of each halo (CSR by position).
__global__ void merge_round(int64_t N, const int64_t* __restrict__ in_start,
                            const int32_t* __restrict__ in_src, uint8_t* __restrict__ state,
                            unsigned long long* __restrict__ undecided) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N || state[j] != 0) return;
    bool all_dead = true;
    for (int64_t e = in_start[j]; e < in_start[j + 1]; ++e) {
        const uint8_t s = state[in_src[e]];
        if (s == 1) {
            state[j] = 2;
            return;
        }
        if (s != 2) all_dead = false;
    }
    if (all_dead) state[j] = 1;
    else atomicAdd(undecided, 1ull);
}

// start[j] = first entry of halo j in entries sorted by halo (key >> shift), j = 0 .. N
template <class K>
__global__ void merge_csr(const K* __restrict__ sorted, int64_t E, int64_t N, int shift,
                          int64_t* __restrict__ start) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > N) return;
    start[j] = j == N ? E : lower_bound(sorted, E, (K)((K)j << shift));
}

// sphere_overlap (Exclusion.jl:86-101): caps (v1, v2) cut from spheres r1, r2 at distance d
__device__ __forceinline__ void sphere_overlap(double d, double r1, double r2, double& v1,
                                               double& v2) {
    v1 = v2 = 0.0;
    if (d >= r1 + r2) return;
    if (d <= fabs(r1 - r2)) return;
    const double h1 = (r2 - r1 + d) * (r2 + r1 - d) / (2.0 * d);
    const double h2 = (r1 - r2 + d) * (r1 + r2 - d) / (2.0 * d);
    v1 = (M_PI / 3.0) * (h1 * h1) * (3.0 * r1 - h1);
    v2 = (M_PI / 3.0) * (h2 * h2) * (3.0 * r2 - h2);
}

// Overlapping survivor pairs (p larger by rank) of volume_reduction! (Exclusion.jl:163-180).
// Every pair gives two caps: the one cut from the smaller halo q, keyed (q << 32 | rank p),
// and p's own, keyed (p << 32 | N + q).  Sorted by key, a halo's caps come in the reference's
// order: those cut by larger halos in their rank order, then its own in its cell-scan order
// (cells by dz, dy, dx, descending catalog index within a cell = ascending position q).
// mode 0 counts two entries per pair; mode 1 writes them from off[p].
__global__ void merge_pairs(Grid g, int mode, const uint8_t* __restrict__ state,
                            int fortran_ties, int64_t* __restrict__ cnt,
                            const int64_t* __restrict__ off, uint64_t* __restrict__ keys,
                            double* __restrict__ vals) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.N) return;
    if (state[p] != 1) {
        if (mode == 0) cnt[p] = 0;  // the scan reads every count: none may be left over
        return;
    }
    const float4 me = g.pos[p];
    const double xi = me.x, yi = me.y, zi = me.z, ri = me.w;
    const int dcell = (int)ceil(2.0 * ri / g.cell);
    const int32_t rkp = g.prank[p];
    int64_t n = 0, w = mode == 1 ? off[p] : 0;
    for_neighbours(g, xi, yi, zi, dcell, [&](int64_t q) {
        if (g.prank[q] <= rkp || state[q] != 1) return;
        const float4 o = g.pos[q];
        const double rj = o.w;
        const double ex = (double)o.x - xi, ey = (double)o.y - yi, ez = (double)o.z - zi;
        const double d = sqrt(ex * ex + ey * ey + ez * ez);
        if (!(d < ri + rj)) return;
        if (mode == 0) {
            n += 2;
            return;
        }
        double v1, v2;
        sphere_overlap(d, ri, rj, v1, v2);
        const double wgt = (fortran_ties && rj == ri) ? 2.0 : 1.0;
        keys[w] = ((uint64_t)q << 32) | (uint32_t)rkp;
        vals[w] = wgt * v2;
        keys[w + 1] = ((uint64_t)p << 32) | (uint32_t)(g.N + q);
        vals[w + 1] = wgt * v1;
        w += 2;
    });
    if (mode == 0) cnt[p] = n;
}

// dV[j] = halo j's caps summed in the reference's order (entries sorted by key; one thread
// per halo walks its run).
__global__ void merge_incoming(const double* __restrict__ vals, const int64_t* __restrict__ start,
                               int64_t N, double* __restrict__ dV) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    double s = 0.0;
    for (int64_t e = start[j]; e < start[j + 1]; ++e) s += vals[e];
    dV[j] = s;
}

// New radii (Exclusion.jl:192-201) and survivor flags, written back in catalog order.
__global__ void merge_radii(Grid g, const int32_t* __restrict__ idx, const uint8_t* __restrict__ state,
                            const double* __restrict__ dV, int reduce, float* __restrict__ newr,
                            int32_t* __restrict__ keep) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.N) return;
    const double r = g.pos[p].w;
    float nr = (float)r;
    bool alive = state[p] == 1;
    if (reduce && alive) {
        const double vol_new = 4.0 * M_PI / 3.0 * (r * r * r) - dV[p];
        if (vol_new <= 0.0) {
            alive = false;
            nr = 0.0f;
        } else {
            nr = (float)cbrt(vol_new * 3.0 / (4.0 * M_PI));
        }
    }
    const int64_t m = idx[p];
    newr[m] = nr;
    keep[m] = alive ? 1 : 0;
}

// Survivors in catalog order, one thread per float of the input (coalesced reads).
__global__ void merge_scatter(const float* __restrict__ rec, int64_t N, int nf,
                              const int32_t* __restrict__ keep, const int64_t* __restrict__ slot,
                              const float* __restrict__ newr, float* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * nf) return;
    const int64_t i = t / nf;
    if (!keep[i]) return;
    const int q = (int)(t - i * nf);
    out[slot[i] * nf + q] = q == 6 ? newr[i] : rec[t];
}

// Extrema of x, y, z and r and a non-finite flag: per-workgroup partials, then one workgroup.
// part[b * 8 + c]: c = 0..2 min x y z, 3..5 max x y z, 6 max r, 7 = 1 if any x, y, z or r is
// not finite.
__device__ __forceinline__ float ext_op(int c, float a, float b) {
    return c < 3 ? fminf(a, b) : fmaxf(a, b);
}

__device__ void ext_block_reduce(float (&v)[8], float* __restrict__ dst) {
    __shared__ float s[8][MB];
    for (int c = 0; c < 8; ++c) s[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int w = MB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 8; ++c)
                s[c][threadIdx.x] = ext_op(c, s[c][threadIdx.x], s[c][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 8) dst[threadIdx.x] = s[threadIdx.x][0];
}

__global__ void merge_extrema_part(const float* __restrict__ rec, int64_t N, int nf,
                                   float* __restrict__ part) {
    float v[8] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0f};
    for (int64_t m = (int64_t)blockIdx.x * MB + threadIdx.x; m < N; m += (int64_t)gridDim.x * MB) {
        const float* h = rec + m * nf;
        for (int c = 0; c < 3; ++c) {
            const float p = h[c];
            v[c] = fminf(v[c], p);
            v[3 + c] = fmaxf(v[3 + c], p);
            if (!isfinite(p)) v[7] = 1.0f;
        }
        const float r = h[6];
        v[6] = fmaxf(v[6], r);
        if (!isfinite(r)) v[7] = 1.0f;
    }
    ext_block_reduce(v, part + 8 * (int64_t)blockIdx.x);
}

__global__ void merge_extrema_final(const float* __restrict__ part, int nblocks,
                                    float* __restrict__ ext) {
    float v[8] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0f};
    for (int b = threadIdx.x; b < nblocks; b += MB)
        for (int c = 0; c < 8; ++c) v[c] = ext_op(c, v[c], part[8 * b + c]);
    ext_block_reduce(v, ext);
}

unsigned nblk(int64_t n) { return (unsigned)((n + MB - 1) / MB); }

}  // namespace

extern "C" hnt_status hnt_merge_catalog(hnt_ctx* ctx, const float* records, int64_t nhalo,
                                        int32_t nfields, int32_t volume_reduction,
                                        int32_t fortran_ties, float* out, int64_t* nout) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(nout && (nhalo == 0 || (records && out)), "hnt_merge_catalog: NULL argument");
    HNT_REQUIRE(nfields == 11 || nfields == 33, "hnt_merge_catalog: records of %d floats "
                "(11 or 33)", nfields);
    HNT_REQUIRE(nhalo >= 0 && nhalo < (1LL << 31), "hnt_merge_catalog: %lld halos",
                (long long)nhalo);
    *nout = 0;
    if (nhalo == 0) return HNT_OK;
    hipStream_t st = ctx->stream;
    const int64_t N = nhalo;
    MergeBufs& B = ctx->mg;

    // ---- domain and grid (Merger.jl:63-73, auto_hash_nc Exclusion.jl:66) ----
    const int nparts = (int)std::min<int64_t>(1024, (N + MB - 1) / MB);
    HNT_TRY(reserve(ctx, B.ext, (size_t)(nparts + 1) * 8 * sizeof(float)));
    float* part = (float*)B.ext.p;
    float* extd = part + 8 * nparts;
    hipLaunchKernelGGL(merge_extrema_part, dim3(nparts), dim3(MB), 0, st, records, N, nfields,
                       part);
    hipLaunchKernelGGL(merge_extrema_final, dim3(1), dim3(MB), 0, st, (const float*)part,
                       nparts, extd);
    float ext[8];
    HNT_HIP(hipMemcpyAsync(ext, extd, sizeof(ext), hipMemcpyDeviceToHost, st));
    HNT_HIP(hipStreamSynchronize(st));
    HNT_REQUIRE(ext[7] == 0.0f, "hnt_merge_catalog: non-finite position or radius");
    const double rmax = ext[6];
    const double pad = 2.0 * rmax;
    const double dmin[3] = {(double)ext[0] - pad, (double)ext[1] - pad, (double)ext[2] - pad};
    const double dmax[3] = {(double)ext[3] + pad, (double)ext[4] + pad, (double)ext[5] + pad};
    const double L = std::max(dmax[0] - dmin[0], std::max(dmax[1] - dmin[1], dmax[2] - dmin[2]));
    // a zero extent (one halo, or coincident halos, without a positive radius) has no grid;
    // the reference fails there too (floor of NaN)
    HNT_REQUIRE(L > 0.0, "hnt_merge_catalog: empty domain (extent %g)", L);
    int nc = (int)std::nearbyint(std::cbrt((double)N));
    nc = nc < 16 ? 16 : (nc > 1024 ? 1024 : nc);
    Grid G;
    memset(&G, 0, sizeof(G));
    G.nc = nc;
    G.cell = L / nc;
    G.ox = dmin[0];
    G.oy = dmin[1];
    G.oz = dmin[2];
    G.N = N;
    const int64_t ncell = (int64_t)nc * nc * nc;

    // ---- rank: stable LSD sorts by z, y, x, then -r (Merger.jl:59) ----
    HNT_TRY(reserve(ctx, B.a, (size_t)N * 4));   // keys
    HNT_TRY(reserve(ctx, B.b, (size_t)N * 4));   // keys alt
    HNT_TRY(reserve(ctx, B.c, (size_t)N * 4));   // perm
    HNT_TRY(reserve(ctx, B.d, (size_t)N * 4));   // perm alt
    HNT_TRY(reserve(ctx, B.rank, (size_t)N * 4));
    uint32_t* ka = (uint32_t*)B.a.p;
    uint32_t* kb = (uint32_t*)B.b.p;
    int32_t* pa = (int32_t*)B.c.p;
    int32_t* pb = (int32_t*)B.d.p;
    size_t tmp = 0;
    HNT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, ka, kb, pa, pb, (int)N, 0, 32, st));
    size_t tmp2 = 0;
    HNT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp2, (uint64_t*)nullptr,
                                               (uint64_t*)nullptr, (double*)nullptr,
                                               (double*)nullptr, (int)N, 0, 64, st));
    HNT_TRY(reserve(ctx, ctx->cub_tmp, std::max(tmp, tmp2) + 256));
    hipLaunchKernelGGL(merge_iota, dim3(nblk(N)), dim3(MB), 0, st, pa, N, 0);
    const int fields[4] = {2, 1, 0, 6};
    for (int q = 0; q < 4; ++q) {
        hipLaunchKernelGGL(merge_keys, dim3(nblk(N)), dim3(MB), 0, st, records, N, nfields,
                           fields[q], q == 3 ? 1 : 0, pa, ka);
        size_t t = ctx->cub_tmp.bytes;
        HNT_HIP(hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp.p, t, ka, kb, pa, pb, (int)N, 0,
                                                   32, st));
        std::swap(pa, pb);
    }
    hipLaunchKernelGGL(merge_rank, dim3(nblk(N)), dim3(MB), 0, st, pa, N, (int32_t*)B.rank.p);

    // ---- hash: halos sorted by cell, descending index within a cell (build_hash), and the
    // compact cell-sorted copy ----
    HNT_TRY(reserve(ctx, B.cidx, (size_t)N * 4));
    HNT_TRY(reserve(ctx, B.cstart, (size_t)(ncell + 1) * 4));
    HNT_TRY(reserve(ctx, B.pos, (size_t)N * sizeof(float4)));
    HNT_TRY(reserve(ctx, B.prank, (size_t)N * 4));
    int32_t* idx = (int32_t*)B.cidx.p;
    int32_t* desc = pb;  // free perm buffer
    hipLaunchKernelGGL(merge_iota, dim3(nblk(N)), dim3(MB), 0, st, desc, N, 1);
    hipLaunchKernelGGL(merge_cell_keys, dim3(nblk(N)), dim3(MB), 0, st, records, N, nfields, G,
                       (const int32_t*)desc, ka);
    {
        size_t t = ctx->cub_tmp.bytes;
        HNT_HIP(hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp.p, t, ka, kb, desc, idx, (int)N,
                                                   0, 32, st));
    }
    hipLaunchKernelGGL(merge_cell_start, dim3(nblk(ncell + 1)), dim3(MB), 0, st,
                       (const uint32_t*)kb, N, ncell, (int32_t*)B.cstart.p);
    hipLaunchKernelGGL(merge_gather, dim3(nblk(N)), dim3(MB), 0, st, records, N, nfields,
                       (const int32_t*)idx, (const int32_t*)B.rank.p, (float4*)B.pos.p,
                       (int32_t*)B.prank.p);
    G.pos = (const float4*)B.pos.p;
    G.prank = (const int32_t*)B.prank.p;
    G.cell_start = (const int32_t*)B.cstart.p;

    // ---- exclusion: containment edges, CSR by contained halo, decision rounds ----
    HNT_TRY(reserve(ctx, B.cnt, (size_t)(N + 1) * 8));
    HNT_TRY(reserve(ctx, B.off, (size_t)(N + 1) * 8));
    int64_t* cnt = (int64_t*)B.cnt.p;
    int64_t* off = (int64_t*)B.off.p;
    hipLaunchKernelGGL(merge_edges, dim3(nblk(N)), dim3(MB), 0, st, G, 0, cnt, nullptr, nullptr,
                       nullptr);
    size_t tmp3 = 0;
    HNT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp3, cnt, off, (int)(N + 1), st));
    if (tmp3 > ctx->cub_tmp.bytes) HNT_TRY(reserve(ctx, ctx->cub_tmp, tmp3));
    HNT_HIP(hipMemsetAsync(cnt + N, 0, 8, st));
    {
        size_t t = ctx->cub_tmp.bytes;
        HNT_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp.p, t, cnt, off, (int)(N + 1), st));
    }
    int64_t E = 0;
    HNT_HIP(hipMemcpyAsync(&E, off + N, 8, hipMemcpyDeviceToHost, st));
    HNT_HIP(hipStreamSynchronize(st));
    HNT_REQUIRE(E < (1LL << 31), "hnt_merge_catalog: %lld containment pairs", (long long)E);
    const size_t ebytes = (size_t)std::max<int64_t>(E, 1) * 4;
    HNT_TRY(reserve(ctx, B.e1, ebytes));
    HNT_TRY(reserve(ctx, B.e2, ebytes));
    HNT_TRY(reserve(ctx, B.e3, ebytes));
    HNT_TRY(reserve(ctx, B.e4, ebytes));
    HNT_TRY(reserve(ctx, B.state, (size_t)N));
    HNT_TRY(reserve(ctx, B.instart, (size_t)(N + 1) * 8));
    uint8_t* state = (uint8_t*)B.state.p;
    HNT_HIP(hipMemsetAsync(state, 0, (size_t)N, st));
    int64_t* in_start = (int64_t*)B.instart.p;
    int32_t* in_src = (int32_t*)B.e3.p;
    if (E > 0) {
        hipLaunchKernelGGL(merge_edges, dim3(nblk(N)), dim3(MB), 0, st, G, 1, nullptr,
                           (const int64_t*)off, (int32_t*)B.e1.p, (int32_t*)B.e2.p);
        size_t t = 0;
        HNT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t, (int32_t*)nullptr,
                                                   (int32_t*)nullptr, (int32_t*)nullptr,
                                                   (int32_t*)nullptr, (int)E, 0, 32, st));
        if (t > ctx->cub_tmp.bytes) HNT_TRY(reserve(ctx, ctx->cub_tmp, t));
        t = ctx->cub_tmp.bytes;
        // keys: contained halo, values: container
        HNT_HIP(hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp.p, t, (int32_t*)B.e2.p,
                                                   (int32_t*)B.e4.p, (int32_t*)B.e1.p, in_src,
                                                   (int)E, 0, 32, st));
        hipLaunchKernelGGL(merge_csr<int32_t>, dim3(nblk(N + 1)), dim3(MB), 0, st,
                           (const int32_t*)B.e4.p, E, N, 0, in_start);
    } else {
        HNT_HIP(hipMemsetAsync(in_start, 0, (size_t)(N + 1) * 8, st));
    }
    HNT_TRY(reserve(ctx, ctx->small, 64));
    unsigned long long* und = (unsigned long long*)ctx->small.p;
    for (int round = 0;; ++round) {
        HNT_REQUIRE(round <= N, "hnt_merge_catalog: exclusion did not settle");
        HNT_HIP(hipMemsetAsync(und, 0, 8, st));
        hipLaunchKernelGGL(merge_round, dim3(nblk(N)), dim3(MB), 0, st, N,
                           (const int64_t*)in_start, (const int32_t*)in_src, state, und);
        unsigned long long left = 0;
        HNT_HIP(hipMemcpyAsync(&left, und, 8, hipMemcpyDeviceToHost, st));
        HNT_HIP(hipStreamSynchronize(st));
        if (left == 0) break;
    }

    // ---- volume reduction (volume_reduction!) ----
    HNT_TRY(reserve(ctx, B.dv, (size_t)N * 8));
    double* dV = (double*)B.dv.p;
    if (volume_reduction) {
        hipLaunchKernelGGL(merge_pairs, dim3(nblk(N)), dim3(MB), 0, st, G, 0,
                           (const uint8_t*)state, fortran_ties, cnt, nullptr, nullptr, nullptr);
        HNT_HIP(hipMemsetAsync(cnt + N, 0, 8, st));
        {
            size_t t = ctx->cub_tmp.bytes;
            HNT_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp.p, t, cnt, off, (int)(N + 1),
                                                     st));
        }
        int64_t P = 0;
        HNT_HIP(hipMemcpyAsync(&P, off + N, 8, hipMemcpyDeviceToHost, st));
        HNT_HIP(hipStreamSynchronize(st));
        HNT_REQUIRE(P < (1LL << 31), "hnt_merge_catalog: %lld cap entries (2 per overlapping "
                    "pair)", (long long)P);
        if (P > 0) {
            HNT_TRY(reserve(ctx, B.pk, (size_t)P * 8 * 2));
            HNT_TRY(reserve(ctx, B.pv, (size_t)P * 8 * 2));
            uint64_t* pk0 = (uint64_t*)B.pk.p;
            uint64_t* pk1 = pk0 + P;
            double* pv0 = (double*)B.pv.p;
            double* pv1 = pv0 + P;
            hipLaunchKernelGGL(merge_pairs, dim3(nblk(N)), dim3(MB), 0, st, G, 1,
                               (const uint8_t*)state, fortran_ties, nullptr,
                               (const int64_t*)off, pk0, pv0);
            size_t t = 0;
            HNT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t, pk0, pk1, pv0, pv1, (int)P, 0,
                                                       64, st));
            if (t > ctx->cub_tmp.bytes) HNT_TRY(reserve(ctx, ctx->cub_tmp, t));
            t = ctx->cub_tmp.bytes;
            HNT_HIP(hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp.p, t, pk0, pk1, pv0, pv1,
                                                       (int)P, 0, 64, st));
            hipLaunchKernelGGL(merge_csr<uint64_t>, dim3(nblk(N + 1)), dim3(MB), 0, st,
                               (const uint64_t*)pk1, P, N, 32, in_start);
            hipLaunchKernelGGL(merge_incoming, dim3(nblk(N)), dim3(MB), 0, st,
                               (const double*)pv1, (const int64_t*)in_start, N, dV);
        } else {
            HNT_HIP(hipMemsetAsync(dV, 0, (size_t)N * 8, st));
        }
    }

    // ---- survivors, in catalog order, with their new radii ----
    HNT_TRY(reserve(ctx, B.newr, (size_t)N * 4));
    int32_t* keep = (int32_t*)B.a.p;  // the key buffers are free now
    hipLaunchKernelGGL(merge_radii, dim3(nblk(N)), dim3(MB), 0, st, G, (const int32_t*)idx,
                       (const uint8_t*)state, (const double*)dV, volume_reduction,
                       (float*)B.newr.p, keep);
    {
        size_t t = 0;
        HNT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t, keep, off, (int)N, st));
        if (t > ctx->cub_tmp.bytes) HNT_TRY(reserve(ctx, ctx->cub_tmp, t));
        t = ctx->cub_tmp.bytes;
        HNT_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp.p, t, keep, off, (int)N, st));
    }
    hipLaunchKernelGGL(merge_scatter, dim3(nblk(N * nfields)), dim3(MB), 0, st, records, N,
                       nfields, (const int32_t*)keep, (const int64_t*)off,
                       (const float*)B.newr.p, out);
    int64_t last_off = 0;
    int32_t last_keep = 0;
    HNT_HIP(hipMemcpyAsync(&last_off, off + (N - 1), 8, hipMemcpyDeviceToHost, st));
    HNT_HIP(hipMemcpyAsync(&last_keep, keep + (N - 1), 4, hipMemcpyDeviceToHost, st));
    HNT_TRY(finish_call(ctx, "hnt_merge_catalog"));
    *nout = last_off + last_keep;
    return HNT_OK;
}

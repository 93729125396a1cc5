// hnt_features.h -- the 18 inner-shell features of a peak (bracket network inputs), shared
// by the standalone feature entry point and the learned peak schedule.
#pragma once

#include "hnt_math.h"

namespace hnt {

constexpr int kFeatShells = 12;

// e, p of the symmetrised tensor T (row-major 3x3), as ev_from_lam
__device__ inline void feat_ep(const double T[9], float* e, float* p) {
    double lam[3];
    eig3_sym(T[0], 0.5 * (T[1] + T[3]), 0.5 * (T[2] + T[6]), T[4], 0.5 * (T[5] + T[7]), T[8],
             lam);
    double F, ev, pv;
    ev_from_lam(lam, &F, &ev, &pv);
    *e = (float)ev;
    *p = (float)pv;
}

// f[0..11]  mean delta of shell s = 1..12 (in-bounds cells; 0 if none)
// f[12..15] e, p of the delta quadrupole sum_cells delta d_L d_K over shells 2..6, then
//           over shells 2..12 (d the cell's offset from the peak)
// f[16]     mean delta over shells 1..12
// f[17]     Rf / alatt
// load(off) reads delta at one cell: only the delta field is read (4 B per cell, lines of
// 32 cells along x that a peak's inner shells revisit), not the 32-B cell records.
template <class Load>
__device__ inline void peak_features(const int4* __restrict__ cells,
                                     const int32_t* __restrict__ shell_start,
                                     const int32_t* __restrict__ shell_count, int ci, int cj,
                                     int ck, int64_t c0, int n, int64_t s1, int64_t s2,
                                     double rf_over_alatt, Load load, float* __restrict__ f) {
    // the quadrupole is symmetric: Q00 Q01 Q02 Q11 Q12 Q22
    double Q[6] = {0, 0, 0, 0, 0, 0};
    double tot = 0.0;
    int64_t ntot = 0;
    // shells 1..12 reach 3 cells (r^2 <= 12): a peak at least 3 cells from every face reads
    // them without bounds checks, four cells in flight at a time
    const bool inside = ci - 3 >= 1 && ci + 3 <= n && cj - 3 >= 1 && cj + 3 <= n &&
                        ck - 3 >= 1 && ck + 3 <= n;
    auto add = [&](const int4& o, double d, double& F) {
        F += d;
        const double x = (double)o.x, y = (double)o.y, z = (double)o.z;
        Q[0] += d * (x * x);
        Q[1] += d * (x * y);
        Q[2] += d * (x * z);
        Q[3] += d * (y * y);
        Q[4] += d * (y * z);
        Q[5] += d * (z * z);
    };
    auto ep = [&](float* e, float* p) {
        const double T[9] = {Q[0], Q[1], Q[2], Q[1], Q[3], Q[4], Q[2], Q[4], Q[5]};
        feat_ep(T, e, p);
    };
    for (int s = 1; s <= kFeatShells; ++s) {
        const int st = shell_start[s - 1], nc = shell_count[s - 1];
        double F = 0.0;
        int nin = 0;
        if (inside) {
            int c = 0;
            for (; c + 4 <= nc; c += 4) {
                int4 o[4];
                float v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    o[u] = cells[st + c + u];
                    v[u] = load(c0 + o[u].x + o[u].y * s1 + o[u].z * s2);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) add(o[u], (double)v[u], F);
            }
            for (; c < nc; ++c) {
                const int4 o = cells[st + c];
                add(o, (double)load(c0 + o.x + o.y * s1 + o.z * s2), F);
            }
            nin = nc;
        } else {
            for (int c = 0; c < nc; ++c) {
                const int4 o = cells[st + c];
                const int i1 = ci + o.x, i2 = cj + o.y, i3 = ck + o.z;
                if (i1 < 1 || i1 > n || i2 < 1 || i2 > n || i3 < 1 || i3 > n) continue;
                add(o, (double)load(c0 + o.x + o.y * s1 + o.z * s2), F);
                ++nin;
            }
        }
        f[s - 1] = nin > 0 ? (float)(F / (double)nin) : 0.0f;
        tot += F;
        ntot += nin;
        if (s == 6) ep(&f[12], &f[13]);
    }
    ep(&f[14], &f[15]);
    f[16] = ntot > 0 ? (float)(tot / (double)ntot) : 0.0f;
    f[17] = (float)rf_over_alatt;
}

}  // namespace hnt

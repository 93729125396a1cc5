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
// f[12..15] e, p of sym(sum_cells psi1_L d_K) over shells 2..6, then over shells 2..12
// f[16]     mean delta over shells 1..12
// f[17]     Rf / alatt
// load(off, delta, psi1x, psi1y, psi1z) reads one cell.
template <class Load>
__device__ inline void peak_features(const int4* __restrict__ cells,
                                     const int32_t* __restrict__ shell_start,
                                     const int32_t* __restrict__ shell_count, int ci, int cj,
                                     int ck, int64_t c0, int n, int64_t s1, int64_t s2,
                                     double rf_over_alatt, Load load, float* __restrict__ f) {
    double T[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double tot = 0.0;
    int64_t ntot = 0;
    for (int s = 1; s <= kFeatShells; ++s) {
        const int st = shell_start[s - 1], nc = shell_count[s - 1];
        double F = 0.0;
        int nin = 0;
        for (int c = 0; c < nc; ++c) {
            const int4 o = cells[st + c];
            const int i1 = ci + o.x, i2 = cj + o.y, i3 = ck + o.z;
            if (i1 < 1 || i1 > n || i2 < 1 || i2 > n || i3 < 1 || i3 > n) continue;
            float d, ex, ey, ez;
            load(c0 + o.x + o.y * s1 + o.z * s2, d, ex, ey, ez);
            F += (double)d;
            ++nin;
            const double e3[3] = {(double)ex, (double)ey, (double)ez};
            const double d3[3] = {(double)o.x, (double)o.y, (double)o.z};
#pragma unroll
            for (int L = 0; L < 3; ++L)
#pragma unroll
                for (int K = 0; K < 3; ++K) T[L * 3 + K] += e3[L] * d3[K];
        }
        f[s - 1] = nin > 0 ? (float)(F / (double)nin) : 0.0f;
        tot += F;
        ntot += nin;
        if (s == 6) feat_ep(T, &f[12], &f[13]);
    }
    feat_ep(T, &f[14], &f[15]);
    f[16] = ntot > 0 ? (float)(tot / (double)ntot) : 0.0f;
    f[17] = (float)rf_over_alatt;
}

}  // namespace hnt

/*
 * huntian_oracle.h -- TEST INFRASTRUCTURE.  FP64 CPU restatement of Huntian.jl's
 * peak-finding + radial-shell collapse path, used only as the parity checker by tests/,
 * by __graft_entry__.smoke() and as the CPU baseline leg of bench.py.  It is never the
 * product: libhuntian_hip.so does not link or call it.
 *
 * Each function cites the reference lines it restates (paths relative to the Huntian.jl
 * source tree).  Types come from the public ABI header; here every pointer is host memory.
 */
#ifndef HUNTIAN_ORACLE_H
#define HUNTIAN_ORACLE_H

#include "../include/huntian_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RadialShell.precompute_shells + ShellAnalysisGPU.build_shell_tables */
int64_t hro_shell_cells(int32_t rmax);
int32_t hro_shell_count(int32_t rmax);
int32_t hro_build_shell_tables(int32_t rmax, int32_t* di, int32_t* dj, int32_t* dk,
                               int32_t* shell_start, int32_t* shell_count, int32_t* shell_r2);

double hro_hRinteg(double u0, double u12);          /* RadialShell.jl:95-121 */
void hro_atab4(double* akk /* [1101] */);            /* RadialShell.jl:125-147 */
/* RadialShell.kernel_strain (RadialShell.jl:225-274); 1-based mp/mlow/mupp; G, Gf
 * [3*nshells] (shell-major: m*3 + L), SR [9*nshells] (m*9 + L*3 + K).  out: Ebar 11,12,13,
 * 22,23,33, grad xyz, gradf xyz. */
void hro_kernel_strain(const double* rad, const double* G, const double* Gf, const double* SR,
                       int32_t mp, int32_t mlow, int32_t mupp, double wRnor, double aRnor,
                       double hlatt_1, double hlatt_2, double* out);
/* Ascending eigenvalues of a symmetric 3x3 (column-major, upper triangle used) */
void hro_eig3(const double* A, double* lam);
double hro_interp3(const hnt_ctab* ct, double x, double y, double z);
double hro_fsc_of_z(const hnt_ctab* ct, double z);   /* RadialShell.jl:181-195 */

/* PeakFind.find_peaks (PeakFind.jl:15-78), one filter.  Returns the number of peaks
 * written (<= max_peaks); 1-based indices. */
int64_t hro_find_peaks(const float* delta, int8_t* mask, int32_t n1, int32_t n2, int32_t n3,
                       int32_t nbuff, float fcrit, int64_t max_peaks,
                       int32_t* pi, int32_t* pj, int32_t* pk, float* pdelta);
/* The per-filter loop of the drivers (MultiResolution.jl:1046-1066): filters in the given
 * order, mask carried over.  Returns the total; writes at most cap entries. */
int64_t hro_peak_find_multi(const float* const* fields, const float* fcrit, int32_t nfilt,
                            int32_t n, int32_t nbuff, int8_t* mask, int64_t cap,
                            int32_t* pi, int32_t* pj, int32_t* pk, int32_t* filt, float* pdelta);
int64_t hro_peak_find_multi_par(const float* const* fields, const float* fcrit, int32_t nfilt,
                                int32_t n, int32_t nbuff, int8_t* mask, int64_t cap,
                                int32_t* pi, int32_t* pj, int32_t* pk, int32_t* filt,
                                float* pdelta, int32_t nthreads);

/* Per-shell sums of every shell for every peak (the brute-force loop of
 * test/test_shell_gpu.jl:241-300, in the FP64 arithmetic of analyse_peak_gpu). */
void hro_shell_gather_full(const hnt_fields* f, int64_t npeaks, const int32_t* pi,
                           const int32_t* pj, const int32_t* pk, const hnt_shell_tables* t,
                           double* F, int32_t* n, double* S, double* S2, double* G, double* Gf,
                           double* SR, double* lapd);

/* analyse_peak (variant A) / analyse_peak_gpu (B) / B + zvir_half (C) over a batch.
 * nthreads > 1 uses OpenMP over peaks (as Threads.@threads, src/Pipeline.jl:235).
 * Returns 0 on success, nonzero on invalid input. */
int hro_analyse_peaks(const hnt_fields* f, const hnt_batch* b, const hnt_shell_tables* t,
                      const hnt_ctab* ct, const hnt_params* p, hnt_results* r, int8_t* mask,
                      int32_t nthreads);

/* The record loop of the tile driver after the shell analysis (MultiResolution.jl:1159-1219)
 * with the layouts of src/IO/Catalog.jl:4-46, Dlinear_ab, omega_m_a, chi_to_z and
 * peak_redshift (src/Cosmology/Cosmology.jl:50,70-130,287-307,357-385).  Inputs are host
 * arrays: the 1-based peak cells, Rf, the FP64 results (used through their Float32 values,
 * as the reference's GPU results are Float32), FcRf, d2Rf (or NULL).  Writes the records of
 * the first `capacity` halos and returns the number of halos. */
int64_t hro_pack_records(int64_t npeaks, const int32_t* pi, const int32_t* pj,
                         const int32_t* pk, const double* Rf, const hnt_results* r,
                         const float* FcRf, const float* d2Rf, const hnt_cosmo_tables* t,
                         const hnt_pack_params* pp, float* out, int64_t capacity);
double hro_growth_D(const hnt_cosmo_tables* t, double a);   /* Dlinear_ab(a)[1] */
double hro_growth_f(const hnt_cosmo_tables* t, double a);   /* Dlinear_ab(a)[2] */
double hro_pow_m3(double x);
void hro_finalize_eulerian(const float* rec, int64_t nhalo, int32_t nf, const hnt_cosmo_tables* t,
                           double Om, double OL, const double* obs, int32_t ievol, double z_out,
                           float* out);

/* Merger.merge_catalog (src/Merger/Merger.jl:23-89) with build_hash, auto_hash_nc,
 * lagrangian_exclusion!, sphere_overlap, volume_reduction! (src/Merger/Exclusion.jl), over
 * host records [nhalo * nf]; writes the survivors (RTHL replaced) and returns their number,
 * -1 if memory runs out or -2 for an empty domain. */
int64_t hro_merge_catalog(const float* rec, int64_t nhalo, int32_t nf, int32_t volume_reduction,
                          int32_t fortran_ties, float* out);
/* the exclusion and reduction passes alone: survived flags and FP64 radii per halo */
int32_t hro_merge_passes(const float* rec, int64_t nhalo, int32_t nf, int32_t volume_reduction,
                         int32_t fortran_ties, uint8_t* survived, double* new_r);
/* sphere_overlap (Exclusion.jl:86-101) */
void hro_sphere_overlap(double d, double r1, double r2, double* v1, double* v2);
double hro_chi_to_z(const hnt_cosmo_tables* t, double chi);  /* chi_to_z */

/* tile white noise: Threefry-2x64-20 (key k0 k1, counter c0 c1), the Gaussian of a global
 * cell, _generate_tile_noise, _downsample_noise, _generate_extended_residual
 * (src/InitialConditions/RandomField.jl:45-53, src/MultiResolution.jl:35-87,190-227) */
void hro_threefry2x64_20(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1, uint64_t* out);
double hro_threefry_gauss(uint64_t seed, uint64_t idx0);
void hro_tile_noise(int32_t it, int32_t jt, int32_t kt, int32_t nsub, int32_t nmesh, int64_t N,
                    uint64_t seed, float* out);
void hro_downsample_noise(int64_t N, int32_t M, uint64_t seed, float* coarse);
void hro_extended_residual(int32_t it, int32_t jt, int32_t kt, int32_t nsub, int32_t nmesh,
                           int64_t N, uint64_t seed, const float* coarse, int32_t M,
                           int32_t nshell, float* out);

#ifdef __cplusplus
}
#endif

#endif

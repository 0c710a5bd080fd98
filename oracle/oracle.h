/*
 * FP64 CPU oracle for AMBER's 4_fbe estimator hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (amber_amd/, libamber_mi355x.so, the amber/amber_v2 drivers) never
 * links, loads or calls anything here.
 *
 * It is a clean-room restatement of the reference's algorithm; every function
 * cites the reference lines it follows (paths relative to the reference root).
 * How each function is pinned (not by retyping its formula):
 *   orc_fft_c2c       direct O(n^2) DFT on small n and numpy.fft (pocketfft)
 *   orc_ring_c2r      numpy.fft.irfft2 * Ng^2 on the same masked half plane,
 *                     and a brute-force direct 2-D sum on tiny grids
 *   orc_normalize     V = S c gives back c where S > 0 (S := 1), cells with
 *                     S <= 0 untouched, idempotent
 *   orc_bispec_cells  equal to orc_bispec cell for cell (same triplets, same
 *                     products) on flagged cubes with lag windows
 *   orc_uavg          (2/3) H at L = 0 and the area-weighted mean radius by
 *                     quadrature
 *   orc_ng, orc_uavg,
 *   orc_triangles,
 *   orc_annular_rings closed forms and hand-computed cases (triangle counts of
 *                     the shipped 22-ring setup, equilateral borderline)
 *   orc_calcr         Num / N^2 equals the brute-force count of closed
 *                     sampled triangles k1+k2+k3 = 0 on tiny grids
 *   orc_bispec        triple-product invariant sum_x D1 D2 D3 =
 *                     Ng^2 * sum_{k1+k2+k3=0} V1 V2 V3 by brute force, the
 *                     alpha^3 scaling, channel-constant cubes, and the
 *                     closed-form triplet counts
 *   orc_refalgo_*     the reference-algorithm FP32 path; pinned against
 *                     orc_bispec (FP32-level agreement) and by identical counts
 *   orc_maps_bin      hand-computed bins of chosen cells (3-4-5 radii,
 *                     bin edges), the row fold ia = N - i
 *   orc_cldnu         closed forms: a cube V(x,n) = c(x) e^{i phi n} gives
 *                     cldnu[b][d] = (nc - d) cos(phi d) <|c|^2>_b, with
 *                     channel masks sum_n1 sum_x [S1>0][S2>0]|c|^2 / no_b(n1);
 *                     numdnu = nc - d; counts and mean |U| per bin by brute
 *                     enumeration of the cells
 *   orc_grid_setup    the shipped inputparameters (Umax 602, FWHM 1480.73)
 *                     worked by hand
 *   orc_grid          hand-computed cells of chosen (u, v), the v < 0 flip
 *                     equal to the conjugate group, constant visibilities
 *                     summing to count x value, flagged weights, the
 *                     Umax boundary, collapse sums
 * No function is left "parity unpinned".
 */
#ifndef AMBER_ORACLE_H
#define AMBER_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Unnormalised in-place complex DFT of length n; sign=+1 is exp(+2 pi i jk/n)
 * (FFTW_BACKWARD, the direction of every c2r in amber_funcs.c:402,990,1116).
 * re/im have n entries each. */
void orc_fft_c2c(int n, double *re, double *im, int sign);

/* Direct O(n^2) DFT used to validate orc_fft_c2c. */
void orc_dft_direct(int n, const double *re, const double *im, double *ore, double *oim, int sign);

/* normalize_data (amber_funcs.c:1946-1978) in place on file-layout arrays:
 * where NTiG > 0, V /= NTiG and then NTiG := 1. */
void orc_normalize(long cells, long nc, double *re_im, double *samp);

/* Reduced grid side for a triangle whose largest ring has upper edge uhigh
 * (amber.c:193-197): Nmax=(long)((uhigh+3)/Du); Ng=3*Nmax; Ng+=2 if even else 3. */
long orc_ng(double uhigh, double du);

/* Ring-average |U| (amber_funcs.c:2065). */
void orc_uavg(const double *ulow, const double *uhigh, int nr, double *uavg);

/* The reference's hard-coded piecewise ring edges (amber_funcs.c:2022-2053). */
void orc_annular_rings(int nr, double *ulow, double *uhigh);

/* Closed-triangle rule r1>=r2>=r3 (amber_funcs.c:443-449).  Writes up to cap
 * triples (r1,r2,r3) into tri and returns the total count. */
int orc_triangles(const double *uavg, int nr, int *tri, int cap);

/* Ring gather + unnormalised c2r on the Ng grid (Dfg/Ifg, amber_funcs.c:573-621).
 * Source cube in the reference's pixel-major file layout:
 *   re_im  [N*(N/2+1)][nc][2]  (NULL for a sampling plane)
 *   samp   [N*(N/2+1)][nc]     (used when re_im == NULL; real part only)
 * out is Ng*Ng doubles, row-major (row = first FFT axis). */
void orc_ring_c2r(long N, long nc, long Ng, double ulow, double uhigh,
                  const double *re_im, const double *samp, long plane, double *out);

/* calCR (amber_funcs.c:389-491): per ring the c2r of the sampling plane `plane`
 * (5 in the reference) on the full N grid, then for each closed triangle the
 * FP64 triple sum.  rdata gets {r1,r2,r3,Num} per triangle; returns rcount. */
int orc_calcr(long N, long nc, const double *samp, int plane,
              const double *ulow, const double *uhigh, int nr,
              double *rdata, int cap);

/* FP64 bispectrum accumulators for one triangle (amber.c:187-304 v1,
 * amber_v2.c:187-309 v2).  chan_mask[nc] is 1 for channels in the triplet set
 * (all channels when flag_key=0, uf_chans when flag_key=1;
 * amber_funcs.c:2207-2310).  W = 2*nc10+1.
 *   bl64[W*W]   sum_{p1} sum_x D_r1(p1) D_r2(p1+d2) D_r3(p1+d3)      (data)
 *   norm64[W*W] same on the sampling images (v2 numdnu; may be NULL)
 *   cnt[W*W]    number of valid p1 per lag cell (v1 numdnu; may be NULL)
 * Layout [d2+nc10][d3+nc10], as amber_funcs.c:2293-2304.
 * nthreads <= 0 uses OpenMP's default. Returns 0 or a negative error. */
int orc_bispec(long N, long nc, double Du, const double *re_im, const double *samp,
               const double *ulow, const double *uhigh, int r1, int r2, int r3,
               const uint8_t *chan_mask, int nc10,
               double *bl64, double *norm64, int32_t *cnt, int nthreads);

/* Same sums evaluated only at ncell lag cells (d2[i], d3[i]) (lags, not
 * offsets), for large configurations.  out_bl/out_norm/out_abs may be NULL;
 * out_abs gets sum |D1 D2 D3| over the same terms (the scale of the sum). */
int orc_bispec_cells(long N, long nc, double Du, const double *re_im, const double *samp,
                     const double *ulow, const double *uhigh, int r1, int r2, int r3,
                     const uint8_t *chan_mask, int nc10, int ncell,
                     const int *d2, const int *d3, double *out_bl, double *out_norm,
                     int32_t *out_cnt, double *out_abs, int nthreads);

/* Reference-algorithm CPU path (amber.c:187-304 / amber_v2.c:187-309):
 * FP64 c2r per plane cast to float, the 6-int triplet table
 * (amber_funcs.c:2207-2310) and the FP32 sequential triple loop under
 * `omp for schedule(dynamic)` with private accumulators and a critical merge.
 * Restricted to p1 in [p1_lo, p1_hi) so a bounded sample can be timed.
 * bl/numf are W*W floats, cnt W*W ints (any may be NULL except bl).
 * Returns the number of triplets evaluated; *seconds_loop gets the time of
 * the triple loop alone and *seconds_fill the time of the plane fills. */
long orc_refalgo(long N, long nc, double Du, const double *re_im, const double *samp,
                 const double *ulow, const double *uhigh, int r1, int r2, int r3,
                 const uint8_t *chan_mask, int nc10, int p1_lo, int p1_hi, int v2,
                 float *bl, float *numf, int32_t *cnt, int nthreads,
                 double *seconds_fill, double *seconds_loop);

/* MAPS C_l(dnu) (4_fbe/maps.c:100-229 with cl_dnu, amber_funcs.c:1418-1488).
 * Bin of half-plane cell (i, j): -1 outside [umin, umax); *uu_out gets |U|. */
int orc_maps_bin(long N, double Du, long i, long j, int NB, double umin, double umax, double *uu_out);

/* The whole MAPS run on a cube in the file layout (re_im [N(N/2+1)][nc][2],
 * samp [N(N/2+1)][nc], normalised here as maps.c:100-125 does):
 * cldnu[NB][nc] = sum over nn1 of binned C(nn1, nn1 + d), numdnu[nc] the
 * number of pairs per lag d, binu[NB] / num[NB] the mean |U| and cell count
 * of the last pair (nc-1, nc-1).  Returns 0, or -1 when out of memory. */
int orc_cldnu(long N, long nc, double Du, const double *re_im, const double *samp, int NB, double umin,
              double umax, double *cldnu, double *numdnu, double *binu, int32_t *num);

/* Upstream gridder vcg (3_vcg/tgefuncs.c:290-376, 377-626, 816-839).
 * orc_grid_setup: dU and NN from Umax and the beam FWHM (arcmin).
 * orc_grid: ranpar [ngroup][npar] (u at 0, v at 1, seconds), data
 * [ngroup][nchan][nstokes][3] (re, im, wt); channels chan1 .. chan1+nc-1
 * (1-based) of stokes 0 are gridded; gv [NN(NN/2+1)][nc/n_collapse][2],
 * samp [NN(NN/2+1)][nc/n_collapse].  Returns the number of groups gridded,
 * -1 out of memory, -2 a group outside the grid. */
void orc_grid_setup(double Umax, double FWHM, double *dU, long *NN);
long orc_grid(long NN, long nc, long nchan, long nstokes, int chan1, int npar, int flag_key, int n_collapse,
              float nu_chan0, double dU, double Umax, long ngroup, const float *ranpar, const float *data,
              double *gv, double *samp);

#ifdef __cplusplus
}
#endif
#endif

/*
 * FP64 CPU oracle for AMBER's 4_fbe hot path -- TEST INFRASTRUCTURE ONLY.
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker; never by the product.  See oracle.h for how each function
 * is pinned.  Reference citations are relative to the reference root.
 */
#include "oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Complex DFT.  A recursive mixed-radix decimation-in-time Cooley-Tukey     */
/* with an O(p^2) combine for every prime factor p.  It has nothing in       */
/* common with the GPU's Stockham kernels except the definition it computes.*/
/* ------------------------------------------------------------------------ */

static int smallest_factor(int n)
{
    if (n % 4 == 0 && n > 4) return 4;
    for (int p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

static void fft_rec(int n, int tstride, const double complex *x, int xs,
                    double complex *y, double complex *tmp,
                    const double complex *tw, int nt)
{
    if (n == 1) { y[0] = x[0]; return; }
    int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; ++r)
        fft_rec(m, tstride * p, x + (long)r * xs, xs * p, y + (long)r * m, tmp, tw, nt);
    /* X[k + q m] = sum_r w_n^{r (k + q m)} Y_r[k], w_n = tw[tstride] */
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            double complex acc = 0;
            for (int r = 0; r < p; ++r) {
                long e = ((long)r * (k + (long)q * m)) % n;
                acc += y[(long)r * m + k] * tw[(e * tstride) % nt];
            }
            tmp[q] = acc;
        }
        for (int q = 0; q < p; ++q) y[k + (long)q * m] = tmp[q];
    }
}

void orc_fft_c2c(int n, double *re, double *im, int sign)
{
    if (n <= 1) return;
    double complex *tw = malloc(sizeof(double complex) * n);
    double complex *x = malloc(sizeof(double complex) * n);
    double complex *y = malloc(sizeof(double complex) * n);
    double complex *tmp = malloc(sizeof(double complex) * n);
    for (int j = 0; j < n; ++j) {
        double a = 2.0 * M_PI * (double)j / (double)n;
        tw[j] = cos(a) + (sign > 0 ? 1.0 : -1.0) * I * sin(a);
        x[j] = re[j] + I * im[j];
    }
    fft_rec(n, 1, x, 1, y, tmp, tw, n);
    for (int j = 0; j < n; ++j) { re[j] = creal(y[j]); im[j] = cimag(y[j]); }
    free(tw); free(x); free(y); free(tmp);
}

void orc_dft_direct(int n, const double *re, const double *im, double *ore, double *oim, int sign)
{
    for (int k = 0; k < n; ++k) {
        double sr = 0, si = 0;
        for (int j = 0; j < n; ++j) {
            long e = ((long)j * k) % n;
            double a = 2.0 * M_PI * (double)e / (double)n;
            double c = cos(a), s = (sign > 0 ? 1.0 : -1.0) * sin(a);
            sr += re[j] * c - im[j] * s;
            si += re[j] * s + im[j] * c;
        }
        ore[k] = sr; oim[k] = si;
    }
}

/* ------------------------------------------------------------------------ */
/* Integer / FP64 rules                                                      */
/* ------------------------------------------------------------------------ */

/* normalize_data (4_fbe/amber_funcs.c:1946-1978): at every (cell, channel)
 * whose count NTiG > 0 the gridded visibility is divided by NTiG (re and im,
 * FP64), then NTiG := 1; cells with NTiG <= 0 are left as they are.  Arrays
 * in the file layout: re_im[cells][nc][2], samp[cells][nc]. */
void orc_normalize(long cells, long nc, double *re_im, double *samp)
{
    for (long i = 0; i < cells * nc; ++i)
        if (samp[i] > 0) {
            re_im[2 * i] /= samp[i];
            re_im[2 * i + 1] /= samp[i];
        }
    for (long i = 0; i < cells * nc; ++i)
        if (samp[i] > 0) samp[i] = 1;
}

long orc_ng(double uhigh, double du)
{   /* amber.c:193-197 */
    double umax = uhigh + 3;
    long nmax = (long)(umax / du);
    long ng = (long)(3 * nmax);
    ng = (ng % 2 == 0) ? ng + 2 : ng + 3;
    return ng;
}

void orc_uavg(const double *ulow, const double *uhigh, int nr, double *uavg)
{   /* amber_funcs.c:2065 */
    for (int i = 0; i < nr; ++i)
        uavg[i] = (2. / 3.) * (pow(uhigh[i], 3.) - pow(ulow[i], 3.)) /
                  (pow(uhigh[i], 2.) - pow(ulow[i], 2.));
}

void orc_annular_rings(int nr, double *ulow, double *uhigh)
{   /* amber_funcs.c:2025-2053: piecewise-linear edges, then ULow[0]=1,
       UHigh[Nr-1]=270. */
    static const int iv[4][3] = {{0, 50, 5}, {50, 100, 10}, {100, 200, 20}, {200, 270, 35}};
    long idx = 0;
    for (int i = 0; i < nr; ++i) { ulow[i] = 0; uhigh[i] = 0; }
    for (int i = 0; i < 4; ++i)
        for (int j = iv[i][0]; j < iv[i][1]; j += iv[i][2])
            if (idx < nr) {
                ulow[idx] = j;
                if (idx > 0) uhigh[idx - 1] = ulow[idx];
                ++idx;
            }
    ulow[0] = 1.;
    uhigh[nr - 1] = 270.;
}

int orc_triangles(const double *uavg, int nr, int *tri, int cap)
{   /* amber_funcs.c:443-449 with the same double expression order */
    int n = 0;
    for (int r1 = 0; r1 < nr; ++r1)
        for (int r2 = 0; r2 <= r1; ++r2) {
            double tt = uavg[r2] / uavg[r1];
            if (tt >= 0.5 && tt <= 1.0)
                for (int r3 = 0; r3 <= r2; ++r3) {
                    double mu = 0.5 * (tt + 1. / tt - (uavg[r3] / uavg[r1]) * (uavg[r3] / uavg[r2]));
                    if (mu >= 0.5 / tt && mu <= 1.0) {
                        if (n < cap && tri) { tri[3 * n] = r1; tri[3 * n + 1] = r2; tri[3 * n + 2] = r3; }
                        ++n;
                    }
                }
        }
    return n;
}

/* ------------------------------------------------------------------------ */
/* c2r of a half-plane H[rows][cols_h] (cols_h = Ng/2+1): column IDFTs along */
/* the first axis, then a Hermitian-extended row IDFT whose real part is     */
/* the output.  A j=0 column that is not Hermitian enters through its real   */
/* projection, i.e. as the Hermitian-symmetrised column.                     */
/* ------------------------------------------------------------------------ */
static void c2r_half(long Ng, double complex *H, double *out)
{
    long Ngh = Ng / 2 + 1;
    double *re = malloc(sizeof(double) * Ng), *im = malloc(sizeof(double) * Ng);
    for (long j = 0; j < Ngh; ++j) {
        int any = 0;
        for (long i = 0; i < Ng; ++i) {
            re[i] = creal(H[i * Ngh + j]); im[i] = cimag(H[i * Ngh + j]);
            any |= (re[i] != 0.0 || im[i] != 0.0);
        }
        if (!any) continue;
        orc_fft_c2c((int)Ng, re, im, +1);
        for (long i = 0; i < Ng; ++i) H[i * Ngh + j] = re[i] + I * im[i];
    }
    for (long x = 0; x < Ng; ++x) {
        double complex *g = H + x * Ngh;
        re[0] = creal(g[0]); im[0] = 0;
        for (long j = 1; j < Ng; ++j) { re[j] = 0; im[j] = 0; }
        for (long j = 1; j < Ngh; ++j) {
            if (2 * j == Ng) { re[j] = creal(g[j]); im[j] = 0; continue; }
            re[j] = creal(g[j]); im[j] = cimag(g[j]);
            re[Ng - j] = creal(g[j]); im[Ng - j] = -cimag(g[j]);
        }
        orc_fft_c2c((int)Ng, re, im, +1);
        for (long y = 0; y < Ng; ++y) out[x * Ng + y] = re[y];
    }
    free(re); free(im);
}

void orc_ring_c2r(long N, long nc, long Ng, double ulow, double uhigh,
                  const double *re_im, const double *samp, long plane, double *out)
{   /* Dfg amber_funcs.c:573-595 (re_im) / Ifg :599-621 (samp) */
    long Nh = N / 2 + 1, Ngh = Ng / 2 + 1;
    double complex *H = calloc((size_t)(Ng * Ngh), sizeof(double complex));
    for (long ia = -Ng / 2 + 1; ia < Ng / 2 + 1; ++ia)
        for (long j = 0; j < Ng / 3; ++j) {
            long iN = (ia < 0) ? (N + ia) : ia;
            long iNg = (ia < 0) ? (Ng + ia) : ia;
            double uu = sqrt(1. * (double)(ia * ia + j * j));
            if (uu >= ulow && uu < uhigh && uu > 0.0) {
                long src = iN * Nh + j;
                if (iN < 0 || iN >= N || j >= Nh) continue; /* outside the source cube */
                if (re_im)
                    H[iNg * Ngh + j] = re_im[(src * nc + plane) * 2] + I * re_im[(src * nc + plane) * 2 + 1];
                else
                    H[iNg * Ngh + j] = samp[src * nc + plane];
            }
        }
    c2r_half(Ng, H, out);
    free(H);
}

int orc_calcr(long N, long nc, const double *samp, int plane,
              const double *ulow, const double *uhigh, int nr,
              double *rdata, int cap)
{   /* amber_funcs.c:389-491 */
    long Nh = N / 2 + 1, P = N * N;
    double *II = malloc(sizeof(double) * (size_t)P * nr);
    double complex *H = malloc(sizeof(double complex) * (size_t)(N * Nh));
    for (int r = 0; r < nr; ++r) {
        memset(H, 0, sizeof(double complex) * (size_t)(N * Nh));
        for (long i = 0; i < N; ++i) {               /* :419-429 */
            long ia = (i > N / 2) ? (N - i) : i;
            for (long j = 0; j < N / 3; ++j) {
                long index = i * Nh + j;
                double uu = sqrt(1.0 * (double)(ia * ia + j * j));
                if (uu >= ulow[r] && uu < uhigh[r] && uu > 0.0)
                    H[index] = samp[index * nc + plane];
            }
        }
        c2r_half(N, H, II + (size_t)r * P);
    }
    free(H);
    double *uavg = malloc(sizeof(double) * nr);
    orc_uavg(ulow, uhigh, nr, uavg);
    int n = 0;
    for (int r1 = 0; r1 < nr; ++r1)                   /* :443-470 */
        for (int r2 = 0; r2 <= r1; ++r2) {
            double tt = uavg[r2] / uavg[r1];
            if (tt >= 0.5 && tt <= 1.0)
                for (int r3 = 0; r3 <= r2; ++r3) {
                    double mu = 0.5 * (tt + 1. / tt - (uavg[r3] / uavg[r1]) * (uavg[r3] / uavg[r2]));
                    if (mu >= 0.5 / tt && mu <= 1.0) {
                        const double *a = II + (size_t)r1 * P, *b = II + (size_t)r2 * P, *c = II + (size_t)r3 * P;
                        double num = 0.0;
                        for (long x = 0; x < P; ++x) num += a[x] * b[x] * c[x];
                        if (n < cap && rdata) {
                            rdata[4 * n] = r1; rdata[4 * n + 1] = r2; rdata[4 * n + 2] = r3; rdata[4 * n + 3] = num;
                        }
                        ++n;
                    }
                }
        }
    free(uavg); free(II);
    return n;
}

/* ------------------------------------------------------------------------ */
/* Bispectrum accumulators in FP64                                           */
/* ------------------------------------------------------------------------ */

/* Planes of one ring at Ng for the channels flagged in need[] (others zero). */
static double *ring_planes(long N, long nc, long Ng, double lo, double hi,
                           const double *re_im, const double *samp, const uint8_t *need, int nthreads)
{
    long P = Ng * Ng;
    double *d = calloc((size_t)P * nc, sizeof(double));
    if (!d) return NULL;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (long p = 0; p < nc; ++p)
        if (need[p]) orc_ring_c2r(N, nc, Ng, lo, hi, re_im, samp, p, d + (size_t)p * P);
    return d;
}

static int threads_or_default(int nthreads)
{
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads; return 1;
#endif
}

int orc_bispec(long N, long nc, double Du, const double *re_im, const double *samp,
               const double *ulow, const double *uhigh, int r1, int r2, int r3,
               const uint8_t *chan_mask, int nc10,
               double *bl64, double *norm64, int32_t *cnt, int nthreads)
{
    int W = 2 * nc10 + 1;
    long Ng = orc_ng(uhigh[r1], Du), P = Ng * Ng;
    nthreads = threads_or_default(nthreads);
    int rr[3] = {r1, r2, r3};
    double *D[3] = {0}, *S[3] = {0};
    for (int k = 0; k < 3; ++k) {
        int dup = -1;
        for (int q = 0; q < k; ++q) if (rr[q] == rr[k]) dup = q;
        if (dup >= 0) { D[k] = D[dup]; S[k] = S[dup]; continue; }
        D[k] = ring_planes(N, nc, Ng, ulow[rr[k]], uhigh[rr[k]], re_im, NULL, chan_mask, nthreads);
        if (!D[k]) return -1;
        if (norm64) {
            S[k] = ring_planes(N, nc, Ng, ulow[rr[k]], uhigh[rr[k]], NULL, samp, chan_mask, nthreads);
            if (!S[k]) return -1;
        }
    }
    /* Each thread owns whole d2 rows, so the p1 order of every sum is fixed
       (ascending, as the triplet table orders p1 first: amber_funcs.c:2290). */
#pragma omp parallel num_threads(nthreads)
    {
        double *A = malloc(sizeof(double) * P), *AS = norm64 ? malloc(sizeof(double) * P) : NULL;
#pragma omp for schedule(dynamic)
        for (int i2 = 0; i2 < W; ++i2) {
            int d2 = i2 - nc10;
            for (int i3 = 0; i3 < W; ++i3) {
                bl64[(long)i2 * W + i3] = 0;
                if (norm64) norm64[(long)i2 * W + i3] = 0;
                if (cnt) cnt[(long)i2 * W + i3] = 0;
            }
            for (long p1 = 0; p1 < nc; ++p1) {
                long p2 = p1 + d2;
                if (!chan_mask[p1] || p2 < 0 || p2 >= nc || !chan_mask[p2]) continue;
                const double *a1 = D[0] + p1 * P, *a2 = D[1] + p2 * P;
                for (long x = 0; x < P; ++x) A[x] = a1[x] * a2[x];
                if (norm64) {
                    const double *s1 = S[0] + p1 * P, *s2 = S[1] + p2 * P;
                    for (long x = 0; x < P; ++x) AS[x] = s1[x] * s2[x];
                }
                for (int i3 = 0; i3 < W; ++i3) {
                    long p3 = p1 + (i3 - nc10);
                    if (p3 < 0 || p3 >= nc || !chan_mask[p3]) continue;
                    const double *a3 = D[2] + p3 * P;
                    double s = 0;
                    for (long x = 0; x < P; ++x) s += A[x] * a3[x];
                    bl64[(long)i2 * W + i3] += s;
                    if (norm64) {
                        const double *s3 = S[2] + p3 * P;
                        double t = 0;
                        for (long x = 0; x < P; ++x) t += AS[x] * s3[x];
                        norm64[(long)i2 * W + i3] += t;
                    }
                    if (cnt) cnt[(long)i2 * W + i3] += 1;
                }
            }
        }
        free(A); free(AS);
    }
    for (int k = 0; k < 3; ++k) {
        int dup = 0;
        for (int q = 0; q < k; ++q) if (rr[q] == rr[k]) dup = 1;
        if (!dup) { free(D[k]); free(S[k]); }
    }
    return 0;
}

int orc_bispec_cells(long N, long nc, double Du, const double *re_im, const double *samp,
                     const double *ulow, const double *uhigh, int r1, int r2, int r3,
                     const uint8_t *chan_mask, int nc10, int ncell,
                     const int *d2s, const int *d3s, double *out_bl, double *out_norm,
                     int32_t *out_cnt, double *out_abs, int nthreads)
{
    long Ng = orc_ng(uhigh[r1], Du), P = Ng * Ng;
    nthreads = threads_or_default(nthreads);
    int rr[3] = {r1, r2, r3};
    double *D[3] = {0}, *S[3] = {0};
    for (int k = 0; k < 3; ++k) {
        int dup = -1;
        for (int q = 0; q < k; ++q) if (rr[q] == rr[k]) dup = q;
        if (dup >= 0) { D[k] = D[dup]; S[k] = S[dup]; continue; }
        if (out_bl) {
            D[k] = ring_planes(N, nc, Ng, ulow[rr[k]], uhigh[rr[k]], re_im, NULL, chan_mask, nthreads);
            if (!D[k]) return -1;
        }
        if (out_norm) {
            S[k] = ring_planes(N, nc, Ng, ulow[rr[k]], uhigh[rr[k]], NULL, samp, chan_mask, nthreads);
            if (!S[k]) return -1;
        }
    }
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int c = 0; c < ncell; ++c) {
        long d2 = d2s[c], d3 = d3s[c];
        double sb = 0, sn = 0, sa = 0;
        int n = 0;
        if (d2 < -nc10 || d2 > nc10 || d3 < -nc10 || d3 > nc10) goto done;
        for (long p1 = 0; p1 < nc; ++p1) {
            long p2 = p1 + d2, p3 = p1 + d3;
            if (!chan_mask[p1] || p2 < 0 || p2 >= nc || !chan_mask[p2] ||
                p3 < 0 || p3 >= nc || !chan_mask[p3]) continue;
            if (out_bl) {
                const double *a = D[0] + p1 * P, *b = D[1] + p2 * P, *e = D[2] + p3 * P;
                double s = 0, t = 0;
                for (long x = 0; x < P; ++x) {
                    const double v = a[x] * b[x] * e[x];
                    s += v;
                    t += fabs(v);
                }
                sb += s;
                sa += t;
            }
            if (out_norm) {
                const double *a = S[0] + p1 * P, *b = S[1] + p2 * P, *e = S[2] + p3 * P;
                double s = 0;
                for (long x = 0; x < P; ++x) s += a[x] * b[x] * e[x];
                sn += s;
            }
            ++n;
        }
    done:
        if (out_bl) out_bl[c] = sb;
        if (out_abs) out_abs[c] = sa;
        if (out_norm) out_norm[c] = sn;
        if (out_cnt) out_cnt[c] = n;
    }
    for (int k = 0; k < 3; ++k) {
        int dup = 0;
        for (int q = 0; q < k; ++q) if (rr[q] == rr[k]) dup = 1;
        if (!dup) { free(D[k]); free(S[k]); }
    }
    return 0;
}

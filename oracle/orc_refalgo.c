/*
 * Reference-algorithm CPU path -- TEST / BASELINE INFRASTRUCTURE ONLY.
 * Used by bench.py's cpu_baseline leg (timed on the host cores) and by tests
 * as the reference's own FP32 numerics; never by the product.
 *
 * Restates the loop structure of amber.c:187-304 (v1) and amber_v2.c:187-309
 * (v2): per ring an FP64 c2r for every plane cast to float (fill_s2d_float,
 * amber_funcs.c:967-1010; fill_s2I :1109-1131), the 6-int triplet table
 * (precompute_unflg_vcp :2207-2259 / precompute_vcp_dnu :2264-2310), and the
 * FP32 sequential triple loop under `omp for schedule(dynamic)` with private
 * accumulators and a critical merge (amber.c:233-283, amber_v2.c:229-288).
 * Built with -O3 -ffast-math like the reference's own 4_fbe/makefile:6-7.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#include <omp.h>

long orc_refalgo(long N, long nc, double Du, const double *re_im, const double *samp,
                 const double *ulow, const double *uhigh, int r1, int r2, int r3,
                 const uint8_t *chan_mask, int nc10, int p1_lo, int p1_hi, int v2,
                 float *bl, float *numf, int32_t *cnt, int nthreads,
                 double *seconds_fill, double *seconds_loop)
{
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    long Ng = orc_ng(uhigh[r1], Du), P = Ng * Ng;
    int W = 2 * nc10 + 1;
    long ncsize = (long)W * W;
    if (p1_lo < 0) p1_lo = 0;
    if (p1_hi > nc) p1_hi = (int)nc;

    /* planes touched by the p1 sample: p in [p1_lo-nc10, p1_hi+nc10) */
    long plo = p1_lo - nc10 < 0 ? 0 : p1_lo - nc10;
    long phi = p1_hi + nc10 > nc ? nc : p1_hi + nc10;
    int rr[3] = {r1, r2, r3};
    float *sd[3], *sI[3] = {0, 0, 0};
    double t0 = omp_get_wtime();
    for (int k = 0; k < 3; ++k) {
        sd[k] = calloc((size_t)P * nc, sizeof(float));
        if (v2) sI[k] = calloc((size_t)P * nc, sizeof(float));
#pragma omp parallel num_threads(nthreads)
        {
            double *D = malloc(sizeof(double) * P);
#pragma omp for schedule(dynamic)
            for (long p = plo; p < phi; ++p) {
                orc_ring_c2r(N, nc, Ng, ulow[rr[k]], uhigh[rr[k]], re_im, NULL, p, D);
                for (long x = 0; x < P; ++x) sd[k][p * P + x] = (float)D[x];
                if (v2) {
                    orc_ring_c2r(N, nc, Ng, ulow[rr[k]], uhigh[rr[k]], NULL, samp, p, D);
                    for (long x = 0; x < P; ++x) sI[k][p * P + x] = (float)D[x];
                }
            }
            free(D);
        }
    }
    double t1 = omp_get_wtime();

    /* triplet table: p1 order, then p2, then p3 (amber_funcs.c:2236-2256) */
    long count = 0;
    for (long p1 = p1_lo; p1 < p1_hi; ++p1) {
        if (!chan_mask[p1]) continue;
        for (long p2 = 0; p2 < nc; ++p2) {
            if (!chan_mask[p2] || labs(p2 - p1) > nc10) continue;
            for (long p3 = 0; p3 < nc; ++p3)
                if (chan_mask[p3] && labs(p3 - p1) <= nc10) ++count;
        }
    }
    int *vcp = malloc(sizeof(int) * 6 * (size_t)(count ? count : 1));
    long ix = 0;
    for (long p1 = p1_lo; p1 < p1_hi; ++p1) {
        if (!chan_mask[p1]) continue;
        for (long p2 = 0; p2 < nc; ++p2) {
            if (!chan_mask[p2] || labs(p2 - p1) > nc10) continue;
            int dnu2 = (int)(p2 - p1 + nc10);
            for (long p3 = 0; p3 < nc; ++p3) {
                if (!chan_mask[p3] || labs(p3 - p1) > nc10) continue;
                int dnu3 = (int)(p3 - p1 + nc10);
                vcp[ix++] = (int)p1; vcp[ix++] = (int)p2; vcp[ix++] = (int)p3;
                vcp[ix++] = dnu2; vcp[ix++] = dnu3; vcp[ix++] = dnu2 * W + dnu3;
            }
        }
    }
    memset(bl, 0, sizeof(float) * ncsize);
    if (numf) memset(numf, 0, sizeof(float) * ncsize);
    if (cnt) memset(cnt, 0, sizeof(int32_t) * ncsize);

    double t2 = omp_get_wtime();
#pragma omp parallel num_threads(nthreads)
    {
        float *pb = calloc(ncsize, sizeof(float));
        float *pn = v2 ? calloc(ncsize, sizeof(float)) : NULL;
        int32_t *pc = calloc(ncsize, sizeof(int32_t));
#pragma omp for schedule(dynamic) nowait
        for (long ipc = 0; ipc < count; ++ipc) {
            const int *t = vcp + 6 * ipc;
            const float *restrict a = sd[0] + (long)t[0] * P;
            const float *restrict b = sd[1] + (long)t[1] * P;
            const float *restrict c = sd[2] + (long)t[2] * P;
            float bis = 0.f;
            if (v2) {
                const float *restrict ia = sI[0] + (long)t[0] * P;
                const float *restrict ib = sI[1] + (long)t[1] * P;
                const float *restrict ic = sI[2] + (long)t[2] * P;
                float nt = 0.f;
                for (long x = 0; x < P; ++x) {
                    bis += a[x] * b[x] * c[x];
                    nt += ia[x] * ib[x] * ic[x];
                }
                pn[t[5]] += nt;
            } else {
                for (long x = 0; x < P; ++x) bis += a[x] * b[x] * c[x];
            }
            pb[t[5]] += bis;
            pc[t[5]] += 1;
        }
#pragma omp critical
        {
            for (long i = 0; i < ncsize; ++i) {
                bl[i] += pb[i];
                if (numf && pn) numf[i] += pn[i];
                if (cnt) cnt[i] += pc[i];
            }
        }
        free(pb); free(pn); free(pc);
    }
    double t3 = omp_get_wtime();
    if (seconds_fill) *seconds_fill = t1 - t0;
    if (seconds_loop) *seconds_loop = t3 - t2;
    for (int k = 0; k < 3; ++k) { free(sd[k]); free(sI[k]); }
    free(vcp);
    return count;
}

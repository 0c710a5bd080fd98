/*
 * FP64 CPU oracle for the MAPS estimator C_l(dnu).  TEST INFRASTRUCTURE ONLY
 * (the checker of amb_maps_cldnu; the product never links it).
 *
 * Restates 4_fbe/maps.c:100-125 (normalisation, identical to normalize_data),
 * maps.c:151-229 (the pair loop nn1 <= nn2, cldnu[b][nn2-nn1] += binps[b],
 * numdnu[nn2-nn1] += 1, binu/num of the last call) and cl_dnu
 * (4_fbe/amber_funcs.c:1418-1488): per pair, for every half-plane cell with
 * umin <= uu < umax, uu = sqrt(ia^2 + j^2) Du (ia = i or N - i),
 *   bin = floor(log10(uu/umin)/binsiz), binsiz = log10(umax/umin)/NB,
 *   pksum[bin] += (Re1 Re2 + Im1 Im2) * NTiG1, no[bin] += NTiG1 (int),
 *   usum[bin] += uu * NTiG1,
 * then pksum and usum divided by no where no != 0.  Sums keep the written
 * order (cells ascending, pairs nn1 ascending); compiled without fast-math.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

int orc_maps_bin(long N, double Du, long i, long j, int NB, double umin, double umax, double *uu_out)
{
    const long ia = (i > N / 2) ? (N - i) : i;
    const double uu = sqrt(1. * (ia * ia + j * j)) * Du;
    const double binsiz = log10(umax / umin) / (1. * NB);
    if (uu_out) *uu_out = uu;
    if (!(uu >= umin && uu < umax)) return -1;
    return (int)floor(log10(uu / umin) / binsiz);
}

int orc_cldnu(long N, long nc, double Du, const double *re_im, const double *samp, int NB, double umin,
              double umax, double *cldnu, double *numdnu, double *binu, int32_t *num)
{
    const long Nh = N / 2 + 1, npix = N * Nh;
    double *v = malloc(sizeof(double) * npix * nc * 2), *s = malloc(sizeof(double) * npix * nc);
    int *bin = malloc(sizeof(int) * npix);
    double *uu = malloc(sizeof(double) * npix), *pk = malloc(sizeof(double) * NB), *us = malloc(sizeof(double) * NB);
    int *no = malloc(sizeof(int) * NB);
    if (!v || !s || !bin || !uu || !pk || !us || !no) {
        free(v); free(s); free(bin); free(uu); free(pk); free(us); free(no);
        return -1;
    }
    memcpy(v, re_im, sizeof(double) * npix * nc * 2);
    memcpy(s, samp, sizeof(double) * npix * nc);
    for (long x = 0; x < npix * nc; ++x)
        if (s[x] > 0) {
            v[2 * x] /= s[x];
            v[2 * x + 1] /= s[x];
        }
    for (long x = 0; x < npix * nc; ++x)
        if (s[x] > 0) s[x] = 1;
    for (long i = 0; i < N; ++i)
        for (long j = 0; j < Nh; ++j) {
            int b = orc_maps_bin(N, Du, i, j, NB, umin, umax, &uu[i * Nh + j]);
            bin[i * Nh + j] = b >= NB ? NB - 1 : b;
        }
    memset(cldnu, 0, sizeof(double) * NB * nc);
    memset(numdnu, 0, sizeof(double) * nc);
    for (long n1 = 0; n1 < nc; ++n1)
        for (long n2 = n1; n2 < nc; ++n2) {
            for (int b = 0; b < NB; ++b) { pk[b] = 0; us[b] = 0; no[b] = 0; }
            for (long x = 0; x < npix; ++x) {
                const int b = bin[x];
                if (b < 0) continue;
                const double *a = v + 2 * (x * nc + n1), *c = v + 2 * (x * nc + n2);
                const double s1 = s[x * nc + n1];
                double numer = a[0] * c[0] + a[1] * c[1];
                numer *= s1;
                no[b] += s1;
                pk[b] += numer;
                us[b] += uu[x] * s1;
            }
            for (int b = 0; b < NB; ++b)
                if (no[b] != 0) {
                    us[b] /= (1. * no[b]);
                    pk[b] /= (1. * no[b]);
                }
            numdnu[n2 - n1] += 1.;
            for (int b = 0; b < NB; ++b) cldnu[b * nc + (n2 - n1)] += pk[b];
        }
    /* binu and num hold the last call's values (nn1 = nn2 = nc - 1) */
    for (int b = 0; b < NB; ++b) {
        binu[b] = us[b];
        num[b] = no[b];
    }
    free(v); free(s); free(bin); free(uu); free(pk); free(us); free(no);
    return 0;
}

/*
 * CPU oracle for the upstream gridder vcg.  TEST INFRASTRUCTURE ONLY (the
 * checker of amb_grid; the product never links it).
 *
 * Restates 3_vcg/tgefuncs.c: initialize_gridding (:290-376: dU from the
 * primary-beam FWHM, NN = 2 ceil(Umax/dU)), group_loop_gridding (:377-509,
 * flag_key 1: a visibility enters only when its weight is > 0 and then
 * counts 1 in NTiG) and group_loop_gridding_without_flag (:511-626, flag_key
 * 0: every visibility enters and counts 1), and collapse_channels
 * (:816-839).  Per group, in group order:
 *   signv = (v < 0) ? -1 : 1;  u1 = nu0 signv u;  v1 = nu0 signv v   (float)
 *   Uval = sqrt(u1 u1 + v1 v1) (float sum, double sqrt); skip if > Umax
 *   ii = (int) round(u1/dU), ii += NN if < 0;  jj = (int) round(v1/dU)
 *   GV[ii (NN/2+1) + jj][n] += (re, signv im) (float values into double sums)
 * Only the RR stokes (index 0) reaches the output files (WriteGVnAC_gridding
 * :628-691, Write_NTiG :728-748); the LL sums are never written and are not
 * formed here.  Compiled without fast-math: float and double operations in
 * the written order, as the reference's objects (built with plain -c).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_grid_setup(double Umax, double FWHM, double *dU, long *NN)
{
    const double theta_0 = 0.6 * FWHM * M_PI / (180. * 60.);
    *dU = sqrt(log(2)) / (M_PI * theta_0);
    const int Ng = (int)ceil(Umax / *dU);
    *NN = (long)2 * Ng;
}

long orc_grid(long NN, long nc, long nchan, long nstokes, int chan1, int npar, int flag_key, int n_collapse,
              float nu_chan0, double dU, double Umax, long ngroup, const float *ranpar, const float *data,
              double *gv, double *samp)
{
    const long Nh = NN / 2 + 1, ngrid = NN * Nh;
    const long ncout = nc / n_collapse;
    double *g = calloc((size_t)ngrid * nc * 2, sizeof(double));
    double *s = calloc((size_t)ngrid * nc, sizeof(double));
    if (!g || !s) {
        free(g);
        free(s);
        return -1;
    }
    long count = 0;
    for (long grp = 0; grp < ngroup; ++grp) {
        const float *rp = ranpar + grp * npar;
        const float signv = (rp[1] < 0.) ? -1. : 1.;
        const float u1 = nu_chan0 * signv * rp[0];
        const float v1 = nu_chan0 * signv * rp[1];
        const double Uval = sqrt(u1 * u1 + v1 * v1);
        if (!(Uval <= Umax)) continue;
        int ii = (int)round(u1 / dU);
        ii = (ii < 0) ? NN + ii : ii;
        const int jj = (int)round(v1 / dU);
        if (ii < 0 || ii >= NN || jj < 0 || jj >= Nh) {
            free(g);
            free(s);
            return -2;
        }
        const long index = (long)ii * Nh + jj;
        const float *v = data + (size_t)grp * nchan * nstokes * 3 + (size_t)(chan1 - 1) * nstokes * 3;
        for (long nn = 0; nn < nc; ++nn) {
            const float *RR = v + nn * nstokes * 3;
            float re, im, ng;
            if (flag_key == 1) {
                if (RR[2] > 0.) {
                    re = RR[0];
                    im = signv * RR[1];
                    ng = 1.;
                } else {
                    re = 0.;
                    im = 0.;
                    ng = 0.;
                }
            } else {
                re = RR[0];
                im = signv * RR[1];
                ng = 1.;
            }
            g[2 * (index * nc + nn)] += re;
            g[2 * (index * nc + nn) + 1] += im;
            s[index * nc + nn] += ng;
        }
        ++count;
    }
    if (n_collapse <= 1) {
        memcpy(gv, g, sizeof(double) * ngrid * nc * 2);
        memcpy(samp, s, sizeof(double) * ngrid * nc);
    } else {
        for (long idx = 0; idx < ngrid; ++idx)
            for (long col = 0; col < ncout; ++col) {
                double sumre = 0.0, sumim = 0.0, sumNg = 0.0;
                for (long ki = 0; ki < n_collapse; ++ki) {
                    sumre += g[2 * (idx * nc + col * n_collapse + ki)];
                    sumim += g[2 * (idx * nc + col * n_collapse + ki) + 1];
                    sumNg += s[idx * nc + col * n_collapse + ki];
                }
                gv[2 * (idx * ncout + col)] = sumre;
                gv[2 * (idx * ncout + col) + 1] = sumim;
                samp[idx * ncout + col] = sumNg;
            }
    }
    free(g);
    free(s);
    return count;
}

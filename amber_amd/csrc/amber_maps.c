/*
 * Drop-in driver `amber_maps` for the MAPS program (4_fbe/maps.c:44-229) on
 * top of the amber_mi355x C ABI.  Same working-directory contract: run from a
 * directory holding input_params (read_inputs, amber_funcs.c:159-172) and
 * results/, read <data_path>/GV_data_<Rlzn> and <data_path>/sampling, write
 * results/cldnu in the reference's text layout.
 *
 * The reference's second argument (a GV file name) is accepted and unused, as
 * in maps.c, where the cube path comes from data_path and Rlzn.  The cube's
 * own header gives N, nc and Du (initialize, amber_funcs.c:1736-1739,
 * overrides both input_params and maps.c's hard-coded nc = 768).  The bin
 * setup NB = 15, umin = 1, umax = 500 (maps.c:161-163) is the default and
 * can be changed with --nb, --umin and --umax.
 */
#include "amber_mi355x.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void usage(void)
{
    fprintf(stderr, "Usage: amber_maps <Rlzn> [GV] [--params FILE] [--nb N] [--umin U] [--umax U]\n"
                    "          [--out FILE] [--device N]\n");
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        usage();
        return 2;
    }
    const int rlzn = atoi(argv[1]);
    const char *params = "input_params", *out = "results/cldnu";
    int NB = 15, device = 0;
    double umin = 1., umax = 500.;
    for (int a = 2; a < argc; ++a) {
        const char *k = argv[a];
        const char *v = a + 1 < argc ? argv[a + 1] : NULL;
        if (k[0] != '-' && a == 2) continue; /* the reference's GV argument */
        if (!v) { usage(); return 2; }
        if (!strcmp(k, "--params")) params = v;
        else if (!strcmp(k, "--nb")) NB = atoi(v);
        else if (!strcmp(k, "--umin")) umin = atof(v);
        else if (!strcmp(k, "--umax")) umax = atof(v);
        else if (!strcmp(k, "--out")) out = v;
        else if (!strcmp(k, "--device")) device = atoi(v);
        else { usage(); return 2; }
        ++a;
    }
    const double T = now();
    int64_t NNp, ncp;
    double Dup, nu0, dnu, chan0;
    char data_path[512];
    int rc = amb_read_params(params, &NNp, &Dup, &nu0, &dnu, &chan0, &ncp, data_path, sizeof data_path);
    if (rc) {
        fprintf(stderr, "amber_maps: %s: %s\n", params, amb_strerror(rc));
        return 1;
    }
    int64_t N, nc;
    double Du;
    int32_t fk;
    double *gv = NULL, *samp = NULL;
    rc = amb_read_cube_files(data_path, rlzn, &N, &nc, &Du, &fk, &gv, &samp);
    if (rc) {
        fprintf(stderr, "amber_maps: %s/GV_data_%d: %s\n", data_path, rlzn, amb_strerror(rc));
        return 1;
    }
    printf("Read gridded visibilities: %lld x %lld x %lld, Du = %f\n", (long long)N, (long long)(N / 2 + 1),
           (long long)nc, Du);
    amb_opts o = {device, AMB_PREC_FP32, 0, 0};
    amb_ctx *ctx = NULL;
    if ((rc = amb_create(&o, &ctx))) {
        fprintf(stderr, "amber_maps: device: %s\n", amb_strerror(rc));
        return 1;
    }
    double *cl = malloc(sizeof(double) * NB * nc), *nd = malloc(sizeof(double) * nc);
    double *bu = malloc(sizeof(double) * NB);
    int32_t *nm = malloc(sizeof(int32_t) * NB);
    if (!cl || !nd || !bu || !nm) {
        fprintf(stderr, "amber_maps: out of host memory\n");
        return 1;
    }
    printf("Computing the MAPS. The result will be in %s.\n", out);
    rc = amb_maps_cldnu(ctx, N, nc, Du, gv, samp, NB, umin, umax, cl, nd, bu, nm);
    if (rc) {
        fprintf(stderr, "amber_maps: %s\n", amb_last_error(ctx));
        return 1;
    }
    if ((rc = amb_write_cldnu(out, NB, nc, bu, cl, nd, nm))) {
        fprintf(stderr, "amber_maps: %s: %s\n", out, amb_strerror(rc));
        return 1;
    }
    amb_stats st;
    amb_get_stats(ctx, &st);
    printf("MAPS device time %.3f ms (pair kernel %.3f ms), total %.3f s\n", st.ms_maps_total, st.ms_maps,
           now() - T);
    free(cl); free(nd); free(bu); free(nm);
    amb_free(gv);
    amb_free(samp);
    amb_destroy(ctx);
    return 0;
}

/*
 * Drop-in drivers `amber` (v1) and `amber_v2` (built with -DAMB_V2) on top of
 * the amber_mi355x C ABI.  Same working-directory contract as the reference
 * (4_fbe/amber.c:21-335, 4_fbe/amber_v2.c:21-345): run from a directory
 * holding input_params_ess and results/v1 (results/v2), read
 * <data_path>/GV_data_<Rlzn>, <data_path>/sampling and, for flagged data,
 * <data_path>/uf_chans.txt, and write the same result files byte for byte in
 * layout.
 *
 * The reference hard-codes Nr=22, nc10=383, uf_nc=552 and the first
 * triangle (amber.c:43-49,187; amber_v2.c:187).  They are options here with
 * those defaults, except the first triangle, which defaults to 0 (a full
 * run); pass --rp-start 365 (v1) or 635 (v2) to reproduce the reference's
 * partial runs.  --rings FILE replaces the built-in piecewise edges.
 * --wire hardened (or both) also writes each triangle as a self-describing
 * bsw1_/bsw2_ file (versioned header with nc10, Ng, precision, triangle id
 * and a CRC-32 of the payload); the legacy files stay the default.
 */
#include "amber_mi355x.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef AMB_V2
#define PROG "amber_v2"
#define RES_PATH "results/v2"
#else
#define PROG "amber"
#define RES_PATH "results/v1"
#endif

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void usage(void)
{
    fprintf(stderr,
            "Usage: %s <Rlzn> [--params FILE] [--rings FILE] [--nr N] [--nc10 N] [--uf-nc N]\n"
            "          [--rp-start N] [--rp-end N] [--precision fp32|bf16x3|fp16x3|lagft] [--device N]\n"
            "          [--calcr-plane N] [--res-path DIR] [--wire legacy|hardened|both]\n"
            "          [--cert exact|strict] [--tune FILE]\n",
            PROG);
}

static int read_rings_file(const char *path, double **lo, double **hi, int *nr)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return AMB_EIO;
    int cap = 64, n = 0;
    double *l = malloc(sizeof(double) * cap), *h = malloc(sizeof(double) * cap);
    double a, b;
    while (fscanf(fp, "%lf %lf", &a, &b) == 2) {
        if (n == cap) {
            cap *= 2;
            l = realloc(l, sizeof(double) * cap);
            h = realloc(h, sizeof(double) * cap);
        }
        l[n] = a; h[n] = b; ++n;
    }
    fclose(fp);
    if (n == 0) { free(l); free(h); return AMB_EFORMAT; }
    *lo = l; *hi = h; *nr = n;
    return AMB_OK;
}

#define CHECK(call)                                                                        \
    do {                                                                                   \
        int rc_ = (call);                                                                  \
        if (rc_ != AMB_OK) {                                                               \
            fprintf(stderr, "%s: %s: %s (%s)\n", PROG, #call, amb_strerror(rc_),            \
                    ctx ? amb_last_error(ctx) : "");                                       \
            exit(EXIT_FAILURE);                                                            \
        }                                                                                  \
    } while (0)

int main(int argc, char **argv)
{
    amb_ctx *ctx = NULL;
    if (argc < 2) { usage(); return 1; }
    int Rlzn = atoi(argv[1]);
    const char *params = "input_params_ess", *rings_file = NULL, *res_path = RES_PATH;
    int nr = 22, nc10 = 383, uf_nc = 552, rp_start = 0, rp_end = -1, device = 0, calcr_plane = 5;
    int prec = AMB_PREC_LAGFT; /* the fastest and most accurate form */
    int wire_legacy = 1, wire_hardened = 0; /* --wire: which result files to write */
    int cert_strict = 0; /* --cert: uncertified fast v2 normalisation sums recomputed exactly, or an error */
    const char *tune_file = NULL; /* --tune: launch-configuration table, loaded, or tuned and saved */
    for (int i = 2; i < argc; ++i) {
        const char *a = argv[i];
        const char *v = i + 1 < argc ? argv[i + 1] : NULL;
        if (!v) { usage(); return 1; }
        if (!strcmp(a, "--params")) params = v;
        else if (!strcmp(a, "--rings")) rings_file = v;
        else if (!strcmp(a, "--nr")) nr = atoi(v);
        else if (!strcmp(a, "--nc10")) nc10 = atoi(v);
        else if (!strcmp(a, "--uf-nc")) uf_nc = atoi(v);
        else if (!strcmp(a, "--rp-start")) rp_start = atoi(v);
        else if (!strcmp(a, "--rp-end")) rp_end = atoi(v);
        else if (!strcmp(a, "--device")) device = atoi(v);
        else if (!strcmp(a, "--calcr-plane")) calcr_plane = atoi(v);
        else if (!strcmp(a, "--res-path")) res_path = v;
        else if (!strcmp(a, "--precision")) {
            if (!strcmp(v, "fp32")) prec = AMB_PREC_FP32;
            else if (!strcmp(v, "bf16x3")) prec = AMB_PREC_BF16X3;
            else if (!strcmp(v, "fp16x3")) prec = AMB_PREC_FP16X3;
            else if (!strcmp(v, "lagft")) prec = AMB_PREC_LAGFT;
            else { usage(); return 1; }
        } else if (!strcmp(a, "--wire")) {
            if (!strcmp(v, "legacy")) { wire_legacy = 1; wire_hardened = 0; }
            else if (!strcmp(v, "hardened")) { wire_legacy = 0; wire_hardened = 1; }
            else if (!strcmp(v, "both")) { wire_legacy = 1; wire_hardened = 1; }
            else { usage(); return 1; }
        } else if (!strcmp(a, "--tune")) {
            tune_file = v;
        } else if (!strcmp(a, "--cert")) {
            if (!strcmp(v, "exact") || !strcmp(v, "count")) cert_strict = 0;
            else if (!strcmp(v, "strict")) cert_strict = 1;
            else { usage(); return 1; }
        } else { usage(); return 1; }
        ++i;
    }
    double T = now();
    char data_path[512];
    double nu0, dnu, chan0;
    int rc = amb_read_params_ess(params, &nu0, &dnu, &chan0, data_path, sizeof data_path);
    if (rc) { fprintf(stderr, "%s: reading %s: %s\n", PROG, params, amb_strerror(rc)); return 1; }
    printf("\nInput parameters:\n\nnu0=%.2lf\ndeltanu=%.2lf\nchan0=%.2lf\nData Path=%s\n\n", nu0, dnu, chan0, data_path);

    amb_opts o = {device, prec, 0, 1};
    rc = amb_create(&o, &ctx);
    if (rc) { fprintf(stderr, "%s: amb_create: %s\n", PROG, amb_strerror(rc)); return 1; }
    CHECK(amb_set_cert_strict(ctx, cert_strict));
    int cube;
    CHECK(amb_load_cube_files(ctx, data_path, Rlzn, &cube));
    char gvp[1024];
    int64_t NN, nc;
    double Du;
    int32_t flag_key;
    snprintf(gvp, sizeof gvp, "%s/GV_data_%d", data_path, Rlzn);
    CHECK(amb_read_gv_header(gvp, &NN, &nc, &Du, &flag_key));
    printf("NN=%lld, nc=%lld, Du=%lf, Flag=%d, nc10=%d, Unflagged Channles=%d\n", (long long)NN,
           (long long)nc, Du, flag_key, nc10, uf_nc);
    printf("result will be saved in %s\n", res_path);

    double *lo, *hi;
    if (rings_file) {
        rc = read_rings_file(rings_file, &lo, &hi, &nr);
        if (rc) { fprintf(stderr, "%s: reading %s: %s\n", PROG, rings_file, amb_strerror(rc)); return 1; }
    } else {
        lo = malloc(sizeof(double) * nr);
        hi = malloc(sizeof(double) * nr);
        amb_annular_rings(nr, lo, hi);
    }
    double *uavg = malloc(sizeof(double) * nr);
    amb_uavg(lo, hi, nr, uavg);
    printf("\t%s\t\t%s\t\t\t%s\n", "ULow", "UHigh", "Uavg");
    for (int i = 0; i < nr; ++i) printf("\t%f\t%f\t%f\n", lo[i], hi[i], uavg[i]);
    CHECK(amb_set_rings(ctx, lo, hi, nr));

    int *chans = NULL;
    if (flag_key == 1) {
        char ufp[1024];
        snprintf(ufp, sizeof ufp, "%s/uf_chans.txt", data_path);
        chans = malloc(sizeof(int) * (uf_nc > 0 ? uf_nc : 1));
        rc = amb_read_uf_chans(ufp, uf_nc, nc, chans);
        if (rc) { fprintf(stderr, "%s: reading %s (%d ids): %s\n", PROG, ufp, uf_nc, amb_strerror(rc)); return 1; }
    }
    CHECK(amb_set_channels(ctx, cube, chans, uf_nc, nc10));

#ifndef AMB_V2
    /* calCR writes closed_rings/rcount (amber.c:127) */
    int ntri_max = amb_triangles(uavg, nr, NULL, 0);
    double *rd0 = malloc(sizeof(double) * 4 * (ntri_max ? ntri_max : 1));
    int rcount0 = 0;
    CHECK(amb_closed_rings(ctx, cube, calcr_plane, rd0, ntri_max, &rcount0));
    CHECK(amb_write_closed_rings(res_path, rd0, rcount0));
    printf("\nTotal combinations of rings on a plane forming closed triangle = %d\n", rcount0);
    free(rd0);
#endif
    int rcount;
    double *rdata;
    rc = amb_read_rings(res_path, &rcount, &rdata);
    if (rc) { fprintf(stderr, "%s: reading %s/{rcount,closed_rings}: %s\n", PROG, res_path, amb_strerror(rc)); return 1; }
    printf("Successfully read %d ring configurations.\n", rcount);

    const int W = 2 * nc10 + 1;
    const size_t ncsize = (size_t)W * W;
    /* the rp loop (amber.c:187-304) as batches of triangles: each
       amb_bispec_batch call runs grouped waves; blocks keep the host
       buffers near 1 GiB */
    int blk = (int)((1u << 30) / (ncsize * 8));
    if (blk < 1) blk = 1;
    if (blk > 64) blk = 64;
    float *bl = malloc(sizeof(float) * ncsize * blk);
#ifndef AMB_V2
    int32_t *cnt = malloc(sizeof(int32_t) * ncsize * blk);
#else
    float *norm = malloc(sizeof(float) * ncsize * blk);
#endif
    int *tri = malloc(sizeof(int) * 3 * blk);
    if (rp_end < 0 || rp_end > rcount) rp_end = rcount;
    if (tune_file && prec == AMB_PREC_LAGFT) {
        printf("--tune ignored: the lag-Fourier form has no launch-configuration table\n");
        tune_file = NULL;
    }
    if (tune_file) {
        FILE *tf = fopen(tune_file, "r");
        if (tf) {
            fclose(tf);
            CHECK(amb_tune_load(ctx, tune_file));
            printf("Launch configurations loaded from %s\n", tune_file);
        } else if (rp_end > rp_start) {
            /* tuned on up to 16 of this run's triangles, evenly spaced */
            const int ns = rp_end - rp_start < 16 ? rp_end - rp_start : 16;
            amb_tri_req *rq = calloc(ns, sizeof *rq);
            float *sb = malloc(sizeof(float) * ncsize * ns);
            if (!rq || !sb) { fprintf(stderr, "%s: out of host memory\n", PROG); return 1; }
            for (int k = 0; k < ns; ++k) {
                const int rp = rp_start + (int)((long)k * (rp_end - rp_start) / ns);
                rq[k].cube_id = cube;
                rq[k].r1 = (int)rdata[rp * 4];
                rq[k].r2 = (int)rdata[rp * 4 + 1];
                rq[k].r3 = (int)rdata[rp * 4 + 2];
                rq[k].bl = sb + (size_t)k * ncsize;
            }
            const double t0 = now();
            CHECK(amb_autotune(ctx, rq, ns, 120.0, tune_file));
            printf("Launch configurations tuned on %d triangles in %.1f s, saved to %s\n", ns, now() - t0, tune_file);
            free(rq);
            free(sb);
        }
    }
    for (int rp0 = rp_start; rp0 < rp_end; rp0 += blk) {
        const int nb = rp_end - rp0 < blk ? rp_end - rp0 : blk;
        for (int k = 0; k < nb; ++k)
            for (int e = 0; e < 3; ++e) tri[3 * k + e] = (int)rdata[(rp0 + k) * 4 + e];
        double t0 = now();
#ifndef AMB_V2
        CHECK(amb_bispec_batch(ctx, cube, tri, nb, bl, cnt, NULL));
#else
        CHECK(amb_bispec_batch(ctx, cube, tri, nb, bl, NULL, norm));
#endif
        const double dt = now() - t0;
        for (int k = 0; k < nb; ++k) {
            const int rp = rp0 + k, r1 = tri[3 * k], r2 = tri[3 * k + 1], r3 = tri[3 * k + 2];
            const long Ng1 = amb_ng(hi[r1], Du), Ng2 = Ng1;
#ifndef AMB_V2
            /* amber.c:198-201,287 with the same double expression order */
            double gfact = (1. * Ng1) / (1. * NN);
            double gfact_sq = gfact * gfact;
            double Num = gfact_sq * rdata[rp * 4 + 3];
            double inv = 1.0 / (Ng1 * Ng2 * 1.);
            if (wire_legacy)
                CHECK(amb_write_bsout_v1(res_path, r1, r2, r3, Rlzn, W, Num * inv, Num, bl + ncsize * k,
                                         cnt + ncsize * k));
            amb_bsw_header h = {{0}, 0, 0, AMB_BSW_KIND_V1, 0, nc10, W, r1, r2, r3, Rlzn, Ng1, nc, prec, 0,
                                Num * inv, Num, 0, {0}};
            if (wire_hardened) CHECK(amb_write_bsw(res_path, &h, bl + ncsize * k, cnt + ncsize * k));
#else
            double inv = 1.0 / (Ng1 * Ng2 * 1.);
            if (wire_legacy)
                CHECK(amb_write_bsout_v2(res_path, r1, r2, r3, Rlzn, W, inv, bl + ncsize * k, norm + ncsize * k));
            amb_bsw_header h = {{0}, 0, 0, AMB_BSW_KIND_V2, 0, nc10, W, r1, r2, r3, Rlzn, Ng1, nc, prec, 0,
                                inv, 0.0, 0, {0}};
            if (wire_hardened) CHECK(amb_write_bsw(res_path, &h, bl + ncsize * k, norm + ncsize * k));
#endif
            printf("rp=%d (r1=%d r2=%d r3=%d, Ng=%ld)\n", rp, r1, r2, r3, Ng1);
        }
        printf("triangles %d..%d: %.3f s\n", rp0, rp0 + nb - 1, dt);
    }
    printf("Total time taken = %.3f s\n", now() - T);
#ifdef AMB_V2
    amb_stats st;
    if (amb_get_stats(ctx, &st) == AMB_OK && st.n_norm_exact > 0)
        printf("normalisation of %lld triangle(s) recomputed in FP64 (fast sums %g from the 1/8 grid of Ng^2)\n",
               (long long)st.n_norm_exact, st.cert_margin_max);
#endif
    amb_destroy(ctx);
    free(bl); free(tri); free(rdata); free(lo); free(hi); free(uavg); free(chans);
#ifndef AMB_V2
    free(cnt);
#else
    free(norm);
#endif
    return 0;
}

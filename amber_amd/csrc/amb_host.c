/*
 * Host-side pieces of amber_mi355x: file formats, integer/FP64 rules and the
 * deterministic synthetic-cube generator.  Plain C11 + OpenMP.
 * Byte layouts follow the reference's writers and readers (cited per
 * function); every error is returned, never exit().
 */
#include "amber_mi355x.h"
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void amb_free(void *p) { free(p); }

const char *amb_strerror(int code)
{
    switch (code) {
    case AMB_OK: return "ok";
    case AMB_EARG: return "invalid argument";
    case AMB_EIO: return "file missing or unreadable";
    case AMB_EFORMAT: return "truncated or inconsistent file";
    case AMB_EHIP: return "HIP runtime error";
    case AMB_ENOMEM: return "host allocation failed";
    case AMB_EBUDGET: return "request exceeds the HBM budget";
    case AMB_ESTATE: return "context state not set (rings, channels or cube)";
    case AMB_ECERT: return "integer certificate failed";
    case AMB_EBUSY: return "queue full";
    default: return "unknown error";
    }
}

/* ---------------------------------------------------------------- params */
int amb_read_params_ess(const char *path, double *nu0, double *dnu, double *chan0,
                        char *data_path, size_t cap)
{   /* amber_funcs.c:175-188 */
    if (!path || !data_path || cap < 2) return AMB_EARG;
    FILE *fp = fopen(path, "r");
    if (!fp) return AMB_EIO;
    char buf[512];
    double a, b, c;
    int n = fscanf(fp, "%lf%lf%lf%511s", &a, &b, &c, buf);
    fclose(fp);
    if (n != 4) return AMB_EFORMAT;
    if (nu0) *nu0 = a;
    if (dnu) *dnu = b;
    if (chan0) *chan0 = c;
    if (strlen(buf) + 1 > cap) return AMB_EARG;
    strcpy(data_path, buf);
    return AMB_OK;
}

/* ---------------------------------------------------------------- cube files */
static int read_header(FILE *fp, int64_t *NN, int64_t *nc, double *Du, int32_t *fk)
{   /* amber_funcs.c:1736-1739 / 1877-1880: four separate freads, 28 bytes */
    if (fread(NN, sizeof(int64_t), 1, fp) != 1) return AMB_EFORMAT;
    if (fread(nc, sizeof(int64_t), 1, fp) != 1) return AMB_EFORMAT;
    if (fread(Du, sizeof(double), 1, fp) != 1) return AMB_EFORMAT;
    if (fread(fk, sizeof(int32_t), 1, fp) != 1) return AMB_EFORMAT;
    if (*NN <= 0 || *nc <= 0 || *NN > (1 << 16) || *nc > (1 << 16) || !(*Du > 0)) return AMB_EFORMAT;
    return AMB_OK;
}

int amb_read_gv_header(const char *path, int64_t *NN, int64_t *nc, double *Du, int32_t *flag_key)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return AMB_EIO;
    int rc = read_header(fp, NN, nc, Du, flag_key);
    fclose(fp);
    return rc;
}

int amb_read_cube_files(const char *data_path, int rlzn, int64_t *NN, int64_t *nc, double *Du,
                        int32_t *flag_key, double **gv, double **samp)
{   /* initialize + ReadGVnAC_gridding, amber_funcs.c:1720-1779, 1861-1911 */
    char f[1024];
    snprintf(f, sizeof f, "%s/GV_data_%d", data_path, rlzn);
    FILE *fp = fopen(f, "rb");
    if (!fp) return AMB_EIO;
    int rc = read_header(fp, NN, nc, Du, flag_key);
    if (rc) { fclose(fp); return rc; }
    size_t cells = (size_t)(*NN) * (size_t)(*NN / 2 + 1);
    size_t nv = cells * (size_t)(*nc);
    double *g = malloc(sizeof(double) * 2 * nv);
    double *s = malloc(sizeof(double) * nv);
    if (!g || !s) { free(g); free(s); fclose(fp); return AMB_ENOMEM; }
    size_t got = fread(g, sizeof(double), 2 * nv, fp);
    fclose(fp);
    if (got != 2 * nv) { free(g); free(s); return AMB_EFORMAT; }
    snprintf(f, sizeof f, "%s/sampling", data_path);
    fp = fopen(f, "rb");
    if (!fp) { free(g); free(s); return AMB_EIO; }
    got = fread(s, sizeof(double), nv, fp);
    fclose(fp);
    if (got != nv) { free(g); free(s); return AMB_EFORMAT; }
    *gv = g;
    *samp = s;
    return AMB_OK;
}

int amb_write_cube_files(const char *data_path, int rlzn, int64_t NN, int64_t nc, double Du,
                         int32_t flag_key, const double *gv, const double *samp)
{   /* WriteGVnAC_gridding 3_vcg/tgefuncs.c:639-655, Write_NTiG :728-748 */
    char f[1024];
    size_t nv = (size_t)NN * (size_t)(NN / 2 + 1) * (size_t)nc;
    snprintf(f, sizeof f, "%s/GV_data_%d", data_path, rlzn);
    FILE *fp = fopen(f, "wb");
    if (!fp) return AMB_EIO;
    int ok = fwrite(&NN, sizeof NN, 1, fp) == 1 && fwrite(&nc, sizeof nc, 1, fp) == 1 &&
             fwrite(&Du, sizeof Du, 1, fp) == 1 && fwrite(&flag_key, sizeof flag_key, 1, fp) == 1 &&
             fwrite(gv, sizeof(double), 2 * nv, fp) == 2 * nv;
    if (fclose(fp) != 0 || !ok) return AMB_EIO;
    snprintf(f, sizeof f, "%s/sampling", data_path);
    fp = fopen(f, "wb");
    if (!fp) return AMB_EIO;
    ok = fwrite(samp, sizeof(double), nv, fp) == nv;
    if (fclose(fp) != 0 || !ok) return AMB_EIO;
    return AMB_OK;
}

int amb_read_uf_chans(const char *path, int uf_nc, int64_t nc, int *chans)
{   /* amber_funcs.c:2485-2508; a short read or an out-of-range id is an
       error here (the reference prints and continues, docs/REVIEW.md:15). */
    if (uf_nc < 0 || !chans) return AMB_EARG;
    FILE *fp = fopen(path, "r");
    if (!fp) return AMB_EIO;
    int i = 0;
    while (i < uf_nc && fscanf(fp, "%d", &chans[i]) == 1) ++i;
    fclose(fp);
    if (i < uf_nc) return AMB_EFORMAT;
    for (int k = 0; k < uf_nc; ++k)
        if (chans[k] < 0 || chans[k] >= nc) return AMB_EFORMAT;
    return AMB_OK;
}

int amb_write_uf_chans(const char *path, int uf_nc, const int *chans)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return AMB_EIO;
    for (int i = 0; i < uf_nc; ++i) fprintf(fp, "%d\n", chans[i]);
    return fclose(fp) == 0 ? AMB_OK : AMB_EIO;
}

/* ---------------------------------------------------------------- rings files */
int amb_write_closed_rings(const char *res_path, const double *rdata, int rcount)
{   /* amber_funcs.c:406-411,459-463 (binary, 4 doubles per triangle) and
       :476-483 (text "%d") */
    char f[1024];
    snprintf(f, sizeof f, "%s/closed_rings", res_path);
    FILE *fp = fopen(f, "wb");
    if (!fp) return AMB_EIO;
    int ok = rcount == 0 || fwrite(rdata, sizeof(double), (size_t)4 * rcount, fp) == (size_t)4 * rcount;
    if (fclose(fp) != 0 || !ok) return AMB_EIO;
    snprintf(f, sizeof f, "%s/rcount", res_path);
    fp = fopen(f, "w");
    if (!fp) return AMB_EIO;
    fprintf(fp, "%d", rcount);
    return fclose(fp) == 0 ? AMB_OK : AMB_EIO;
}

int amb_read_rings(const char *res_path, int *rcount, double **rdata)
{   /* read_rings, amber_funcs.c:499-545 */
    char f[1024];
    snprintf(f, sizeof f, "%s/rcount", res_path);
    FILE *fp = fopen(f, "r");
    if (!fp) return AMB_EIO;
    int n;
    int got = fscanf(fp, "%d", &n);
    fclose(fp);
    if (got != 1 || n < 0) return AMB_EFORMAT;
    snprintf(f, sizeof f, "%s/closed_rings", res_path);
    fp = fopen(f, "rb");
    if (!fp) return AMB_EIO;
    double *d = malloc(sizeof(double) * 4 * (size_t)(n ? n : 1));
    if (!d) { fclose(fp); return AMB_ENOMEM; }
    size_t r = fread(d, sizeof(double), (size_t)4 * n, fp);
    fclose(fp);
    if (r != (size_t)4 * n) { free(d); return AMB_EFORMAT; }
    *rcount = n;
    *rdata = d;
    return AMB_OK;
}

/* ---------------------------------------------------------------- bsout files */
int amb_write_bsout_v1(const char *res_path, int r1, int r2, int r3, int rlzn, int W,
                       double ntri, double num, const float *bl, const int32_t *cnt)
{   /* amber.c:289-304 */
    char f[1024];
    size_t n = (size_t)W * W;
    snprintf(f, sizeof f, "%s/bsout_final_ra%d_rb%d_rc%d_R%d", res_path, r1, r2, r3, rlzn);
    FILE *fp = fopen(f, "wb");
    if (!fp) return AMB_EIO;
    int ok = fwrite(&ntri, sizeof(double), 1, fp) == 1 && fwrite(&num, sizeof(double), 1, fp) == 1 &&
             fwrite(bl, sizeof(float), n, fp) == n && fwrite(cnt, sizeof(int32_t), n, fp) == n;
    if (fclose(fp) != 0 || !ok) return AMB_EIO;
    return AMB_OK;
}

int amb_write_bsout_v2(const char *res_path, int r1, int r2, int r3, int rlzn, int W,
                       double inv_ng2, const float *bl, const float *norm)
{   /* amber_v2.c:295-309 */
    char f[1024];
    size_t n = (size_t)W * W;
    snprintf(f, sizeof f, "%s/bsout_final2_ra%d_rb%d_rc%d_R%d", res_path, r1, r2, r3, rlzn);
    FILE *fp = fopen(f, "wb");
    if (!fp) return AMB_EIO;
    int ok = fwrite(&inv_ng2, sizeof(double), 1, fp) == 1 && fwrite(bl, sizeof(float), n, fp) == n &&
             fwrite(norm, sizeof(float), n, fp) == n;
    if (fclose(fp) != 0 || !ok) return AMB_EIO;
    return AMB_OK;
}

/* ---------------------------------------------------------------- hardened result files */
_Static_assert(sizeof(amb_bsw_header) == 104, "amb_bsw_header layout");
static const char BSW_MAGIC[8] = {'A', 'M', 'B', 'R', 'B', 'S', 'W', 0};

/* CRC-32 of the concatenation a||b: reflected polynomial 0xEDB88320, init
 * and final xor ~0 (the IEEE 802.3 / zlib checksum).  The 1 KB table is
 * built per call, which keeps the function reentrant at negligible cost
 * next to the file I/O it guards. */
static uint32_t crc32_two(const void *a, size_t na, const void *b, size_t nb)
{
    uint32_t table[256];
    for (uint32_t i = 0; i < 256; ++i) {
        uint32_t c = i;
        for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        table[i] = c;
    }
    uint32_t c = 0xFFFFFFFFu;
    const unsigned char *p = (const unsigned char *)a;
    for (size_t i = 0; i < na; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    p = (const unsigned char *)b;
    for (size_t i = 0; i < nb; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

uint32_t amb_crc32(const void *p, size_t n) { return crc32_two(p, n, NULL, 0); }

int amb_write_bsw(const char *res_path, const amb_bsw_header *h, const float *bl, const void *den)
{
    if (!res_path || !h || !bl || !den) return AMB_EARG;
    if ((h->kind != AMB_BSW_KIND_V1 && h->kind != AMB_BSW_KIND_V2) || h->nc10 < 0 || h->W != 2 * h->nc10 + 1)
        return AMB_EARG;
    amb_bsw_header o = *h;
    memcpy(o.magic, BSW_MAGIC, 8);
    o.version = AMB_BSW_VERSION;
    o.header_bytes = (uint32_t)sizeof(amb_bsw_header);
    o.endian = 0x01020304u;
    const size_t n = (size_t)h->W * h->W;
    o.payload_bytes = 8 * (uint64_t)n;
    o.payload_crc = crc32_two(bl, 4 * n, den, 4 * n);
    memset(o.reserved, 0, sizeof o.reserved);
    char f[1024];
    snprintf(f, sizeof f, "%s/bsw%u_ra%d_rb%d_rc%d_R%d", res_path, o.kind, o.r1, o.r2, o.r3, o.rlzn);
    FILE *fp = fopen(f, "wb");
    if (!fp) return AMB_EIO;
    int ok = fwrite(&o, sizeof o, 1, fp) == 1 && fwrite(bl, 4, n, fp) == n && fwrite(den, 4, n, fp) == n;
    if (fclose(fp) != 0 || !ok) return AMB_EIO;
    return AMB_OK;
}

static long file_size(FILE *fp)
{
    if (fseek(fp, 0, SEEK_END) != 0) return -1;
    long s = ftell(fp);
    if (fseek(fp, 0, SEEK_SET) != 0) return -1;
    return s;
}

int amb_read_bsw(const char *path, amb_bsw_header *h, float *bl, void *den, int64_t cap)
{
    if (!path || !h) return AMB_EARG;
    FILE *fp = fopen(path, "rb");
    if (!fp) return AMB_EIO;
    const long size = file_size(fp);
    amb_bsw_header o;
    int rc = AMB_OK;
    if (size < (long)sizeof o || fread(&o, sizeof o, 1, fp) != 1) rc = AMB_EFORMAT;
    else if (memcmp(o.magic, BSW_MAGIC, 8) != 0 || o.version != AMB_BSW_VERSION ||
             o.header_bytes != sizeof o || o.endian != 0x01020304u ||
             (o.kind != AMB_BSW_KIND_V1 && o.kind != AMB_BSW_KIND_V2) || o.nc10 < 0 || o.W != 2 * o.nc10 + 1 ||
             o.payload_bytes != 8 * (uint64_t)o.W * o.W || (uint64_t)size != sizeof o + o.payload_bytes)
        rc = AMB_EFORMAT;
    if (rc) {
        fclose(fp);
        return rc;
    }
    /* the CRC covers the payload: read it whole (into the caller's arrays when given) */
    const size_t n = (size_t)o.W * o.W;
    const int want = bl && den && cap >= (int64_t)n;
    float *a = want ? bl : malloc(4 * n);
    void *b = want ? den : malloc(4 * n);
    if (!a || !b) rc = AMB_ENOMEM;
    else if (fread(a, 4, n, fp) != n || fread(b, 4, n, fp) != n) rc = AMB_EFORMAT;
    else if (crc32_two(a, 4 * n, b, 4 * n) != o.payload_crc) rc = AMB_EFORMAT;
    if (!want) {
        free(a);
        free(b);
    }
    fclose(fp);
    if (rc == AMB_OK || rc == AMB_EFORMAT) *h = o;
    return rc;
}

int amb_read_bsout(const char *path, int *kind, int *W, double *h0, double *h1, float *bl, void *den, int64_t cap)
{   /* legacy layouts of amber.c:289-304 and amber_v2.c:297-309 */
    if (!path || !kind || !W) return AMB_EARG;
    FILE *fp = fopen(path, "rb");
    if (!fp) return AMB_EIO;
    const long size = file_size(fp);
    int k = 0, w = 0;
    for (int kk = 1; kk <= 2 && !k; ++kk) {
        const long head = kk == 1 ? 16 : 8;
        if (size < head + 8 || (size - head) % 8) continue;
        const long sq = (size - head) / 8;
        long r = (long)sqrt((double)sq);
        while (r * r > sq) --r;
        while ((r + 1) * (r + 1) <= sq) ++r;
        if (r * r == sq && r % 2 == 1) {
            k = kk;
            w = (int)r;
        }
    }
    if (!k) {
        fclose(fp);
        return AMB_EFORMAT;
    }
    double a = 0, b = 0;
    int ok = fread(&a, 8, 1, fp) == 1 && (k == 2 || fread(&b, 8, 1, fp) == 1);
    const size_t n = (size_t)w * w;
    if (ok && bl && den && cap >= (int64_t)n) ok = fread(bl, 4, n, fp) == n && fread(den, 4, n, fp) == n;
    fclose(fp);
    if (!ok) return AMB_EFORMAT;
    *kind = k;
    *W = w;
    if (h0) *h0 = a;
    if (h1) *h1 = b;
    return AMB_OK;
}

/* ---------------------------------------------------------------- gridder (host part) */
int amb_read_vcg_inputs(const char *path, int *nc, double *Umax, double *FWHM, double *f)
{   /* read_inputs, 3_vcg/tgefuncs.c:28-37 */
    if (!path) return AMB_EARG;
    FILE *fp = fopen(path, "r");
    if (!fp) return AMB_EIO;
    int n_;
    double a, b, c;
    int n = fscanf(fp, "%d%lf%lf%lf", &n_, &a, &b, &c);
    fclose(fp);
    if (n != 4) return AMB_EFORMAT;
    if (nc) *nc = n_;
    if (Umax) *Umax = a;
    if (FWHM) *FWHM = b;
    if (f) *f = c;
    return AMB_OK;
}

int amb_grid_setup(double Umax, double FWHM, double *dU, int64_t *NN)
{   /* initialize_gridding, tgefuncs.c:305-308 */
    if (!dU || !NN || !(Umax > 0) || !(FWHM > 0)) return AMB_EARG;
    const double theta_0 = 0.6 * FWHM * M_PI / (180. * 60.);
    *dU = sqrt(log(2)) / (M_PI * theta_0);
    const int Ng = (int)ceil(Umax / *dU);
    *NN = (int64_t)2 * Ng;
    return AMB_OK;
}

int64_t amb_grid_cells(const amb_grid_cfg *cfg, int64_t ngroup, const float *ranpar, int64_t *cell, float *signv)
{   /* tgefuncs.c:386-391 and 448-451, float arithmetic as written; returns
       the number of groups inside Umax, or AMB_EARG for a bad config or a
       cell outside the grid */
    if (!cfg || ngroup < 0 || (ngroup && (!ranpar || !cell)) || cfg->npar < 2 || cfg->NN < 2 || !(cfg->dU > 0))
        return AMB_EARG;
    const int64_t NN = cfg->NN, Nh = NN / 2 + 1;
    int64_t in = 0;
    for (int64_t g = 0; g < ngroup; ++g) {
        const float *rp = ranpar + g * cfg->npar;
        const float sv = (rp[1] < 0.) ? -1. : 1.;
        const float u1 = cfg->nu_chan0 * sv * rp[0];
        const float v1 = cfg->nu_chan0 * sv * rp[1];
        const float uu = u1 * u1 + v1 * v1;
        const double Uval = sqrt(uu);
        if (signv) signv[g] = sv;
        if (!(Uval <= cfg->Umax)) {
            cell[g] = -1;
            continue;
        }
        int ii = (int)round(u1 / cfg->dU);
        ii = (ii < 0) ? (int)(NN + ii) : ii;
        const int jj = (int)round(v1 / cfg->dU);
        if (ii < 0 || ii >= NN || jj < 0 || jj >= Nh) return AMB_EARG;
        cell[g] = (int64_t)ii * Nh + jj;
        ++in;
    }
    return in;
}

int amb_read_params(const char *path, int64_t *NN, double *Du, double *nu0, double *dnu, double *chan0,
                    int64_t *nc, char *data_path, size_t cap)
{   /* read_inputs, amber_funcs.c:159-172: seven fields, the path at most 511 chars */
    if (!path || !data_path || cap < 2) return AMB_EARG;
    FILE *fp = fopen(path, "r");
    if (!fp) return AMB_EIO;
    char buf[512];
    long nn, c;
    double du, a, b, ch;
    int n = fscanf(fp, "%ld%lf%lf%lf%lf%ld%511s", &nn, &du, &a, &b, &ch, &c, buf);
    fclose(fp);
    if (n != 7) return AMB_EFORMAT;
    if (NN) *NN = nn;
    if (Du) *Du = du;
    if (nu0) *nu0 = a;
    if (dnu) *dnu = b;
    if (chan0) *chan0 = ch;
    if (nc) *nc = c;
    if (strlen(buf) + 1 > cap) return AMB_EARG;
    strcpy(data_path, buf);
    return AMB_OK;
}

int amb_write_cldnu(const char *path, int NB, int64_t nc, const double *binu, const double *cldnu,
                    const double *numdnu, const int32_t *num)
{   /* maps.c:213-222: bin-major, one line per (bin, lag) */
    if (!path || NB < 0 || nc < 0 || (NB && nc && (!binu || !cldnu || !numdnu || !num))) return AMB_EARG;
    FILE *fp = fopen(path, "w");
    if (!fp) return AMB_EIO;
    int ok = 1;
    for (int i = 0; i < NB && ok; ++i)
        for (int64_t j = 0; j < nc && ok; ++j)
            ok = fprintf(fp, "%d\t%e\t%d\t%e\t%e\t%d\n", i, binu[i], (int)j, cldnu[(size_t)i * nc + j], numdnu[j],
                         num[i]) > 0;
    if (fclose(fp) != 0 || !ok) return AMB_EIO;
    return AMB_OK;
}

/* ---------------------------------------------------------------- rules */
long amb_ng(double uhigh, double du)
{   /* amber.c:193-197, same integer truncation and parity step */
    double umax = uhigh + 3;
    long nmax = (long)(umax / du);
    long ng = (long)(3 * nmax);
    return (ng % 2 == 0) ? ng + 2 : ng + 3;
}

long amb_ring_reach(double uhigh)
{   /* a ring cell (ia, j) has ia^2 + j^2 < s_hi, s_hi the least s with
       sqrt(s) >= UHigh (sqrt in double, as the ring test); so |ia|, |j| are
       at most isqrt(s_hi - 1) */
    if (!(uhigh > 0)) return 0;
    long s = (long)floor(uhigh * uhigh) - 4;
    if (s < 0) s = 0;
    while (sqrt((double)s) < uhigh) ++s;
    if (s == 0) return 0;
    long m = (long)floor(sqrt((double)(s - 1)));
    while (m * m > s - 1) --m;
    while ((m + 1) * (m + 1) <= s - 1) ++m;
    return m;
}

static int smooth7(long n)
{
    static const long pr[4] = {2, 3, 5, 7};
    for (int i = 0; i < 4; ++i)
        while (n % pr[i] == 0) n /= pr[i];
    return n == 1;
}

long amb_compute_grid(double uhigh1, double uhigh2, double uhigh3, double du)
{
    const long ng = amb_ng(uhigh1, du);
    const long m1 = amb_ring_reach(uhigh1), m2 = amb_ring_reach(uhigh2), m3 = amb_ring_reach(uhigh3);
    long mx = m1 > m2 ? m1 : m2;
    if (m3 > mx) mx = m3;
    /* n > m1 + m2 + m3: a sum of three ring frequencies is a multiple of n
       only when it is zero; n >= 2 mx + 2: every ring row and column is a
       distinct index of the half spectrum, clear of its Nyquist row and
       column */
    long need = m1 + m2 + m3 + 1;
    if (2 * mx + 2 > need) need = 2 * mx + 2;
    if (need >= ng) return ng; /* the reference grid aliases, or nothing smaller holds the rings */
    long n = need + (need & 1);
    while (!smooth7(n)) n += 2;
    return n < ng ? n : ng;
}

void amb_uavg(const double *ulow, const double *uhigh, int nr, double *uavg)
{   /* amber_funcs.c:2065: (2/3)(H^3-L^3)/(H^2-L^2) */
    for (int i = 0; i < nr; ++i) {
        double h = uhigh[i], l = ulow[i];
        uavg[i] = (2. / 3.) * (pow(h, 3.) - pow(l, 3.)) / (pow(h, 2.) - pow(l, 2.));
    }
}

void amb_annular_rings(int nr, double *ulow, double *uhigh)
{   /* amber_funcs.c:2025-2053: edges step 5 to 50, 10 to 100, 20 to 200,
       35 to 270; each UHigh is the next ULow; ULow[0]=1, UHigh[last]=270 */
    static const int seg[4][3] = {{0, 50, 5}, {50, 100, 10}, {100, 200, 20}, {200, 270, 35}};
    int k = 0;
    for (int i = 0; i < nr; ++i) ulow[i] = uhigh[i] = 0;
    for (int s = 0; s < 4; ++s)
        for (int e = seg[s][0]; e < seg[s][1] && k < nr; e += seg[s][2]) {
            ulow[k] = e;
            if (k > 0) uhigh[k - 1] = e;
            ++k;
        }
    if (nr > 0) { ulow[0] = 1.; uhigh[nr - 1] = 270.; }
}

int amb_triangles(const double *uavg, int nr, int *tri, int cap)
{   /* amber_funcs.c:443-449: r1>=r2>=r3, t=U2/U1 in [0.5,1],
       mu=0.5(t+1/t-(U3/U1)(U3/U2)) in [0.5/t,1]; expression order kept so
       borderline equalities decide the same way */
    int n = 0;
    for (int a = 0; a < nr; ++a)
        for (int b = 0; b <= a; ++b) {
            double t = uavg[b] / uavg[a];
            if (!(t >= 0.5 && t <= 1.0)) continue;
            for (int c = 0; c <= b; ++c) {
                double mu = 0.5 * (t + 1. / t - (uavg[c] / uavg[a]) * (uavg[c] / uavg[b]));
                if (mu >= 0.5 / t && mu <= 1.0) {
                    if (tri && n < cap) { tri[3 * n] = a; tri[3 * n + 1] = b; tri[3 * n + 2] = c; }
                    ++n;
                }
            }
        }
    return n;
}

/* ---------------------------------------------------------------- generator */
static inline uint64_t mix64(uint64_t z)
{   /* splitmix64 finaliser */
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t *s)
{   /* (0,1]: 53 random bits, never 0 so log() is finite */
    *s += 0x9e3779b97f4a7c15ULL;
    return ((mix64(*s) >> 11) + 1) * (1.0 / 9007199254740992.0);
}

int amb_generate(const amb_gen_cfg *cfg, double *gv, double *samp, int *uf_chans, int *uf_nc)
{
    if (!cfg || !gv || !samp || cfg->N < 4 || cfg->N % 2 || cfg->nc < 1 || cfg->nflag < 0 ||
        cfg->nflag >= cfg->nc || (cfg->nc > 5 && cfg->nflag > cfg->nc - 1))
        return AMB_EARG;
    const int64_t N = cfg->N, nc = cfg->nc, Nh = N / 2 + 1;
    const double p0 = cfg->p0 > 0 ? cfg->p0 : 0.4, u0 = cfg->u0 > 0 ? cfg->u0 : 40.0;
    uint8_t *flag = calloc(nc, 1);
    if (!flag) return AMB_ENOMEM;
    /* flagged channels: a seeded Fisher-Yates draw that never picks channel 5
       (calCR reads sampling plane 5, amber_funcs.c:425) */
    {
        int64_t *ids = malloc(sizeof(int64_t) * nc);
        int64_t m = 0;
        for (int64_t c = 0; c < nc; ++c) if (c != 5) ids[m++] = c;
        uint64_t s = cfg->seed ^ 0xf1a9f1a9ULL;
        for (int64_t i = 0; i < cfg->nflag && i < m; ++i) {
            int64_t j = i + (int64_t)(u01(&s) * (double)(m - i));
            if (j >= m) j = m - 1;
            int64_t t = ids[i]; ids[i] = ids[j]; ids[j] = t;
            flag[ids[i]] = 1;
        }
        free(ids);
    }
    int n_uf = 0;
    for (int64_t c = 0; c < nc; ++c)
        if (!flag[c]) { if (uf_chans) uf_chans[n_uf] = (int)c; ++n_uf; }
    if (uf_nc) *uf_nc = n_uf;

#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        int64_t ia = i <= N / 2 ? i : i - N;
        for (int64_t j = 0; j < Nh; ++j) {
            int64_t cell = i * Nh + j;
            double pm = p0 * exp(-sqrt((double)(ia * ia + j * j)) / u0);
            for (int64_t c = 0; c < nc; ++c) {
                uint64_t s = mix64(cfg->seed * 0x632be59bd9b4e019ULL + (uint64_t)(cell * nc + c) + 1);
                double a = u01(&s), b = u01(&s), mk = u01(&s), cn = u01(&s);
                if (cfg->shared_mask) {
                    uint64_t s2 = mix64(cfg->seed * 0x632be59bd9b4e019ULL + (uint64_t)cell * 0x9e3779b1ULL + 7);
                    mk = u01(&s2);
                    cn = u01(&s2);
                }
                double m = (!flag[c] && mk < pm) ? 1.0 : 0.0;
                double cnt = m;
                if (m > 0 && cfg->count_max > 1) {
                    cnt = 1.0 + floor(cn * cfg->count_max);
                    if (cnt > cfg->count_max) cnt = cfg->count_max;
                }
                double r = sqrt(-2.0 * log(a));
                gv[(cell * nc + c) * 2] = r * cos(2 * M_PI * b) * cnt;
                gv[(cell * nc + c) * 2 + 1] = r * sin(2 * M_PI * b) * cnt;
                samp[cell * nc + c] = cnt;
            }
        }
    }
    /* Hermitian symmetry of the self-conjugate columns j=0 and j=N/2:
       (N-i, j) = conj(i, j); rows 0 and N/2 are real there. */
    int64_t cols[2] = {0, N / 2};
    for (int k = 0; k < 2; ++k) {
        int64_t j = cols[k];
        for (int64_t i = N / 2 + 1; i < N; ++i) {
            int64_t dst = i * Nh + j, src = (N - i) * Nh + j;
            for (int64_t c = 0; c < nc; ++c) {
                gv[(dst * nc + c) * 2] = gv[(src * nc + c) * 2];
                gv[(dst * nc + c) * 2 + 1] = -gv[(src * nc + c) * 2 + 1];
                samp[dst * nc + c] = samp[src * nc + c];
            }
        }
        int64_t rows[2] = {0, N / 2};
        for (int q = 0; q < 2; ++q)
            for (int64_t c = 0; c < nc; ++c) gv[((rows[q] * Nh + j) * nc + c) * 2 + 1] = 0.0;
    }
    free(flag);
    return AMB_OK;
}

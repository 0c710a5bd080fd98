/*
 * amber_mi355x -- MI355X-native (gfx950) implementation of AMBER's 4_fbe
 * estimator hot path behind a plain C ABI.
 *
 * The reference has no plugin or FFI layer: its hot path is the in-process C
 * function set declared in 4_fbe/mabs_funcs.h:19-60 plus the inline loops of
 * 4_fbe/amber.c and 4_fbe/amber_v2.c, all driven through globals
 * (4_fbe/vars.h:7-21).  Each entry point below names the reference function
 * or loop it replaces.  Differences in contract, all deliberate:
 *   - state lives in an amb_ctx instead of globals; every call returns an
 *     int status (AMB_OK or a negative AMB_E*), never exit();
 *   - output buffers are caller-owned host memory laid out exactly like the
 *     reference's result files; device memory is owned by the library;
 *   - a context is internally synchronised (one mutex), so it may be shared
 *     by host threads; calls on one context execute one at a time.
 * No torch or HIP types cross this boundary.
 */
#ifndef AMBER_MI355X_H
#define AMBER_MI355X_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define AMB_OK 0
#define AMB_EARG (-1)     /* bad argument / out-of-range ring, plane, triangle */
#define AMB_EIO (-2)      /* file missing or unreadable */
#define AMB_EFORMAT (-3)  /* truncated or inconsistent file */
#define AMB_EHIP (-4)     /* HIP runtime error */
#define AMB_ENOMEM (-5)   /* host allocation failed */
#define AMB_EBUDGET (-6)  /* request does not fit the HBM budget */
#define AMB_ESTATE (-7)   /* rings/channels not set, unknown cube id */
#define AMB_ECERT (-8)    /* an integer-valued sum failed its rounding certificate */
#define AMB_EBUSY (-9)    /* serving queue full */

/* ---- precision of the triple-product contraction ---- */
#define AMB_PREC_FP32 0   /* FP32 MFMA (v_mfma_f32_32x32x2_f32) */
#define AMB_PREC_BF16X3 1 /* hi+lo BF16 split, 3 MFMA passes, FP32 accumulate */
#define AMB_PREC_FP8 2    /* FP8 e4m3 with per-plane absmax scales */
#define AMB_PREC_FP16X3 3 /* hi+lo FP16 split, 3 MFMA passes, FP32 accumulate; operands
                             scaled by powers of two from each ring's plane-set maximum
                             so that FP16's range holds them (22-bit operands, against
                             16 bits for BF16x3) */
#define AMB_PREC_LAGFT 4  /* lag-Fourier form: the sum over reference planes as a
                             triple correlation along the channel axis, evaluated in
                             the DFT domain of the channel axis (FP32 products of
                             channel spectra on the vector ALUs, FP64 segment sums and
                             2-D inverse); L^2/2 complex products per pixel instead of
                             nc^3 real ones.  Batches run as waves of their own; the
                             serving queue runs it; not available in captured graphs.
                             The v2 normalisation of this form is the exact FP64 one. */

const char *amb_strerror(int code);
const char *amb_version(void);

/* =====================================================================
 * Host-side file formats and rules (no device, no context)
 * ===================================================================== */

/* read_inputs_ess (amber_funcs.c:175-188): "%lf%lf%lf%511s". */
int amb_read_params_ess(const char *path, double *nu0, double *dnu, double *chan0,
                        char *data_path, size_t data_path_cap);

/* GV_data_<R> header (amber_funcs.c:1733-1740): int64 NN, int64 nc, f64 Du,
 * int32 flag_key, 28 bytes, no padding. */
int amb_read_gv_header(const char *path, int64_t *NN, int64_t *nc, double *Du, int32_t *flag_key);

/* initialize + ReadGVnAC_gridding (amber_funcs.c:1720-1779, 1861-1911):
 * reads <data_path>/GV_data_<rlzn> and <data_path>/sampling in bulk.
 * *gv gets [NN*(NN/2+1)][nc][2] doubles and *samp [NN*(NN/2+1)][nc] doubles
 * (pixel-major, the file order); release both with amb_free. */
int amb_read_cube_files(const char *data_path, int rlzn, int64_t *NN, int64_t *nc, double *Du,
                        int32_t *flag_key, double **gv, double **samp);
void amb_free(void *p);

/* Writers for the same two files (3_vcg/tgefuncs.c:628-655, 728-748). */
int amb_write_cube_files(const char *data_path, int rlzn, int64_t NN, int64_t nc, double Du,
                         int32_t flag_key, const double *gv, const double *samp);

/* get_uf_chans (amber_funcs.c:2485-2508) reading <path>; unlike the
 * reference, a short read or an id outside [0,nc) is an error (AMB_EFORMAT). */
int amb_read_uf_chans(const char *path, int uf_nc, int64_t nc, int *chans);
int amb_write_uf_chans(const char *path, int uf_nc, const int *chans);

/* closed_rings + rcount (writer amber_funcs.c:459-483, reader read_rings
 * :499-545).  amb_read_rings allocates *rdata (rcount*4 doubles; amb_free). */
int amb_write_closed_rings(const char *res_path, const double *rdata, int rcount);
int amb_read_rings(const char *res_path, int *rcount, double **rdata);

/* bsout files.  v1 (amber.c:287-304): f64 Ntri, f64 Num, f32 bl[W*W],
 * i32 cnt[W*W] into <res_path>/bsout_final_ra%d_rb%d_rc%d_R%d.
 * v2 (amber_v2.c:295-309): f64 inv_Ng1_Ng2, f32 bl[W*W], f32 norm[W*W] into
 * <res_path>/bsout_final2_ra%d_rb%d_rc%d_R%d. */
int amb_write_bsout_v1(const char *res_path, int r1, int r2, int r3, int rlzn, int W,
                       double ntri, double num, const float *bl, const int32_t *cnt);
int amb_write_bsout_v2(const char *res_path, int r1, int r2, int r3, int rlzn, int W,
                       double inv_ng2, const float *bl, const float *norm);

/* Hardened result files (no counterpart in the reference, whose files carry
 * no dimensions or version tag, README.md:120).  A fixed 104-byte header,
 * then the legacy payload unchanged: f32 bl[W*W], then i32 counts (kind 1)
 * or f32 norm (kind 2).  Written as <res_path>/bsw1_ra%d_rb%d_rc%d_R%d
 * (kind 1) or bsw2_... (kind 2), a prefix the legacy readers' bsout_final*
 * patterns do not match.  The writer fills magic, version, header_bytes,
 * endian, payload_bytes and payload_crc; the reader checks every one of
 * them, W = 2 nc10 + 1 and the file size, and returns AMB_EFORMAT on any
 * mismatch (a truncated, corrupted or foreign-endian file). */
#define AMB_BSW_VERSION 1
#define AMB_BSW_KIND_V1 1
#define AMB_BSW_KIND_V2 2
typedef struct amb_bsw_header {
    char magic[8];          /* "AMBRBSW" + NUL */
    uint32_t version;       /* AMB_BSW_VERSION */
    uint32_t header_bytes;  /* 104 */
    uint32_t kind;          /* AMB_BSW_KIND_V1 / _V2 */
    uint32_t endian;        /* 0x01020304 in the writer's byte order */
    int32_t nc10, W;        /* W = 2 nc10 + 1 */
    int32_t r1, r2, r3, rlzn;
    int64_t ng;             /* reduced grid side of the triangle (amber.c:193-197) */
    int64_t nc;             /* channels of the cube */
    int32_t precision;      /* AMB_PREC_* of the data contraction */
    uint32_t payload_crc;   /* CRC-32 (IEEE 802.3, reflected 0xEDB88320) of the payload */
    double h0, h1;          /* kind 1: Ntri, Num (amber.c:287); kind 2: inv_Ng1_Ng2, 0 */
    uint64_t payload_bytes; /* 8 W^2 */
    uint8_t reserved[8];    /* zero */
} amb_bsw_header;
int amb_write_bsw(const char *res_path, const amb_bsw_header *h, const float *bl, const void *den);
/* Reads a hardened file: *h always (when the header is valid); bl and den
 * (W*W elements each) when non-NULL and cap >= W*W. */
int amb_read_bsw(const char *path, amb_bsw_header *h, float *bl, void *den, int64_t cap);
/* Reads a legacy bsout file of either kind, W inferred from the size
 * (16 + 8 W^2 for v1, 8 + 8 W^2 for v2; exactly one can hold for odd W).
 * *kind, *W always; h0/h1 as in amb_bsw_header; arrays when cap >= W*W. */
int amb_read_bsout(const char *path, int *kind, int *W, double *h0, double *h1, float *bl, void *den,
                   int64_t cap);
uint32_t amb_crc32(const void *p, size_t n);

/* Rules.  amb_ng: amber.c:193-197.  amb_annular_rings: the reference's
 * hard-coded edges, amber_funcs.c:2022-2053.  amb_uavg: :2065.
 * amb_triangles: :443-449 (returns the count; fills up to cap triples). */
long amb_ng(double uhigh, double du);
/* The grid a triangle's sums are computed on (no counterpart in the
 * reference, which computes every triangle on amb_ng(UHigh[r1])): the
 * smallest even 7-smooth n with n > m1 + m2 + m3 and n >= 2 max(m) + 2,
 * m = amb_ring_reach of each ring's UHigh (the largest |component| of a
 * ring cell), or Ng itself when that is not smaller.  On such a grid no sum
 * of three ring frequencies wraps, and neither does one on Ng when n < Ng
 * (n < Ng implies m1 + m2 + m3 < Ng), so the triple sum over n^2 pixels is
 * n^2 times the closed-triangle sum over the ring cells and the reference's
 * sum over Ng^2 pixels is (Ng/n)^2 times the sum over n^2 (amber.c:193-266
 * sums D_r1 D_r2 D_r3 over Ng(r1)^2 pixels). */
long amb_ring_reach(double uhigh);
long amb_compute_grid(double uhigh1, double uhigh2, double uhigh3, double du);
void amb_annular_rings(int nr, double *ulow, double *uhigh);
void amb_uavg(const double *ulow, const double *uhigh, int nr, double *uavg);
int amb_triangles(const double *uavg, int nr, int *tri, int cap);

/* Deterministic synthetic cube (no counterpart in the reference, which ships
 * no data: examples/README.md:3).  Cell values (N(0,1)+iN(0,1))*m with
 * m ~ Bernoulli(p0*exp(-|U|/u0)), |U| in cells, Hermitian on the j=0 and
 * j=N/2 columns; nflag channels (never channel 5) fully flagged.
 * shared_mask=1 uses one mask for every unflagged channel.
 * count_max>1 stores sampling counts in [1,count_max] for sampled cells
 * (values scaled by the count, so normalise_data divides them back).
 * gv/samp as amb_read_cube_files; uf_chans gets the unflagged ids (size nc). */
typedef struct amb_gen_cfg {
    int64_t N, nc;
    int nflag;
    int shared_mask;
    int count_max;
    uint64_t seed;
    double p0, u0;
} amb_gen_cfg;
int amb_generate(const amb_gen_cfg *cfg, double *gv, double *samp, int *uf_chans, int *uf_nc);

/* =====================================================================
 * Device context
 * ===================================================================== */
typedef struct amb_ctx amb_ctx;

typedef struct amb_opts {
    int device;          /* HIP device ordinal */
    int precision;       /* AMB_PREC_* for the data contraction */
    uint64_t hbm_budget; /* bytes the library may hold; 0 = 90% of free HBM */
    int verbose;
} amb_opts;

int amb_create(const amb_opts *opts, amb_ctx **out);
void amb_destroy(amb_ctx *ctx);
const char *amb_last_error(amb_ctx *ctx);
int amb_set_precision(amb_ctx *ctx, int precision);

/* Upload a cube given in the file layout.  Replaces initialize +
 * ReadGVnAC_gridding + normalize_data (amber_funcs.c:1720-1978): values are
 * divided by the sampling count where it is >0 and the count becomes 1, in
 * FP64, then stored plane-major as complex64 / float in HBM. */
int amb_load_cube(amb_ctx *ctx, int64_t N, int64_t nc, double Du, int32_t flag_key,
                  const double *gv, const double *samp, int *cube_id);
int amb_load_cube_files(amb_ctx *ctx, const char *data_path, int rlzn, int *cube_id);
int amb_release_cube(amb_ctx *ctx, int cube_id);

/* Triplet set (precompute_unflg_vcp / precompute_vcp_dnu,
 * amber_funcs.c:2207-2310): with the cube's flag_key=1 the channels are
 * uf_chans[0..uf_nc), with flag_key=0 all channels (uf_chans ignored).
 * Lags |p2-p1|,|p3-p1| <= nc10.  The table itself is never materialised. */
int amb_set_channels(amb_ctx *ctx, int cube_id, const int *uf_chans, int uf_nc, int nc10);

/* Ring edges in cells (replaces annular_rings, amber_funcs.c:2022-2094). */
int amb_set_rings(amb_ctx *ctx, const double *ulow, const double *uhigh, int nr);

/* calCR (amber_funcs.c:389-491) on the GPU: per ring the FP64 c2r of
 * sampling plane `plane` (5 in the reference) on the full grid, per closed
 * triangle the FP64 triple sum, as the reference's FFTW double path.  Num
 * is an integer multiple of N^2/8 (certified within 1/32 of it, else
 * AMB_ECERT) and written exactly.  rdata gets rcount x {r1,r2,r3,Num}.
 * N <= 1365 (two FP64 transform buffers in 64 KiB of LDS). */
int amb_closed_rings(amb_ctx *ctx, int cube_id, int plane, double *rdata, int cap, int *rcount);

/* amb_closed_rings with no device allocation at all.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (FP64 ring images, FFT scratch, twiddles, triangle sums) is a
 * static device array of 64 MiB that every device loading the library
 * holds until the process ends; a call that needs more (nr * N^2 * 8 bytes
 * plus a little) fails with AMB_EBUDGET.  Results equal amb_closed_rings'
 * bit for bit.  All calls on one device share the array: calls that
 * overlap on one GPU (several contexts or host threads at once) race. */
int amb_closed_rings_noalloc(amb_ctx *ctx, int cube_id, int plane, double *rdata, int cap, int *rcount);

/* One triangle, v1 (fill_s2d_float x3 + amber.c:226-304): bl[W*W] f32,
 * cnt[W*W] i32; *ntri and *num as written by the reference (Num =
 * (Ng/N)^2 * rdata_num, Ntri = Num/(Ng*Ng)). */
int amb_bispec_v1(amb_ctx *ctx, int cube_id, int r1, int r2, int r3, double rdata_num,
                  float *bl, int32_t *cnt, double *ntri, double *num);

/* One triangle, v2 (fill_s2d_float x3 + fill_s2I x3 + amber_v2.c:222-309):
 * bl[W*W] f32 and norm[W*W] f32 (exact integer*Ng^2 cast to f32). */
int amb_bispec_v2(amb_ctx *ctx, int cube_id, int r1, int r2, int r3,
                  float *bl, float *norm, double *inv_ng2);

/* Extended single-triangle call used by tests and tools: any output may be
 * NULL.  bl64 receives the compensated accumulator before the f32 cast;
 * norm64 the exact normalisation in FP64. */
typedef struct amb_result {
    float *bl;
    double *bl64;
    int32_t *cnt;
    float *norm;
    double *norm64;
    double inv_ng2;
    long ng;
    double cert_margin; /* worst |S/P - round(S/P)| over normalisation sums */
} amb_result;
int amb_bispec_ex(amb_ctx *ctx, int cube_id, int r1, int r2, int r3, int want_norm, amb_result *res);

/* Many triangles of one cube in one call, grouped by r1 so filtered planes
 * are built once per (ring, Ng) and reused.  tri: ntri x {r1,r2,r3}.
 * bl: ntri*W*W f32 (required); cnt, norm: may be NULL.  This replaces the
 * whole rp loop of amber.c:187-326 / amber_v2.c:187-333 for one cube. */
int amb_bispec_batch(amb_ctx *ctx, int cube_id, const int *tri, int ntri,
                     float *bl, int32_t *cnt, float *norm);

/* One triangle of one resident cube with caller-owned host outputs
 * (bl: W*W f32, required; cnt: W*W i32 and norm: W*W f32, may be NULL; a
 * non-NULL norm asks for the v2 normalisation).  No counterpart in the
 * reference, which runs one cube per process. */
typedef struct amb_tri_req {
    int cube_id, r1, r2, r3;
    float *bl;
    int32_t *cnt;
    float *norm;
} amb_tri_req;

/* Triangles of any resident cubes in one call.  Results are bit-identical
 * to one amb_bispec_ex per triangle in every batch mode. */
int amb_bispec_multi(amb_ctx *ctx, const amb_tri_req *req, int n);

/* How amb_bispec_batch / amb_bispec_multi execute:
 *   AMB_BATCH_GROUPED (default): waves of triangles, each wave one
 *     persistent contraction launch over the work items of all its
 *     triangles (largest first), one grouped reduction and one copy-out;
 *   AMB_BATCH_PER_TRIANGLE: launches and a host sync per triangle. */
#define AMB_BATCH_PER_TRIANGLE 0
#define AMB_BATCH_GROUPED 1
int amb_set_batch_mode(amb_ctx *ctx, int mode);

/* Launch configuration of the contraction for one shape bucket, chosen by
 * the autotuner.  A bucket is (precision, ceil(log2 Ng), ceil(log2 nact)).
 * Every configuration computes each triangle with the same arithmetic in
 * the same order, so results do not depend on the table: it only moves
 * time.  The reference has no counterpart (its triangle loop,
 * amber.c:187-326, has nothing to tune); this table replaces that loop's
 * fixed schedule. */
typedef struct amb_tune_entry {
    int prec;      /* AMB_PREC_* */
    int ng_log2;   /* bucket: ceil(log2 Ng) */
    int nact_log2; /* bucket: ceil(log2 active channels) */
    int variant;   /* contraction kernel: 0 plain k-steps, 1 interleaved
                      A formation, 2 direct-to-LDS pipelined (split forms) */
    int xrun_log2; /* work items per run of an XCD work queue, log2 (0..12) */
    int wave_gib;  /* contraction partials per grouped wave, GiB (1..256) */
    int grid_pct;  /* persistent grid, percent of the resident workgroups */
    double tflops; /* issued MFMA TFLOP/s of the data contraction measured
                      with this entry (0: not measured) */
} amb_tune_entry;

/* Times candidate configurations of each shape bucket present in req (a
 * sample of the caller's workload; the requests' outputs receive their
 * results) by coordinate descent over variant, queue run, wave size and
 * grid, on resident planes, and installs the fastest per bucket.  Every
 * candidate's results are compared with the first run's bit for bit (a
 * difference is an AMB_EHIP error).  max_seconds bounds the search (the
 * best so far is kept).  path (may be NULL): the table is also saved
 * there.  Buckets not tuned fall back to the nearest tuned bucket of the
 * same precision, then to the built-in defaults. */
int amb_autotune(amb_ctx *ctx, const amb_tri_req *req, int n, double max_seconds, const char *path);
/* text table: a header line naming the device architecture and CU count
 * (a table from another device is an AMB_EFORMAT error), one entry a line */
int amb_tune_save(amb_ctx *ctx, const char *path);
int amb_tune_load(amb_ctx *ctx, const char *path);
int amb_tune_clear(amb_ctx *ctx);
/* installed entries, up to cap; *n gets the count */
int amb_tune_entries(amb_ctx *ctx, amb_tune_entry *out, int cap, int *n);
/* the configuration the context would use for (Ng, nact) at its precision */
int amb_tune_lookup(amb_ctx *ctx, int ng, int nact, amb_tune_entry *out);
/* one configuration for every shape, overriding the table (the prec and
 * bucket fields are ignored); NULL returns to the table */
int amb_set_launch_config(amb_ctx *ctx, const amb_tune_entry *cfg);

/* Host wall time (ms) from the start of the last amb_bispec_batch /
 * amb_bispec_multi call to each triangle's results being in host memory,
 * in the order the triangles were given (a triangle of a grouped wave
 * completes with its wave).  *n gets the count (<= cap written). */
int amb_batch_times(amb_ctx *ctx, double *ms, int cap, int *n);

/* Re-upload new values into a resident cube of the same shape (the next
 * realisation, GV_data_<R>); cached planes of the cube are dropped. */
int amb_update_cube(amb_ctx *ctx, int cube_id, const double *gv, const double *samp);

/* A fixed batch captured once into a HIP graph and replayed: the graph
 * rebuilds every filtered plane it reads from the cube buffers (so a
 * replay after amb_update_cube sees the new realisation), runs one grouped
 * contraction per wave, the grouped reduction and one copy-out, and the
 * host scatters the results into the requests' output pointers, which must
 * stay valid for the graph's life.  The graph owns its planes and scratch;
 * its cubes cannot be released or re-channelled while it lives. */
typedef struct amb_graph amb_graph;
int amb_graph_create(amb_ctx *ctx, const amb_tri_req *req, int n, amb_graph **out);
int amb_graph_run(amb_ctx *ctx, amb_graph *g);
void amb_graph_destroy(amb_ctx *ctx, amb_graph *g);

/* Serving queue (no counterpart in the reference, which runs one cube per
 * process, amber.c:32-38): continuous batching at the granularity of
 * work items (the contraction's, or in AMB_PREC_LAGFT the tile kernel's
 * (tile, pixel segment) items, a triangle's v2 normalisation then coming
 * from the exact FP64 path when it completes).  A scheduler thread admits requests' triangles
 * as device memory allows and launches bounded contraction iterations,
 * each taking a fair share of work items from every admitted request, so
 * a request that arrives while others run starts within an iteration or
 * two instead of waiting for them to end; a triangle is reduced and
 * copied out as soon as its last items have run.  Triangles are admitted
 * against the context's HBM budget: their planes come from the LRU plane
 * cache (unpinned planes are evicted as needed, planes of admitted
 * triangles are held until they finish), their partial sums and results
 * from pools reserved at start (a quarter and a sixteenth of the budget,
 * at most 16 GiB and 2 GiB).  max_active bounds the triangles admitted at
 * once.  Results are bit-identical to amb_bispec_multi of the same
 * triangles.  Output pointers of a request must stay valid until amb_wait
 * returns for its ticket.  While serving, the context's precision and
 * rings are fixed, and cubes with admitted triangles cannot be released,
 * updated or re-channelled (AMB_ESTATE); other calls may run between
 * iterations. */
typedef uint64_t amb_ticket;
int amb_serve_start(amb_ctx *ctx, int max_active);
/* AMB_EBUSY: 4096 requests already wait for admission; AMB_EBUDGET: a
 * triangle of the request can never fit the pools */
int amb_submit(amb_ctx *ctx, const amb_tri_req *req, int n, amb_ticket *ticket);

/* A cube carried by a request: uploaded by a thread of its own on a stream
 * of its own (the copies overlap the contraction), then resident under the
 * id returned at submission, which later requests may name (they wait for
 * the upload).  Such a cube is evicted, least recently used first, once no
 * admitted triangle reads it and the budget needs its memory; a request
 * naming an evicted cube fails with AMB_ESTATE.  gv / samp are in the
 * layout of amb_read_cube_files and must stay valid until the ticket is
 * done; chans (uf_nc ids, flag_key = 1) likewise. */
typedef struct amb_cube_src {
    int64_t N, nc;
    double Du;
    int32_t flag_key;
    const double *gv, *samp;
    const int *chans;
    int uf_nc, nc10;
} amb_cube_src;
/* req[i].cube_id of 0 names the carried cube; AMB_EBUDGET when the cube
 * cannot fit the budget even with every evictable byte freed */
int amb_submit_cube(amb_ctx *ctx, const amb_cube_src *cube, const amb_tri_req *req, int n, int *cube_id,
                    amb_ticket *ticket);
/* *done = 1 once every triangle of the ticket is in host memory */
int amb_poll(amb_ctx *ctx, amb_ticket ticket, int *done);
/* blocks until done; *latency_ms = submit to completion (may be NULL);
 * returns the ticket's status (the first error of its triangles) */
int amb_wait(amb_ctx *ctx, amb_ticket ticket, double *latency_ms);
/* finishes every submitted request, then stops */
int amb_serve_stop(amb_ctx *ctx);

/* One ring-filtered real-space image of plane `plane`, as the triple
 * product reads it: kind 0 is the data image of fill_s2d_float
 * (amber_funcs.c:967-1010), kind 1 the sampling image of fill_s2I
 * (amber_funcs.c:1109-1131).  ng <= 0 takes Ng of ring `ring` itself
 * (amber.c:193-197 computes Ng from the triangle's r1).  out: ng*ng f32,
 * row-major [x][y]; cap is its length in floats.  The plane must belong to
 * the channel set (flagged planes are never built). */
int amb_ring_image(amb_ctx *ctx, int cube_id, int ring, int64_t ng, int plane, int kind,
                   float *out, int64_t cap);

/* Lag-space post-processing of result arrays (the consumer of the bsout
 * files): get_cyl_BS (mabs_funcs.py:375-388), or with window=1 cylbs_win
 * (mfunc_winds.py:113-127).  For each of n triangles: fold the W x W arrays
 * (fold_quadrants, mabs_funcs.py:210-222), divide by mabs_norm and the
 * folded denominator, cut ncut rows/columns (cut_nan, :259-260), apply the
 * Blackman-Nuttall window if asked, and transform to k_par
 * (Bdnu2Bk_fft, :325-334).  den: W x W int32 counts (v1, den_is_int = 1)
 * or float32 norms (v2).  out: n x N x N float64, N = W - 2 ncut. */
typedef struct amb_cyl_params {
    int ncut;         /* n_nancut, mabs_funcs.py:16 */
    double mabs_norm; /* mabs_funcs.py:22 */
    double rc;        /* mabs_funcs.py:24 */
    double rc_prime;  /* mabs_funcs.py:25 */
    double dnu;       /* dnu_mwa, mabs_funcs.py:27 */
    int window;
} amb_cyl_params;
int amb_cyl_bispec(amb_ctx *ctx, int n, int W, const float *bl, const void *den, int den_is_int,
                   const amb_cyl_params *p, double *out);

/* MAPS C_l(dnu): replaces the maps program's pair loop (4_fbe/maps.c:100-229)
 * over cl_dnu (4_fbe/amber_funcs.c:1418-1488) for one cube given in the file
 * layout (gv [N(N/2+1)][nc][2], samp [N(N/2+1)][nc], normalised inside as
 * maps.c:100-125 does).  NB log bins of |U| = sqrt(ia^2+j^2) Du over
 * [umin, umax) (the reference uses 15, 1, 500).  Outputs, as the reference
 * leaves them before writing results/cldnu:
 *   cldnu[NB][nc]  sum over n1 of the binned C(n1, n1 + d) / no_b(n1)
 *   numdnu[nc]     number of channel pairs at lag d (nc - d)
 *   binu[NB]       mean |U| of the bin's sampled cells of channel nc-1
 *   num[NB]        number of those cells
 * FP64 throughout; counts exact.  Sampling counts must be >= 0 (AMB_EARG
 * otherwise; the gridder never writes a negative count). */
int amb_maps_cldnu(amb_ctx *ctx, int64_t N, int64_t nc, double Du, const double *gv, const double *samp,
                   int NB, double umin, double umax, double *cldnu, double *numdnu, double *binu,
                   int32_t *num);

/* Upstream gridder vcg on the GPU (3_vcg/tgefuncs.c).  The reference reads
 * a UVFITS random-groups file through CFITSIO (read_ranpar / read_data,
 * read_fits_funcmfs.c:99-120); these entry points take what those calls
 * return, for every group: pcount random parameters (u at 0, v at 1, in
 * seconds) and nchan x nstokes x 3 floats (re, im, wt).
 *   amb_read_vcg_inputs  read_inputs (tgefuncs.c:28-37): nc Umax FWHM f
 *   amb_grid_setup       initialize_gridding (:290-376): dU from the beam
 *                        FWHM (arcmin), NN = 2 ceil(Umax/dU)
 *   amb_grid_cells       the cell of every group (-1 beyond Umax), with the
 *                        reference's float arithmetic (:386-391, 448-451)
 *   amb_grid             group_loop_gridding (flag_key 1, :377-509) or
 *                        group_loop_gridding_without_flag (0, :511-626),
 *                        then collapse_channels (:816-839) when
 *                        n_collapse > 1: gv [NN(NN/2+1)][nc/n_collapse][2]
 *                        and samp [..][nc/n_collapse] in the GV_data /
 *                        sampling layout, bit-identical to the reference's
 *                        sums (each cell adds its groups in group order)
 *   amb_grid_cube        the same, kept resident in HBM as a cube (as if
 *                        written and read back by amb_load_cube), so the
 *                        bispectrum path needs no file round trip
 * Only the RR stokes (index 0) reaches the reference's output files; the LL
 * sums it also forms are never written and are not formed here.  A group
 * whose cell falls outside the grid (possible only when NN or dU do not come
 * from amb_grid_setup) is AMB_EARG. */
typedef struct amb_grid_cfg {
    int64_t NN;           /* grid side */
    int64_t nc;           /* channels gridded, from chan1 */
    int64_t nchan;        /* channels in the data (NAXIS4) */
    int64_t nstokes;      /* stokes in the data (NAXIS3); 3 floats each */
    int chan1;            /* first channel, 1-based (tgefuncs.c:19) */
    int npar;             /* random parameters per group (PCOUNT) */
    int flag_key;         /* 1: weight > 0 only; 0: every visibility */
    int n_collapse;       /* 1: none; k: sums of k adjacent channels */
    float nu_chan0;       /* CRVAL4, Hz (float as in read_fits_funcmfs.c:27) */
    double dU, Umax;      /* wavelengths */
} amb_grid_cfg;
int amb_read_vcg_inputs(const char *path, int *nc, double *Umax, double *FWHM, double *f);
int amb_grid_setup(double Umax, double FWHM, double *dU, int64_t *NN);
int64_t amb_grid_cells(const amb_grid_cfg *cfg, int64_t ngroup, const float *ranpar, int64_t *cell, float *signv);
int amb_grid(amb_ctx *ctx, const amb_grid_cfg *cfg, int64_t ngroup, const float *ranpar, const float *data,
             double *gv, double *samp, int64_t *ngridded);
int amb_grid_cube(amb_ctx *ctx, const amb_grid_cfg *cfg, int64_t ngroup, const float *ranpar, const float *data,
                  int *cube_id, int64_t *ngridded);

/* results/cldnu text (maps.c:213-222): NB*nc lines "b binu d cldnu numdnu num". */
int amb_write_cldnu(const char *path, int NB, int64_t nc, const double *binu, const double *cldnu,
                    const double *numdnu, const int32_t *num);
/* input_params of the maps program (read_inputs, amber_funcs.c:159-172):
 * NN Du nu0 dnu chan0 nc data_path; data_path gets at most cap-1 chars. */
int amb_read_params(const char *path, int64_t *NN, double *Du, double *nu0, double *dnu, double *chan0,
                    int64_t *nc, char *data_path, size_t cap);

/* Drop every cached filtered plane of a cube (they are rebuilt on demand).
 * Planes stay resident across calls otherwise, as the serving loop wants. */
int amb_drop_planes(amb_ctx *ctx, int cube_id);

/* Per-stage device time and work since the last reset (HIP events on the
 * library's stream). */
typedef struct amb_stats {
    double ms_fft;          /* ring gather + c2r kernels */
    double ms_contract;     /* triple-product GEMM kernels */
    double ms_reduce;       /* lag reduction + counts */
    double ms_norm;         /* normalisation contraction (v2) */
    double ms_total;        /* wall of device work incl. the above */
    double flop_contract;   /* algorithmic 2*cnt*Ng^2 per eval, data only */
    double bytes_fft;       /* algorithmic bytes of the fft stage */
    int64_t n_contract_launch, n_fft_launch;
    int64_t evals;          /* (triangle, lag cell) pairs with cnt > 0 */
    double flop_contract_exec; /* MFMA work issued by the data contraction:
                                  whole 128x128x4 tiles over the padded pixel
                                  pitch, times the passes of the form (3 for
                                  BF16x3); below flop_contract x passes when
                                  r1 == r2 tiles are skipped */
    double ms_graph;        /* device time of amb_graph_run replays */
    int64_t n_graph_launch;
    double ms_maps;         /* MAPS pair-contraction kernel */
    double ms_maps_total;   /* all MAPS kernels (normalise .. lag sums) */
    double flop_maps;       /* algorithmic 4 flops per (channel pair n1<=n2, binned cell) */
    int64_t n_maps_launch;
    double ms_grid;         /* gridding kernels (sums and collapse) */
    int64_t n_grid_groups;  /* groups gridded */
    double ms_post;         /* lag-space transform matrix products */
    double flop_post;       /* their algorithmic 8 M N^2 flop per triangle */
    double cert_margin_max; /* worst normalisation grid distance (see amb_set_cert_strict) */
    int64_t n_uncertified;  /* triangle calls or waves whose normalisation was not certified */
    double flop_contract_distinct; /* flop_contract over the distinct sums only: where
                                      r1 == r2 (or r2 == r3) makes two triplets' sums the
                                      same, the pair counts once */
    int64_t n_norm_exact;   /* triangles whose v2 normalisation was recomputed in FP64 */
    int64_t n_evict_planes; /* filtered-plane sets evicted (LRU) for HBM budget */
    int64_t n_evict_cubes;  /* cubes carried by requests evicted (LRU, idle) for HBM budget */
    int64_t n_cube_uploads; /* cubes carried by requests uploaded by the serving queue */
    int64_t n_serve_iters;  /* contraction iterations the serving queue launched */
    double flop_lagft;      /* lag-Fourier form: 14 flops per distinct (k2, k3, pixel) of the
                               half spectrum k3 <= L/2, the algorithmic work of its products
                               (ms_contract times them) */
    int64_t plane_pool_bytes; /* device memory the plane pool holds (in use or free for
                                 reuse), when read */
} amb_stats;
int amb_get_stats(amb_ctx *ctx, amb_stats *st);
int amb_reset_stats(amb_ctx *ctx);

/* The v2 normalisation sum of a triplet is a whole number of eighths of
 * Ng^2 (the sampling images are 0/1 masks, half-weighted on the kx = 0
 * column).  The fast contraction (BF16x3, or FP32 in the FP32 form) rounds
 * each sum to the nearest 1/8; the distance of the triangle's worst sum
 * from that grid is its certificate.  Within 1/32 the rounded sums are
 * exact.  A triangle beyond it (large Ng: the sums reach ~1e5 Ng^2, past
 * what FP32 accumulation resolves) is recomputed with FP64 sampling images
 * and FP64 MFMA products, whose sums sit within ~1e-9 of the grid: the
 * normalisation is exact in every case (stats: n_uncertified counts the
 * fast path's misses, n_norm_exact the recomputations).  strict = 0
 * (default): recompute.  strict = 1: fail with AMB_ECERT instead, for
 * callers that must not pay the FP64 cost.  No counterpart in the
 * reference, whose FP32 sums (amber_v2.c:262-266) are never exact at these
 * sizes.  A sum beyond the int32 range of eighths fails in either mode. */
int amb_set_cert_strict(amb_ctx *ctx, int strict);

/* Grids of the lag-Fourier, FP32 and FP16x3 forms.  reduced = 1 (default):
 * each triangle's images (and channel spectra) and sums are computed on
 * amb_compute_grid of its rings and the sums scaled by (Ng/n)^2 in FP64 (the
 * normalisation rounded to its 1/8 grid of n^2 first, so it stays exact);
 * the lag arrays are the reference's sums over Ng(r1)^2 pixels
 * (amber.c:193-266) to rounding.  reduced = 0: every triangle on Ng(r1), as
 * the reference.  The split-BF16 and FP8 forms always use Ng(r1): their
 * per-product representation error does not shrink with the grid, so their
 * relative error would grow as Ng/n.  No counterpart in the reference. */
int amb_set_grid(amb_ctx *ctx, int reduced);

#ifdef __cplusplus
}
#endif
#endif

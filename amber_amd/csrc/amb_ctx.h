/*
 * Host-side state of a device context, shared by the C-ABI translation
 * units (amb_api.hip: context, cubes, planes, one triangle; amb_wave.hip:
 * grouped waves of many triangles, captured graphs and the serving queue).
 * Not part of the public interface.
 */
#pragma once
#include "amb_internal.h"
#include "amber_mi355x.h"

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

struct AmbGraph;
struct AmbServer;

namespace amb {

struct Cube {
    int64_t N = 0, nc = 0;
    double Du = 1.0;
    int flag_key = 0;
    float2 *vis_t = nullptr;
    float *samp_t = nullptr;
    size_t bytes = 0;
    bool chans = false;
    int nc10 = 0;
    int graphs = 0;            /* captured graphs that read this cube's buffers */
    int serving = 0;           /* triangles of the serving queue admitted on this cube */
    bool uploaded = false;     /* carried by a request: evictable once idle */
    int holds = 0;             /* requests carrying it not yet admitted (not evictable) */
    uint64_t last_use = 0;     /* tick of its last admitted triangle (LRU of uploaded cubes) */
    std::vector<int> act, inv; /* compact <-> channel */
    int *d_act = nullptr;      /* [nact] */
    int *d_inv = nullptr;      /* inv[nc] then first_ge[nc+1] */
    int *d_wlo = nullptr, *d_whi = nullptr;
    std::vector<int> wlo, whi; /* host copies of the per-plane windows */
    /* lag-Fourier form: channel transform length L (0 until first use), row
       pitch of the spectra (L/2 + 1 rounded up to 4), the channel set's lag
       counts on the device and the host, their cells with cnt > 0 and sum */
    int lag_L = 0, lag_lhp = 0;
    int *d_lagcnt = nullptr;
    int *d_lagpos = nullptr; /* digit-reversed sequence position of each active plane (in-place channel DFT) */
    std::vector<int32_t> h_lagcnt;
    int64_t lag_evals = 0, lag_pairs = 0;
    int nact() const { return (int)act.size(); }
};

/* One set of filtered planes: every active plane of (cube, kind, ring, Ng).
 * kind 0: data images (FP32), 1: sampling images (FP32),
 * 2 / 3: hi and lo BF16 halves of kind 0 / 1, stored hi planes then lo,
 * 4: e4m3 data images times their B-side scales,
 * 5: per-plane FP8 scales of kind 0 (A side [nact], then B side [nact]),
 * 7: hi and lo FP16 halves of kind 0 times 2^8, stored hi planes then lo,
 * 8: FP64 sampling images (the exact v2 normalisation),
 * 9: channel spectra of kind 11 (lag-Fourier form): [Ng^2][lag_lhp] complex
 *    FP32, the DFT over channels of each pixel of the set times 2^(7 - E),
 *    then the [max, E] words,
 * 10: channel spectra of kind 8 (the lag-Fourier normalisation), complex FP64,
 * 11: data images (FP32) as computed, not normalised, then the max word
 *    (the source of kind 9, which scales while it reads).
 * Kind 0 sets are stored normalised by a power of two (max |D| <= 2^7) and
 * end in two words: max |D| as computed and its exponent E. */
struct PlaneSet {
    int cube, kind, ng, jcnt;
    double lo, hi;
    void *d = nullptr;
    size_t bytes = 0;
    uint64_t last = 0;
    int pins = 0;  /* held by the call in progress (cleared by unpin_all) */
    int spins = 0; /* held by triangles the serving queue has admitted */
};

/* What fills one plane set, recorded so it can be launched now or inside
 * a captured graph. */
struct BuildOp {
    int kind = 0, cube = 0;
    RingGeom g{};
    void *dst = nullptr;
    const void *src = nullptr; /* kinds 2-7, 9: the FP32 images; 10: the FP64 sampling images */
    const void *sc = nullptr;  /* kind 4: the FP8 scales */
};

/* Bump allocation inside one device arena (a graph wave's planes); with
 * base == nullptr it only measures. */
struct Bump {
    char *base = nullptr;
    size_t off = 0, cap = ~(size_t)0;
};

/* Where planes come from: the context's LRU cache (builds launched at
 * once, memory from the budget) or a graph wave's list (builds recorded
 * for capture, memory from a bump arena). */
struct PlaneSrc {
    std::vector<PlaneSet> *list;
    std::vector<BuildOp> *record = nullptr; /* null: launch now */
    Bump *bump = nullptr;                   /* null: dev_alloc */
};

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    double *acc = nullptr;
};

/* One triangle resolved against a cube: planes pinned, kernel arguments
 * filled except the scratch pointers an executor assigns. */
struct TriJob {
    int cube = 0, r1 = 0, r2 = 0, r3 = 0, want_norm = 0;
    long ng = 0, P = 0, pitch = 0; /* Ng(r1); pixels and pitch of the compute grid */
    long ngc = 0;                  /* the compute grid (compute_grid) */
    int nact = 0, W = 0, sym12 = 0, sym23 = 0;
    size_t cells = 0;
    ContractArgs a{}; /* data contraction */
    ContractArgs b{}; /* normalisation contraction (want_norm) */
    ReduceArgs ra{};
    double flop_exec = 0; /* MFMA work the data contraction issues */
};

} // namespace amb

struct amb_ctx {
    int device = 0, prec = AMB_PREC_BF16X3, verbose = 0, sync_debug = 0, no_sym = 0, no_shear = 0;
    int batch_mode = AMB_BATCH_GROUPED;
    int cert_strict = 0; /* uncertified normalisation sums: count (0) or fail (1) */
    int grid_reduced = 1; /* lag-Fourier, FP32, FP16x3 sums on amb_compute_grid (1) or on Ng(r1) (0) */
    uint64_t budget = 0, used = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    std::string err;
    std::vector<double> ulow, uhigh, uavg;
    std::map<int, amb::Cube> cubes;
    int next_cube = 1;
    std::vector<amb::PlaneSet> planes;
    uint64_t tick = 0;
    std::map<int, float2 *> tw;
    std::map<int, FftPlan> plans;
    std::map<int, double2 *> tw64; /* FP64 twiddles (exact normalisation) */
    int grouped_blocks = 0; /* resident workgroups of the grouped contraction */
    /* launch configurations by (prec, ng_log2, nact_log2); `forced` (when
       has_forced) overrides them for every shape */
    std::map<std::tuple<int, int, int>, amb_tune_entry> tune;
    amb_tune_entry forced{};
    bool has_forced = false;
    std::string arch; /* gcnArchName of the device */
    std::vector<void *> *pin_log = nullptr; /* when set, get_planes notes every cache plane it pins */
    /* arena of the cached planes: device chunks allocated once and kept
       (until the device or the budget needs them back), planes placed
       first-fit in their free ranges.  Every reader of a plane runs on the
       library stream, so a freed range is reused in stream order and
       eviction never waits for the device.  plane_live: bytes of live
       planes (counted in used); arena_held: bytes of all chunks. */
    struct ArenaChunk {
        char *base = nullptr;
        size_t size = 0, live = 0;
        std::map<size_t, size_t> fr; /* offset -> length of free ranges */
    };
    std::vector<ArenaChunk> arena;
    uint64_t arena_held = 0, plane_live = 0;
    int wave_max = 1024;    /* triangles per grouped wave */
    void *na_scratch = nullptr; /* device address of the static scratch (allocation-free entry points) */
    /* scratch (grow-only) */
    float2 *G = nullptr; size_t G_bytes = 0;
    float *C = nullptr; size_t C_bytes = 0;
    int *Cq = nullptr; size_t Cq_bytes = 0;
    int *Cqx = nullptr; size_t Cqx_bytes = 0;   /* exact normalisation partials */
    void *outx = nullptr; size_t outx_bytes = 0; /* exact normalisation results */
    unsigned *margin = nullptr;
    void *out = nullptr; size_t out_bytes = 0;
    void *out2 = nullptr; size_t out2_bytes = 0; /* second result buffer of grouped waves */
    /* grouped waves copy results to the host on their own stream (default
       priority): ev_red marks a wave's reduction, ev_cp[k] the end of the
       copy out of result buffer k */
    hipStream_t cstream = nullptr;
    hipEvent_t ev_red = nullptr, ev_cp[2] = {nullptr, nullptr};
    void *wdesc = nullptr; size_t wdesc_bytes = 0; /* wave descriptors */
    /* lag-Fourier form: tile partials, the FP64 spectrum of one triangle,
       lag arrays of a wave (two buffers, as the grouped waves) */
    void *lag_part = nullptr; size_t lag_part_bytes = 0;
    void *lag_st = nullptr; size_t lag_st_bytes = 0;
    void *lag_out = nullptr; size_t lag_out_bytes = 0;
    void *lag_out2 = nullptr; size_t lag_out2_bytes = 0; /* second result buffer (copy overlap) */
    float lag_margin = 0.f; /* worst normalisation grid distance of the last lag-Fourier call */
    /* r1 == r2 tile tables of the lag-Fourier form by transform length:
       device block [tiles (I, J)] [nt x nt tile index] [k2 list of the
       k3 = L/2 column that no listed tile holds], tile count, bytes, list
       length (-1: the list was too long, every k2 is summed apart) */
    std::map<int, std::tuple<void *, int, size_t, int>> lag_pairs_tab;
    /* r2 == r3 tile tables by transform length: [tiles (a, b)] [nt2 x nt3
       tile index], tile count, bytes */
    std::map<int, std::tuple<void *, int, size_t>> lag_r23_tab;
    /* r1 == r2 == r3 sheared tiles per L: device list + map, tile count, bytes, fixed-point cells */
    std::map<int, std::tuple<void *, int, size_t, int>> lag_equi_tab;
    /* pinned host staging, two buffers: a grouped wave's results are copied
       out of one while the next wave's land in the other; the lag-space
       post-processing stages alternate triangles in them the same way */
    void *hstage[2] = {nullptr, nullptr}; size_t hstage_bytes[2] = {0, 0};
    /* stats */
    amb_stats st{};
    std::vector<double> batch_ms; /* per-triangle completion time of the last batch */
    std::vector<amb::Timer> timers;
    std::vector<hipEvent_t> evpool;
    std::vector<AmbGraph *> graphs;
    AmbServer *server = nullptr;
};

namespace amb {

int fail(amb_ctx *c, int code, const char *fmt, ...);
void timer_start(amb_ctx *c, double *acc);
void timer_stop(amb_ctx *c);
void timers_collect(amb_ctx *c);
int dev_alloc(amb_ctx *c, void **p, size_t bytes, const char *what);
void dev_free(amb_ctx *c, void *p, size_t bytes);
/* cached plane sets: from the stream-ordered pool on the context's stream */
int plane_alloc(amb_ctx *c, void **p, size_t bytes);
void plane_free(amb_ctx *c, void *p, size_t bytes);
int grow(amb_ctx *c, void **p, size_t *have, size_t need, const char *what);
int grow_host(amb_ctx *c, void **p, size_t *have, size_t need);
/* creates the copy stream and its events on first use */
int copy_stream(amb_ctx *c);
/* memcpy of several (dst, src, bytes) ranges split over host threads */
void par_copy(const std::vector<std::tuple<void *, const void *, size_t>> &r);
int twiddles(amb_ctx *c, int n, float2 **out, FftPlan *pl);
int find_cube(amb_ctx *c, int id, Cube **cb);
long plane_pitch(long ng);
void unpin_all(amb_ctx *c);
/* moves one call-pin of each logged plane to the serving queue (spin) or
   drops it (spin = false); unspin releases serving pins */
void pins_to_server(amb_ctx *c, const std::vector<void *> &log, bool spin);
void unspin(amb_ctx *c, const std::vector<void *> &log);
/* upload of host cube arrays into a cube's device buffers on stream s,
   staged through stg (stg_bytes); s is synchronised */
int upload_cube(amb_ctx *c, Cube &cb, const double *gv, const double *samp, hipStream_t s, void *stg, size_t stg_bytes);
/* frees a cube's device buffers and its planes and forgets it */
void release_cube_buffers(amb_ctx *c, int id);
/* amb_set_channels without its lock and state checks */
int set_channels_core(amb_ctx *c, int id, const int *uf, int uf_nc, int nc10);
/* planes of (cube, kind, ring r) at grid ng from src, pinned on return */
/* ngref (0: ng): the reference grid whose column cut j < ngref/3 selects
   the ring cells (amber_funcs.c:582); the set is computed on ng */
int get_planes(amb_ctx *c, const PlaneSrc &src, int cube_id, int kind, int r, int ng, void **out, int ngref = 0);
/* launches one recorded build on stream s with FFT scratch G */
hipError_t launch_build(amb_ctx *c, hipStream_t s, const BuildOp &op, float2 *G);
size_t build_scratch_bytes(amb_ctx *c, const BuildOp &op);
/* resolve a triangle: validate, pin its planes, fill kernel arguments */
int prepare_tri(amb_ctx *c, const PlaneSrc &src, int cube_id, int r1, int r2, int r3, int want_norm, TriJob &j);
/* evals and algorithmic flops of one finished triangle from its counts;
 * count_tri is thread-safe, account_counts adds to the context's stats */
void account_tri(amb_ctx *c, const TriJob &j, const int32_t *cnt);
void count_tri(const TriJob &j, const int32_t *cnt, int64_t *evals, int64_t *pairs, int64_t *diag);
void account_counts(amb_ctx *c, const TriJob &j, int64_t evals, int64_t pairs, int64_t diag);
/* the normalisation certificate of one call or wave (worst grid distance m):
 * recorded in the stats; an error if out of the int32 range of eighths, or
 * beyond 1/32 with cert_strict set */
int check_cert(amb_ctx *c, float m, bool *recompute);
/* the exact (FP64) v2 normalisation of one triangle into host arrays */
int norm_exact(amb_ctx *c, int id, int r1, int r2, int r3, float *norm, double *norm64, float *margin);
/* post-processing teardown at amb_destroy (nothing held today) */
void post_release(amb_ctx *c);
/* algorithmic HBM bytes of one c2r build (0 for derived kinds) */
double build_fft_bytes(amb_ctx *c, const BuildOp &op);
/* one triangle with its own launches, results in host memory */
int bispec_core(amb_ctx *c, int id, int r1, int r2, int r3, int want_norm, amb_result *res);
/* the launch configuration for a triangle of grid ng over nact planes at
 * the context's precision: the forced one, else the table's bucket, else
 * the nearest tuned bucket of that precision, else the defaults */
amb_tune_entry tune_cfg(amb_ctx *c, long ng, int nact);
/* kernel variant a launch over (nact, pitch) runs under configuration t */
int tune_variant(const amb_tune_entry &t, int prec, int nact, long pitch);
/* grouped waves of the requests (results into their buffers); done_ms may be null */
int run_grouped(amb_ctx *c, const amb_tri_req *req, int n, double t0, double *done_ms);
/* FP64 twiddles of length n, built on the device once per context */
int twiddles64(amb_ctx *c, int n, double2 **out);
/* lag-Fourier form (amb_lagft.hip): transform length and counts of a
 * cube's channel set (once), their release, the bytes of a spectra set at
 * grid ng, its build from the kind 0 set op.src, and triangles in waves
 * (bl64: per-request FP64 lag arrays, may be null) */
int lag_setup(amb_ctx *c, Cube &cb);
void lag_release(amb_ctx *c, Cube &cb);
size_t lag_spectra_bytes(const Cube &cb, long ng, bool fp64);
hipError_t lag_launch_spectra(amb_ctx *c, hipStream_t s, const BuildOp &op);
int lagft_run(amb_ctx *c, const amb_tri_req *req, int n, double t0, double *done_ms, double *const *bl64,
              double *const *norm64);
/* one triangle of the lag-Fourier form as the serving queue runs it: its
 * tiling, spectra sets (pinned when resolved) and work items */
struct LagTri {
    int cube = 0, r1 = 0, r2 = 0, r3 = 0;
    long ng = 0, P = 0; /* Ng(r1) of the reference; pixels of the compute grid */
    long ngc = 0;       /* the compute grid (amb_compute_grid, or ng) */
    int L = 0, nseg = 0, mode = 0, ntile = 0, nyq_sep = 0;
    void *X[3] = {nullptr, nullptr, nullptr};
    const int2 *tlist = nullptr;
    const int *tmap = nullptr;
    const int *nyql = nullptr;
    int nnyq = 0;
    size_t cells = 0;
};
int lag_prepare(amb_ctx *c, int cube, int r1, int r2, int r3, LagTri *t);
size_t lag_part_bytes(const LagTri &t);
long lag_items(const LagTri &t);
/* pixels of one lag-Fourier work item (a tile over one pixel segment) */
long lag_item_px();
/* the grid a triangle's sums run on: amb_compute_grid of its three rings
   when reduced grids are on (the default), else Ng(r1) */
long compute_grid(const amb_ctx *c, const Cube &cb, int r1, int r2, int r3);
size_t lag_desc_bytes();
size_t lag_st_cap(); /* the spectrum buffer of any transform length */
void lag_fill_range(amb_ctx *c, const LagTri &t, void *part, long i0, long item0, void *h);
hipError_t lag_launch_tiles(hipStream_t s, const void *d, int n, long total);
hipError_t lag_launch_finish(amb_ctx *c, hipStream_t s, const LagTri &t, const void *d_whole, void *part, void *St,
                             float *bl);
void lag_account(amb_ctx *c, const LagTri &t, int32_t *cnt);
/* calCR; noalloc: all device scratch from the static array, no allocation */
int closed_rings(amb_ctx *c, int id, int plane, double *rdata, int cap, int *rcount, bool noalloc);

} // namespace amb

#define HIPCHK(ctx, expr)                                                                    \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return amb::fail(ctx, AMB_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                            \
    } while (0)

/*
 * C-ABI implementation: device context, resident cubes, the filtered-plane
 * cache and the per-triangle orchestration that replaces the rp loop of
 * amber.c:187-326 / amber_v2.c:187-333.  Grouped waves, captured graphs
 * and the serving queue live in amb_wave.hip.
 *
 * HBM layout (one cube, nc channels, N grid, Nh = N/2+1):
 *   vis_t [nc][Nh][N] complex64 and samp_t [nc][Nh][N] float -- plane-major,
 *     column-contiguous so the c2r column pass gathers a ring column with
 *     unit stride;
 *   filtered planes [nact][pitch] float per (cube, kind, ring edges, Ng),
 *     pitch = Ng*Ng rounded up to 64 with a zero tail, nact = channels of
 *     the triplet set (flagged planes are never built: they never enter a
 *     triplet, amber_funcs.c:2207-2259);
 *   C [nact][nact][nact] float: per-p1 contraction results, reduced over p1
 *     into the [W][W] lag arrays of the result files.
 */
#include "amb_ctx.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

using Lock = std::lock_guard<std::recursive_mutex>;

/* Scratch of the allocation-free entry points: one array of
 * AMB_NOALLOC_SCRATCH_BYTES on every device that loads the library, held
 * until the process ends.  Host code never uses the array's name as an
 * address: its device address comes from hipGetSymbolAddress, once per
 * device (noalloc_scratch). */
static __device__ __align__(256) char scratch[AMB_NOALLOC_SCRATCH_BYTES];

namespace amb {

/* the device address of `scratch` on the current device `dev`, looked up
 * once per device; null if the lookup failed */
static void *noalloc_scratch(int dev)
{
    static std::mutex mu;
    static std::map<int, void *> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
    }
    addr[dev] = p;
    return p;
}

/* ------------------------------------------------------------------ helpers */
int fail(amb_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
    if (c->verbose) fprintf(stderr, "amber_mi355x: %s\n", buf);
    return code;
}

static hipEvent_t ev_get(amb_ctx *c)
{
    if (!c->evpool.empty()) {
        hipEvent_t e = c->evpool.back();
        c->evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void timer_start(amb_ctx *c, double *acc)
{
    Timer t;
    t.a = ev_get(c);
    t.b = nullptr;
    t.acc = acc;
    if (t.a) hipEventRecord(t.a, c->stream);
    c->timers.push_back(t);
}

void timer_stop(amb_ctx *c)
{
    for (auto it = c->timers.rbegin(); it != c->timers.rend(); ++it)
        if (!it->b) {
            it->b = ev_get(c);
            if (it->b) hipEventRecord(it->b, c->stream);
            return;
        }
}

/* fold recorded intervals whose end has passed into the stats (after a
   stream sync: all of them); intervals still running stay recorded */
void timers_collect(amb_ctx *c)
{
    std::vector<Timer> keep;
    for (auto &t : c->timers) {
        if (t.b && hipEventQuery(t.b) == hipErrorNotReady) {
            keep.push_back(t);
            continue;
        }
        if (t.a && t.b) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess && t.acc) *t.acc += ms;
        }
        if (t.a) c->evpool.push_back(t.a);
        if (t.b) c->evpool.push_back(t.b);
    }
    (void)hipGetLastError();
    c->timers.swap(keep);
}

void dev_free(amb_ctx *c, void *p, size_t bytes)
{
    if (p) {
        hipFree(p);
        c->used -= bytes;
    }
}

/* evict unpinned filtered planes, least recently used first */
static bool evict_one(amb_ctx *c)
{
    int best = -1;
    for (size_t i = 0; i < c->planes.size(); ++i)
        if (c->planes[i].pins == 0 && c->planes[i].spins == 0 &&
            (best < 0 || c->planes[i].last < c->planes[best].last))
            best = (int)i;
    if (best < 0) return false;
    plane_free(c, c->planes[best].d, c->planes[best].bytes);
    c->planes.erase(c->planes.begin() + best);
    c->st.n_evict_planes++;
    return true;
}

namespace {

constexpr size_t ARENA_CHUNK = (size_t)4 << 30; /* default chunk of the plane arena */
constexpr size_t ARENA_ALIGN = 4096;

size_t arena_round(size_t b) { return (b + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN; }

/* device memory the context holds: live allocations other than planes, and
   every arena chunk */
uint64_t held_bytes(amb_ctx *c) { return c->used - c->plane_live + c->arena_held; }

/* returns the arena's empty chunks to the device (once the stream has
   drained: their last readers may still run) */
bool arena_release_empty(amb_ctx *c)
{
    bool any = false;
    for (size_t i = 0; i < c->arena.size();) {
        if (c->arena[i].live == 0) {
            if (!any) hipStreamSynchronize(c->stream);
            hipFree(c->arena[i].base);
            c->arena_held -= c->arena[i].size;
            c->arena.erase(c->arena.begin() + i);
            any = true;
        } else {
            ++i;
        }
    }
    return any;
}

/* first fit over the chunks' free ranges */
bool arena_take(amb_ctx *c, size_t bytes, void **p)
{
    for (auto &k : c->arena)
        for (auto it = k.fr.begin(); it != k.fr.end(); ++it)
            if (it->second >= bytes) {
                const size_t off = it->first, rest = it->second - bytes;
                k.fr.erase(it);
                if (rest) k.fr[off + bytes] = rest;
                k.live += bytes;
                *p = k.base + off;
                return true;
            }
    return false;
}

} // namespace

int dev_alloc(amb_ctx *c, void **p, size_t bytes, const char *what)
{
    *p = nullptr;
    if (bytes == 0) return AMB_OK;
    while (held_bytes(c) + bytes > c->budget)
        if (!arena_release_empty(c) && !evict_one(c))
            return fail(c, AMB_EBUDGET, "%s: %zu bytes exceed the HBM budget (%llu live, %llu held of %llu)", what,
                        bytes, (unsigned long long)c->used, (unsigned long long)held_bytes(c),
                        (unsigned long long)c->budget);
    hipError_t e = hipMalloc(p, bytes);
    while (e != hipSuccess) { /* the device is full outside the budget's view: give back what the arena holds */
        (void)hipGetLastError();
        if (!arena_release_empty(c) && !evict_one(c)) break;
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(c, AMB_EBUDGET, "%s: hipMalloc(%zu) failed: %s", what, bytes, hipGetErrorString(e));
    }
    c->used += bytes;
    return AMB_OK;
}

int plane_alloc(amb_ctx *c, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return AMB_OK;
    bytes = arena_round(bytes);
    while (c->used + bytes > c->budget)
        if (!evict_one(c))
            return fail(c, AMB_EBUDGET, "filtered planes: %zu bytes exceed the HBM budget (%llu used of %llu)", bytes,
                        (unsigned long long)c->used, (unsigned long long)c->budget);
    while (!arena_take(c, bytes, p)) {
        /* a new chunk (the default size, else the request alone) within the
           budget; else empty chunks go back and planes are evicted */
        bool grown = false;
        for (size_t cs : {std::max(ARENA_CHUNK, bytes), bytes}) {
            if (held_bytes(c) + cs > c->budget) continue;
            void *q = nullptr;
            if (hipMalloc(&q, cs) != hipSuccess) {
                (void)hipGetLastError();
                continue;
            }
            amb_ctx::ArenaChunk k;
            k.base = (char *)q;
            k.size = cs;
            k.fr[0] = cs;
            c->arena.push_back(std::move(k));
            c->arena_held += cs;
            grown = true;
            break;
        }
        if (grown || arena_release_empty(c) || evict_one(c)) continue;
        return fail(c, AMB_EBUDGET, "filtered planes: %zu bytes fit neither the budget nor the device (%llu held of %llu)",
                    bytes, (unsigned long long)held_bytes(c), (unsigned long long)c->budget);
    }
    c->used += bytes;
    c->plane_live += bytes;
    return AMB_OK;
}

void plane_free(amb_ctx *c, void *p, size_t bytes)
{
    if (!p) return;
    bytes = arena_round(bytes);
    for (auto &k : c->arena)
        if ((char *)p >= k.base && (char *)p < k.base + k.size) {
            auto it = k.fr.emplace((size_t)((char *)p - k.base), bytes).first;
            auto nx = std::next(it);
            if (nx != k.fr.end() && it->first + it->second == nx->first) {
                it->second += nx->second;
                k.fr.erase(nx);
            }
            if (it != k.fr.begin()) {
                auto pv = std::prev(it);
                if (pv->first + pv->second == it->first) {
                    pv->second += it->second;
                    k.fr.erase(it);
                }
            }
            k.live -= bytes;
            break;
        }
    c->used -= bytes;
    c->plane_live -= bytes;
}

/* grow-only device scratch */
int grow(amb_ctx *c, void **p, size_t *have, size_t need, const char *what)
{
    if (*have >= need) return AMB_OK;
    dev_free(c, *p, *have);
    *p = nullptr;
    *have = 0;
    int rc = dev_alloc(c, p, need, what);
    if (rc) return rc;
    *have = need;
    return AMB_OK;
}

/* grow-only pinned host staging */
int copy_stream(amb_ctx *c)
{
    if (c->cstream) return AMB_OK;
    HIPCHK(c, hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_red, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_cp[0], hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_cp[1], hipEventDisableTiming));
    return AMB_OK;
}

void par_copy(const std::vector<std::tuple<void *, const void *, size_t>> &r)
{
    size_t tot = 0;
    for (const auto &x : r) tot += std::get<2>(x);
    const int nth = tot < ((size_t)4 << 20) ? 1 : (int)std::min<size_t>(16, std::max<size_t>(2, tot >> 22));
    auto part = [&](int t) {
        for (const auto &x : r) {
            const size_t b = std::get<2>(x), lo = b * t / nth, hi = b * (t + 1) / nth;
            if (hi > lo) memcpy((char *)std::get<0>(x) + lo, (const char *)std::get<1>(x) + lo, hi - lo);
        }
    };
    if (nth == 1) {
        part(0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 1; t < nth; ++t) th.emplace_back(part, t);
    part(0);
    for (auto &x : th) x.join();
}

int grow_host(amb_ctx *c, void **p, size_t *have, size_t need)
{
    if (*have >= need) return AMB_OK;
    /* pinning is slow: grow by half again at least, so a run of growing
       waves re-pins a few times only */
    need = std::max(need, *have + *have / 2);
    if (*p) hipHostFree(*p);
    *p = nullptr;
    *have = 0;
    if (hipHostMalloc(p, need, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(c, AMB_ENOMEM, "pinned host staging of %zu bytes", need);
    }
    *have = need;
    return AMB_OK;
}

template <class T>
static int scratch(amb_ctx *c, T **p, size_t *have, size_t need, const char *what)
{
    void *q = *p;
    int rc = grow(c, &q, have, need, what);
    *p = (T *)q;
    return rc;
}

int twiddles(amb_ctx *c, int n, float2 **out, FftPlan *pl)
{
    auto it = c->tw.find(n);
    if (it != c->tw.end()) {
        *out = it->second;
        *pl = c->plans[n];
        return AMB_OK;
    }
    std::vector<float2> h(n);
    for (int m = 0; m < n; ++m) {
        double a = 2.0 * M_PI * (double)m / (double)n;
        h[m] = make_float2((float)cos(a), (float)sin(a));
    }
    void *d = nullptr;
    int rc = dev_alloc(c, &d, sizeof(float2) * n, "twiddle table");
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    FftPlan p;
    amb_make_plan(n, &p);
    c->tw[n] = (float2 *)d;
    c->plans[n] = p;
    *out = (float2 *)d;
    *pl = p;
    return AMB_OK;
}

static int ring_geom(amb_ctx *c, const Cube &cb, int r, int ng, RingGeom *g, int ngref)
{
    if (r < 0 || r >= (int)c->ulow.size()) return fail(c, AMB_EARG, "ring %d outside [0,%zu)", r, c->ulow.size());
    g->lo = c->ulow[r];
    g->hi = c->uhigh[r];
    g->ng = ng;
    if (!(g->hi > g->lo)) return fail(c, AMB_EARG, "ring %d has UHigh <= ULow", r);
    /* cells of the ring must lie in the source half plane: |ia| < UHigh and
       j < UHigh need UHigh <= N/2 (rows beyond N/2 would alias, and the
       reference reads outside the cube there) */
    if (g->hi > (double)(cb.N / 2))
        return fail(c, AMB_EARG, "ring %d UHigh=%g exceeds N/2=%lld", r, g->hi, (long long)(cb.N / 2));
    long jc = (long)ceil(g->hi);
    if (jc > ngref / 3) jc = ngref / 3;
    if (jc < 0) jc = 0;
    g->jcnt = (int)jc;
    /* integer radius-squared bounds of the half-open ring, uu > 0 included:
       sqrt is monotone, so each test is a threshold on s = ia^2 + j^2 */
    auto first_ge = [](double r, long from) {
        long s = std::max(from, (long)floor(r > 0 ? r * r : 0.0) - 4);
        while (sqrt((double)s) < r) ++s;
        return s;
    };
    g->s_lo = first_ge(g->lo, 1);
    g->s_hi = std::max(first_ge(g->hi, 0), g->s_lo);
    return AMB_OK;
}

long plane_pitch(long ng) { return (ng * ng + 63) / 64 * 64; }

static size_t plane_bytes(int kind, int nact, size_t pitch)
{
    if (kind == 4) return pitch * nact;
    if (kind == 5) return sizeof(float) * 2 * nact;
    if (kind == 0 || kind == 11) return sizeof(float) * pitch * nact + AMB_SET_SLOT_BYTES; /* + [max, E] */
    if (kind == 8) return sizeof(double) * pitch * nact;
    return sizeof(float) * pitch * nact; /* FP32, or hi + lo BF16 / FP16 */
}

size_t build_scratch_bytes(amb_ctx *c, const BuildOp &op)
{
    if (op.kind > 1 && op.kind != 8 && op.kind != 11) return 0;
    const Cube &cb = c->cubes[op.cube];
    const size_t cb8 = op.kind == 8 ? sizeof(double2) : sizeof(float2);
    return cb8 * (size_t)cb.nact() * (size_t)std::max(op.g.jcnt, 1) * op.g.ng;
}

/* FP64 twiddles of length n (the exact normalisation's c2r), built on the
 * device once per context */
int twiddles64(amb_ctx *c, int n, double2 **out)
{
    auto it = c->tw64.find(n);
    if (it != c->tw64.end()) {
        if (out) *out = it->second;
        return AMB_OK;
    }
    void *d = nullptr;
    int rc = dev_alloc(c, &d, sizeof(double2) * n, "FP64 twiddle table");
    if (rc) return rc;
    HIPCHK(c, amb_launch_twiddles64(c->stream, (double2 *)d, n));
    c->tw64[n] = (double2 *)d;
    if (out) *out = (double2 *)d;
    return AMB_OK;
}

hipError_t launch_build(amb_ctx *c, hipStream_t s, const BuildOp &op, float2 *G)
{
    const Cube &cb = c->cubes[op.cube];
    const int nact = cb.nact();
    const long pitch = plane_pitch(op.g.ng);
    switch (op.kind) {
    case 0:
    case 1:
    case 11: {
        const float2 *tw = c->tw[op.g.ng];
        const FftPlan &pl = c->plans[op.g.ng];
        /* data sets are normalised in place (kind 11: left as computed), their
           [max, E] words after the planes */
        unsigned *slot = op.kind != 1 ? (unsigned *)((float *)op.dst + (size_t)pitch * nact) : nullptr;
        return amb_launch_ring_c2r(s, op.kind == 11 ? 0 : op.kind, cb.vis_t, cb.samp_t, cb.d_act, nact, (int)cb.N,
                                   (int)(cb.N / 2 + 1), pl, tw, op.g, G, (float *)op.dst, pitch, slot, op.kind == 11);
    }
    case 2:
    case 3: {
        __bf16 *h = (__bf16 *)op.dst;
        return amb_launch_split(s, (const float *)op.src, (size_t)pitch * nact, h, h + (size_t)pitch * nact);
    }
    case 4:
        return amb_launch_fp8_quant(s, (const float *)op.src, pitch, nact, (const float *)op.sc + nact,
                                    (uint8_t *)op.dst);
    case 5:
        return amb_launch_fp8_scales(s, (const float *)op.src, pitch, nact, (float *)op.dst);
    case 9:
    case 10:
        return lag_launch_spectra(c, s, op);
    case 8:
        return amb_launch_ring_c2r64(s, cb.samp_t, cb.d_act, nact, (int)cb.N, (int)(cb.N / 2 + 1),
                                     c->plans[op.g.ng], c->tw64[op.g.ng], op.g, (double2 *)G, (double *)op.dst,
                                     pitch);
    default: {
        _Float16 *h = (_Float16 *)op.dst;
        return amb_launch_split_f16(s, (const float *)op.src, (size_t)pitch * nact, h, h + (size_t)pitch * nact);
    }
    }
}

/* algorithmic bytes of a c2r build: Ng^2 floats written per plane plus the
   ring cells read */
double build_fft_bytes(amb_ctx *c, const BuildOp &op)
{
    const Cube &cb = c->cubes[op.cube];
    if (op.kind == 9 || op.kind == 10) /* the image set read, the half spectra written */
        return (double)op.g.ng * op.g.ng * (op.kind == 10 ? 2.0 : 1.0) * (4.0 * cb.nact() + 8.0 * (cb.lag_L / 2 + 1));
    if (op.kind > 1 && op.kind != 11) return 0.0;
    const RingGeom &g = op.g;
    double cells = 0;
    for (long ia = -(long)ceil(g.hi); ia <= (long)ceil(g.hi); ++ia)
        for (long j = 0; j < g.jcnt; ++j) {
            double uu = sqrt((double)(ia * ia + j * j));
            if (uu >= g.lo && uu < g.hi && uu > 0) cells += 1;
        }
    return (double)cb.nact() * (4.0 * (double)g.ng * g.ng + (op.kind != 1 ? 8.0 : 4.0) * cells);
}

int get_planes(amb_ctx *c, const PlaneSrc &src, int cube_id, int kind, int r, int ng, void **out, int ngref)
{
    Cube *cb;
    int rc = find_cube(c, cube_id, &cb);
    if (rc) return rc;
    if (ngref <= 0) ngref = ng;
    RingGeom g;
    if ((rc = ring_geom(c, *cb, r, ng, &g, ngref))) return rc;
    /* columns j < jcnt <= ng/2: the row pass fills columns k and ng - k of
       its Hermitian extension from column k < jcnt, which must not meet */
    if (g.jcnt > ng / 2) return fail(c, AMB_EARG, "ring %d needs %d columns on a grid of %d", r, g.jcnt, ng);
    for (auto &p : *src.list)
        if (p.cube == cube_id && p.kind == kind && p.ng == ng && p.jcnt == g.jcnt && p.lo == g.lo && p.hi == g.hi) {
            p.last = ++c->tick;
            p.pins++;
            if (c->pin_log && src.list == &c->planes) c->pin_log->push_back(p.d);
            *out = p.d;
            return AMB_OK;
        }
    BuildOp op;
    op.kind = kind;
    op.cube = cube_id;
    op.g = g;
    void *dep = nullptr, *sc = nullptr;
    if (kind == 2 || kind == 3) {
        if ((rc = get_planes(c, src, cube_id, kind - 2, r, ng, &dep, ngref))) return rc;
    } else if (kind == 8) {
        /* two FP64 transform buffers in 64 KiB of LDS */
        if (32L * ng > 65536) return fail(c, AMB_EARG, "grid %d exceeds the FP64 LDS transform size (2048)", ng);
        float2 *tw;
        FftPlan pl;
        double2 *tw64;
        if ((rc = twiddles(c, ng, &tw, &pl)) || (rc = twiddles64(c, ng, &tw64))) return rc;
    } else if (kind == 9 || kind == 10) {
        /* channel spectra of the data set (FP32 images) or of the sampling
           set (FP64 images), with the channel transform's twiddles */
        if (src.record || src.bump) return fail(c, AMB_EARG, "lag-Fourier planes are not built inside graphs");
        if ((rc = lag_setup(c, *cb))) return rc;
        if ((rc = get_planes(c, src, cube_id, kind == 9 ? 11 : 8, r, ng, &dep, ngref))) return rc;
        float2 *tw;
        FftPlan pl;
        if ((rc = twiddles(c, cb->lag_L, &tw, &pl))) return rc;
        if (kind == 10 && (rc = twiddles64(c, cb->lag_L, nullptr))) return rc;
    } else if (kind >= 4 && kind != 11) {
        if ((rc = get_planes(c, src, cube_id, 0, r, ng, &dep, ngref))) return rc;
        if (kind == 4 && (rc = get_planes(c, src, cube_id, 5, r, ng, &sc, ngref))) return rc;
    } else {
        /* one transform of ng complex values plus twiddles and a second
           buffer must fit a workgroup's 64 KiB of LDS */
        if (24L * ng > 65536) return fail(c, AMB_EARG, "grid %d exceeds the LDS transform size (2730)", ng);
        float2 *tw;
        FftPlan pl;
        if ((rc = twiddles(c, ng, &tw, &pl))) return rc;
    }
    op.src = dep;
    op.sc = sc;
    const size_t bytes = kind == 9 || kind == 10 ? lag_spectra_bytes(*cb, ng, kind == 10)
                                                 : plane_bytes(kind, cb->nact(), plane_pitch(ng));
    void *d = nullptr;
    if (src.bump) {
        const size_t need = (bytes + 255) / 256 * 256;
        if (src.bump->off + need > src.bump->cap)
            return fail(c, AMB_EBUDGET, "plane arena: %zu + %zu bytes exceed %zu", src.bump->off, need, src.bump->cap);
        d = (void *)((uintptr_t)src.bump->base + src.bump->off); /* base 0: measuring only */
        src.bump->off += need;
    } else if ((rc = plane_alloc(c, &d, bytes))) {
        return rc;
    }
    op.dst = d;
    if (src.record) {
        src.record->push_back(op);
    } else {
        const bool c2r = kind <= 1 || kind == 11, spectra = kind == 9 || kind == 10;
        if ((c2r || kind == 8) && (rc = scratch(c, &c->G, &c->G_bytes, build_scratch_bytes(c, op), "fft scratch"))) {
            plane_free(c, d, bytes);
            return rc;
        }
        if (c2r || spectra) {
            timer_start(c, &c->st.ms_fft);
            c->st.n_fft_launch += spectra ? 1 : 2;
            c->st.bytes_fft += build_fft_bytes(c, op);
        }
        hipError_t e = launch_build(c, c->stream, op, c->G);
        if (c2r || spectra) timer_stop(c);
        if (e != hipSuccess) {
            plane_free(c, d, bytes);
            return fail(c, AMB_EHIP, "plane build (kind %d): %s", kind, hipGetErrorString(e));
        }
    }
    PlaneSet p;
    p.cube = cube_id; p.kind = kind; p.ng = ng; p.jcnt = g.jcnt; p.lo = g.lo; p.hi = g.hi;
    p.d = d; p.bytes = bytes; p.last = ++c->tick; p.pins = 1;
    src.list->push_back(p);
    if (c->pin_log && src.list == &c->planes) c->pin_log->push_back(d);
    *out = d;
    return AMB_OK;
}

void unpin_all(amb_ctx *c)
{
    for (auto &p : c->planes) p.pins = 0;
}

static PlaneSet *plane_of(amb_ctx *c, const void *d)
{
    for (auto &p : c->planes)
        if (p.d == d) return &p;
    return nullptr;
}

void pins_to_server(amb_ctx *c, const std::vector<void *> &log, bool spin)
{
    for (void *d : log)
        if (PlaneSet *p = plane_of(c, d)) {
            if (p->pins > 0) p->pins--;
            if (spin) p->spins++;
        }
}

void unspin(amb_ctx *c, const std::vector<void *> &log)
{
    for (void *d : log)
        if (PlaneSet *p = plane_of(c, d))
            if (p->spins > 0) p->spins--;
}

int find_cube(amb_ctx *c, int id, Cube **cb)
{
    auto it = c->cubes.find(id);
    if (it == c->cubes.end()) return fail(c, AMB_ESTATE, "unknown cube id %d", id);
    *cb = &it->second;
    return AMB_OK;
}

/* planes the serving queue holds stay (they are evicted once released) */
static void drop_cube_planes(amb_ctx *c, int id)
{
    for (size_t i = 0; i < c->planes.size();)
        if (c->planes[i].cube == id && c->planes[i].spins == 0) {
            plane_free(c, c->planes[i].d, c->planes[i].bytes);
            c->planes.erase(c->planes.begin() + i);
        } else ++i;
}

long compute_grid(const amb_ctx *c, const Cube &cb, int r1, int r2, int r3)
{
    const long ng = amb_ng(c->uhigh[r1], cb.Du);
    /* the split-BF16 and FP8 forms stay on Ng: their error is the operands'
       representation error per product, which does not shrink with the
       grid, so against B (n^2 times the closed-triangle sum) it grows as
       Ng/n; the FP32, FP16x3 and lag-Fourier errors, dominated by the
       accumulation, shrink with it */
    if (!c->grid_reduced || c->prec == AMB_PREC_BF16X3 || c->prec == AMB_PREC_FP8) return ng;
    return amb_compute_grid(c->uhigh[r1], c->uhigh[r2], c->uhigh[r3], cb.Du);
}

int prepare_tri(amb_ctx *c, const PlaneSrc &src, int id, int r1, int r2, int r3, int want_norm, TriJob &j)
{
    Cube *cb;
    int rc = find_cube(c, id, &cb);
    if (rc) return rc;
    if (!cb->chans) return fail(c, AMB_ESTATE, "channels not set for cube %d", id);
    const int nr = (int)c->ulow.size();
    if (nr == 0) return fail(c, AMB_ESTATE, "rings not set");
    if (r1 < 0 || r1 >= nr || r2 < 0 || r2 >= nr || r3 < 0 || r3 >= nr)
        return fail(c, AMB_EARG, "triangle (%d,%d,%d) outside [0,%d)", r1, r2, r3, nr);
    const long ng = amb_ng(c->uhigh[r1], cb->Du);
    if (ng <= 0 || ng > 2730) return fail(c, AMB_EARG, "Ng=%ld out of range", ng);
    j = TriJob{};
    j.cube = id; j.r1 = r1; j.r2 = r2; j.r3 = r3; j.want_norm = want_norm;
    /* planes and sums on the compute grid (the ring cells and the column
       cut of Ng), the sums scaled to Ng by the reduction */
    const long ngc = compute_grid(c, *cb, r1, r2, r3);
    j.ng = ng; j.ngc = ngc; j.P = ngc * ngc; j.pitch = plane_pitch(ngc);
    j.nact = cb->nact();
    j.W = 2 * cb->nc10 + 1;
    j.cells = (size_t)j.W * j.W;
    const int prec = c->prec, nact = j.nact;
    const int nprec = prec == AMB_PREC_FP32 ? AMB_PREC_FP32 : AMB_PREC_BF16X3; /* normalisation form */
    /* D/S: FP32 images of r1, r2, r3; the third operand in the form the
       contraction reads (FP32, the hi/lo BF16 split, or scaled e4m3) */
    void *D[3], *S[3] = {nullptr, nullptr, nullptr}, *SC[3] = {nullptr, nullptr, nullptr}, *D0 = nullptr;
    const int rr[3] = {r1, r2, r3};
    const int d3kind = prec == AMB_PREC_FP32 ? 0 : prec == AMB_PREC_BF16X3 ? 2 : prec == AMB_PREC_FP16X3 ? 7 : 4;
    for (int k = 0; k < 3; ++k) {
        if ((rc = get_planes(c, src, id, k == 2 ? d3kind : 0, rr[k], (int)ngc, &D[k], (int)ng))) return rc;
        if (prec == AMB_PREC_FP8 && (rc = get_planes(c, src, id, 5, rr[k], (int)ngc, &SC[k], (int)ng))) return rc;
    }
    /* the FP32 set of r3, for its exponent (built already as the source of
       the derived third operand) */
    if ((rc = get_planes(c, src, id, 0, r3, (int)ngc, &D0, (int)ng))) return rc;
    if (want_norm)
        for (int k = 0; k < 3; ++k) {
            const int kind = (k == 2 && nprec != AMB_PREC_FP32) ? 3 : 1;
            if ((rc = get_planes(c, src, id, kind, rr[k], (int)ngc, &S[k], (int)ng))) return rc;
        }
    /* r1 == r2 symmetry needs every p3 of p1's window inside p2's window */
    const bool full_window = nact > 0 && cb->act.back() - cb->act.front() <= cb->nc10;
    j.sym12 = (r1 == r2 && full_window && !c->no_sym) ? 1 : 0;
    /* r2 == r3: C[p1][p2][p3] and C[p1][p3][p2] are the same sum with the
       roles of two planes exchanged (equal up to rounding), and the window
       of p1 is the same for p2 and p3; with r1 == r2 as well only the exact
       r1 == r2 symmetry is used */
    j.sym23 = (r2 == r3 && r1 != r2 && !c->no_sym) ? 1 : 0;
    ContractArgs &a = j.a;
    a.P = j.P; a.pitch = j.pitch; a.nact = nact; a.wlo = cb->d_wlo; a.whi = cb->d_whi; a.sym12 = j.sym12;
    a.sym23 = j.sym23;
    a.D1 = (const float *)D[0]; a.D2 = (const float *)D[1];
    if (prec == AMB_PREC_FP32) a.D3 = (const float *)D[2];
    else if (prec == AMB_PREC_BF16X3 || prec == AMB_PREC_FP16X3) {
        a.H3 = (const uint16_t *)D[2];
        a.L3 = a.H3 + (size_t)j.pitch * nact;
    } else {
        a.Q3 = (const uint8_t *)D[2];
        a.s1 = (const float *)SC[0];
        a.s2 = (const float *)SC[1];
        a.t3 = (const float *)SC[2] + nact;
    }
    /* exponents of the normalised data sets: the second word after the planes */
    const size_t setw = (size_t)j.pitch * nact;
    a.ex1 = (const int *)((const float *)D[0] + setw) + 1;
    a.ex2 = (const int *)((const float *)D[1] + setw) + 1;
    a.ex3 = (const int *)((const float *)D0 + setw) + 1;
    if (want_norm) {
        ContractArgs &b = j.b;
        b = a;
        b.D1 = (const float *)S[0]; b.D2 = (const float *)S[1];
        b.D3 = nullptr; b.H3 = b.L3 = nullptr; b.Q3 = nullptr; b.s1 = b.s2 = b.t3 = nullptr;
        b.ex1 = b.ex2 = b.ex3 = nullptr;
        if (nprec == AMB_PREC_FP32) b.D3 = (const float *)S[2];
        else { b.H3 = (const uint16_t *)S[2]; b.L3 = b.H3 + (size_t)j.pitch * nact; }
    }
    ReduceArgs &ra = j.ra;
    ra.act = cb->d_act; ra.inv = cb->d_inv;
    ra.nc = (int)cb->nc; ra.nact = nact; ra.nc10 = cb->nc10; ra.P = j.P; ra.sym12 = j.sym12; ra.sym23 = j.sym23;
    ra.gscale = (double)(ng * ng) / (double)j.P;
    ra.Pout = (double)(ng * ng);
    j.flop_exec = (double)amb_contract_live_items(nact, cb->wlo.data(), cb->whi.data(), j.sym12, j.sym23) *
                  amb_contract_item_flops(prec, j.pitch);
    return AMB_OK;
}

void count_tri(const TriJob &j, const int32_t *cnt, int64_t *evals, int64_t *pairs, int64_t *diag)
{   /* integer sums: exact, and free of a floating-point dependency chain */
    int64_t ev = 0, pr = 0, dg = 0;
    for (size_t i = 0; i < j.cells; ++i) {
        ev += cnt[i] > 0;
        pr += cnt[i];
    }
    /* triplets a symmetry maps onto themselves: p1 == p2 (lag row d2 = 0)
       for r1 == r2, p2 == p3 (the d2 == d3 diagonal) for r2 == r3 */
    const size_t W = (size_t)j.W, mid = W / 2;
    if (j.sym12)
        for (size_t k = 0; k < W; ++k) dg += cnt[mid * W + k];
    else if (j.sym23)
        for (size_t k = 0; k < W; ++k) dg += cnt[k * W + k];
    *evals = ev;
    *pairs = pr;
    *diag = dg;
}

void account_counts(amb_ctx *c, const TriJob &j, int64_t evals, int64_t pairs, int64_t diag)
{
    c->st.evals += evals;
    c->st.flop_contract += 2.0 * (double)pairs * (double)j.P;
    /* distinct sums: under a symmetry the triplets pair up, except the
       diag ones that map onto themselves, leaving (pairs + diag) / 2 */
    const double distinct = (j.sym12 || j.sym23) ? 0.5 * (double)(pairs + diag) : (double)pairs;
    c->st.flop_contract_distinct += 2.0 * distinct * (double)j.P;
    c->st.flop_contract_exec += j.flop_exec;
}

int check_cert(amb_ctx *c, float m, bool *recompute)
{
    *recompute = false;
    if (!(m < AMB_CERT_RANGE)) /* also NaN */
        return fail(c, AMB_ECERT, "normalisation sum beyond the int32 range of eighths");
    if (m > c->st.cert_margin_max) c->st.cert_margin_max = m;
    if (m > 1.0f / 32.0f) {
        ++c->st.n_uncertified;
        if (c->cert_strict) return fail(c, AMB_ECERT, "normalisation sums off their 1/8 grid by %g", (double)m);
        *recompute = true;
    }
    return AMB_OK;
}

/* The exact v2 normalisation of one triangle: FP64 sampling images, FP64
 * MFMA products, whole-range FP64 sums rounded to their 1/8 grid, reduced
 * over p1 as the fast path does.  norm (f32) / norm64 (f64): host arrays
 * of W*W, either may be null.  Uses its own scratch and margin word, so it
 * may run while a grouped wave's buffers are still live. */
int norm_exact(amb_ctx *c, int id, int r1, int r2, int r3, float *norm, double *norm64, float *margin)
{
    Cube *cb;
    int rc = find_cube(c, id, &cb);
    if (rc) return rc;
    const long ng = amb_ng(c->uhigh[r1], cb->Du), ngc = compute_grid(c, *cb, r1, r2, r3);
    const long P = ngc * ngc, pitch = plane_pitch(ngc);
    const int nact = cb->nact();
    const size_t W = (size_t)(2 * cb->nc10 + 1), cells = W * W, nC = (size_t)nact * nact * nact;
    const PlaneSrc src{&c->planes};
    void *S[3];
    const int rr[3] = {r1, r2, r3};
    for (int k = 0; k < 3; ++k)
        if ((rc = get_planes(c, src, id, 8, rr[k], (int)ngc, &S[k], (int)ng))) { unpin_all(c); return rc; }
    void *Cq = c->Cqx, *out = c->outx;
    if ((rc = grow(c, &Cq, &c->Cqx_bytes, sizeof(int) * nC, "exact normalisation partials")) ||
        (rc = grow(c, &out, &c->outx_bytes, cells * (sizeof(float) + sizeof(double)), "exact normalisation"))) {
        c->Cqx = (int *)Cq; c->outx = out;
        unpin_all(c);
        return rc;
    }
    c->Cqx = (int *)Cq; c->outx = out;
    double *d_n64 = (double *)out;
    float *d_n = (float *)(d_n64 + cells);
    const bool full_window = nact > 0 && cb->act.back() - cb->act.front() <= cb->nc10;
    ContractArgs64 a{};
    a.D1 = (const double *)S[0]; a.D2 = (const double *)S[1]; a.D3 = (const double *)S[2];
    a.P = P; a.pitch = pitch; a.nact = nact;
    a.sym12 = (r1 == r2 && full_window && !c->no_sym) ? 1 : 0;
    a.sym23 = (r2 == r3 && r1 != r2 && !c->no_sym) ? 1 : 0;
    a.wlo = cb->d_wlo; a.whi = cb->d_whi;
    a.Cq = c->Cqx;
    a.margin = c->margin + 1; /* the fast path uses word 0 */
    ReduceArgs ra{};
    ra.C = nullptr; ra.Cq = c->Cqx;
    ra.act = cb->d_act; ra.inv = cb->d_inv;
    ra.nc = (int)cb->nc; ra.nact = nact; ra.nc10 = cb->nc10; ra.P = P; ra.sym12 = a.sym12; ra.sym23 = a.sym23;
    ra.Pout = (double)(ng * ng);
    ra.norm = d_n; ra.norm64 = d_n64;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    unsigned mbits = 0;
    hipError_t e = amb_launch_fill_u32(c->stream, a.margin, 1, 0u);
    timer_start(c, &c->st.ms_norm);
    if (e == hipSuccess) e = amb_launch_norm64(c->stream, a, 4 * cus);
    timer_stop(c);
    if (e == hipSuccess) e = amb_launch_reduce(c->stream, ra);
    if (e == hipSuccess && norm64) e = hipMemcpyAsync(norm64, d_n64, sizeof(double) * cells, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && norm) e = hipMemcpyAsync(norm, d_n, sizeof(float) * cells, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&mbits, a.margin, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream);
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = es;
    timers_collect(c);
    unpin_all(c);
    if (e != hipSuccess)
        return fail(c, AMB_EHIP, "exact normalisation (%d,%d,%d) Ng=%ld: %s", r1, r2, r3, ng, hipGetErrorString(e));
    float m;
    memcpy(&m, &mbits, sizeof m);
    if (margin) *margin = m;
    ++c->st.n_norm_exact;
    if (!(m <= 1.0f / 32.0f))
        return fail(c, AMB_ECERT, "exact normalisation (%d,%d,%d): FP64 sums off their 1/8 grid by %g", r1, r2, r3,
                    (double)m);
    return AMB_OK;
}

void account_tri(amb_ctx *c, const TriJob &j, const int32_t *cnt)
{
    int64_t ev, pr, dg;
    count_tri(j, cnt, &ev, &pr, &dg);
    account_counts(c, j, ev, pr, dg);
}

/* one triangle with its own launches; output pointers are host memory */
int bispec_core(amb_ctx *c, int id, int r1, int r2, int r3, int want_norm, amb_result *res)
{
    hipSetDevice(c->device);
    if (c->prec == AMB_PREC_LAGFT) {
        Cube *cb;
        int rc = find_cube(c, id, &cb);
        if (rc) return rc;
        if (!cb->chans) return fail(c, AMB_ESTATE, "channels not set for cube %d", id);
        if (r1 < 0 || r1 >= (int)c->uhigh.size()) return fail(c, AMB_EARG, "ring %d", r1);
        const long ng = amb_ng(c->uhigh[r1], cb->Du);
        res->ng = ng;
        res->inv_ng2 = 1.0 / (ng * ng * 1.);
        res->cert_margin = 0;
        const size_t W = 2 * (size_t)cb->nc10 + 1;
        std::vector<float> bl;
        amb_tri_req q{};
        q.cube_id = id; q.r1 = r1; q.r2 = r2; q.r3 = r3;
        q.bl = res->bl;
        if (!q.bl) { bl.resize(W * W); q.bl = bl.data(); }
        q.cnt = res->cnt;
        std::vector<float> nb;
        if (want_norm) {
            q.norm = res->norm;
            if (!q.norm) { nb.resize(W * W); q.norm = nb.data(); }
        }
        double *b64[1] = {res->bl64}, *n64[1] = {want_norm ? res->norm64 : nullptr};
        rc = lagft_run(c, &q, 1, 0.0, nullptr, b64, n64);
        res->cert_margin = want_norm ? c->lag_margin : 0.f;
        return rc;
    }
    TriJob j;
    const PlaneSrc src{&c->planes};
    int rc = prepare_tri(c, src, id, r1, r2, r3, want_norm, j);
    if (rc) { unpin_all(c); return rc; }
    res->ng = j.ng;
    res->inv_ng2 = 1.0 / (j.ng * j.ng * 1.);
    res->cert_margin = 0;
    const size_t cells = j.cells, nC = (size_t)j.nact * j.nact * j.nact;
    if ((rc = scratch(c, &c->C, &c->C_bytes, sizeof(float) * nC, "contraction partials"))) { unpin_all(c); return rc; }
    if (want_norm && (rc = scratch(c, &c->Cq, &c->Cq_bytes, sizeof(int) * nC, "normalisation partials"))) { unpin_all(c); return rc; }
    /* bl, norm | pad to 8 B | bl64, norm64 | cnt */
    const size_t outb = cells * (sizeof(float) * 2 + sizeof(double) * 2 + sizeof(int)) + sizeof(float) * (cells & 1);
    if ((rc = grow(c, &c->out, &c->out_bytes, outb, "lag arrays"))) { unpin_all(c); return rc; }
    float *d_bl = (float *)c->out;
    float *d_norm = d_bl + cells;
    double *d_bl64 = (double *)(d_norm + cells + (cells & 1));
    double *d_norm64 = d_bl64 + cells;
    int *d_cnt = (int *)(d_norm64 + cells);

    const int prec = c->prec;
    const int nprec = prec == AMB_PREC_FP32 ? AMB_PREC_FP32 : AMB_PREC_BF16X3;
    j.a.C = c->C;
    hipError_t e = hipSuccess;
    const char *step = "none";
#define STEP(name, expr)                                              \
    if (e == hipSuccess) {                                            \
        step = name;                                                  \
        e = (expr);                                                   \
        if (e == hipSuccess && c->sync_debug) e = hipStreamSynchronize(c->stream); \
    }
    const amb_tune_entry tc = tune_cfg(c, j.ngc, j.nact);
    timer_start(c, &c->st.ms_contract);
    STEP("contract", amb_launch_contract(c->stream, prec, 0, tune_variant(tc, prec, j.nact, j.pitch), j.a));
    timer_stop(c);
    c->st.n_contract_launch++;
    if (want_norm) {
        j.b.Cq = c->Cq;
        j.b.margin = c->margin;
        STEP("margin reset", amb_launch_fill_u32(c->stream, c->margin, 1, 0u));
        timer_start(c, &c->st.ms_norm);
        STEP("norm contract", amb_launch_contract(c->stream, nprec, 1, tune_variant(tc, nprec, j.nact, j.pitch), j.b));
        timer_stop(c);
    }
    ReduceArgs ra = j.ra;
    ra.C = c->C; ra.Cq = want_norm ? c->Cq : nullptr;
    ra.bl = d_bl; ra.bl64 = d_bl64; ra.cnt = d_cnt;
    ra.norm = want_norm ? d_norm : nullptr; ra.norm64 = want_norm ? d_norm64 : nullptr;
    timer_start(c, &c->st.ms_reduce);
    STEP("reduce", amb_launch_reduce(c->stream, ra));
    timer_stop(c);
    std::vector<int> hcnt;
    int32_t *cnt_dst = res->cnt;
    if (!cnt_dst) { hcnt.resize(cells); cnt_dst = hcnt.data(); }
    if (res->bl) STEP("copy bl", hipMemcpyAsync(res->bl, d_bl, sizeof(float) * cells, hipMemcpyDeviceToHost, c->stream));
    if (res->bl64) STEP("copy bl64", hipMemcpyAsync(res->bl64, d_bl64, sizeof(double) * cells, hipMemcpyDeviceToHost, c->stream));
    STEP("copy cnt", hipMemcpyAsync(cnt_dst, d_cnt, sizeof(int) * cells, hipMemcpyDeviceToHost, c->stream));
    if (want_norm && res->norm) STEP("copy norm", hipMemcpyAsync(res->norm, d_norm, sizeof(float) * cells, hipMemcpyDeviceToHost, c->stream));
    if (want_norm && res->norm64) STEP("copy norm64", hipMemcpyAsync(res->norm64, d_norm64, sizeof(double) * cells, hipMemcpyDeviceToHost, c->stream));
    unsigned mbits = 0;
    if (want_norm) STEP("copy margin", hipMemcpyAsync(&mbits, c->margin, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    STEP("sync", hipStreamSynchronize(c->stream));
#undef STEP
    timers_collect(c);
    unpin_all(c);
    if (e != hipSuccess)
        return fail(c, AMB_EHIP, "bispectrum (%d,%d,%d) Ng=%ld nact=%d step %s: %s", r1, r2, r3, j.ng, j.nact, step,
                    hipGetErrorString(e));
    account_tri(c, j, cnt_dst);
    if (want_norm) {
        float m;
        memcpy(&m, &mbits, sizeof m);
        res->cert_margin = m;
        bool again = false;
        if ((rc = check_cert(c, m, &again))) return rc;
        /* sums the fast path could not resolve on their grid: recomputed
           exactly in FP64 */
        if (again) {
            float mx = 0.f;
            rc = norm_exact(c, id, r1, r2, r3, res->norm, res->norm64, &mx);
            res->cert_margin = mx;
            return rc;
        }
    }
    return AMB_OK;
}

/* ------------------------------------------------------------- calCR */
/* Device scratch of one calCR call, 256-byte aligned pieces of one block:
 * the FP64 ring images [nr][N*N], the column-pass scratch [jmax][N], the
 * N twiddles, the plane id, the triangle list and the triangle sums. */
struct CrLayout {
    size_t images, G, tw, plane, tri, num, total;
};

static CrLayout cr_layout(int N, int nr, int jmax, int ntri)
{
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    CrLayout L;
    size_t o = 0;
    L.images = o; o += al(sizeof(double) * (size_t)nr * N * N);
    L.G = o; o += al(sizeof(double2) * (size_t)std::max(jmax, 1) * N);
    L.tw = o; o += al(sizeof(double2) * (size_t)N);
    L.plane = o; o += al(sizeof(int));
    L.tri = o; o += al(sizeof(int) * 3 * (size_t)std::max(ntri, 1));
    L.num = o; o += al(sizeof(double) * (size_t)std::max(ntri, 1));
    L.total = o;
    return L;
}

/* calCR (amber_funcs.c:389-491): the ring images of one sampling plane on
 * the full N x N grid in FP64, their triple sums per closed triangle, each
 * Num = N^2 (count of closed sampled triangles) rounded to its 1/8 grid.
 * noalloc: the scratch is the library's static device array (no device
 * allocation at all); otherwise it is allocated and freed here. */
int closed_rings(amb_ctx *c, int id, int plane, double *rdata, int cap, int *rcount, bool noalloc)
{
    hipSetDevice(c->device);
    Cube *cb;
    int rc = find_cube(c, id, &cb);
    if (rc) return rc;
    const int nr = (int)c->ulow.size();
    if (nr == 0) return fail(c, AMB_ESTATE, "rings not set");
    if (plane < 0 || plane >= cb->nc) return fail(c, AMB_EARG, "sampling plane %d outside [0,%lld)", plane, (long long)cb->nc);
    const int N = (int)cb->N;
    const size_t P = (size_t)N * N;
    const int ntri = amb_triangles(c->uavg.data(), nr, nullptr, 0);
    std::vector<int> tri(3 * (size_t)std::max(ntri, 1));
    amb_triangles(c->uavg.data(), nr, tri.data(), ntri);
    *rcount = ntri;
    if (ntri == 0) return AMB_OK;
    /* two transform buffers of N complex doubles in 64 KiB of LDS */
    if (32L * N > 65536) return fail(c, AMB_EARG, "grid %d exceeds the FP64 LDS transform size (2048)", N);
    FftPlan pl;
    amb_make_plan(N, &pl);
    int jmax = 0;
    std::vector<RingGeom> geo(nr);
    for (int r = 0; r < nr; ++r) {
        if ((rc = ring_geom(c, *cb, r, N, &geo[r], N))) return rc;
        jmax = std::max(jmax, geo[r].jcnt);
    }
    const CrLayout L = cr_layout(N, nr, jmax, ntri);
    char *base = nullptr;
    if (noalloc) {
        if (!c->na_scratch) return fail(c, AMB_ESTATE, "no static scratch on device %d", c->device);
        if (L.total > AMB_NOALLOC_SCRATCH_BYTES)
            return fail(c, AMB_EBUDGET, "calCR of %d rings on a %d^2 grid needs %zu bytes of scratch; the "
                        "allocation-free entry point holds %zu (amb_closed_rings has no such limit)",
                        nr, N, L.total, (size_t)AMB_NOALLOC_SCRATCH_BYTES);
        base = (char *)c->na_scratch;
    } else if ((rc = dev_alloc(c, (void **)&base, L.total, "calCR scratch"))) {
        return rc;
    }
    double *II = (double *)(base + L.images);
    double2 *G = (double2 *)(base + L.G), *tw = (double2 *)(base + L.tw);
    int *dp = (int *)(base + L.plane), *dtri = (int *)(base + L.tri);
    double *dnum = (double *)(base + L.num);
    std::vector<double> num(ntri);
    /* the plane id by a fill kernel, the triangle list by one copy */
    hipError_t e = amb_launch_fill_u32(c->stream, (unsigned *)dp, 1, (unsigned)plane);
    if (e == hipSuccess) e = hipMemcpyAsync(dtri, tri.data(), sizeof(int) * 3 * ntri, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = amb_launch_twiddles64(c->stream, tw, N);
    /* the kernel timers' events come from a pool that may create events:
       not on the allocation-free path */
    if (!noalloc) timer_start(c, &c->st.ms_fft);
    /* calCR's gather (amber_funcs.c:419-429) is the Dfg gather at Ng = N */
    for (int r = 0; r < nr && e == hipSuccess; ++r)
        e = amb_launch_ring_c2r64(c->stream, cb->samp_t, dp, 1, N, N / 2 + 1, pl, tw, geo[r], G, II + (size_t)r * P,
                                  (long)P);
    if (!noalloc) timer_stop(c);
    if (e == hipSuccess) e = amb_launch_triple_sum(c->stream, II, (long)P, dtri, ntri, dnum);
    if (e == hipSuccess) e = hipMemcpyAsync(num.data(), dnum, sizeof(double) * ntri, hipMemcpyDeviceToHost, c->stream);
    /* always drain: the copies read host vectors of this frame */
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = es;
    if (!noalloc) {
        timers_collect(c);
        dev_free(c, base, L.total);
    }
    if (e != hipSuccess) return fail(c, AMB_EHIP, "calCR: %s", hipGetErrorString(e));
    /* Num = N^2 * (count of closed sampled triangles), a multiple of N^2/8;
       the FP64 sums sit within ~1e-9 of that grid */
    const double NN2 = (double)N * (double)N;
    for (int t = 0; t < ntri; ++t) {
        double q = num[t] / NN2;
        double q8 = std::rint(q * 8.0);
        if (std::fabs(q - q8 * 0.125) > 1.0 / 32.0)
            return fail(c, AMB_ECERT, "calCR triangle %d: Num/N^2=%.6f is not within 1/32 of a multiple of 1/8", t, q);
        if (rdata && t < cap) {
            rdata[4 * t] = tri[3 * t];
            rdata[4 * t + 1] = tri[3 * t + 1];
            rdata[4 * t + 2] = tri[3 * t + 2];
            rdata[4 * t + 3] = q8 * 0.125 * NN2;
        }
    }
    return AMB_OK;
}

} // namespace amb

using namespace amb;

/* ------------------------------------------------------------------ API */
extern "C" {

const char *amb_version(void) { return "amber_mi355x 0.2 (gfx950)"; }

int amb_create(const amb_opts *o, amb_ctx **out)
{
    if (!out) return AMB_EARG;
    *out = nullptr;
    amb_ctx *c = new (std::nothrow) amb_ctx();
    if (!c) return AMB_ENOMEM;
    if (o) {
        c->device = o->device;
        c->prec = o->precision;
        c->verbose = o->verbose;
    }
    if (c->prec < AMB_PREC_FP32 || c->prec > AMB_PREC_LAGFT) { delete c; return AMB_EARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->device || c->device < 0) {
        (void)hipGetLastError();
        delete c;
        return AMB_EHIP;
    }
    if (hipSetDevice(c->device) != hipSuccess) { delete c; return AMB_EHIP; }
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete c; return AMB_EHIP; }
    c->sync_debug = getenv("AMB_SYNC_DEBUG") != nullptr;
    c->no_sym = getenv("AMB_NO_SYM") != nullptr; /* for A/B checks of the symmetry */
    c->no_shear = getenv("AMB_NO_SHEAR") != nullptr; /* r1 == r2 on the (i, j) pair tiles instead */
    if (const char *v = getenv("AMB_WAVE_TRIS")) { /* smaller grouped waves, for tests */
        const int k = atoi(v);
        if (k > 0 && k < c->wave_max) c->wave_max = k;
    }
    c->budget = (o && o->hbm_budget) ? o->hbm_budget : (uint64_t)(0.9 * (double)fr);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return AMB_EHIP; }
    void *m = nullptr;
    if (hipMalloc(&m, 256) != hipSuccess) { hipStreamDestroy(c->stream); delete c; return AMB_EHIP; }
    c->margin = (unsigned *)m;
    c->used = 256;
    c->grouped_blocks = amb_contract_grouped_blocks(c->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) c->arch = prop.gcnArchName;
    if (c->arch.empty()) c->arch = "unknown";
    for (char &ch : c->arch)
        if (ch == ' ') ch = '_';
    c->na_scratch = noalloc_scratch(c->device); /* with the context's device current */
    *out = c;
    return AMB_OK;
}

void amb_destroy(amb_ctx *c)
{
    if (!c) return;
    amb_serve_stop(c);
    hipSetDevice(c->device);
    while (!c->graphs.empty()) amb_graph_destroy(c, (amb_graph *)c->graphs.back());
    hipStreamSynchronize(c->stream);
    timers_collect(c);
    for (auto &p : c->planes) plane_free(c, p.d, p.bytes);
    hipStreamSynchronize(c->stream);
    for (auto &k : c->arena) hipFree(k.base);
    for (auto &kv : c->cubes) {
        Cube &cb = kv.second;
        hipFree(cb.vis_t); hipFree(cb.samp_t); hipFree(cb.d_act); hipFree(cb.d_inv);
        hipFree(cb.d_wlo); hipFree(cb.d_whi); hipFree(cb.d_lagcnt); hipFree(cb.d_lagpos);
    }
    for (auto &kv : c->tw) hipFree(kv.second);
    for (auto &kv : c->tw64) hipFree(kv.second);
    if (c->cstream) {
        hipStreamSynchronize(c->cstream);
        hipStreamDestroy(c->cstream);
        hipEventDestroy(c->ev_red);
        hipEventDestroy(c->ev_cp[0]);
        hipEventDestroy(c->ev_cp[1]);
    }
    hipFree(c->G); hipFree(c->C); hipFree(c->Cq); hipFree(c->margin); hipFree(c->out); hipFree(c->out2);
    hipFree(c->Cqx); hipFree(c->outx);
    hipFree(c->lag_part); hipFree(c->lag_st); hipFree(c->lag_out); hipFree(c->lag_out2);
    for (auto &kv : c->lag_pairs_tab) hipFree(std::get<0>(kv.second));
    for (auto &kv : c->lag_r23_tab) hipFree(std::get<0>(kv.second));
    for (auto &kv : c->lag_equi_tab) hipFree(std::get<0>(kv.second));
    hipFree(c->wdesc);
    for (void *h : c->hstage)
        if (h) hipHostFree(h);
    for (auto e : c->evpool) hipEventDestroy(e);
    post_release(c);
    hipStreamDestroy(c->stream);
    delete c;
}

const char *amb_last_error(amb_ctx *c) { return c ? c->err.c_str() : "null context"; }

int amb_set_precision(amb_ctx *c, int prec)
{
    if (!c) return AMB_EARG;
    Lock g(c->mu);
    if (prec < AMB_PREC_FP32 || prec > AMB_PREC_LAGFT) return fail(c, AMB_EARG, "precision %d", prec);
    if (c->server && prec != c->prec) return fail(c, AMB_ESTATE, "the precision is fixed while serving");
    c->prec = prec;
    return AMB_OK;
}

int amb_set_batch_mode(amb_ctx *c, int mode)
{
    if (!c) return AMB_EARG;
    Lock g(c->mu);
    if (mode != AMB_BATCH_PER_TRIANGLE && mode != AMB_BATCH_GROUPED) return fail(c, AMB_EARG, "batch mode %d", mode);
    c->batch_mode = mode;
    return AMB_OK;
}

} // extern "C"

namespace amb {

/* normalize_data (amber_funcs.c:1946-1978) fused into the transposing
   upload; rows of cells are staged through the device buffer stg */
int upload_cube(amb_ctx *c, Cube &cb, const double *gv, const double *samp, hipStream_t s, void *stg, size_t stg_bytes)
{
    const int64_t N = cb.N, nc = cb.nc, Nh = N / 2 + 1;
    const size_t row_doubles = (size_t)Nh * nc * 3;
    int rows = (int)std::min<size_t>((size_t)N, stg_bytes / (row_doubles * 8));
    if (rows < 1) return fail(c, AMB_EARG, "upload staging of %zu bytes holds no row of %lld cells", stg_bytes, (long long)Nh);
    double *dg = (double *)stg, *ds = dg + (size_t)rows * Nh * nc * 2;
    hipError_t e = hipSuccess;
    for (int64_t i0 = 0; i0 < N && e == hipSuccess; i0 += rows) {
        int r = (int)std::min<int64_t>(rows, N - i0);
        size_t n1 = (size_t)r * Nh * nc;
        e = hipMemcpyAsync(dg, gv + (size_t)i0 * Nh * nc * 2, n1 * 16, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(ds, samp + (size_t)i0 * Nh * nc, n1 * 8, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = amb_launch_upload(s, dg, ds, (int)i0, r, (int)N, (int)Nh, (int)nc, cb.vis_t, cb.samp_t);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) return fail(c, AMB_EHIP, "cube upload: %s", hipGetErrorString(e));
    return AMB_OK;
}

/* bytes of the staging buffer of an upload: <= ~256 MiB of whole rows */
static size_t upload_staging_bytes(const Cube &cb)
{
    const size_t row = (size_t)(cb.N / 2 + 1) * cb.nc * 3 * 8;
    const size_t rows = std::min<size_t>((size_t)cb.N, std::max<size_t>(1, (256u << 20) / row));
    return row * rows;
}

static int upload_cube(amb_ctx *c, Cube &cb, const double *gv, const double *samp)
{
    const size_t sb = upload_staging_bytes(cb);
    void *stg = nullptr;
    int rc = dev_alloc(c, &stg, sb, "upload staging");
    if (rc) return rc;
    rc = upload_cube(c, cb, gv, samp, c->stream, stg, sb);
    dev_free(c, stg, sb);
    return rc;
}

} // namespace amb

extern "C" {

int amb_load_cube(amb_ctx *c, int64_t N, int64_t nc, double Du, int32_t flag_key, const double *gv,
                  const double *samp, int *cube_id)
{
    if (!c || !gv || !samp || !cube_id) return AMB_EARG;
    Lock lk(c->mu);
    hipSetDevice(c->device);
    if (N < 4 || N % 2 || nc < 1 || !(Du > 0)) return fail(c, AMB_EARG, "bad cube header N=%lld nc=%lld", (long long)N, (long long)nc);
    const size_t nv = (size_t)N * (N / 2 + 1) * nc;
    Cube cb;
    cb.N = N; cb.nc = nc; cb.Du = Du; cb.flag_key = flag_key;
    void *v = nullptr, *s = nullptr;
    int rc = dev_alloc(c, &v, sizeof(float2) * nv, "cube");
    if (rc) return rc;
    rc = dev_alloc(c, &s, sizeof(float) * nv, "sampling cube");
    if (rc) { dev_free(c, v, sizeof(float2) * nv); return rc; }
    cb.vis_t = (float2 *)v;
    cb.samp_t = (float *)s;
    cb.bytes = sizeof(float2) * nv + sizeof(float) * nv;
    if ((rc = upload_cube(c, cb, gv, samp))) {
        dev_free(c, v, sizeof(float2) * nv);
        dev_free(c, s, sizeof(float) * nv);
        return rc;
    }
    int id = c->next_cube++;
    c->cubes[id] = cb;
    *cube_id = id;
    return AMB_OK;
}

int amb_update_cube(amb_ctx *c, int id, const double *gv, const double *samp)
{
    if (!c || !gv || !samp) return AMB_EARG;
    Lock lk(c->mu);
    hipSetDevice(c->device);
    Cube *cb;
    int rc = find_cube(c, id, &cb);
    if (rc) return rc;
    if (cb->serving) return fail(c, AMB_ESTATE, "cube %d has triangles in the serving queue", id);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_cube_planes(c, id); /* filtered from the old values */
    return upload_cube(c, *cb, gv, samp);
}

int amb_load_cube_files(amb_ctx *c, const char *data_path, int rlzn, int *cube_id)
{
    if (!c || !data_path) return AMB_EARG;
    int64_t NN, nc;
    double Du;
    int32_t fk;
    double *gv = nullptr, *samp = nullptr;
    int rc = amb_read_cube_files(data_path, rlzn, &NN, &nc, &Du, &fk, &gv, &samp);
    if (rc) {
        Lock lk(c->mu);
        return fail(c, rc, "reading %s/GV_data_%d and sampling: %s", data_path, rlzn, amb_strerror(rc));
    }
    rc = amb_load_cube(c, NN, nc, Du, fk, gv, samp, cube_id);
    amb_free(gv);
    amb_free(samp);
    return rc;
}

int amb_release_cube(amb_ctx *c, int id)
{
    if (!c) return AMB_EARG;
    Lock lk(c->mu);
    Cube *cb;
    int rc = find_cube(c, id, &cb);
    if (rc) return rc;
    if (cb->graphs) return fail(c, AMB_ESTATE, "cube %d is read by %d captured graph(s)", id, cb->graphs);
    if (cb->serving) return fail(c, AMB_ESTATE, "cube %d has triangles in the serving queue", id);
    hipStreamSynchronize(c->stream);
    release_cube_buffers(c, id);
    return AMB_OK;
}

} // extern "C"

void amb::release_cube_buffers(amb_ctx *c, int id)
{
    Cube *cb = &c->cubes[id];
    drop_cube_planes(c, id);
    const size_t nv = (size_t)cb->N * (cb->N / 2 + 1) * cb->nc;
    dev_free(c, cb->vis_t, sizeof(float2) * nv);
    dev_free(c, cb->samp_t, sizeof(float) * nv);
    const size_t na = cb->act.size(), ncs = (size_t)cb->nc;
    dev_free(c, cb->d_act, sizeof(int) * na);
    dev_free(c, cb->d_inv, sizeof(int) * (2 * ncs + 1));
    dev_free(c, cb->d_wlo, sizeof(int) * na);
    dev_free(c, cb->d_whi, sizeof(int) * na);
    lag_release(c, *cb);
    c->cubes.erase(id);
}

extern "C" {

int amb_set_channels(amb_ctx *c, int id, const int *uf, int uf_nc, int nc10)
{
    if (!c) return AMB_EARG;
    Lock lk(c->mu);
    hipSetDevice(c->device);
    Cube *cb;
    int rc = find_cube(c, id, &cb);
    if (rc) return rc;
    if (cb->graphs) return fail(c, AMB_ESTATE, "cube %d is read by %d captured graph(s)", id, cb->graphs);
    if (cb->serving) return fail(c, AMB_ESTATE, "cube %d has triangles in the serving queue", id);
    return set_channels_core(c, id, uf, uf_nc, nc10);
}

} // extern "C"

int amb::set_channels_core(amb_ctx *c, int id, const int *uf, int uf_nc, int nc10)
{
    Cube *cb;
    int rc = find_cube(c, id, &cb);
    if (rc) return rc;
    const int nc = (int)cb->nc;
    if (nc10 < 0 || nc10 > nc - 1) return fail(c, AMB_EARG, "nc10=%d outside [0,%d]", nc10, nc - 1);
    std::vector<char> on(nc, 0);
    if (cb->flag_key == 1) {
        if (!uf || uf_nc <= 0) return fail(c, AMB_EARG, "flag_key=1 needs uf_chans");
        for (int i = 0; i < uf_nc; ++i) {
            if (uf[i] < 0 || uf[i] >= nc) return fail(c, AMB_EARG, "uf_chans[%d]=%d outside [0,%d)", i, uf[i], nc);
            on[uf[i]] = 1;
        }
    } else {
        std::fill(on.begin(), on.end(), 1);
    }
    /* planes are built per channel set: drop this cube's cached planes (a
       cube without channels has neither planes nor tables in use) */
    if (cb->chans) hipStreamSynchronize(c->stream);
    drop_cube_planes(c, id);
    lag_release(c, *cb);
    size_t na_old = cb->act.size();
    dev_free(c, cb->d_act, sizeof(int) * na_old);
    dev_free(c, cb->d_inv, sizeof(int) * (2 * (size_t)nc + 1));
    dev_free(c, cb->d_wlo, sizeof(int) * na_old);
    dev_free(c, cb->d_whi, sizeof(int) * na_old);
    cb->d_act = cb->d_inv = cb->d_wlo = cb->d_whi = nullptr;
    cb->chans = false;
    cb->act.clear();
    cb->inv.assign(nc, -1);
    for (int p = 0; p < nc; ++p)
        if (on[p]) { cb->inv[p] = (int)cb->act.size(); cb->act.push_back(p); }
    const int na = (int)cb->act.size();
    std::vector<int> invfg(2 * (size_t)nc + 1);
    for (int p = 0; p < nc; ++p) invfg[p] = cb->inv[p];
    /* first_ge[p] = first compact index whose channel is >= p */
    for (int p = nc, k = na; p >= 0; --p) {
        while (k > 0 && cb->act[k - 1] >= p) --k;
        invfg[nc + p] = k;
    }
    std::vector<int> wlo(na), whi(na);
    for (int i = 0, l = 0, h = 0; i < na; ++i) {
        int p1 = cb->act[i];
        while (l < na && cb->act[l] < p1 - nc10) ++l;
        while (h < na && cb->act[h] <= p1 + nc10) ++h;
        wlo[i] = l;
        whi[i] = h;
    }
    void *p;
    if ((rc = dev_alloc(c, &p, sizeof(int) * na, "channel table"))) return rc;
    cb->d_act = (int *)p;
    if ((rc = dev_alloc(c, &p, sizeof(int) * invfg.size(), "channel table"))) return rc;
    cb->d_inv = (int *)p;
    if ((rc = dev_alloc(c, &p, sizeof(int) * na, "window table"))) return rc;
    cb->d_wlo = (int *)p;
    if ((rc = dev_alloc(c, &p, sizeof(int) * na, "window table"))) return rc;
    cb->d_whi = (int *)p;
    HIPCHK(c, hipMemcpy(cb->d_act, cb->act.data(), sizeof(int) * na, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(cb->d_inv, invfg.data(), sizeof(int) * invfg.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(cb->d_wlo, wlo.data(), sizeof(int) * na, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(cb->d_whi, whi.data(), sizeof(int) * na, hipMemcpyHostToDevice));
    cb->wlo = wlo;
    cb->whi = whi;
    cb->nc10 = nc10;
    cb->chans = true;
    return AMB_OK;
}

extern "C" {

int amb_set_rings(amb_ctx *c, const double *lo, const double *hi, int nr)
{
    if (!c || !lo || !hi || nr <= 0) return AMB_EARG;
    Lock lk(c->mu);
    if (c->server) return fail(c, AMB_ESTATE, "the rings are fixed while serving");
    for (int i = 0; i < nr; ++i)
        if (!(hi[i] > lo[i]) || !(lo[i] >= 0)) return fail(c, AMB_EARG, "ring %d edges [%g,%g)", i, lo[i], hi[i]);
    c->ulow.assign(lo, lo + nr);
    c->uhigh.assign(hi, hi + nr);
    c->uavg.resize(nr);
    amb_uavg(lo, hi, nr, c->uavg.data());
    return AMB_OK;
}

int amb_closed_rings(amb_ctx *c, int id, int plane, double *rdata, int cap, int *rcount)
{
    if (!c || !rcount) return AMB_EARG;
    Lock lk(c->mu);
    return closed_rings(c, id, plane, rdata, cap, rcount, false);
}

int amb_closed_rings_noalloc(amb_ctx *c, int id, int plane, double *rdata, int cap, int *rcount)
{
    if (!c || !rcount) return AMB_EARG;
    Lock lk(c->mu);
    return closed_rings(c, id, plane, rdata, cap, rcount, true);
}

int amb_bispec_ex(amb_ctx *c, int id, int r1, int r2, int r3, int want_norm, amb_result *res)
{
    if (!c || !res) return AMB_EARG;
    Lock lk(c->mu);
    return bispec_core(c, id, r1, r2, r3, want_norm, res);
}

int amb_bispec_v1(amb_ctx *c, int id, int r1, int r2, int r3, double rdata_num, float *bl, int32_t *cnt,
                  double *ntri, double *num)
{
    if (!c || !bl || !cnt) return AMB_EARG;
    Lock lk(c->mu);
    amb_result r{};
    r.bl = bl;
    r.cnt = cnt;
    int rc = bispec_core(c, id, r1, r2, r3, 0, &r);
    if (rc) return rc;
    /* amber.c:198-201,287 with the same double expression order */
    const Cube &cb = c->cubes[id];
    long Ng1 = r.ng, Ng2 = r.ng;
    double gfact = (1. * Ng1) / (1. * cb.N);
    double gfact_sq = gfact * gfact;
    double Num = gfact_sq * rdata_num;
    double inv = 1.0 / (Ng1 * Ng2 * 1.);
    if (num) *num = Num;
    if (ntri) *ntri = Num * inv;
    return AMB_OK;
}

int amb_bispec_v2(amb_ctx *c, int id, int r1, int r2, int r3, float *bl, float *norm, double *inv_ng2)
{
    if (!c || !bl || !norm) return AMB_EARG;
    Lock lk(c->mu);
    amb_result r{};
    r.bl = bl;
    r.norm = norm;
    int rc = bispec_core(c, id, r1, r2, r3, 1, &r);
    if (rc) return rc;
    if (inv_ng2) *inv_ng2 = r.inv_ng2;
    return AMB_OK;
}

int amb_batch_times(amb_ctx *c, double *ms, int cap, int *n)
{
    if (!c || !n) return AMB_EARG;
    Lock lk(c->mu);
    *n = (int)c->batch_ms.size();
    for (int i = 0; i < *n && i < cap && ms; ++i) ms[i] = c->batch_ms[i];
    return AMB_OK;
}

int amb_ring_image(amb_ctx *c, int id, int ring, int64_t ng, int plane, int kind, float *out, int64_t cap)
{
    if (!c || !out) return AMB_EARG;
    Lock lk(c->mu);
    hipSetDevice(c->device);
    Cube *cb;
    int rc = find_cube(c, id, &cb);
    if (rc) return rc;
    if (!cb->chans) return fail(c, AMB_ESTATE, "channels not set for cube %d", id);
    if (ring < 0 || ring >= (int)c->ulow.size()) return fail(c, AMB_EARG, "ring %d outside the ring set", ring);
    if (kind != 0 && kind != 1) return fail(c, AMB_EARG, "image kind %d (0 data, 1 sampling)", kind);
    if (plane < 0 || plane >= cb->nc) return fail(c, AMB_EARG, "plane %d outside [0,%lld)", plane, (long long)cb->nc);
    const int ia = cb->inv[plane];
    if (ia < 0) return fail(c, AMB_EARG, "plane %d is not in the channel set", plane);
    if (ng <= 0) ng = amb_ng(c->uhigh[ring], cb->Du);
    if (ng < 1 || ng > 2730) return fail(c, AMB_EARG, "grid %lld outside [1,2730]", (long long)ng);
    const size_t P = (size_t)ng * ng;
    if (cap < (int64_t)P) return fail(c, AMB_EARG, "output holds %lld floats, %zu needed", (long long)cap, P);
    void *d = nullptr;
    rc = get_planes(c, PlaneSrc{&c->planes}, id, kind, ring, (int)ng, &d);
    unpin_all(c);
    if (rc) return rc;
    const float *src = (const float *)d + (size_t)ia * plane_pitch((int)ng);
    int slot[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(out, src, sizeof(float) * P, hipMemcpyDeviceToHost, c->stream));
    if (kind == 0)
        HIPCHK(c, hipMemcpyAsync(slot, (const float *)d + (size_t)plane_pitch((int)ng) * cb->nact(), sizeof slot,
                                 hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    timers_collect(c);
    if (kind == 0) { /* data sets are stored normalised: undo the power of two and the sign */
        const float f = (ia & 1 ? -1.f : 1.f) * ldexpf(1.f, slot[1] - AMB_SET_EXP);
        for (size_t i = 0; i < P; ++i) out[i] *= f;
    }
    return AMB_OK;
}

int amb_drop_planes(amb_ctx *c, int id)
{
    if (!c) return AMB_EARG;
    Lock lk(c->mu);
    Cube *cb;
    int rc = find_cube(c, id, &cb);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_cube_planes(c, id);
    return AMB_OK;
}

int amb_get_stats(amb_ctx *c, amb_stats *st)
{
    if (!c || !st) return AMB_EARG;
    Lock lk(c->mu);
    *st = c->st;
    st->ms_total = st->ms_fft + st->ms_contract + st->ms_reduce + st->ms_norm;
    st->plane_pool_bytes = (int64_t)c->arena_held;
    return AMB_OK;
}

int amb_reset_stats(amb_ctx *c)
{
    if (!c) return AMB_EARG;
    Lock lk(c->mu);
    c->st = amb_stats{};
    return AMB_OK;
}

int amb_set_cert_strict(amb_ctx *c, int strict)
{
    if (!c) return AMB_EARG;
    Lock lk(c->mu);
    c->cert_strict = strict ? 1 : 0;
    return AMB_OK;
}

int amb_set_grid(amb_ctx *c, int reduced)
{
    if (!c) return AMB_EARG;
    Lock lk(c->mu);
    c->grid_reduced = reduced ? 1 : 0;
    return AMB_OK;
}

} // extern "C"

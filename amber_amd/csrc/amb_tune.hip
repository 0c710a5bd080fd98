/*
 * Autotuner of the grouped contraction's launch configuration.
 *
 * The reference runs its triangles one after another with one fixed
 * schedule (amber.c:187-326).  Here the persistent contraction of a
 * grouped wave has knobs that move time but not results: the kernel
 * variant, the run length of the per-XCD work queues, the size of a wave
 * and the share of the resident workgroups it launches.  Every setting
 * computes each triangle's sums with the same arithmetic in the same
 * order, so results are bit-identical under any table (the tuner checks
 * that on every candidate it times).
 *
 * The table is keyed by (precision, ceil(log2 Ng), ceil(log2 nact)): the
 * item count of a triangle grows with nact^3 and its cost per item with
 * Ng^2, so shapes within a bucket stress the kernel alike.  Tuning times
 * the bucket's sample triangles as grouped waves on resident planes by
 * coordinate descent, keeps a candidate only when it is faster by more
 * than the noise margin, and installs the winner.
 */
#include "amb_ctx.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>

using Lock = std::lock_guard<std::recursive_mutex>;
using namespace amb;

namespace {

using Key = std::tuple<int, int, int>;

/* a candidate must beat the incumbent by this fraction of its time */
constexpr double NOISE = 0.02;

int log2_ceil(long x)
{
    int b = 0;
    while ((1l << b) < x) ++b;
    return b;
}

bool split_form(int prec) { return prec == AMB_PREC_BF16X3 || prec == AMB_PREC_FP16X3; }

amb_tune_entry defaults(int prec, int gb, int nb)
{
    amb_tune_entry t{};
    t.prec = prec;
    t.ng_log2 = gb;
    t.nact_log2 = nb;
    t.variant = split_form(prec) ? 2 : 0;
    t.xrun_log2 = 5;
    /* three times this wave size gave 0.3% fewer wave tails and 1.5 times
       the contraction's L2-miss traffic (more triangles' planes in flight) */
    t.wave_gib = 16;
    t.grid_pct = 100;
    return t;
}

bool entry_ok(const amb_tune_entry &t)
{
    return t.variant >= 0 && t.variant < AMB_NVAR && t.xrun_log2 >= 0 && t.xrun_log2 <= 12 && t.wave_gib >= 1 &&
           t.wave_gib <= 256 && t.grid_pct >= 1 && t.grid_pct <= 100;
}

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

/* order-dependent 64-bit digest of n bytes (whole words, then the tail) */
uint64_t digest(uint64_t h, const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *)p;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        memcpy(&w, b + i, 8);
        h = (h ^ w) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 29;
    }
    for (; i < n; ++i) h = (h ^ b[i]) * 0x100000001B3ull;
    return h;
}

const char *prec_name(int p)
{
    static const char *n[] = {"fp32", "bf16x3", "fp8", "fp16x3"};
    return (p >= 0 && p <= 3) ? n[p] : "?";
}

int prec_of(const char *s)
{
    for (int p = 0; p <= 3; ++p)
        if (!strcmp(s, prec_name(p))) return p;
    return -1;
}

} // namespace

namespace amb {

amb_tune_entry tune_cfg(amb_ctx *c, long ng, int nact)
{
    const int gb = log2_ceil(ng), nb = log2_ceil(std::max(nact, 1));
    if (c->has_forced) {
        amb_tune_entry t = c->forced;
        t.prec = c->prec;
        t.ng_log2 = gb;
        t.nact_log2 = nb;
        return t;
    }
    auto it = c->tune.find(Key{c->prec, gb, nb});
    if (it != c->tune.end()) return it->second;
    const amb_tune_entry *best = nullptr;
    int bd = 1 << 30;
    for (auto &kv : c->tune) {
        if (std::get<0>(kv.first) != c->prec) continue;
        const int d = std::abs(std::get<1>(kv.first) - gb) + std::abs(std::get<2>(kv.first) - nb);
        if (d < bd) { bd = d; best = &kv.second; }
    }
    return best ? *best : defaults(c->prec, gb, nb);
}

int tune_variant(const amb_tune_entry &t, int prec, int nact, long pitch)
{
    if (!split_form(prec)) return 0;
    if (t.variant == 2 && !amb_contract_v2_fits(nact, pitch)) return 0;
    return t.variant;
}

} // namespace amb

namespace {

struct Bucket {
    Key key;
    std::vector<amb_tri_req> req;
    double cost = 0;     /* items x pitch, for the order of tuning */
    double c_bytes = 0;  /* contraction partials of all its triangles */
};

/* one grouped run of the bucket's requests under configuration t: the
   contraction's device time (data + normalisation) and the outputs' digest */
int time_run(amb_ctx *c, const Bucket &b, const amb_tune_entry &t, double *ms, double *tflops, uint64_t *dg)
{
    c->forced = t;
    c->has_forced = true;
    c->st = amb_stats{};
    int rc = run_grouped(c, b.req.data(), (int)b.req.size(), now_ms(), nullptr);
    if (rc) return rc;
    *ms = c->st.ms_contract + c->st.ms_norm;
    *tflops = c->st.ms_contract > 0 ? c->st.flop_contract_exec / (c->st.ms_contract * 1e9) : 0;
    uint64_t h = 0xcbf29ce484222325ull;
    for (const amb_tri_req &q : b.req) {
        const Cube &cb = c->cubes[q.cube_id];
        const size_t cells = (size_t)(2 * cb.nc10 + 1) * (2 * cb.nc10 + 1);
        h = digest(h, q.bl, cells * sizeof(float));
        if (q.cnt) h = digest(h, q.cnt, cells * sizeof(int32_t));
        if (q.norm) h = digest(h, q.norm, cells * sizeof(float));
    }
    *dg = h;
    return AMB_OK;
}

int tune_bucket(amb_ctx *c, const Bucket &b, double deadline)
{
    const int prec = std::get<0>(b.key);
    amb_tune_entry best = tune_cfg(c, 1l << std::get<1>(b.key), 1 << std::get<2>(b.key));
    best.prec = prec;
    best.ng_log2 = std::get<1>(b.key);
    best.nact_log2 = std::get<2>(b.key);
    double ms, tf, t_best, tf_best;
    uint64_t d0, d;
    /* the first run builds the planes; its digest is the reference */
    int rc = time_run(c, b, best, &ms, &tf, &d0);
    if (rc) return rc;
    auto measure = [&](const amb_tune_entry &t, double *out, double *otf) -> int {
        double m[2], f[2];
        for (int r = 0; r < 2; ++r) {
            const int e = time_run(c, b, t, &m[r], &f[r], &d);
            if (e) return e;
            if (d != d0)
                return fail(c, AMB_EHIP, "tuning: variant %d xrun 2^%d wave %d GiB grid %d%% changed the results "
                            "of bucket (%s, Ng 2^%d, nact 2^%d)", t.variant, t.xrun_log2, t.wave_gib, t.grid_pct,
                            prec_name(prec), std::get<1>(b.key), std::get<2>(b.key));
        }
        const int r = m[1] < m[0] ? 1 : 0;
        *out = m[r];
        *otf = f[r];
        return AMB_OK;
    };
    if ((rc = measure(best, &t_best, &tf_best))) return rc;
    /* coordinate descent, one pass per knob */
    std::vector<std::vector<int>> cand(4);
    if (split_form(prec)) cand[0] = {2, 1, 0};
    cand[1] = {5, 4, 6, 3, 7};
    /* a wave size above the bucket's whole partials gives the same waves */
    for (int g : {16, 4, 48})
        if ((double)std::min((uint64_t)g << 30, c->budget / 4) < b.c_bytes || g == best.wave_gib) cand[2].push_back(g);
    cand[3] = {100, 88, 75};
    for (int k = 0; k < 4; ++k)
        for (int v : cand[k]) {
            if (now_ms() > deadline) goto done;
            amb_tune_entry t = best;
            int *f = k == 0 ? &t.variant : k == 1 ? &t.xrun_log2 : k == 2 ? &t.wave_gib : &t.grid_pct;
            if (*f == v) continue;
            *f = v;
            double tt, ttf;
            if ((rc = measure(t, &tt, &ttf))) return rc;
            if (c->verbose)
                fprintf(stderr, "amb tune %s Ng 2^%d nact 2^%d: var %d xrun %d wave %d grid %d: %.3f ms (best %.3f)\n",
                        prec_name(prec), t.ng_log2, t.nact_log2, t.variant, t.xrun_log2, t.wave_gib, t.grid_pct, tt,
                        t_best);
            if (tt < (1.0 - NOISE) * t_best) { best = t; t_best = tt; tf_best = ttf; }
        }
done:
    best.tflops = tf_best;
    c->tune[b.key] = best;
    return AMB_OK;
}

} // namespace

extern "C" {

int amb_autotune(amb_ctx *c, const amb_tri_req *req, int n, double max_seconds, const char *path)
{
    if (!c || !req || n <= 0 || !(max_seconds > 0)) return AMB_EARG;
    Lock lk(c->mu);
    if (c->server) return fail(c, AMB_ESTATE, "tuning while the serving queue runs");
    if (c->prec == AMB_PREC_LAGFT)
        return fail(c, AMB_ESTATE, "the lag-Fourier form has no contraction launch configuration to tune");
    std::map<Key, Bucket> bk;
    for (int i = 0; i < n; ++i) {
        const amb_tri_req &q = req[i];
        Cube *cb;
        int rc = find_cube(c, q.cube_id, &cb);
        if (rc) return rc;
        if (!q.bl) return fail(c, AMB_EARG, "request %d has no bl buffer", i);
        if (!cb->chans) return fail(c, AMB_ESTATE, "channels not set for cube %d", q.cube_id);
        if (q.r1 < 0 || q.r1 >= (int)c->uhigh.size()) return fail(c, AMB_EARG, "ring %d out of range", q.r1);
        if (q.r2 < 0 || q.r2 >= (int)c->uhigh.size() || q.r3 < 0 || q.r3 >= (int)c->uhigh.size())
            return fail(c, AMB_EARG, "triangle (%d,%d,%d) out of range", q.r1, q.r2, q.r3);
        const long ng = compute_grid(c, *cb, q.r1, q.r2, q.r3);
        const int nact = cb->nact();
        const Key k{c->prec, log2_ceil(ng), log2_ceil(std::max(nact, 1))};
        Bucket &b = bk[k];
        b.key = k;
        b.req.push_back(q);
        b.cost += (double)amb_contract_items(nact) * (double)plane_pitch(ng);
        b.c_bytes += (double)nact * nact * nact * sizeof(float);
    }
    std::vector<Bucket *> order;
    for (auto &kv : bk) order.push_back(&kv.second);
    std::stable_sort(order.begin(), order.end(), [](const Bucket *a, const Bucket *b) { return a->cost > b->cost; });
    const amb_stats saved = c->st;
    const amb_tune_entry saved_forced = c->forced;
    const bool saved_has = c->has_forced;
    const std::vector<double> saved_ms = c->batch_ms;
    const double deadline = now_ms() + 1e3 * max_seconds;
    int rc = AMB_OK;
    for (Bucket *b : order) {
        if (now_ms() > deadline) break;
        if ((rc = tune_bucket(c, *b, deadline))) break;
    }
    c->forced = saved_forced;
    c->has_forced = saved_has;
    c->st = saved;
    c->batch_ms = saved_ms;
    if (!rc && path) rc = amb_tune_save(c, path);
    return rc;
}

int amb_tune_save(amb_ctx *c, const char *path)
{
    if (!c || !path) return AMB_EARG;
    Lock lk(c->mu);
    FILE *f = fopen(path, "w");
    if (!f) return fail(c, AMB_EIO, "cannot write %s", path);
    fprintf(f, "amber_mi355x-tune 1 %s %d\n", c->arch.c_str(), c->grouped_blocks);
    fprintf(f, "# prec ng_log2 nact_log2 variant xrun_log2 wave_gib grid_pct tflops\n");
    for (auto &kv : c->tune) {
        const amb_tune_entry &t = kv.second;
        fprintf(f, "%s %d %d %d %d %d %d %.17g\n", prec_name(t.prec), t.ng_log2, t.nact_log2, t.variant, t.xrun_log2,
                t.wave_gib, t.grid_pct, t.tflops);
    }
    const bool ok = !ferror(f);
    if (fclose(f) != 0 || !ok) return fail(c, AMB_EIO, "write error on %s", path);
    return AMB_OK;
}

int amb_tune_load(amb_ctx *c, const char *path)
{
    if (!c || !path) return AMB_EARG;
    Lock lk(c->mu);
    FILE *f = fopen(path, "r");
    if (!f) return fail(c, AMB_EIO, "cannot read %s", path);
    char line[512], arch[128];
    int ver = 0, blocks = 0;
    if (!fgets(line, sizeof line, f) || sscanf(line, "amber_mi355x-tune %d %127s %d", &ver, arch, &blocks) != 3 ||
        ver != 1) {
        fclose(f);
        return fail(c, AMB_EFORMAT, "%s: not a tuning table", path);
    }
    if (c->arch != arch || blocks != c->grouped_blocks) {
        fclose(f);
        return fail(c, AMB_EFORMAT, "%s: tuned on %s with %d resident workgroups, this device is %s with %d", path,
                    arch, blocks, c->arch.c_str(), c->grouped_blocks);
    }
    std::map<Key, amb_tune_entry> tab;
    int ln = 1;
    while (fgets(line, sizeof line, f)) {
        ++ln;
        if (line[0] == '#' || line[0] == '\n') continue;
        char pn[16];
        amb_tune_entry t{};
        if (sscanf(line, "%15s %d %d %d %d %d %d %lf", pn, &t.ng_log2, &t.nact_log2, &t.variant, &t.xrun_log2,
                   &t.wave_gib, &t.grid_pct, &t.tflops) != 8 ||
            (t.prec = prec_of(pn)) < 0 || !entry_ok(t) || t.ng_log2 < 0 || t.ng_log2 > 12 || t.nact_log2 < 0 ||
            t.nact_log2 > 16) {
            fclose(f);
            return fail(c, AMB_EFORMAT, "%s:%d: bad entry", path, ln);
        }
        tab[Key{t.prec, t.ng_log2, t.nact_log2}] = t;
    }
    fclose(f);
    c->tune.swap(tab);
    return AMB_OK;
}

int amb_tune_clear(amb_ctx *c)
{
    if (!c) return AMB_EARG;
    Lock lk(c->mu);
    c->tune.clear();
    return AMB_OK;
}

int amb_tune_entries(amb_ctx *c, amb_tune_entry *out, int cap, int *n)
{
    if (!c || !n || cap < 0 || (cap > 0 && !out)) return AMB_EARG;
    Lock lk(c->mu);
    int k = 0;
    for (auto &kv : c->tune) {
        if (k < cap) out[k] = kv.second;
        ++k;
    }
    *n = k;
    return AMB_OK;
}

int amb_tune_lookup(amb_ctx *c, int ng, int nact, amb_tune_entry *out)
{
    if (!c || !out || ng <= 0 || nact <= 0) return AMB_EARG;
    Lock lk(c->mu);
    *out = tune_cfg(c, ng, nact);
    return AMB_OK;
}

int amb_set_launch_config(amb_ctx *c, const amb_tune_entry *t)
{
    if (!c) return AMB_EARG;
    Lock lk(c->mu);
    if (!t) {
        c->has_forced = false;
        return AMB_OK;
    }
    if (!entry_ok(*t)) return fail(c, AMB_EARG, "launch configuration out of range");
    c->forced = *t;
    c->has_forced = true;
    return AMB_OK;
}

} // extern "C"

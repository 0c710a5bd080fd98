/*
 * Serving queue: continuous batching at the granularity of contraction
 * work items (no counterpart in the reference, which loads one cube per
 * process, amber.c:32-38, amber_funcs.c:1861-1941).
 *
 * A scheduler thread keeps a set of admitted triangles.  Each iteration it
 * takes a bounded amount of work (items x pixels, about two rounds of the
 * persistent grid) from them, split fairly between the requests that still
 * have items, launches one grouped contraction over those item ranges (a
 * descriptor names the triangle's first item it covers; every item owns
 * its block of partial sums, so a triangle's items may run in any number
 * of launches), reduces the triangles whose last items ran, and copies
 * their lag arrays out.  Two iterations are in flight; the host admits new
 * triangles and composes the next iteration while the device runs.  A
 * request that arrives mid-run is admitted at once (memory permitting) and
 * gets its share of the next iteration.
 *
 * In the lag-Fourier precision the work items are the (tile, pixel
 * segment) items of the triple-product kernel; a triangle whose last items
 * ran gets its segment sums, 2-D inverse and lag array in the iteration.
 *
 * Memory: planes come from the context's LRU cache and are held by
 * serving pins while their triangles are admitted; partial sums and
 * results come from two pools reserved at start; cubes carried by requests
 * are uploaded by a second thread on its own stream and evicted, least
 * recently used first, once idle and in the way of a new upload.
 */
#include "amb_ctx.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <atomic>
#include <set>

using Lock = std::lock_guard<std::recursive_mutex>;
using namespace amb;

namespace {

constexpr size_t ALIGN = 256;
constexpr int NSLOT = 3;                       /* descriptor slots */
constexpr int INFLIGHT = 2;                    /* iterations on the device at once */
constexpr int MAX_QUEUE = 4096;                /* requests waiting for admission */
constexpr int MAX_HELD = 2;                    /* cubes uploaded ahead of their requests' admission */
constexpr size_t UPLOAD_STAGING = (size_t)256 << 20; /* at most; a sixteenth of the budget */
/* work of one iteration per resident workgroup, in items x pixels: two
   items of a 512^2 grid (about 40 ms of the split form) */
constexpr double ITER_PX_PER_BLOCK = 2.0 * 512 * 512;
/* a lag-Fourier item (one 128 x 128 tile over lag_item_px() pixels) in the
   same units: eight pixel units per item pixel (about 32 items of 2048
   pixels, or 64 of 1024, per workgroup and iteration) */
constexpr double LAG_UNITS_PER_PX = 8.0;

size_t al(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

/* first-fit allocator of byte ranges in [0, cap) */
struct Pool {
    size_t cap = 0;
    std::map<size_t, size_t> fr; /* offset -> length of free ranges */
    void init(size_t c)
    {
        cap = c;
        fr.clear();
        if (c) fr[0] = c;
    }
    bool take(size_t n, size_t *off)
    {
        n = al(n);
        for (auto it = fr.begin(); it != fr.end(); ++it)
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n, o2 = it->first + n;
                fr.erase(it);
                if (rest) fr[o2] = rest;
                return true;
            }
        return false;
    }
    void give(size_t off, size_t n)
    {
        n = al(n);
        auto it = fr.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != fr.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            fr.erase(nx);
        }
        if (it != fr.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                fr.erase(it);
            }
        }
    }
};

/* results of one triangle in the out pool: bl | cnt | norm | margin word */
struct OutLayout {
    size_t cnt, norm, margin, total;
};
OutLayout out_layout(size_t cells, bool norm)
{
    OutLayout o;
    o.cnt = al(cells * sizeof(float));
    o.norm = o.cnt + al(cells * sizeof(int));
    o.margin = o.norm + (norm ? al(cells * sizeof(float)) : 0);
    o.total = o.margin + ALIGN;
    return o;
}

} // namespace

struct AmbServer {
    struct Ticket {
        std::vector<amb_tri_req> req;
        int up_cube = 0;          /* the cube this request carries (0: none) */
        amb_cube_src up{};
        std::vector<int> up_chans;
        size_t admitted = 0;      /* triangles admitted (or failed) so far */
        size_t left = 0;          /* triangles not finished */
        int status = AMB_OK;
        double t_submit = 0, t_done = 0;
        bool done = false;
        bool busy = false;        /* the scheduler holds a pointer to it (not erasable) */
    };
    struct Job {
        uint64_t tk = 0;
        int idx = 0;
        TriJob j;                 /* contraction; for lag-Fourier jobs only cube, rings, cells, want_norm */
        bool lag = false;
        LagTri lt;                /* lag-Fourier jobs */
        std::vector<void *> pins;
        size_t c_off = 0, q_off = 0, o_off = 0, c_bytes = 0, c_total = 0, o_bytes = 0;
        OutLayout ol{};
        long items = 0, next = 0, nitems = 0, nnext = 0;
        double px = 0;            /* work of one item, pixel units */
    };
    struct Iter {
        int slot = 0;
        std::vector<int> finishing; /* jobs reduced and copied out in it */
    };
    amb_ctx *c = nullptr;
    std::thread sched, upl;
    std::mutex m;
    std::condition_variable cv_work, cv_done, cv_up;
    std::map<uint64_t, Ticket> tickets;
    std::deque<uint64_t> waiting;  /* not fully admitted, arrival order */
    std::deque<uint64_t> uploads;  /* tickets whose cube is to be uploaded */
    std::set<int> pending;         /* cube ids reserved for uploads not yet resident */
    uint64_t next = 1;
    bool stop = false;
    bool broken = false;           /* a launch failed: every later triangle fails */
    std::atomic<int> n_jobs{0};    /* admitted triangles not finished */
    int max_active = 256;
    bool lag = false;              /* the lag-Fourier form (fixed while serving) */
    /* the scheduler's own state */
    std::map<int, Job> jobs;
    int next_job = 1;
    std::vector<int> active;       /* jobs with items left to issue, admission order */
    std::deque<Iter> inflight;
    int slot_next = 0;
    /* owned device and host memory */
    Pool cpool, opool;
    char *cbuf = nullptr, *obuf = nullptr, *ohost = nullptr;
    size_t slot_bytes = 0;
    char *dslot[NSLOT] = {}, *hslot[NSLOT] = {};
    hipEvent_t ev[NSLOT] = {};
    hipStream_t ustream = nullptr;
    void *ustg = nullptr;
    size_t ustg_bytes = 0;
    void *lag_st = nullptr;        /* the lag-Fourier 2-D spectrum, shared by finishing triangles in stream order */
    size_t lag_st_bytes = 0;
};

using Srv = AmbServer;

namespace {

/* ticket bookkeeping: one triangle of ticket id ended with status rc
   (call with s->m held) */
void tri_done(Srv *s, uint64_t id, int rc)
{
    auto it = s->tickets.find(id);
    if (it == s->tickets.end()) return;
    Srv::Ticket &t = it->second;
    if (rc && t.status == AMB_OK) t.status = rc;
    if (t.left > 0 && --t.left == 0) {
        t.done = true;
        t.t_done = now_ms();
        s->cv_done.notify_all();
    }
}

/* the triangles of ticket id not yet admitted end with status rc */
void fail_rest(Srv *s, uint64_t id, int rc)
{
    std::lock_guard<std::mutex> lk(s->m);
    Srv::Ticket &t = s->tickets[id];
    while (t.admitted < t.req.size()) {
        ++t.admitted;
        tri_done(s, id, rc);
    }
}

/* evicts the least recently used idle uploaded cube (context lock held) */
bool evict_cube(amb_ctx *c, Srv *s)
{
    int best = 0;
    uint64_t bl = ~(uint64_t)0;
    for (auto &kv : c->cubes) {
        const Cube &cb = kv.second;
        if (cb.uploaded && cb.serving == 0 && cb.holds == 0 && cb.graphs == 0 && cb.last_use < bl) {
            best = kv.first;
            bl = cb.last_use;
        }
    }
    if (!best) return false;
    (void)s;
    hipStreamSynchronize(c->stream);
    release_cube_buffers(c, best);
    c->st.n_evict_cubes++;
    return true;
}

/* cubes uploaded for requests not yet admitted (context lock held) */
int held_cubes(amb_ctx *c)
{
    int n = 0;
    for (auto &kv : c->cubes) n += kv.second.holds > 0;
    return n;
}

/* admits waiting triangles while memory allows, in arrival order: a request
   that waits for memory that admitted triangles will free holds back the
   later ones; with nothing admitted, a request that does not fit because
   cubes carried by later requests take the memory lets those go first, and
   one that fits nowhere fails with AMB_EBUDGET */
void admit(Srv *s)
{
    amb_ctx *c = s->c;
    std::vector<uint64_t> ready;
    {
        std::lock_guard<std::mutex> lk(s->m);
        for (uint64_t id : s->waiting) {
            const Srv::Ticket &t = s->tickets[id];
            bool ok = true;
            for (size_t k = t.admitted; k < t.req.size() && ok; ++k) ok = !s->pending.count(t.req[k].cube_id);
            if (ok) ready.push_back(id);
        }
    }
    if (ready.empty()) return;
    Lock lk(c->mu);
    hipSetDevice(c->device);
    std::vector<uint64_t> finished;
    uint64_t first_stuck = 0;
    bool progress = false;
    for (uint64_t id : ready) {
        Srv::Ticket *t;
        {
            std::lock_guard<std::mutex> g(s->m);
            t = &s->tickets[id];
            t->busy = true;
        }
        bool blocked = false, stuck = false;
        while (t->admitted < t->req.size()) {
            if ((int)s->jobs.size() >= s->max_active) { blocked = true; break; }
            const amb_tri_req &q = t->req[t->admitted];
            Srv::Job jb;
            std::vector<void *> log;
            c->pin_log = &log;
            int rc = s->lag ? lag_prepare(c, q.cube_id, q.r1, q.r2, q.r3, &jb.lt)
                            : prepare_tri(c, PlaneSrc{&c->planes}, q.cube_id, q.r1, q.r2, q.r3, q.norm != nullptr, jb.j);
            c->pin_log = nullptr;
            if (rc == AMB_EBUDGET && !s->jobs.empty()) { /* finishing triangles free their planes */
                pins_to_server(c, log, false);
                blocked = true;
                break;
            }
            if (rc == AMB_EBUDGET && evict_cube(c, s)) { /* nothing in flight: idle uploaded cubes go */
                pins_to_server(c, log, false);
                continue;
            }
            if (rc == AMB_EBUDGET && held_cubes(c) > 0) { /* later requests' cubes hold the memory */
                pins_to_server(c, log, false);
                stuck = true;
                break;
            }
            if (rc) {
                pins_to_server(c, log, false);
                std::lock_guard<std::mutex> g(s->m);
                ++t->admitted;
                tri_done(s, id, rc);
                continue;
            }
            pins_to_server(c, log, true);
            jb.pins.swap(log);
            TriJob &j = jb.j;
            if (s->lag) {
                jb.lag = true;
                j.cube = q.cube_id;
                j.r1 = q.r1;
                j.r2 = q.r2;
                j.r3 = q.r3;
                j.cells = jb.lt.cells;
                j.want_norm = q.norm != nullptr; /* the exact FP64 path when the triangle completes */
                jb.c_bytes = jb.c_total = al(lag_part_bytes(jb.lt));
                jb.o_bytes = al(jb.lt.cells * sizeof(float));
            } else {
                const size_t n3 = (size_t)j.nact * j.nact * j.nact;
                jb.c_bytes = al(n3 * sizeof(float));
                jb.c_total = jb.c_bytes * (j.want_norm ? 2 : 1);
                jb.ol = out_layout(j.cells, j.want_norm);
                jb.o_bytes = jb.ol.total;
            }
            if (jb.c_total > s->cpool.cap) {
                unspin(c, jb.pins);
                std::lock_guard<std::mutex> g(s->m);
                ++t->admitted;
                tri_done(s, id, fail(c, AMB_EBUDGET, "triangle needs %zu bytes of partial sums, the serving pool holds %zu",
                                     jb.c_total, s->cpool.cap));
                continue;
            }
            bool got_c = s->cpool.take(jb.c_total, &jb.c_off);
            bool got_o = got_c && s->opool.take(jb.o_bytes, &jb.o_off);
            if (!got_o) {
                if (got_c) s->cpool.give(jb.c_off, jb.c_total);
                unspin(c, jb.pins);
                if (!s->jobs.empty()) { blocked = true; break; }
                /* submission checked the pools: this cannot happen with nothing admitted */
                std::lock_guard<std::mutex> g(s->m);
                ++t->admitted;
                tri_done(s, id, fail(c, AMB_EBUDGET, "triangle does not fit the serving pools"));
                continue;
            }
            jb.q_off = jb.c_off + jb.c_bytes;
            if (jb.lag) {
                jb.items = lag_items(jb.lt);
                jb.nitems = 0;
                jb.px = LAG_UNITS_PER_PX * (double)lag_item_px();
            } else {
                jb.items = amb_contract_items(j.nact);
                jb.nitems = j.want_norm ? jb.items : 0;
                jb.px = (double)j.pitch;
            }
            jb.tk = id;
            jb.idx = (int)t->admitted;
            /* the margin word collects atomicMax over the triangle's items */
            if (!jb.lag && j.want_norm &&
                amb_launch_fill_u32(c->stream, (unsigned *)(s->obuf + jb.o_off + jb.ol.margin), 1, 0u) != hipSuccess)
                s->broken = true;
            Cube &cb = c->cubes[q.cube_id];
            cb.serving++;
            cb.last_use = ++c->tick;
            const int jid = s->next_job++;
            s->jobs.emplace(jid, std::move(jb));
            s->active.push_back(jid);
            s->n_jobs++;
            progress = true;
            std::lock_guard<std::mutex> g(s->m);
            ++t->admitted;
        }
        if (t->admitted == t->req.size()) {
            finished.push_back(id);
            /* the carried cube was held for this request until now */
            auto cit = t->up_cube ? c->cubes.find(t->up_cube) : c->cubes.end();
            if (cit != c->cubes.end() && cit->second.holds > 0) cit->second.holds--;
        }
        {
            std::lock_guard<std::mutex> g(s->m);
            t->busy = false;
            s->cv_done.notify_all();
        }
        if (blocked) break; /* arrival order: later requests wait behind this one */
        if (stuck && !first_stuck) first_stuck = id;
    }
    if (first_stuck && !progress && s->jobs.empty()) {
        /* nothing runs and nothing was admitted: the oldest stuck triangle fits nowhere */
        std::lock_guard<std::mutex> g(s->m);
        Srv::Ticket &t = s->tickets[first_stuck];
        ++t.admitted;
        tri_done(s, first_stuck, fail(c, AMB_EBUDGET, "triangle does not fit the HBM budget beside the cubes "
                                                       "carried by waiting requests"));
        if (t.admitted == t.req.size()) {
            finished.push_back(first_stuck);
            auto cit = t.up_cube ? c->cubes.find(t.up_cube) : c->cubes.end();
            if (cit != c->cubes.end() && cit->second.holds > 0) cit->second.holds--;
        }
    }
    std::lock_guard<std::mutex> g(s->m);
    for (uint64_t id : finished) s->waiting.erase(std::find(s->waiting.begin(), s->waiting.end(), id));
    if (!finished.empty()) s->cv_up.notify_all(); /* an upload gated on held cubes may go */
}

/* one descriptor range: items [i0, i0 + n) of job jid, data or norm */
struct Range {
    int jid;
    bool norm;
    long i0, n;
    long pitch;
};

/* composes and launches one iteration (context lock held); false when no
   job has items left */
bool launch_iter(Srv *s)
{
    amb_ctx *c = s->c;
    if (s->active.empty()) return false;
    const double quota = ITER_PX_PER_BLOCK * c->grouped_blocks;
    /* fair shares: every request with items left gets quota / requests,
       spent on its triangles in admission order; what is left over goes to
       the oldest triangles */
    std::map<uint64_t, double> share;
    for (int jid : s->active) share[s->jobs[jid].tk] = 0;
    for (auto &kv : share) kv.second = quota / share.size();
    std::map<int, std::pair<long, long>> take; /* jid -> (data items, norm items) */
    double spent = 0;
    auto grant = [&](Srv::Job &jb, int jid, double budget) -> double {
        auto &tk = take[jid];
        const long d_left = jb.items - jb.next - tk.first, n_left = jb.nitems - jb.nnext - tk.second;
        const double px = jb.px;
        long want = (long)std::max(1.0, std::floor(budget / px));
        const long d = std::min(want, d_left);
        tk.first += d;
        want -= d;
        const long n = std::min(std::max(want, 0l), n_left);
        tk.second += n;
        return (double)(d + n) * px;
    };
    for (int jid : s->active) {
        Srv::Job &jb = s->jobs[jid];
        double &sh = share[jb.tk];
        if (sh <= 0) continue;
        const double used = grant(jb, jid, sh);
        sh -= used;
        spent += used;
    }
    for (int jid : s->active) {
        if (spent >= quota) break;
        spent += grant(s->jobs[jid], jid, quota - spent);
    }
    std::vector<Range> rd, rn, rl;
    std::vector<int> finishing;
    for (int jid : s->active) {
        auto it = take.find(jid);
        if (it == take.end()) continue;
        Srv::Job &jb = s->jobs[jid];
        if (it->second.first > 0) (jb.lag ? rl : rd).push_back({jid, false, jb.next, it->second.first, jb.j.pitch});
        if (it->second.second > 0) rn.push_back({jid, true, jb.nnext, it->second.second, jb.j.pitch});
        jb.next += it->second.first;
        jb.nnext += it->second.second;
        if (jb.next == jb.items && jb.nnext == jb.nitems) finishing.push_back(jid);
    }
    if (rd.empty() && rn.empty() && rl.empty() && finishing.empty()) return false;
    std::vector<int> fin_c, fin_l; /* finishing contraction and lag-Fourier jobs */
    for (int jid : finishing) (s->jobs[jid].lag ? fin_l : fin_c).push_back(jid);
    /* largest items first: the tail of the launch is made of small ones */
    auto by_pitch = [](const Range &a, const Range &b) { return a.pitch > b.pitch; };
    std::stable_sort(rd.begin(), rd.end(), by_pitch);
    std::stable_sort(rn.begin(), rn.end(), by_pitch);
    /* descriptor block of this iteration */
    const int slot = s->slot_next;
    s->slot_next = (s->slot_next + 1) % NSLOT;
    char *h = s->hslot[slot], *d = s->dslot[slot];
    size_t o = 0;
    const size_t off_da = o; o = al(o + sizeof(ContractArgs) * rd.size());
    const size_t off_na = o; o = al(o + sizeof(ContractArgs) * rn.size());
    const size_t off_ra = o; o = al(o + sizeof(ReduceArgs) * fin_c.size());
    const size_t off_b = o; o = al(o + sizeof(long) * (rd.size() + 1));
    const size_t off_nb = o; o = al(o + sizeof(long) * (rn.size() + 1));
    const size_t off_cb = o; o = al(o + sizeof(long) * (fin_c.size() + 1));
    const size_t off_ctr = o; o = al(o + sizeof(unsigned long long) * 2 * AMB_GROUPED_COUNTERS);
    const size_t ld = lag_desc_bytes();
    const size_t off_la = o; o = al(o + ld * rl.size());
    const size_t off_lw = o; o = al(o + ld * fin_l.size());
    if (o > s->slot_bytes) { /* cannot happen below max_active triangles; split the work instead */
        for (const Range &r : rd) s->jobs[r.jid].next -= r.n;
        for (const Range &r : rn) s->jobs[r.jid].nnext -= r.n;
        for (const Range &r : rl) s->jobs[r.jid].next -= r.n;
        return false;
    }
    ContractArgs *da = (ContractArgs *)(h + off_da), *na = (ContractArgs *)(h + off_na);
    ReduceArgs *ra = (ReduceArgs *)(h + off_ra);
    long *bg = (long *)(h + off_b), *nb = (long *)(h + off_nb), *cb = (long *)(h + off_cb);
    long it = 0, nit = 0, cl = 0;
    const Srv::Job *big = nullptr;
    for (size_t k = 0; k < rd.size(); ++k) {
        const Srv::Job &jb = s->jobs[rd[k].jid];
        if (!big || jb.j.pitch * (double)jb.items > big->j.pitch * (double)big->items) big = &jb;
        da[k] = jb.j.a;
        da[k].C = (float *)(s->cbuf + jb.c_off);
        da[k].item0 = rd[k].i0;
        bg[k] = it;
        it += rd[k].n;
    }
    bg[rd.size()] = it;
    for (size_t k = 0; k < rn.size(); ++k) {
        const Srv::Job &jb = s->jobs[rn[k].jid];
        if (!big) big = &jb;
        na[k] = jb.j.b;
        na[k].C = nullptr;
        na[k].Cq = (int *)(s->cbuf + jb.q_off);
        na[k].margin = (unsigned *)(s->obuf + jb.o_off + jb.ol.margin);
        na[k].item0 = rn[k].i0;
        nb[k] = nit;
        nit += rn[k].n;
    }
    nb[rn.size()] = nit;
    for (size_t k = 0; k < fin_c.size(); ++k) {
        const Srv::Job &jb = s->jobs[fin_c[k]];
        char *out = s->obuf + jb.o_off;
        ra[k] = jb.j.ra;
        ra[k].C = (const float *)(s->cbuf + jb.c_off);
        ra[k].Cq = jb.j.want_norm ? (const int *)(s->cbuf + jb.q_off) : nullptr;
        ra[k].bl = (float *)out;
        ra[k].bl64 = nullptr;
        ra[k].cnt = (int *)(out + jb.ol.cnt);
        ra[k].norm = jb.j.want_norm ? (float *)(out + jb.ol.norm) : nullptr;
        ra[k].norm64 = nullptr;
        cb[k] = cl;
        cl += (long)jb.j.cells;
    }
    cb[fin_c.size()] = cl;
    /* lag-Fourier item ranges (each descriptor names the launch's first item
       of its range and the triangle's item that is), and one whole-triangle
       descriptor per finishing triangle for its k3 = L/2 column */
    long lit = 0;
    for (size_t k = 0; k < rl.size(); ++k) {
        const Srv::Job &jb = s->jobs[rl[k].jid];
        lag_fill_range(c, jb.lt, s->cbuf + jb.c_off, rl[k].i0, lit, h + off_la + k * ld);
        lit += rl[k].n;
    }
    for (size_t k = 0; k < fin_l.size(); ++k) {
        const Srv::Job &jb = s->jobs[fin_l[k]];
        lag_fill_range(c, jb.lt, s->cbuf + jb.c_off, 0, 0, h + off_lw + k * ld);
    }
    /* launch configuration of the costliest triangle, a variant all support */
    const int prec = c->prec, nprec = prec == AMB_PREC_FP32 ? AMB_PREC_FP32 : AMB_PREC_BF16X3;
    amb_tune_entry t{};
    int var = 0, nvar = 0, blocks = 8;
    if (big) {
        t = tune_cfg(c, big->j.ngc, big->j.nact);
        var = nvar = t.variant;
        for (const Range &r : rd) var = std::min(var, tune_variant(t, prec, s->jobs[r.jid].j.nact, r.pitch));
        for (const Range &r : rn) nvar = std::min(nvar, tune_variant(t, nprec, s->jobs[r.jid].j.nact, r.pitch));
        blocks = std::max(8, (int)((long)c->grouped_blocks * t.grid_pct / 100) / 8 * 8);
    }
    unsigned long long *ctr = (unsigned long long *)(d + off_ctr);
    hipError_t e = hipMemcpyAsync(d, h, o, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && !rd.empty()) {
        timer_start(c, &c->st.ms_contract);
        e = amb_launch_contract_grouped(c->stream, prec, 0, var, (const ContractArgs *)(d + off_da),
                                        (const long *)(d + off_b), (int)rd.size(), it, ctr, blocks, t.xrun_log2);
        timer_stop(c);
    }
    if (e == hipSuccess && !rn.empty()) {
        timer_start(c, &c->st.ms_norm);
        e = amb_launch_contract_grouped(c->stream, nprec, 1, nvar, (const ContractArgs *)(d + off_na),
                                        (const long *)(d + off_nb), (int)rn.size(), nit, ctr + AMB_GROUPED_COUNTERS,
                                        blocks, t.xrun_log2);
        timer_stop(c);
    }
    if (e == hipSuccess && !fin_c.empty()) {
        timer_start(c, &c->st.ms_reduce);
        e = amb_launch_reduce_grouped(c->stream, (const ReduceArgs *)(d + off_ra), (const long *)(d + off_cb),
                                      (int)fin_c.size(), cl);
        timer_stop(c);
    }
    if (e == hipSuccess && !rl.empty()) {
        timer_start(c, &c->st.ms_contract);
        e = lag_launch_tiles(c->stream, d + off_la, (int)rl.size(), lit);
        timer_stop(c);
    }
    if (e == hipSuccess && !fin_l.empty()) {
        timer_start(c, &c->st.ms_reduce);
        for (size_t k = 0; k < fin_l.size() && e == hipSuccess; ++k) {
            const Srv::Job &jb = s->jobs[fin_l[k]];
            e = lag_launch_finish(c, c->stream, jb.lt, d + off_lw + k * ld, s->cbuf + jb.c_off, s->lag_st,
                                  (float *)(s->obuf + jb.o_off));
        }
        timer_stop(c);
    }
    for (int jid : finishing) {
        const Srv::Job &jb = s->jobs[jid];
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->ohost + jb.o_off, s->obuf + jb.o_off, jb.o_bytes, hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(s->ev[slot], c->stream);
    c->st.n_contract_launch++;
    c->st.n_serve_iters++;
    s->active.erase(std::remove_if(s->active.begin(), s->active.end(),
                                   [&](int jid) {
                                       const Srv::Job &jb = s->jobs[jid];
                                       return jb.next == jb.items && jb.nnext == jb.nitems;
                                   }),
                    s->active.end());
    Srv::Iter itr;
    itr.slot = slot;
    itr.finishing = finishing;
    if (e != hipSuccess) {
        /* a launch error leaves the stream unusable: every triangle of the
           iteration fails, the event is not waited on */
        fail(c, AMB_EHIP, "serving iteration: %s", hipGetErrorString(e));
        (void)hipGetLastError();
        itr.slot = -1;
        s->broken = true;
    }
    s->inflight.push_back(itr);
    return true;
}

/* waits for the oldest iteration and delivers its finished triangles */
void complete_oldest(Srv *s)
{
    amb_ctx *c = s->c;
    Srv::Iter itr = s->inflight.front();
    s->inflight.pop_front();
    int erc = s->broken ? AMB_EHIP : AMB_OK;
    if (itr.slot < 0) erc = AMB_EHIP;
    else if (hipEventSynchronize(s->ev[itr.slot]) != hipSuccess) {
        (void)hipGetLastError();
        erc = AMB_EHIP;
    }
    std::vector<std::pair<uint64_t, int>> done;
    {
        Lock lk(c->mu);
        timers_collect(c);
        for (int jid : itr.finishing) {
            Srv::Job &jb = s->jobs[jid];
            const TriJob &j = jb.j;
            int rc = erc;
            const amb_tri_req *q = nullptr;
            {
                std::lock_guard<std::mutex> g(s->m);
                q = &s->tickets[jb.tk].req[jb.idx];
            }
            if (!rc && jb.lag) {
                memcpy(q->bl, s->ohost + jb.o_off, sizeof(float) * j.cells);
                lag_account(c, jb.lt, q->cnt);
                if (q->norm) rc = norm_exact(c, j.cube, j.r1, j.r2, j.r3, q->norm, nullptr, nullptr);
            } else if (!rc) {
                const char *o = s->ohost + jb.o_off;
                const int32_t *cnt = (const int32_t *)(o + jb.ol.cnt);
                memcpy(q->bl, o, sizeof(float) * j.cells);
                if (q->cnt) memcpy(q->cnt, cnt, sizeof(int32_t) * j.cells);
                if (j.want_norm && q->norm) memcpy(q->norm, o + jb.ol.norm, sizeof(float) * j.cells);
                int64_t ev = 0, pr = 0, dg = 0;
                count_tri(j, cnt, &ev, &pr, &dg);
                account_counts(c, j, ev, pr, dg);
                if (j.want_norm) {
                    float mg;
                    memcpy(&mg, o + jb.ol.margin, sizeof mg);
                    bool again = false;
                    rc = check_cert(c, mg, &again);
                    if (!rc && again) rc = norm_exact(c, j.cube, j.r1, j.r2, j.r3, q->norm, nullptr, nullptr);
                }
            }
            unspin(c, jb.pins);
            auto cit = c->cubes.find(j.cube);
            if (cit != c->cubes.end() && cit->second.serving > 0) cit->second.serving--;
            s->cpool.give(jb.c_off, jb.c_total);
            s->opool.give(jb.o_off, jb.o_bytes);
            done.emplace_back(jb.tk, rc);
            s->jobs.erase(jid);
            s->n_jobs--;
        }
    }
    std::lock_guard<std::mutex> g(s->m);
    for (auto &d : done) tri_done(s, d.first, d.second);
    s->cv_up.notify_all(); /* an upload waiting for memory may retry */
}

void sched_loop(Srv *s)
{
    amb_ctx *c = s->c;
    hipSetDevice(c->device);
    for (;;) {
        admit(s);
        bool launched = false;
        if ((int)s->inflight.size() < INFLIGHT && !s->active.empty()) {
            Lock lk(c->mu);
            launched = launch_iter(s);
        }
        if (!s->inflight.empty() && (!launched || (int)s->inflight.size() >= INFLIGHT)) {
            complete_oldest(s);
            continue;
        }
        if (launched) continue;
        std::unique_lock<std::mutex> lk(s->m);
        if (s->stop && s->waiting.empty() && s->uploads.empty() && s->jobs.empty()) return;
        /* idle: woken by a submission or a finished upload */
        s->cv_work.wait_for(lk, std::chrono::milliseconds(20));
    }
}

/* uploads cubes carried by requests, in arrival order */
void upload_loop(Srv *s)
{
    amb_ctx *c = s->c;
    hipSetDevice(c->device);
    for (;;) {
        uint64_t id;
        amb_cube_src src;
        std::vector<int> chans;
        int cid;
        {
            std::unique_lock<std::mutex> lk(s->m);
            s->cv_up.wait(lk, [&] { return s->stop || !s->uploads.empty(); });
            if (s->uploads.empty()) return;
            id = s->uploads.front();
            Srv::Ticket &t = s->tickets[id];
            src = t.up;
            chans = t.up_chans;
            cid = t.up_cube;
        }
        /* look-ahead: at most MAX_HELD cubes uploaded for requests not yet admitted */
        for (;;) {
            {
                Lock lk(c->mu);
                if (held_cubes(c) < MAX_HELD) break;
            }
            std::unique_lock<std::mutex> g(s->m);
            s->cv_up.wait_for(g, std::chrono::milliseconds(20));
        }
        Cube cb;
        cb.N = src.N; cb.nc = src.nc; cb.Du = src.Du; cb.flag_key = src.flag_key;
        const size_t nv = (size_t)src.N * (src.N / 2 + 1) * src.nc;
        int rc = AMB_OK;
        for (;;) { /* device buffers, evicting idle uploaded cubes, else waiting for finishing triangles */
            Lock lk(c->mu);
            void *v = nullptr, *sm = nullptr;
            rc = dev_alloc(c, &v, sizeof(float2) * nv, "uploaded cube");
            if (!rc) {
                rc = dev_alloc(c, &sm, sizeof(float) * nv, "uploaded sampling cube");
                if (rc) dev_free(c, v, sizeof(float2) * nv);
            }
            if (!rc) {
                cb.vis_t = (float2 *)v;
                cb.samp_t = (float *)sm;
                cb.bytes = sizeof(float2) * nv + sizeof(float) * nv;
                break;
            }
            if (rc != AMB_EBUDGET) break;
            if (evict_cube(c, s)) continue;
            if (s->n_jobs.load() == 0) break; /* nothing will free memory */
            std::unique_lock<std::mutex> g(s->m);
            s->cv_up.wait_for(g, std::chrono::milliseconds(20));
        }
        if (!rc) rc = upload_cube(c, cb, src.gv, src.samp, s->ustream, s->ustg, s->ustg_bytes);
        if (!rc) {
            Lock lk(c->mu);
            c->cubes[cid] = cb;
            rc = set_channels_core(c, cid, chans.empty() ? nullptr : chans.data(), (int)chans.size(), src.nc10);
            Cube &r = c->cubes[cid];
            r.uploaded = true;
            r.holds = 1; /* until its request is admitted */
            r.last_use = ++c->tick;
            c->st.n_cube_uploads++;
            if (rc) {
                release_cube_buffers(c, cid);
            }
        } else if (cb.vis_t) {
            Lock lk(c->mu);
            dev_free(c, cb.vis_t, sizeof(float2) * nv);
            dev_free(c, cb.samp_t, sizeof(float) * nv);
        }
        if (rc) fail_rest(s, id, rc);
        std::lock_guard<std::mutex> lk(s->m);
        s->uploads.pop_front();
        s->pending.erase(cid);
        s->cv_work.notify_all();
    }
}

void server_free(amb_ctx *c, Srv *s)
{
    if (s->ustream) { hipStreamSynchronize(s->ustream); hipStreamDestroy(s->ustream); }
    hipStreamSynchronize(c->stream);
    dev_free(c, s->cbuf, s->cpool.cap);
    dev_free(c, s->obuf, s->opool.cap);
    dev_free(c, s->ustg, s->ustg_bytes);
    dev_free(c, s->lag_st, s->lag_st_bytes);
    for (int k = 0; k < NSLOT; ++k) {
        dev_free(c, s->dslot[k], s->slot_bytes);
        if (s->hslot[k]) hipHostFree(s->hslot[k]);
        if (s->ev[k]) hipEventDestroy(s->ev[k]);
    }
    if (s->ohost) hipHostFree(s->ohost);
    delete s;
}

/* checks a request against the pools and the queue (context lock held) */
int check_request(amb_ctx *c, Srv *s, const amb_tri_req *req, int n, const amb_cube_src *up)
{
    for (int i = 0; i < n; ++i) {
        if (!req[i].bl) return fail(c, AMB_EARG, "request %d has no bl buffer", i);
        int nact, W;
        if (up && req[i].cube_id == 0) {
            nact = up->flag_key == 1 ? up->uf_nc : (int)up->nc;
            W = 2 * up->nc10 + 1;
        } else {
            auto it = c->cubes.find(req[i].cube_id);
            if (it == c->cubes.end()) {
                if (s->pending.count(req[i].cube_id)) continue; /* sized when its upload was submitted */
                return fail(c, AMB_ESTATE, "unknown cube id %d", req[i].cube_id);
            }
            nact = it->second.nact();
            W = 2 * it->second.nc10 + 1;
        }
        const size_t n3 = (size_t)nact * nact * nact;
        /* lag-Fourier partial sums depend on the grid and are checked at admission */
        const size_t cb = s->lag ? 0 : al(n3 * sizeof(float)) * (req[i].norm ? 2 : 1);
        const size_t ob = s->lag ? al((size_t)W * W * sizeof(float)) : out_layout((size_t)W * W, req[i].norm != nullptr).total;
        if (cb > s->cpool.cap || ob > s->opool.cap)
            return fail(c, AMB_EBUDGET, "triangle %d needs %zu + %zu bytes, the serving pools hold %zu + %zu", i, cb,
                        ob, s->cpool.cap, s->opool.cap);
    }
    return AMB_OK;
}

int submit(amb_ctx *c, const amb_cube_src *up, const amb_tri_req *req, int n, int *cube_id, amb_ticket *ticket)
{
    if (!c || !req || n <= 0 || !ticket || (up && !cube_id)) return AMB_EARG;
    Srv *s;
    Srv::Ticket t;
    {
        Lock lk(c->mu);
        s = c->server;
        if (!s) return fail(c, AMB_ESTATE, "serving not started");
        if (up) {
            if (!up->gv || !up->samp || up->N < 4 || up->N % 2 || up->nc < 1 || !(up->Du > 0) || up->nc10 < 0 ||
                up->nc10 > up->nc - 1 || (up->flag_key == 1 && (!up->chans || up->uf_nc <= 0)))
                return fail(c, AMB_EARG, "carried cube: bad header or channels");
            const size_t nv = (size_t)up->N * (up->N / 2 + 1) * up->nc;
            const uint64_t reserved = s->cpool.cap + 2 * s->opool.cap + s->ustg_bytes;
            if (nv * 12 + reserved > c->budget)
                return fail(c, AMB_EBUDGET, "carried cube of %zu bytes does not fit the HBM budget", nv * 12);
        }
        std::lock_guard<std::mutex> g(s->m);
        int rc = check_request(c, s, req, n, up);
        if (rc) return rc;
        if (s->stop) return fail(c, AMB_ESTATE, "serving is stopping");
        if ((int)s->waiting.size() >= MAX_QUEUE) return fail(c, AMB_EBUSY, "%d requests wait for admission", MAX_QUEUE);
        t.req.assign(req, req + n);
        if (up) {
            t.up = *up;
            if (up->flag_key == 1) t.up_chans.assign(up->chans, up->chans + up->uf_nc);
            t.up_cube = c->next_cube++;
            for (auto &q : t.req)
                if (q.cube_id == 0) q.cube_id = t.up_cube;
            *cube_id = t.up_cube;
        }
        t.left = (size_t)n;
        t.t_submit = now_ms();
        const uint64_t id = s->next++;
        if (up) {
            s->pending.insert(t.up_cube);
            s->uploads.push_back(id);
        }
        s->tickets[id] = std::move(t);
        s->waiting.push_back(id);
        *ticket = id;
    }
    s->cv_work.notify_all();
    if (up) s->cv_up.notify_all();
    return AMB_OK;
}

} // namespace

extern "C" {

int amb_serve_start(amb_ctx *c, int max_active)
{
    if (!c || max_active <= 0) return AMB_EARG;
    Lock lk(c->mu);
    hipSetDevice(c->device);
    if (c->server) return fail(c, AMB_ESTATE, "serving already started");
    Srv *s = new (std::nothrow) Srv();
    if (!s) return AMB_ENOMEM;
    s->c = c;
    s->max_active = std::min(max_active, 4096);
    s->lag = c->prec == AMB_PREC_LAGFT;
    const size_t ccap = al(std::min<size_t>((size_t)16 << 30, c->budget / 4));
    const size_t ocap = al(std::min<size_t>((size_t)2 << 30, c->budget / 16));
    s->slot_bytes = al((sizeof(ContractArgs) * 2 + sizeof(ReduceArgs) + 3 * sizeof(long) + 2 * lag_desc_bytes()) *
                           (size_t)(s->max_active + 1) +
                       16 * ALIGN);
    void *p = nullptr;
    int rc = dev_alloc(c, &p, ccap, "serving partial-sum pool");
    if (!rc) { s->cbuf = (char *)p; s->cpool.init(ccap); }
    if (!rc && !(rc = dev_alloc(c, &p, ocap, "serving result pool"))) { s->obuf = (char *)p; s->opool.init(ocap); }
    const size_t ucap = al(std::min<size_t>(UPLOAD_STAGING, c->budget / 16));
    if (!rc && !(rc = dev_alloc(c, &p, ucap, "upload staging"))) { s->ustg = p; s->ustg_bytes = ucap; }
    if (!rc && s->lag && !(rc = dev_alloc(c, &p, lag_st_cap(), "lag-Fourier spectrum"))) {
        s->lag_st = p;
        s->lag_st_bytes = lag_st_cap();
    }
    for (int k = 0; k < NSLOT && !rc; ++k) {
        if ((rc = dev_alloc(c, &p, s->slot_bytes, "serving descriptors"))) break;
        s->dslot[k] = (char *)p;
        if (hipHostMalloc(&p, s->slot_bytes, 0) != hipSuccess || hipEventCreateWithFlags(&s->ev[k], hipEventDisableTiming) != hipSuccess)
            rc = fail(c, AMB_EHIP, "serving descriptors: pinned memory or event");
        else s->hslot[k] = (char *)p;
    }
    if (!rc && hipHostMalloc(&p, ocap, 0) != hipSuccess) rc = fail(c, AMB_EHIP, "serving result staging (%zu bytes)", ocap);
    else if (!rc) s->ohost = (char *)p;
    /* default priority, like every stream of the library */
    if (!rc && hipStreamCreateWithFlags(&s->ustream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(c, AMB_EHIP, "upload stream");
    if (rc) {
        (void)hipGetLastError();
        server_free(c, s);
        return rc;
    }
    c->server = s;
    s->sched = std::thread(sched_loop, s);
    s->upl = std::thread(upload_loop, s);
    return AMB_OK;
}

int amb_submit(amb_ctx *c, const amb_tri_req *req, int n, amb_ticket *ticket)
{
    return submit(c, nullptr, req, n, nullptr, ticket);
}

int amb_submit_cube(amb_ctx *c, const amb_cube_src *cube, const amb_tri_req *req, int n, int *cube_id,
                    amb_ticket *ticket)
{
    if (!cube) return AMB_EARG;
    return submit(c, cube, req, n, cube_id, ticket);
}

int amb_poll(amb_ctx *c, amb_ticket ticket, int *done)
{
    if (!c || !done) return AMB_EARG;
    Srv *s = c->server;
    if (!s) return AMB_ESTATE;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return AMB_EARG;
    *done = it->second.done ? 1 : 0;
    return AMB_OK;
}

int amb_wait(amb_ctx *c, amb_ticket ticket, double *latency_ms)
{
    if (!c) return AMB_EARG;
    Srv *s = c->server;
    if (!s) return AMB_ESTATE;
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return AMB_EARG;
    s->cv_done.wait(lk, [&] { return it->second.done && !it->second.busy; });
    const int rc = it->second.status;
    if (latency_ms) *latency_ms = it->second.t_done - it->second.t_submit;
    s->tickets.erase(it);
    return rc;
}

int amb_serve_stop(amb_ctx *c)
{
    if (!c) return AMB_EARG;
    Srv *s;
    {
        Lock lk(c->mu);
        s = c->server;
        if (!s) return AMB_OK;
    }
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_up.notify_all();
    s->cv_work.notify_all();
    if (s->upl.joinable()) s->upl.join();     /* finishes the uploads queued */
    if (s->sched.joinable()) s->sched.join(); /* drains every admitted and waiting request */
    Lock lk(c->mu);
    c->server = nullptr;
    server_free(c, s);
    return AMB_OK;
}

} // extern "C"

/*
 * Many triangles per launch: grouped waves (amb_bispec_batch /
 * amb_bispec_multi), captured graphs of a fixed batch, and the serving
 * queue.  The reference runs triangles one after another inside one
 * process per cube (amber.c:187-326); here a wave hands the work items of
 * all its triangles to one persistent contraction launch (largest
 * triangles first, so the tail is made of small items), reduces every
 * triangle's lag cells in one launch and returns all results in one copy.
 * Every triangle's arithmetic is the same as in a launch of its own, so
 * results do not depend on what else shares the wave.
 */
#include "amb_ctx.h"

#include <functional>

#include <algorithm>
#include <chrono>
#include <cstring>

using Lock = std::lock_guard<std::recursive_mutex>;
using namespace amb;

namespace {

constexpr size_t ALIGN = 256;
constexpr size_t OUT_CAP = (size_t)1 << 30;  /* result bytes per wave */
constexpr int MAX_WAVE = 1024;               /* triangles per wave */

size_t al(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

/* One wave: jobs in launch order with their offsets in the arenas. */
struct Wave {
    std::vector<TriJob> jobs;
    std::vector<int> req;       /* index of each job in the caller's list */
    std::vector<size_t> c_off;  /* element offset of the job's C / Cq block */
    std::vector<size_t> o_off;  /* byte offset of the job's results: bl | cnt | norm */
    size_t c_total = 0, o_total = 0;
    int nnorm = 0;
    /* descriptor block: data args, norm args, reduce args, item begins
       (data, norm), cell begins, two counters, one margin word per
       normalised triangle */
    size_t off_da = 0, off_na = 0, off_ra = 0, off_b = 0, off_nb = 0, off_cb = 0, off_ctr = 0, off_mg = 0;
    size_t d_total = 0;
    long items = 0, nitems = 0, cells = 0;
};

size_t job_out_bytes(const TriJob &j)
{
    return al(j.cells * sizeof(float)) + al(j.cells * sizeof(int)) + (j.want_norm ? al(j.cells * sizeof(float)) : 0);
}

size_t job_c_elems(const TriJob &j) { return al((size_t)j.nact * j.nact * j.nact * sizeof(float)) / sizeof(float); }

/* order jobs largest first and lay out arenas and descriptors */
void layout(Wave &w)
{
    const int n = (int)w.jobs.size();
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        const TriJob &a = w.jobs[x], &b = w.jobs[y];
        return a.pitch != b.pitch ? a.pitch > b.pitch : a.nact > b.nact;
    });
    std::vector<TriJob> jobs(n);
    std::vector<int> req(n);
    for (int i = 0; i < n; ++i) { jobs[i] = w.jobs[ord[i]]; req[i] = w.req[ord[i]]; }
    w.jobs.swap(jobs);
    w.req.swap(req);
    w.c_off.assign(n, 0);
    w.o_off.assign(n, 0);
    w.c_total = w.o_total = 0;
    w.nnorm = 0;
    w.items = w.nitems = w.cells = 0;
    for (int i = 0; i < n; ++i) {
        w.c_off[i] = w.c_total;
        w.c_total += job_c_elems(w.jobs[i]);
        w.o_off[i] = w.o_total;
        w.o_total += job_out_bytes(w.jobs[i]);
        w.nnorm += w.jobs[i].want_norm;
        w.items += amb_contract_items(w.jobs[i].nact);
        if (w.jobs[i].want_norm) w.nitems += amb_contract_items(w.jobs[i].nact);
        w.cells += (long)w.jobs[i].cells;
    }
    size_t o = 0;
    w.off_da = o; o = al(o + sizeof(ContractArgs) * n);
    w.off_na = o; o = al(o + sizeof(ContractArgs) * std::max(w.nnorm, 1));
    w.off_ra = o; o = al(o + sizeof(ReduceArgs) * n);
    w.off_b = o; o = al(o + sizeof(long) * (n + 1));
    w.off_nb = o; o = al(o + sizeof(long) * (w.nnorm + 1));
    w.off_cb = o; o = al(o + sizeof(long) * (n + 1));
    w.off_ctr = o; o = al(o + sizeof(unsigned long long) * 2 * AMB_GROUPED_COUNTERS);
    w.off_mg = o; o = al(o + sizeof(unsigned) * std::max(w.nnorm, 1));
    w.d_total = o;
}

struct WaveBufs {
    float *C;
    int *Cq;
    char *out;  /* device results */
    char *desc; /* device descriptors */
    char *host; /* pinned host copy of out, then the margin word */
};

/* host image of the descriptor block for the given buffers */
std::vector<char> descriptors(const Wave &w, const WaveBufs &b)
{
    const int n = (int)w.jobs.size();
    std::vector<char> h(w.d_total, 0);
    ContractArgs *da = (ContractArgs *)(h.data() + w.off_da);
    ContractArgs *na = (ContractArgs *)(h.data() + w.off_na);
    ReduceArgs *ra = (ReduceArgs *)(h.data() + w.off_ra);
    long *bg = (long *)(h.data() + w.off_b), *nb = (long *)(h.data() + w.off_nb), *cb = (long *)(h.data() + w.off_cb);
    unsigned *mg = (unsigned *)(b.desc + w.off_mg);
    long it = 0, nit = 0, cl = 0;
    int k = 0;
    for (int i = 0; i < n; ++i) {
        const TriJob &j = w.jobs[i];
        char *o = b.out + w.o_off[i];
        float *bl = (float *)o;
        int *cnt = (int *)(o + al(j.cells * sizeof(float)));
        float *norm = j.want_norm ? (float *)(o + al(j.cells * sizeof(float)) + al(j.cells * sizeof(int))) : nullptr;
        da[i] = j.a;
        da[i].C = b.C + w.c_off[i];
        bg[i] = it;
        it += amb_contract_items(j.nact);
        if (j.want_norm) {
            na[k] = j.b;
            na[k].C = nullptr;
            na[k].Cq = b.Cq + w.c_off[i];
            na[k].margin = mg + k; /* per triangle: an uncertified one is recomputed alone */
            nb[k] = nit;
            nit += amb_contract_items(j.nact);
            ++k;
        }
        ra[i] = j.ra;
        ra[i].C = b.C + w.c_off[i];
        ra[i].Cq = j.want_norm ? b.Cq + w.c_off[i] : nullptr;
        ra[i].bl = bl;
        ra[i].bl64 = nullptr;
        ra[i].cnt = cnt;
        ra[i].norm = norm;
        ra[i].norm64 = nullptr;
        cb[i] = cl;
        cl += (long)j.cells;
    }
    bg[n] = it;
    nb[k] = nit;
    cb[n] = cl;
    return h;
}

/* contraction partials (floats) of a wave that starts with triangle j:
   the tuned wave size of j's shape, at most a quarter of the HBM budget */
size_t wave_c_cap(amb_ctx *c, const TriJob &j)
{
    const amb_tune_entry t = tune_cfg(c, j.ngc, j.nact);
    return std::min<size_t>((size_t)t.wave_gib << 30, c->budget / 4) / sizeof(float);
}

/* the launch configuration of a wave: its costliest triangle's (the
   largest items x pitch), one kernel variant every triangle supports */
amb_tune_entry wave_cfg(amb_ctx *c, const Wave &w, int prec, int *var, int *blocks)
{
    size_t k = 0;
    double best = -1;
    for (size_t i = 0; i < w.jobs.size(); ++i) {
        const double cost = (double)amb_contract_items(w.jobs[i].nact) * (double)w.jobs[i].pitch;
        if (cost > best) { best = cost; k = i; }
    }
    const amb_tune_entry t = tune_cfg(c, w.jobs[k].ngc, w.jobs[k].nact);
    *var = t.variant;
    for (const TriJob &j : w.jobs) *var = std::min(*var, tune_variant(t, prec, j.nact, j.pitch));
    /* a multiple of the 8 XCDs, so that every work queue has takers */
    *blocks = std::max(8, (int)((long)c->grouped_blocks * t.grid_pct / 100) / 8 * 8);
    return t;
}

/* the launches of one wave on stream s (descriptors already on device);
   copy: the lag arrays' copy to b.host on s too (else the caller makes it) */
hipError_t launch_wave(amb_ctx *c, hipStream_t s, const Wave &w, const WaveBufs &b, bool timed, bool copy = true)
{
    const int n = (int)w.jobs.size();
    const int prec = c->prec, nprec = prec == AMB_PREC_FP32 ? AMB_PREC_FP32 : AMB_PREC_BF16X3;
    unsigned long long *ctr = (unsigned long long *)(b.desc + w.off_ctr);
    unsigned *mg = (unsigned *)(b.desc + w.off_mg);
    int var, nvar, blocks;
    const amb_tune_entry t = wave_cfg(c, w, prec, &var, &blocks);
    wave_cfg(c, w, nprec, &nvar, &blocks);
    if (timed) timer_start(c, &c->st.ms_contract);
    hipError_t e = amb_launch_contract_grouped(s, prec, 0, var, (const ContractArgs *)(b.desc + w.off_da),
                                               (const long *)(b.desc + w.off_b), n, w.items, ctr, blocks,
                                               t.xrun_log2);
    if (timed) timer_stop(c);
    if (e == hipSuccess && w.nnorm) {
        e = amb_launch_fill_u32(s, mg, w.nnorm, 0u);
        if (timed) timer_start(c, &c->st.ms_norm);
        if (e == hipSuccess)
            e = amb_launch_contract_grouped(s, nprec, 1, nvar, (const ContractArgs *)(b.desc + w.off_na),
                                            (const long *)(b.desc + w.off_nb), w.nnorm, w.nitems, ctr + AMB_GROUPED_COUNTERS,
                                            blocks, t.xrun_log2);
        if (timed) timer_stop(c);
    }
    if (timed) timer_start(c, &c->st.ms_reduce);
    if (e == hipSuccess)
        e = amb_launch_reduce_grouped(s, (const ReduceArgs *)(b.desc + w.off_ra), (const long *)(b.desc + w.off_cb), n,
                                      w.cells);
    if (timed) timer_stop(c);
    if (e == hipSuccess && copy) e = hipMemcpyAsync(b.host, b.out, w.o_total, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && w.nnorm)
        e = hipMemcpyAsync(b.host + w.o_total, mg, sizeof(unsigned) * w.nnorm, hipMemcpyDeviceToHost, s);
    return e;
}

/* results of a finished wave from the host copy into the requests; the
   requests whose normalisation sums missed their certificate are appended
   to *redo (recomputed exactly once every wave is done) */
int finish_wave(amb_ctx *c, const Wave &w, const char *host, const amb_tri_req *req, std::vector<int> *redo)
{
    /* the triangles' copies out of the pinned staging and their count sums
       are independent: spread them over a few host threads */
    const size_t nj = w.jobs.size();
    std::vector<int64_t> ev(nj), pr(nj), dg(nj);
    auto one = [&](size_t i) {
        const TriJob &j = w.jobs[i];
        const amb_tri_req &q = req[w.req[i]];
        const char *o = host + w.o_off[i];
        const int32_t *cnt = (const int32_t *)(o + al(j.cells * sizeof(float)));
        memcpy(q.bl, o, sizeof(float) * j.cells);
        if (q.cnt) memcpy(q.cnt, cnt, sizeof(int32_t) * j.cells);
        if (j.want_norm && q.norm)
            memcpy(q.norm, o + al(j.cells * sizeof(float)) + al(j.cells * sizeof(int)), sizeof(float) * j.cells);
        count_tri(j, cnt, &ev[i], &pr[i], &dg[i]);
    };
    const size_t nth = std::min<size_t>(nj, 8);
    if (nth <= 1) {
        for (size_t i = 0; i < nj; ++i) one(i);
    } else {
        std::vector<std::thread> th;
        th.reserve(nth);
        for (size_t t = 0; t < nth; ++t)
            th.emplace_back([&, t] {
                for (size_t i = t; i < nj; i += nth) one(i);
            });
        for (auto &x : th) x.join();
    }
    for (size_t i = 0; i < nj; ++i) account_counts(c, w.jobs[i], ev[i], pr[i], dg[i]);
    int k = 0;
    for (size_t i = 0; i < nj; ++i) {
        if (!w.jobs[i].want_norm) continue;
        float m;
        memcpy(&m, host + w.o_total + sizeof(unsigned) * k++, sizeof m);
        bool again = false;
        const int rc = check_cert(c, m, &again);
        if (rc) return rc;
        if (again) redo->push_back(w.req[i]);
    }
    return AMB_OK;
}

/* eager execution of one wave with the context's scratch, its results in
   device buffer hb, copied to host staging buffer hb on the copy stream
   (ev_cp[hb] marks the end): `overlap` (the previous wave's copy-out, from
   the other buffers) runs on the host while this wave's kernels run. On
   return the kernels are done; the copy may still run. */
int run_wave(amb_ctx *c, Wave &w, int hb, const std::function<int()> &overlap)
{
    layout(w);
    void *C = c->C, *Cq = c->Cq;
    int rc = grow(c, &C, &c->C_bytes, sizeof(float) * w.c_total, "contraction partials");
    c->C = (float *)C;
    if (!rc && w.nnorm) {
        rc = grow(c, &Cq, &c->Cq_bytes, sizeof(int) * w.c_total, "normalisation partials");
        c->Cq = (int *)Cq;
    }
    void **outp = hb ? &c->out2 : &c->out;
    /* the last copy out of buffer hb ended before its wave was finished */
    if (!rc) rc = grow(c, outp, hb ? &c->out2_bytes : &c->out_bytes, w.o_total, "lag arrays");
    if (!rc) rc = copy_stream(c);
    if (!rc) rc = grow(c, &c->wdesc, &c->wdesc_bytes, w.d_total, "wave descriptors");
    if (!rc) rc = grow_host(c, &c->hstage[hb], &c->hstage_bytes[hb], w.o_total + sizeof(unsigned) * std::max(w.nnorm, 1));
    if (rc) return rc;
    WaveBufs b{c->C, c->Cq, (char *)*outp, (char *)c->wdesc, (char *)c->hstage[hb]};
    std::vector<char> h = descriptors(w, b);
    HIPCHK(c, hipMemcpyAsync(b.desc, h.data(), h.size(), hipMemcpyHostToDevice, c->stream));
    hipError_t e = launch_wave(c, c->stream, w, b, true, false);
    if (e == hipSuccess) e = hipEventRecord(c->ev_red, c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->cstream, c->ev_red, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(b.host, b.out, w.o_total, hipMemcpyDeviceToHost, c->cstream);
    if (e == hipSuccess) e = hipEventRecord(c->ev_cp[hb], c->cstream);
    c->st.n_contract_launch++;
    const int orc = overlap();
    /* always drain: the descriptor copy reads h, which dies with this frame */
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = es;
    timers_collect(c);
    if (e != hipSuccess) {
        hipStreamSynchronize(c->cstream);
        return fail(c, AMB_EHIP, "grouped wave of %zu triangles: %s", w.jobs.size(), hipGetErrorString(e));
    }
    return orc;
}

int run_grouped_waves(amb_ctx *c, const amb_tri_req *req, int n, double t0, double *done_ms, std::vector<int> *redo);

/* the exact normalisation of the requests whose fast sums missed their
   certificate (after every wave: it has buffers of its own but shares the
   stream) */
int redo_norms(amb_ctx *c, const amb_tri_req *req, const std::vector<int> &redo, double t0, double *done_ms)
{
    for (int i : redo) {
        const amb_tri_req &q = req[i];
        int rc = norm_exact(c, q.cube_id, q.r1, q.r2, q.r3, q.norm, nullptr, nullptr);
        if (rc) return rc;
        if (done_ms) done_ms[i] = now_ms() - t0;
    }
    return AMB_OK;
}

/* all requests as grouped waves; done_ms[i] gets request i's completion
   time relative to t0 (may be null). No result copy is left running on
   return, also after an error. */
int grouped_run(amb_ctx *c, const amb_tri_req *req, int n, double t0, double *done_ms)
{
    std::vector<int> redo;
    int rc = run_grouped_waves(c, req, n, t0, done_ms, &redo);
    if (c->cstream) hipStreamSynchronize(c->cstream);
    if (!rc) rc = redo_norms(c, req, redo, t0, done_ms);
    return rc;
}

int run_grouped_waves(amb_ctx *c, const amb_tri_req *req, int n, double t0, double *done_ms, std::vector<int> *redo)
{
    hipSetDevice(c->device);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    /* (cube, r1) together: one Ng and shared planes */
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return req[x].cube_id != req[y].cube_id ? req[x].cube_id < req[y].cube_id : req[x].r1 < req[y].r1;
    });
    const PlaneSrc src{&c->planes};
    /* w: the wave being assembled; prev: a wave whose kernels finished and
       whose results are copied (ev_cp[hb ^ 1]) into host staging buffer hb ^ 1 */
    Wave w, prev;
    int hb = 0;
    size_t c_tot = 0, o_tot = 0;
    auto finish_prev = [&]() -> int {
        if (prev.jobs.empty()) return AMB_OK;
        const hipError_t e = hipEventSynchronize(c->ev_cp[hb ^ 1]);
        if (e != hipSuccess) {
            prev = Wave{};
            return fail(c, AMB_EHIP, "copy of a grouped wave's results: %s", hipGetErrorString(e));
        }
        const int rc = finish_wave(c, prev, (const char *)c->hstage[hb ^ 1], req, redo);
        const double t = now_ms() - t0;
        if (done_ms)
            for (int r : prev.req) done_ms[r] = t;
        prev = Wave{};
        return rc;
    };
    auto flush = [&]() -> int {
        if (w.jobs.empty()) return AMB_OK;
        int rc = run_wave(c, w, hb, finish_prev);
        unpin_all(c);
        if (rc) return rc;
        prev = std::move(w);
        hb ^= 1;
        w = Wave{};
        c_tot = o_tot = 0;
        return AMB_OK;
    };
    for (int k = 0; k < n; ++k) {
        const amb_tri_req &q = req[order[k]];
        if (!q.bl) return fail(c, AMB_EARG, "request %d has no bl buffer", order[k]);
        Cube *cb;
        int rc = find_cube(c, q.cube_id, &cb);
        if (rc) { unpin_all(c); return rc; }
        const size_t nc3 = al((size_t)cb->nact() * cb->nact() * cb->nact() * sizeof(float)) / sizeof(float);
        if (!w.jobs.empty() && (c_tot + nc3 > wave_c_cap(c, w.jobs[0]) || (int)w.jobs.size() >= c->wave_max))
            if ((rc = flush())) return rc;
        TriJob j;
        rc = prepare_tri(c, src, q.cube_id, q.r1, q.r2, q.r3, q.norm != nullptr, j);
        if (rc == AMB_EBUDGET && !w.jobs.empty()) { /* planes of the wave so far hold the memory */
            if ((rc = flush())) return rc;
            rc = prepare_tri(c, src, q.cube_id, q.r1, q.r2, q.r3, q.norm != nullptr, j);
        }
        if (rc) { unpin_all(c); return rc; }
        const size_t ob = job_out_bytes(j);
        if (!w.jobs.empty() && o_tot + ob > OUT_CAP) {
            /* this job's planes stay pinned across the flush: unpin_all
               would release them, so re-resolve after it */
            if ((rc = flush())) return rc;
            if ((rc = prepare_tri(c, src, q.cube_id, q.r1, q.r2, q.r3, q.norm != nullptr, j))) { unpin_all(c); return rc; }
        }
        w.jobs.push_back(j);
        w.req.push_back(order[k]);
        c_tot += nc3;
        o_tot += ob;
    }
    const int rc = flush();
    return rc ? rc : finish_prev();
}

} // namespace

int amb::run_grouped(amb_ctx *c, const amb_tri_req *req, int n, double t0, double *done_ms)
{
    return grouped_run(c, req, n, t0, done_ms);
}

/* ------------------------------------------------------------------ graphs */
struct AmbGraph {
    std::vector<amb_tri_req> req;
    std::vector<Wave> waves;
    std::vector<std::vector<BuildOp>> builds; /* per wave, launched before its contraction */
    std::vector<int> cubes;
    std::vector<size_t> w_out; /* byte offset of each wave's results in out / host */
    std::vector<size_t> w_desc;
    /* the planes of one wave at a time live in `arena`: a wave's builds
       overwrite the previous wave's planes only after its kernels ran */
    void *arena = nullptr, *C = nullptr, *Cq = nullptr, *out = nullptr, *desc = nullptr, *G = nullptr,
         *host = nullptr;
    size_t arena_bytes = 0, C_bytes = 0, Cq_bytes = 0, out_bytes = 0, desc_bytes = 0, G_bytes = 0, host_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double fft_bytes = 0;
    long n_fft = 0;
};

static void graph_free(amb_ctx *c, AmbGraph *g)
{
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    dev_free(c, g->arena, g->arena_bytes);
    dev_free(c, g->C, g->C_bytes);
    dev_free(c, g->Cq, g->Cq_bytes);
    dev_free(c, g->out, g->out_bytes);
    dev_free(c, g->desc, g->desc_bytes);
    dev_free(c, g->G, g->G_bytes);
    if (g->host) hipHostFree(g->host);
    for (int id : g->cubes) {
        auto it = c->cubes.find(id);
        if (it != c->cubes.end()) it->second.graphs--;
    }
    delete g;
}

/* Split a fixed batch into waves by C size, plane bytes and count, by
   resolving every triangle against a measuring arena. */
static int plan_graph_waves(amb_ctx *c, const amb_tri_req *req, const std::vector<int> &order, size_t p_cap,
                            std::vector<std::vector<int>> *plan, size_t *arena)
{
    std::vector<int> cur;
    size_t c_cap = 0; /* the tuned wave size of the wave's first triangle */
    std::vector<PlaneSet> list;
    std::vector<BuildOp> rec;
    Bump mb;
    size_t c_tot = 0;
    *arena = 0;
    for (int i : order) {
        const amb_tri_req &q = req[i];
        if (!q.bl) return fail(c, AMB_EARG, "request %d has no bl buffer", i);
        std::vector<PlaneSet> l2 = list;
        std::vector<BuildOp> r2 = rec;
        Bump b2 = mb;
        TriJob j;
        int rc = prepare_tri(c, PlaneSrc{&l2, &r2, &b2}, q.cube_id, q.r1, q.r2, q.r3, q.norm != nullptr, j);
        if (rc) return rc;
        const size_t nc3 = job_c_elems(j);
        if (!cur.empty() && (c_tot + nc3 > c_cap || b2.off > p_cap || (int)cur.size() >= MAX_WAVE)) {
            plan->push_back(cur);
            *arena = std::max(*arena, mb.off);
            cur.clear();
            l2.clear();
            r2.clear();
            b2 = Bump{};
            c_tot = 0;
            if ((rc = prepare_tri(c, PlaneSrc{&l2, &r2, &b2}, q.cube_id, q.r1, q.r2, q.r3, q.norm != nullptr, j)))
                return rc;
        }
        if (b2.off > p_cap)
            return fail(c, AMB_EBUDGET, "triangle (%d,%d,%d): planes of %zu bytes exceed the graph arena cap %zu",
                        q.r1, q.r2, q.r3, b2.off, p_cap);
        if (cur.empty()) c_cap = wave_c_cap(c, j);
        cur.push_back(i);
        list.swap(l2);
        rec.swap(r2);
        mb = b2;
        c_tot += nc3;
    }
    if (!cur.empty()) {
        plan->push_back(cur);
        *arena = std::max(*arena, mb.off);
    }
    return AMB_OK;
}

extern "C" {

int amb_bispec_multi(amb_ctx *c, const amb_tri_req *req, int n)
{
    if (!c || (n > 0 && !req) || n < 0) return AMB_EARG;
    Lock lk(c->mu);
    const double t0 = now_ms();
    c->batch_ms.assign(n, 0.0);
    if (c->prec == AMB_PREC_LAGFT) return lagft_run(c, req, n, t0, c->batch_ms.data(), nullptr, nullptr);
    if (c->batch_mode == AMB_BATCH_PER_TRIANGLE) {
        std::vector<int> order(n);
        for (int i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return req[x].cube_id != req[y].cube_id ? req[x].cube_id < req[y].cube_id : req[x].r1 < req[y].r1;
        });
        for (int k = 0; k < n; ++k) {
            const amb_tri_req &q = req[order[k]];
            if (!q.bl) return fail(c, AMB_EARG, "request %d has no bl buffer", order[k]);
            amb_result r{};
            r.bl = q.bl;
            r.cnt = q.cnt;
            r.norm = q.norm;
            int rc = bispec_core(c, q.cube_id, q.r1, q.r2, q.r3, q.norm != nullptr, &r);
            if (rc) return rc;
            c->batch_ms[order[k]] = now_ms() - t0;
        }
        return AMB_OK;
    }
    return run_grouped(c, req, n, t0, c->batch_ms.data());
}

int amb_bispec_batch(amb_ctx *c, int id, const int *tri, int ntri, float *bl, int32_t *cnt, float *norm)
{
    if (!c || !tri || ntri < 0 || !bl) return AMB_EARG;
    Lock lk(c->mu);
    Cube *cb;
    int rc = find_cube(c, id, &cb);
    if (rc) return rc;
    if (!cb->chans) return fail(c, AMB_ESTATE, "channels not set for cube %d", id);
    const size_t cells = (size_t)(2 * cb->nc10 + 1) * (2 * cb->nc10 + 1);
    std::vector<amb_tri_req> req(ntri);
    for (int i = 0; i < ntri; ++i) {
        req[i].cube_id = id;
        req[i].r1 = tri[3 * i];
        req[i].r2 = tri[3 * i + 1];
        req[i].r3 = tri[3 * i + 2];
        req[i].bl = bl + (size_t)i * cells;
        req[i].cnt = cnt ? cnt + (size_t)i * cells : nullptr;
        req[i].norm = norm ? norm + (size_t)i * cells : nullptr;
    }
    return amb_bispec_multi(c, req.data(), ntri);
}

int amb_graph_create(amb_ctx *c, const amb_tri_req *req, int n, amb_graph **out)
{
    if (!c || !req || n <= 0 || !out) return AMB_EARG;
    Lock lk(c->mu);
    hipSetDevice(c->device);
    *out = nullptr;
    if (c->prec == AMB_PREC_LAGFT)
        return fail(c, AMB_ESTATE, "the lag-Fourier form runs its own waves; it is not captured in graphs");
    AmbGraph *g = new (std::nothrow) AmbGraph();
    if (!g) return AMB_ENOMEM;
    g->req.assign(req, req + n);
    for (int i = 0; i < n; ++i)
        if (c->cubes.count(req[i].cube_id) &&
            std::find(g->cubes.begin(), g->cubes.end(), req[i].cube_id) == g->cubes.end()) {
            g->cubes.push_back(req[i].cube_id);
            c->cubes[req[i].cube_id].graphs++;
        }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return req[x].cube_id != req[y].cube_id ? req[x].cube_id < req[y].cube_id : req[x].r1 < req[y].r1;
    });
    const size_t p_cap = std::min<size_t>((size_t)96 << 30, c->budget / 2);
    std::vector<std::vector<int>> plan;
    size_t arena = 0;
    int rc = plan_graph_waves(c, req, order, p_cap, &plan, &arena);
    if (!rc) rc = grow(c, &g->arena, &g->arena_bytes, std::max<size_t>(arena, 256), "graph plane arena");
    /* resolve each wave against the real arena, recording its builds */
    for (size_t k = 0; k < plan.size() && !rc; ++k) {
        Wave w;
        std::vector<PlaneSet> list;
        std::vector<BuildOp> rec;
        Bump bump{(char *)g->arena, 0, g->arena_bytes};
        for (int i : plan[k]) {
            TriJob j;
            const amb_tri_req &q = req[i];
            if ((rc = prepare_tri(c, PlaneSrc{&list, &rec, &bump}, q.cube_id, q.r1, q.r2, q.r3, q.norm != nullptr, j)))
                break;
            w.jobs.push_back(j);
            w.req.push_back(i);
        }
        g->waves.push_back(w);
        g->builds.push_back(rec);
    }
    if (rc) { graph_free(c, g); return rc; }
    /* arenas: C and Cq reused wave after wave, results and descriptors per wave */
    size_t cmax = 0, otot = 0, dtot = 0, gmax = 0;
    bool any_norm = false;
    for (auto &wv : g->waves) {
        layout(wv);
        cmax = std::max(cmax, wv.c_total);
        g->w_out.push_back(otot);
        g->w_desc.push_back(dtot);
        otot += al(wv.o_total + sizeof(unsigned) * std::max(wv.nnorm, 1));
        dtot += al(wv.d_total);
        any_norm |= wv.nnorm > 0;
    }
    for (auto &wb : g->builds)
        for (auto &op : wb) {
            gmax = std::max(gmax, build_scratch_bytes(c, op));
            if (op.kind <= 1) { g->fft_bytes += build_fft_bytes(c, op); g->n_fft += 2; }
        }
    if ((rc = grow(c, &g->C, &g->C_bytes, sizeof(float) * cmax, "graph contraction partials")) ||
        (any_norm && (rc = grow(c, &g->Cq, &g->Cq_bytes, sizeof(int) * cmax, "graph normalisation partials"))) ||
        (rc = grow(c, &g->out, &g->out_bytes, otot, "graph lag arrays")) ||
        (rc = grow(c, &g->desc, &g->desc_bytes, dtot, "graph descriptors")) ||
        (gmax && (rc = grow(c, &g->G, &g->G_bytes, gmax, "graph fft scratch"))) ||
        (rc = grow_host(c, &g->host, &g->host_bytes, otot))) {
        graph_free(c, g);
        return rc;
    }
    hipError_t e = hipSuccess;
    for (size_t k = 0; k < g->waves.size() && e == hipSuccess; ++k) {
        WaveBufs b{(float *)g->C, (int *)g->Cq, (char *)g->out + g->w_out[k], (char *)g->desc + g->w_desc[k],
                   (char *)g->host + g->w_out[k]};
        std::vector<char> h = descriptors(g->waves[k], b);
        e = hipMemcpy(b.desc, h.data(), h.size(), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    /* capture: per wave its plane builds, then its launches */
    if (e == hipSuccess) e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = hipSuccess;
        for (size_t k = 0; k < g->waves.size() && le == hipSuccess; ++k) {
            for (auto &op : g->builds[k])
                if (le == hipSuccess) le = launch_build(c, c->stream, op, (float2 *)g->G);
            WaveBufs b{(float *)g->C, (int *)g->Cq, (char *)g->out + g->w_out[k], (char *)g->desc + g->w_desc[k],
                       (char *)g->host + g->w_out[k]};
            if (le == hipSuccess) le = launch_wave(c, c->stream, g->waves[k], b, false);
        }
        e = hipStreamEndCapture(c->stream, &g->graph);
        if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(c, AMB_EHIP, "graph capture of %d triangles: %s", n, hipGetErrorString(e));
        graph_free(c, g);
        return rc;
    }
    c->graphs.push_back(g);
    *out = (amb_graph *)g;
    return AMB_OK;
}

int amb_graph_run(amb_ctx *c, amb_graph *gh)
{
    if (!c || !gh) return AMB_EARG;
    Lock lk(c->mu);
    AmbGraph *g = (AmbGraph *)gh;
    if (std::find(c->graphs.begin(), c->graphs.end(), g) == c->graphs.end()) return fail(c, AMB_EARG, "unknown graph");
    hipSetDevice(c->device);
    timer_start(c, &c->st.ms_graph);
    hipError_t e = hipGraphLaunch(g->exec, c->stream);
    timer_stop(c);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    timers_collect(c);
    if (e != hipSuccess) return fail(c, AMB_EHIP, "graph launch: %s", hipGetErrorString(e));
    c->st.n_graph_launch++;
    c->st.bytes_fft += g->fft_bytes;
    c->st.n_fft_launch += g->n_fft;
    int rc = AMB_OK;
    std::vector<int> redo;
    for (size_t k = 0; k < g->waves.size(); ++k) {
        int r = finish_wave(c, g->waves[k], (const char *)g->host + g->w_out[k], g->req.data(), &redo);
        if (r && !rc) rc = r;
    }
    /* outside the graph: the exact normalisation of uncertified triangles */
    if (!rc) rc = redo_norms(c, g->req.data(), redo, now_ms(), nullptr);
    return rc;
}

void amb_graph_destroy(amb_ctx *c, amb_graph *gh)
{
    if (!c || !gh) return;
    Lock lk(c->mu);
    AmbGraph *g = (AmbGraph *)gh;
    auto it = std::find(c->graphs.begin(), c->graphs.end(), g);
    if (it == c->graphs.end()) return;
    c->graphs.erase(it);
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    graph_free(c, g);
}

} // extern "C"

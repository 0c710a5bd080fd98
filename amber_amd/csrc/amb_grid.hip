/*
 * The upstream gridder vcg on the GPU (3_vcg/tgefuncs.c:377-626 and the
 * channel collapse :816-839).  The reference walks the visibility groups in
 * file order and, for each group inside Umax, adds its RR visibilities of
 * every channel into one uv cell: GV[cell][n] += (re, signv im) and
 * NTiG[cell][n] += 1 (flag_key 0) or [wt > 0] (flag_key 1), with float
 * values summed into doubles starting from zero.
 *
 * Each cell's sums involve only the groups that land in it, in group order.
 * The host computes every group's cell with the reference's float arithmetic
 * (amb_grid_cells) and sorts the groups by cell, keeping group order inside a
 * cell.  k_grid then gives each non-empty cell a block and each channel a
 * thread, which adds that cell's groups in order: the same additions in the
 * same order as the reference, so the result is bit-identical, with no
 * atomics.  k_collapse sums k adjacent channels from zero in channel order
 * as collapse_channels does.
 *
 * Layout: the visibility groups stay as read ([group][nchan][nstokes][3]
 * floats); the output is the GV_data / sampling file layout,
 * [NN (NN/2+1)][nc][2] and [..][nc] doubles, or, for amb_grid_cube, the
 * resident plane-major complex64 cube the bispectrum path reads.
 * Roofline: HBM; per gridded (group, channel) 12 bytes read, per (cell,
 * channel) 24 bytes written.
 */
#include "amb_ctx.h"

#include <algorithm>
#include <vector>

namespace {

__global__ __launch_bounds__(256) void k_grid(const float *__restrict__ data, const long long *__restrict__ gidx,
                                              const float *__restrict__ gsign, const int *__restrict__ cstart,
                                              const long long *__restrict__ ccell, long long gstride, int cofs,
                                              int nstokes, int nc, int flag_key, double2 *__restrict__ gv,
                                              double *__restrict__ samp)
{
    const int n = blockIdx.y * blockDim.x + threadIdx.x;
    if (n >= nc) return;
    const int k = blockIdx.x;
    double re = 0.0, im = 0.0, ng = 0.0;
    for (int i = cstart[k]; i < cstart[k + 1]; ++i) {
        const float *RR = data + gidx[i] * gstride + (long long)(cofs + n) * nstokes * 3;
        const float sv = gsign[i];
        float r = RR[0], m = sv * RR[1], w = 1.f;
        if (flag_key == 1 && !(RR[2] > 0.f)) {
            r = 0.f;
            m = 0.f;
            w = 0.f;
        }
        re += (double)r;
        im += (double)m;
        ng += (double)w;
    }
    const size_t o = (size_t)ccell[k] * nc + n;
    gv[o] = make_double2(re, im);
    samp[o] = ng;
}

__global__ void k_collapse(const double2 *__restrict__ gv, const double *__restrict__ samp, long long ngrid, int nc,
                           int k, int ncout, double2 *__restrict__ gvo, double *__restrict__ so)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)ngrid * ncout) return;
    const size_t idx = e / ncout, col = e - idx * ncout;
    double sr = 0.0, si = 0.0, sn = 0.0;
    for (int ki = 0; ki < k; ++ki) {
        const size_t i = idx * nc + col * k + ki;
        sr += gv[i].x;
        si += gv[i].y;
        sn += samp[i];
    }
    gvo[e] = make_double2(sr, si);
    so[e] = sn;
}

} // namespace

using namespace amb;

/* grid on the device; the result (file layout, nc_out channels) is left at
 * *out_gv / *out_samp inside the block *blk of *blk_bytes */
static int grid_core(amb_ctx *c, const amb_grid_cfg *cfg, int64_t ngroup, const float *ranpar, const float *data,
                     void **blk, size_t *blk_bytes, double2 **out_gv, double **out_samp, int64_t *ngridded)
{
    if (!cfg || ngroup < 0 || (ngroup && (!ranpar || !data))) return fail(c, AMB_EARG, "grid: bad arguments");
    if (cfg->nc < 1 || cfg->chan1 < 1 || cfg->chan1 - 1 + cfg->nc > cfg->nchan || cfg->nstokes < 1 ||
        cfg->n_collapse < 1 || cfg->nc / cfg->n_collapse < 1 || cfg->NN < 4 || cfg->NN % 2 ||
        cfg->NN > (1 << 16) || cfg->nc > (1 << 16))
        return fail(c, AMB_EARG, "grid: channels %lld from %d of %lld, collapse %d, NN %lld", (long long)cfg->nc,
                    cfg->chan1, (long long)cfg->nchan, cfg->n_collapse, (long long)cfg->NN);
    const int64_t NN = cfg->NN, Nh = NN / 2 + 1, ngrid = NN * Nh, nc = cfg->nc;
    const int64_t ncout = nc / cfg->n_collapse;
    std::vector<int64_t> cell(std::max<int64_t>(ngroup, 1));
    std::vector<float> sign(std::max<int64_t>(ngroup, 1));
    const int64_t in = amb_grid_cells(cfg, ngroup, ranpar, cell.data(), sign.data());
    if (in < 0) return fail(c, AMB_EARG, "grid: a group falls outside the %lld x %lld grid", (long long)NN,
                            (long long)Nh);
    if (in > INT32_MAX) return fail(c, AMB_EARG, "grid: %lld groups exceed the cell tables", (long long)in);
    /* the gridded groups sorted by cell, group order kept inside a cell */
    std::vector<long long> ccell;
    std::vector<int> cstart;
    std::vector<long long> gidx(std::max<int64_t>(in, 1));
    std::vector<float> gsign(std::max<int64_t>(in, 1));
    {
        std::vector<int64_t> order(std::max<int64_t>(in, 1));
        int64_t m = 0;
        for (int64_t g = 0; g < ngroup; ++g)
            if (cell[g] >= 0) order[m++] = g;
        std::stable_sort(order.begin(), order.begin() + m, [&](int64_t a, int64_t b) { return cell[a] < cell[b]; });
        for (int64_t i = 0; i < m; ++i) {
            const int64_t g = order[i];
            if (i == 0 || cell[g] != cell[order[i - 1]]) {
                cstart.push_back((int)i);
                ccell.push_back(cell[g]);
            }
            gidx[i] = g;
            gsign[i] = sign[g];
        }
        cstart.push_back((int)m);
    }
    const int nonempty = (int)ccell.size();
    const int64_t gstride = cfg->nchan * cfg->nstokes * 3;
    const size_t data_bytes = sizeof(float) * (size_t)ngroup * gstride;
    const size_t full = (size_t)ngrid * nc;
    size_t off = 0;
    auto place = [&](size_t b) { size_t o = off; off += (b + 255) / 256 * 256; return o; };
    const size_t o_d = place(std::max<size_t>(data_bytes, 4));
    const size_t o_gi = place(sizeof(long long) * gidx.size()), o_gs = place(sizeof(float) * gsign.size());
    const size_t o_cs = place(sizeof(int) * cstart.size());
    const size_t o_cc = place(sizeof(long long) * std::max<size_t>(ccell.size(), 1));
    const size_t o_gv = place(sizeof(double2) * full), o_s = place(sizeof(double) * full);
    size_t o_gv2 = 0, o_s2 = 0;
    if (cfg->n_collapse > 1) {
        o_gv2 = place(sizeof(double2) * (size_t)ngrid * ncout);
        o_s2 = place(sizeof(double) * (size_t)ngrid * ncout);
    }
    void *base = nullptr;
    int rc = dev_alloc(c, &base, off, "gridder buffers");
    if (rc) return rc;
    char *d = (char *)base;
    hipStream_t st = c->stream;
    hipError_t e = hipSuccess;
    if (data_bytes) e = hipMemcpyAsync(d + o_d, data, data_bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && in) e = hipMemcpyAsync(d + o_gi, gidx.data(), sizeof(long long) * in, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && in) e = hipMemcpyAsync(d + o_gs, gsign.data(), sizeof(float) * in, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d + o_cs, cstart.data(), sizeof(int) * cstart.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nonempty)
        e = hipMemcpyAsync(d + o_cc, ccell.data(), sizeof(long long) * nonempty, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(d + o_gv, 0, sizeof(double2) * full, st);
    if (e == hipSuccess) e = hipMemsetAsync(d + o_s, 0, sizeof(double) * full, st);
    if (e == hipSuccess) {
        timer_start(c, &c->st.ms_grid);
        if (nonempty)
            hipLaunchKernelGGL(k_grid, dim3(nonempty, (unsigned)((nc + 255) / 256)), dim3(256), 0, st,
                               (const float *)(d + o_d), (const long long *)(d + o_gi), (const float *)(d + o_gs),
                               (const int *)(d + o_cs), (const long long *)(d + o_cc), (long long)gstride,
                               cfg->chan1 - 1, (int)cfg->nstokes, (int)nc, cfg->flag_key, (double2 *)(d + o_gv),
                               (double *)(d + o_s));
        if (cfg->n_collapse > 1) {
            const size_t tot = (size_t)ngrid * ncout;
            hipLaunchKernelGGL(k_collapse, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                               (const double2 *)(d + o_gv), (const double *)(d + o_s), (long long)ngrid, (int)nc,
                               cfg->n_collapse, (int)ncout, (double2 *)(d + o_gv2), (double *)(d + o_s2));
        }
        timer_stop(c);
        e = hipGetLastError();
    }
    /* the host tables above are copied from this frame: finish before it ends */
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        hipStreamSynchronize(st);
        timers_collect(c);
        dev_free(c, base, off);
        return fail(c, AMB_EHIP, "grid: %s", hipGetErrorString(e));
    }
    *blk = base;
    *blk_bytes = off;
    *out_gv = (double2 *)(d + (cfg->n_collapse > 1 ? o_gv2 : o_gv));
    *out_samp = (double *)(d + (cfg->n_collapse > 1 ? o_s2 : o_s));
    if (ngridded) *ngridded = in;
    c->st.n_grid_groups += in;
    return AMB_OK;
}

extern "C" int amb_grid(amb_ctx *c, const amb_grid_cfg *cfg, int64_t ngroup, const float *ranpar, const float *data,
                        double *gv, double *samp, int64_t *ngridded)
{
    if (!c || !gv || !samp) return AMB_EARG;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    hipSetDevice(c->device);
    void *blk = nullptr;
    size_t bytes = 0;
    double2 *dg = nullptr;
    double *ds = nullptr;
    int rc = grid_core(c, cfg, ngroup, ranpar, data, &blk, &bytes, &dg, &ds, ngridded);
    if (rc) return rc;
    const size_t n = (size_t)cfg->NN * (cfg->NN / 2 + 1) * (cfg->nc / cfg->n_collapse);
    hipError_t e = hipMemcpyAsync(gv, dg, sizeof(double2) * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(samp, ds, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    timers_collect(c);
    dev_free(c, blk, bytes);
    if (e != hipSuccess) return fail(c, AMB_EHIP, "grid: %s", hipGetErrorString(e));
    return AMB_OK;
}

extern "C" int amb_grid_cube(amb_ctx *c, const amb_grid_cfg *cfg, int64_t ngroup, const float *ranpar,
                             const float *data, int *cube_id, int64_t *ngridded)
{
    if (!c || !cube_id) return AMB_EARG;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    hipSetDevice(c->device);
    void *blk = nullptr;
    size_t bytes = 0;
    double2 *dg = nullptr;
    double *ds = nullptr;
    int rc = grid_core(c, cfg, ngroup, ranpar, data, &blk, &bytes, &dg, &ds, ngridded);
    if (rc) return rc;
    const int64_t N = cfg->NN, Nh = N / 2 + 1, ncout = cfg->nc / cfg->n_collapse;
    const size_t nv = (size_t)N * Nh * ncout;
    Cube cb;
    cb.N = N;
    cb.nc = ncout;
    cb.Du = cfg->dU;
    cb.flag_key = cfg->flag_key;
    void *v = nullptr, *s = nullptr;
    rc = dev_alloc(c, &v, sizeof(float2) * nv, "cube");
    if (!rc) {
        rc = dev_alloc(c, &s, sizeof(float) * nv, "sampling cube");
        if (rc) dev_free(c, v, sizeof(float2) * nv);
    }
    if (rc) {
        hipStreamSynchronize(c->stream);
        timers_collect(c);
        dev_free(c, blk, bytes);
        return rc;
    }
    cb.vis_t = (float2 *)v;
    cb.samp_t = (float *)s;
    cb.bytes = sizeof(float2) * nv + sizeof(float) * nv;
    /* the file-layout rows go through the same normalise + transpose as an
       uploaded cube */
    hipError_t e = amb_launch_upload(c->stream, (const double *)dg, ds, 0, (int)N, (int)N, (int)Nh, (int)ncout,
                                     cb.vis_t, cb.samp_t);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    timers_collect(c);
    dev_free(c, blk, bytes);
    if (e != hipSuccess) {
        dev_free(c, v, sizeof(float2) * nv);
        dev_free(c, s, sizeof(float) * nv);
        return fail(c, AMB_EHIP, "grid to cube: %s", hipGetErrorString(e));
    }
    const int id = c->next_cube++;
    c->cubes[id] = cb;
    *cube_id = id;
    return AMB_OK;
}

/*
 * MAPS C_l(dnu), the power-spectrum sibling of the bispectrum estimator
 * (4_fbe/maps.c:100-229 with cl_dnu, 4_fbe/amber_funcs.c:1418-1488).  For
 * every channel pair nn1 <= nn2 and every |U| bin b the reference sums, over
 * the half-plane cells of the bin in cell order,
 *   pksum_b(nn1, nn2) = sum_x (Re1 Re2 + Im1 Im2) NTiG1 / no_b(nn1)
 * and accumulates cldnu[b][nn2 - nn1] += pksum_b over nn1 ascending.  That is
 * one Gram matrix per bin, C_b = A_b^T B_b over the cells x of the bin, with
 * A_b[x][n] = NTiG(x,n) V(x,n) and B_b[x][n] = V(x,n) in complex FP64 (the
 * real part of A conj(B) is two real products per cell).
 *
 * Device layout: the cube stays pixel-major as in the files,
 * V[x][n] complex double and NTiG[x][n] double, so one cell's channels are
 * contiguous.  The host sorts the binned cells by bin (cell order kept
 * inside a bin) and cuts every bin into chunks of at most MAPS_KCH cells.
 *
 * Kernels, in stream order:
 *   k_maps_norm   V /= NTiG and NTiG := 1 where NTiG > 0 (maps.c:100-125),
 *                 in place, FP64 division as written;
 *   k_maps_pairs  per (chunk, 128x128 channel tile with tile(n1) <=
 *                 tile(n2)) the partial Gram block over the chunk's cells:
 *                 8-cell stages through LDS (the next stage's loads in
 *                 flight during the products), 8x8 FP64 outputs per thread,
 *                 written to its own slice of the partial array (no atomics);
 *   k_maps_counts per (chunk, channel) the number of sampled cells (exact);
 *   k_maps_sum    per bin, the chunk partials and counts added in chunk
 *                 order into the bin's first chunk;
 *   k_maps_last   per bin, in cell order, no and usum of the last pair's
 *                 channel nc-1, as the reference's last call leaves them in
 *                 binu/num (terms fetched in parallel, added by one thread);
 *   k_maps_lag    per (bin, lag d) the sum over n1 ascending of
 *                 C_b(n1, n1 + d) / no_b(n1), skipping the division where
 *                 no_b(n1) = 0 as the reference does.
 * Every sum has a fixed order, so results are deterministic run to run; the
 * Gram sums differ from the reference's sequential cell order by FP64
 * rounding only.  Counts are exact.
 *
 * Bins are computed on the host with the reference's expression and libm
 * (sqrt(ia^2 + j^2) Du, floor(log10(uu/umin)/binsiz)), so the cell-to-bin
 * map is the reference's exactly; a cell whose rounded index reaches NB
 * (possible only within an ulp of umax, an out-of-bounds write in the
 * reference) is counted in bin NB - 1.
 *
 * Roofline: the pair kernel does 4 FP64 flops per (pair, cell) on the
 * VALU, whose FP64 rate on gfx950 equals the FP64 MFMA rate (78.6 TFLOP/s),
 * and reads each cell's channels once per tile row and column from cache.
 */
#include "amb_ctx.h"

#include <cmath>
#include <vector>

namespace {

constexpr int MT = 128;         /* channel tile */
constexpr int MKP = 8;          /* cells per LDS stage */
constexpr int MAPS_KCH = 2048;  /* cells per chunk */

struct MapsChunk {
    int bin, start, count, pad;
};

__global__ void k_maps_norm(double2 *__restrict__ v, double *__restrict__ s, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double sv = s[i];
        if (sv > 0) {
            double2 x = v[i];
            x.x /= sv;
            x.y /= sv;
            v[i] = x;
            s[i] = 1.0;
        }
    }
}

__global__ __launch_bounds__(256) void k_maps_pairs(const double2 *__restrict__ v, const double *__restrict__ s,
                                                    const int *__restrict__ order,
                                                    const MapsChunk *__restrict__ chunks, int nc, int ntile,
                                                    double *__restrict__ part)
{
    __shared__ double2 As[MKP][MT], Bs[MKP][MT];
    int p = blockIdx.x, ti = 0;
    while (p >= ntile - ti) {
        p -= ntile - ti;
        ++ti;
    }
    const int tj = ti + p;
    const MapsChunk ch = chunks[blockIdx.y];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int lp = tid >> 5, lc = (tid & 31) * 4; /* loader: cell lp, channels lc..lc+3 */
    const int r0 = ti * MT, c0 = tj * MT;
    double acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;

    /* stage k0's operands into registers: A = NTiG V, B = V, zero past the
     * chunk or the channel count */
    double2 a[4], b[4];
    auto fetch = [&](int k0) {
        if (k0 + lp < ch.count) {
            const size_t x = (size_t)order[ch.start + k0 + lp] * nc;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int n1 = r0 + lc + q, n2 = c0 + lc + q;
                if (n1 < nc) {
                    const double2 t = v[x + n1];
                    const double w = s[x + n1];
                    a[q] = make_double2(t.x * w, t.y * w);
                } else {
                    a[q] = make_double2(0.0, 0.0);
                }
                b[q] = n2 < nc ? v[x + n2] : make_double2(0.0, 0.0);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = b[q] = make_double2(0.0, 0.0);
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < ch.count; k0 += MKP) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            As[lp][lc + q] = a[q];
            Bs[lp][lc + q] = b[q];
        }
        __syncthreads();
        if (k0 + MKP < ch.count) fetch(k0 + MKP); /* in flight during this stage's products */
#pragma unroll
        for (int kk = 0; kk < MKP; ++kk) {
            double2 ra[8], rb[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) ra[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 8; ++j) rb[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    acc[i][j] = fma(ra[i].x, rb[j].x, acc[i][j]);
                    acc[i][j] = fma(ra[i].y, rb[j].y, acc[i][j]);
                }
        }
    }
    double *out = part + (size_t)blockIdx.y * nc * nc;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int n1 = r0 + ty + 16 * i;
        if (n1 >= nc) continue;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int n2 = c0 + tx + 16 * j;
            if (n2 < nc) out[(size_t)n1 * nc + n2] = acc[i][j];
        }
    }
}

__global__ void k_maps_counts(const double *__restrict__ s, const int *__restrict__ order,
                              const MapsChunk *__restrict__ chunks, int nc, int *__restrict__ nop)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nc) return;
    const MapsChunk ch = chunks[blockIdx.y];
    int no = 0;
    for (int k = 0; k < ch.count; ++k) no = (int)(no + s[(size_t)order[ch.start + k] * nc + n]);
    nop[(size_t)blockIdx.y * nc + n] = no;
}

/* one block per bin: 256 cells' terms are fetched in parallel into LDS and
 * thread 0 adds them in cell order, the reference's sequential order and
 * expressions (no += NTiG as int, usum += uu * NTiG) */
__global__ __launch_bounds__(256) void k_maps_last(const double *__restrict__ s, const int *__restrict__ order,
                                                   const double *__restrict__ uu, const int *__restrict__ bstart,
                                                   int nc, int *__restrict__ no_last, double *__restrict__ us_last)
{
#pragma clang fp contract(off)
    __shared__ double wv[256], pv[256];
    const int b = blockIdx.x, t = threadIdx.x;
    const int k0 = bstart[b], k1 = bstart[b + 1];
    int no = 0;
    double us = 0.0;
    for (int base = k0; base < k1; base += 256) {
        const int m = min(256, k1 - base);
        if (t < m) {
            const double w = s[(size_t)order[base + t] * nc + (nc - 1)];
            wv[t] = w;
            pv[t] = uu[base + t] * w;
        }
        __syncthreads();
        if (t == 0)
            for (int i = 0; i < m; ++i) {
                no = (int)(no + wv[i]);
                us += pv[i];
            }
        __syncthreads();
    }
    if (t == 0) {
        no_last[b] = no;
        us_last[b] = us;
    }
}

/* fold a bin's chunk partials into its first chunk, chunk order kept:
 * part[c0] = part[c0] + part[c0+1] + ..., upper triangle only, and the
 * counts nop[c0][n1] likewise */
__global__ void k_maps_sum(double *__restrict__ part, int *__restrict__ nop, const int *__restrict__ bchunk, int nc)
{
    const int b = blockIdx.y;
    const int c0 = bchunk[b], c1 = bchunk[b + 1];
    if (c1 - c0 < 2) return;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nn = (size_t)nc * nc;
    if (e >= nn) return;
    const int n1 = (int)(e / nc), n2 = (int)(e - (size_t)n1 * nc);
    if (n2 >= n1) {
        double v = part[c0 * nn + e];
        for (int c = c0 + 1; c < c1; ++c) v += part[c * nn + e];
        part[c0 * nn + e] = v;
    }
    if (e < (size_t)nc) {
        int no = nop[(size_t)c0 * nc + e];
        for (int c = c0 + 1; c < c1; ++c) no += nop[(size_t)c * nc + e];
        nop[(size_t)c0 * nc + e] = no;
    }
}

__global__ void k_maps_lag(const double *__restrict__ part, const int *__restrict__ nop,
                           const int *__restrict__ bchunk, int nc, double *__restrict__ cl)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (d >= nc) return;
    const int c0 = bchunk[b], c1 = bchunk[b + 1];
    double acc = 0.0;
    if (c1 > c0) {
        const double *P = part + (size_t)c0 * nc * nc;
        const int *no = nop + (size_t)c0 * nc;
#pragma unroll 8
        for (int n1 = 0; n1 + d < nc; ++n1) {
            double v = P[(size_t)n1 * nc + n1 + d];
            const int k = no[n1];
            if (k != 0) v /= (1. * k);
            acc += v;
        }
    }
    cl[(size_t)b * nc + d] = acc;
}

} // namespace

using namespace amb;

extern "C" int amb_maps_cldnu(amb_ctx *c, int64_t N, int64_t nc, double Du, const double *gv, const double *samp,
                              int NB, double umin, double umax, double *cldnu, double *numdnu, double *binu,
                              int32_t *num)
{
    if (!c || !gv || !samp || !cldnu || !numdnu || !binu || !num) return AMB_EARG;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    hipSetDevice(c->device);
    if (N < 2 || nc < 1 || nc > (1 << 20) || !(Du > 0) || NB < 1 || !(umin > 0) || !(umax > umin))
        return fail(c, AMB_EARG, "MAPS: bad arguments N=%lld nc=%lld NB=%d umin=%g umax=%g", (long long)N,
                    (long long)nc, NB, umin, umax);
    const int64_t Nh = N / 2 + 1, npix = N * Nh;
    if (npix > (int64_t)1 << 30) return fail(c, AMB_EARG, "MAPS: grid N=%lld too large", (long long)N);

    /* cell -> bin with the reference's expression (amber_funcs.c:1446-1460) */
    const double binsiz = log10(umax / umin) / (1. * NB);
    std::vector<int> cbin(npix);
    std::vector<double> cuu(npix);
    std::vector<int> bcount(NB, 0);
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = 0; j < Nh; ++j) {
            const int64_t ia = (i > N / 2) ? (N - i) : i;
            const double uu = sqrt(1. * (ia * ia + j * j)) * Du;
            int b = -1;
            if ((uu >= umin) && (uu < umax)) {
                b = (int)floor(log10(uu / umin) / binsiz);
                if (b >= NB) b = NB - 1;
                if (b < 0) b = 0;
                ++bcount[b];
            }
            cbin[i * Nh + j] = b;
            cuu[i * Nh + j] = uu;
        }
    std::vector<int> bstart(NB + 1, 0);
    for (int b = 0; b < NB; ++b) bstart[b + 1] = bstart[b] + bcount[b];
    const int ncell = bstart[NB];
    std::vector<int> order(std::max(ncell, 1));
    std::vector<double> ouu(std::max(ncell, 1));
    {
        std::vector<int> fill(bstart.begin(), bstart.end() - 1);
        for (int64_t x = 0; x < npix; ++x)
            if (cbin[x] >= 0) {
                const int k = fill[cbin[x]]++;
                order[k] = (int)x;
                ouu[k] = cuu[x];
            }
    }
    std::vector<MapsChunk> chunks;
    std::vector<int> bchunk(NB + 1, 0);
    for (int b = 0; b < NB; ++b) {
        for (int k = bstart[b]; k < bstart[b + 1]; k += MAPS_KCH)
            chunks.push_back({b, k, std::min(MAPS_KCH, bstart[b + 1] - k), 0});
        bchunk[b + 1] = (int)chunks.size();
    }
    const int nchunk = (int)chunks.size();
    const int ntile = (int)((nc + MT - 1) / MT), npair = ntile * (ntile + 1) / 2;
    if (nchunk > 65535) return fail(c, AMB_EARG, "MAPS: %d cell chunks exceed the launch grid", nchunk);

    /* one device block: cube, sampling, tables, partials, results */
    const size_t nv = (size_t)npix * nc;
    size_t off = 0;
    auto place = [&](size_t bytes) { size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
    const size_t o_v = place(sizeof(double2) * nv), o_s = place(sizeof(double) * nv);
    const size_t o_ord = place(sizeof(int) * order.size()), o_uu = place(sizeof(double) * ouu.size());
    const size_t o_ch = place(sizeof(MapsChunk) * std::max(nchunk, 1));
    const size_t o_bs = place(sizeof(int) * (NB + 1)), o_bc = place(sizeof(int) * (NB + 1));
    const size_t o_part = place(sizeof(double) * (size_t)std::max(nchunk, 1) * nc * nc);
    const size_t o_nop = place(sizeof(int) * (size_t)std::max(nchunk, 1) * nc);
    const size_t o_cl = place(sizeof(double) * NB * nc);
    const size_t o_nl = place(sizeof(int) * NB), o_ul = place(sizeof(double) * NB);
    void *base = nullptr;
    int rc = dev_alloc(c, &base, off, "MAPS cube and partials");
    if (rc) return rc;
    char *d = (char *)base;
    double2 *d_v = (double2 *)(d + o_v);
    double *d_s = (double *)(d + o_s), *d_uu = (double *)(d + o_uu), *d_part = (double *)(d + o_part);
    double *d_cl = (double *)(d + o_cl), *d_ul = (double *)(d + o_ul);
    int *d_ord = (int *)(d + o_ord), *d_bs = (int *)(d + o_bs), *d_bc = (int *)(d + o_bc);
    int *d_nop = (int *)(d + o_nop), *d_nl = (int *)(d + o_nl);
    MapsChunk *d_ch = (MapsChunk *)(d + o_ch);

    std::vector<double> cl_h((size_t)NB * nc), ul(NB);
    std::vector<int> nl(NB);
    hipStream_t st = c->stream;
    hipError_t e = hipMemcpyAsync(d_v, gv, sizeof(double2) * nv, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_s, samp, sizeof(double) * nv, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && ncell)
        e = hipMemcpyAsync(d_ord, order.data(), sizeof(int) * ncell, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && ncell)
        e = hipMemcpyAsync(d_uu, ouu.data(), sizeof(double) * ncell, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nchunk)
        e = hipMemcpyAsync(d_ch, chunks.data(), sizeof(MapsChunk) * nchunk, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_bs, bstart.data(), sizeof(int) * (NB + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_bc, bchunk.data(), sizeof(int) * (NB + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        timer_start(c, &c->st.ms_maps_total);
        const int nb_norm = (int)std::min<size_t>((nv + 255) / 256, 8192);
        hipLaunchKernelGGL(k_maps_norm, dim3(nb_norm), dim3(256), 0, st, d_v, d_s, nv);
        if (nchunk) {
            timer_start(c, &c->st.ms_maps);
            hipLaunchKernelGGL(k_maps_pairs, dim3(npair, nchunk), dim3(256), 0, st, d_v, d_s, d_ord, d_ch, (int)nc,
                               ntile, d_part);
            timer_stop(c);
            hipLaunchKernelGGL(k_maps_counts, dim3((unsigned)((nc + 255) / 256), nchunk), dim3(256), 0, st, d_s,
                               d_ord, d_ch, (int)nc, d_nop);
            hipLaunchKernelGGL(k_maps_sum, dim3((unsigned)(((size_t)nc * nc + 255) / 256), NB), dim3(256), 0, st,
                               d_part, d_nop, d_bc, (int)nc);
        }
        hipLaunchKernelGGL(k_maps_last, dim3(NB), dim3(256), 0, st, d_s, d_ord, d_uu, d_bs, (int)nc, d_nl, d_ul);
        hipLaunchKernelGGL(k_maps_lag, dim3((unsigned)((nc + 255) / 256), NB), dim3(256), 0, st, d_part, d_nop,
                           d_bc, (int)nc, d_cl);
        timer_stop(c);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(cl_h.data(), d_cl, sizeof(double) * NB * nc, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(nl.data(), d_nl, sizeof(int) * NB, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ul.data(), d_ul, sizeof(double) * NB, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    timers_collect(c);
    dev_free(c, base, off);
    if (e != hipSuccess) return fail(c, AMB_EHIP, "MAPS: %s", hipGetErrorString(e));
    c->st.flop_maps += 4.0 * ((double)nc * (nc + 1) / 2) * ncell;
    c->st.n_maps_launch += nchunk ? 1 : 0;
    for (size_t k = 0; k < cl_h.size(); ++k) cldnu[k] = cl_h[k];
    for (int64_t dl = 0; dl < nc; ++dl) numdnu[dl] = (double)(nc - dl); /* numdnu[d] += 1 per pair */
    for (int b = 0; b < NB; ++b) {
        num[b] = nl[b];
        binu[b] = nl[b] != 0 ? ul[b] / (1. * nl[b]) : ul[b];
    }
    return AMB_OK;
}

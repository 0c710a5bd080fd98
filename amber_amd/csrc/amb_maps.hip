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
 *   k_maps_prep   per (chunk, 64 channels): V /= NTiG where NTiG > 0
 *                 (maps.c:100-125), in place, FP64 division as written, for
 *                 the binned cells (the only ones read later); one ballot
 *                 word of sampled cells per (cell, 64 channels), from which
 *                 the A operand NTiG V is formed on load; counts the
 *                 sampled cells per channel (exact); flags NTiG < 0 (a
 *                 sampling count is never negative; the call then fails);
 *   k_maps_pairs  per (chunk, 128x128 channel tile with tile(n1) <=
 *                 tile(n2)) the partial Gram block over the chunk's cells on
 *                 the FP64 matrix cores, 16-cell stages through LDS with the
 *                 next stage's loads (and the stage after's cell index) in
 *                 flight during the products; each block writes its own
 *                 slice of the partial array (no atomics).  All tiles of a
 *                 chunk run on one XCD, so a cell's channels come from that
 *                 XCD's L2 after the first tile reads them.  The first
 *                 blocks of the same launch compute, per bin and in cell
 *                 order, no and usum of channel nc-1 (what the reference's
 *                 last call leaves in binu/num) alongside;
 *   k_maps_sum    per bin, the chunk partials and counts added in chunk
 *                 order into the bin's first chunk;
 *   k_maps_lag    per (bin, lag d) the sum over n1 of C_b(n1, n1 + d) /
 *                 no_b(n1), skipping the division where no_b(n1) = 0 as the
 *                 reference does; 8 fixed n1 segments added in order.
 * Every sum has a fixed order, so results are deterministic run to run; the
 * Gram and lag sums differ from the reference's sequential order by FP64
 * rounding only.  Counts, binu and num are the reference's exactly.
 *
 * Bins are computed on the host with the reference's expression and libm
 * (sqrt(ia^2 + j^2) Du, floor(log10(uu/umin)/binsiz)), so the cell-to-bin
 * map is the reference's exactly; a cell whose rounded index reaches NB
 * (possible only within an ulp of umax, an out-of-bounds write in the
 * reference) is counted in bin NB - 1.
 *
 * Roofline: 4 FP64 flops per (pair n1 <= n2, binned cell); the FP64 MFMA
 * peak of gfx950 is 78.6 TFLOP/s.
 */
#include "amb_ctx.h"

#include <cmath>
#include <vector>

namespace {

constexpr int MT = 128;           /* channel tile */
constexpr int MKC = 16;           /* cells per LDS stage */
constexpr int MROW = MT + 16;     /* padded LDS row, doubles */
constexpr int MAPS_KCH = 2048;    /* cells per chunk */
constexpr int NXCD = 8;           /* XCDs; workgroup i runs on XCD i % 8 */
typedef double f64x4 __attribute__((ext_vector_type(4)));

struct MapsChunk {
    int bin, start, count, pad;
};

/* per (64-channel slab, chunk): normalise the chunk's cells in place as
 * maps.c:100-125 does (V /= NTiG where NTiG > 0, FP64 division as written),
 * record which cells are sampled as one 64-bit ballot word per (cell, slab)
 * (the A operand is NTiG V = V where the bit is set, else 0), count the
 * sampled cells per channel (exact; 4 waves reduced in LDS) and flag a
 * negative count.  The chunk's cell indices are staged in LDS first so the
 * data loads of consecutive cells are independent.  Only binned cells are
 * read afterwards, so only they are normalised; NTiG itself is left as
 * read (readers test NTiG > 0). */
__global__ __launch_bounds__(256) void k_maps_prep(double2 *__restrict__ v, const double *__restrict__ s,
                                                   unsigned long long *__restrict__ mask, int nslab,
                                                   const int *__restrict__ order,
                                                   const MapsChunk *__restrict__ chunks, int nc, int *__restrict__ nop,
                                                   unsigned *__restrict__ bad)
{
    __shared__ int idx[MAPS_KCH];
    __shared__ int cs[4][64];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int slab = blockIdx.x, n = slab * 64 + lane;
    const MapsChunk ch = chunks[blockIdx.y];
    for (int k = threadIdx.x; k < ch.count; k += 256) idx[k] = order[ch.start + k];
    __syncthreads();
    int cnt = 0;
    const int nn = min(n, nc - 1);
#pragma unroll 8
    for (int k = g; k < ch.count; k += 4) {
        const int x = idx[k];
        const size_t i = (size_t)x * nc + nn;
        const double sv = s[i];
        const bool on = n < nc && sv > 0;
        if (on) {
            double2 t = v[i];
            t.x /= sv;
            t.y /= sv;
            v[i] = t;
            ++cnt;
        }
        if (n < nc && sv < 0) atomicOr(bad, 1u);
        const unsigned long long m = __ballot(on);
        if (lane == 0) mask[(size_t)x * nslab + slab] = m;
    }
    cs[g][lane] = cnt;
    __syncthreads();
    if (g == 0 && n < nc) nop[(size_t)blockIdx.y * nc + n] = cs[0][lane] + cs[1][lane] + cs[2][lane] + cs[3][lane];
}

struct MapsSmem {
    double A[2 * MKC][MROW], B[2 * MKC][MROW];
};

/* no and usum of channel nc-1 over bin b in cell order, the reference's
 * expressions (no += NTiG as int, usum += uu * NTiG): 256 terms at a time
 * fetched in parallel into LDS, added by one thread */
__device__ void maps_last(MapsSmem &L, const double *__restrict__ s, const int *__restrict__ order,
                          const double *__restrict__ uu, const int *__restrict__ bstart, int nc, int b,
                          int *__restrict__ no_last, double *__restrict__ us_last)
{
#pragma clang fp contract(off)
    double *wv = &L.A[0][0], *pv = &L.B[0][0];
    const int t = threadIdx.x;
    const int k0 = bstart[b], k1 = bstart[b + 1];
    int no = 0;
    double us = 0.0;
    for (int base = k0; base < k1; base += 256) {
        const int m = min(256, k1 - base);
        if (t < m) {
            const double r = s[(size_t)order[base + t] * nc + (nc - 1)];
            const double w = r > 0 ? 1.0 : r; /* NTiG after normalisation */
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

/* Wave w owns the 64x64 quarter (w>>1, w&1) of the 128x128 tile as 4x4
 * blocks of v_mfma_f64_16x16x4f64: A[m][k] = lane m + 16k, B[k][n] = lane
 * n + 16k, and the four results of lane l are rows 4r + l/16 (r = 0..3) of
 * column l%16 (rows interleaved across the lane groups, unlike the
 * contiguous rows of the FP32 16x16x4 form).  The K dimension is (cell,
 * re/im): LDS holds 16 cells as 32 rows of 128 channels per operand (row 2c
 * real parts of cell c, row 2c+1 imaginary parts), rows padded by 16
 * doubles so the two half-waves of a read hit disjoint banks. */
__device__ void maps_tile(MapsSmem &L, const double2 *__restrict__ v, const unsigned long long *__restrict__ mask,
                          int nslab, const int *__restrict__ order, const MapsChunk ch, int chunk, int ti, int tj,
                          int nc, double *__restrict__ part)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = (wv >> 1) * 64, wc = (wv & 1) * 64;
    const int lc = tid >> 4, lch = (tid & 15) * 8; /* loader: cell lc, channels lch..lch+7 */
    const int r0 = ti * MT, c0 = tj * MT;
    f64x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (f64x4){0.0, 0.0, 0.0, 0.0};

    /* raw operands of one cell, loaded without branches (indices clamped
     * into the cube, out-of-range values zeroed at the LDS store) so every
     * load of a stage is in flight at once and none is waited for before
     * the next stage's store */
    double2 tv[8], tb[8];
    unsigned long long tm = 0; /* sample bits of the A channels (bit lch % 64 + q) */
    bool okc = false;
    /* k: the cell's place in the chunk; idx: its cube cell (any valid cell
     * when k is past the chunk) */
    auto fetch = [&](int k, int idx) {
        okc = k < ch.count;
        const size_t xb = (size_t)idx * nc;
        tm = mask[(size_t)idx * nslab + min(r0 + lch, nc - 1) / 64];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int n1 = min(r0 + lch + q, nc - 1), n2 = min(c0 + lch + q, nc - 1);
            tv[q] = v[xb + n1];
            tb[q] = v[xb + n2];
        }
    };
    /* the index load is not waited for until the fetch that uses it */
    auto cell = [&](int k) { return order[ch.start + min(k, ch.count - 1)]; };
    fetch(lc, cell(lc));
    int xn = cell(MKC + lc);
    const int fm = lane & 15, fk = lane >> 4;
    for (int k0 = 0; k0 < ch.count; k0 += MKC) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; q += 2) {
            double2 a0, a1, b0, b1;
            const unsigned bits = (unsigned)(tm >> (lch & 63));
            const bool o1a = okc && r0 + lch + q < nc && (bits >> q & 1u);
            const bool o1b = okc && r0 + lch + q + 1 < nc && (bits >> (q + 1) & 1u);
            const bool o2a = okc && c0 + lch + q < nc, o2b = okc && c0 + lch + q + 1 < nc;
            a0 = o1a ? tv[q] : make_double2(0.0, 0.0);
            a1 = o1b ? tv[q + 1] : make_double2(0.0, 0.0);
            b0 = o2a ? tb[q] : make_double2(0.0, 0.0);
            b1 = o2b ? tb[q + 1] : make_double2(0.0, 0.0);
            *(double2 *)&L.A[2 * lc][lch + q] = make_double2(a0.x, a1.x);
            *(double2 *)&L.A[2 * lc + 1][lch + q] = make_double2(a0.y, a1.y);
            *(double2 *)&L.B[2 * lc][lch + q] = make_double2(b0.x, b1.x);
            *(double2 *)&L.B[2 * lc + 1][lch + q] = make_double2(b0.y, b1.y);
        }
        __syncthreads();
        if (k0 + MKC < ch.count) { /* in flight during this stage's products */
            fetch(k0 + MKC + lc, xn);
            xn = cell(k0 + 2 * MKC + lc);
        }
#pragma unroll
        for (int ks = 0; ks < 2 * MKC / 4; ++ks) {
            const int row = 4 * ks + fk;
            double fa[4], fb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) fa[i] = L.A[row][wr + 16 * i + fm];
#pragma unroll
            for (int j = 0; j < 4; ++j) fb[j] = L.B[row][wc + 16 * j + fm];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
    }
    double *out = part + (size_t)chunk * nc * nc;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int n1 = r0 + wr + 16 * i + 4 * r + fk;
            if (n1 >= nc) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n2 = c0 + wc + 16 * j + fm;
                if (n2 < nc) out[(size_t)n1 * nc + n2] = acc[i][j][r];
            }
        }
}

/* grid: nlast blocks (a multiple of 8) for the per-bin last-channel sums,
 * then 8 * ceil(nchunk / 8) * npair tile blocks; tile block w runs on XCD
 * w % 8 and takes chunk w % 8 + 8 * (slot / npair), pair slot % npair,
 * slot = w / 8 */
__global__ __launch_bounds__(256, 2) void k_maps_pairs(const double2 *__restrict__ v,
                                                       const unsigned long long *__restrict__ mask, int nslab,
                                                       const double *__restrict__ s, const int *__restrict__ order,
                                                       const MapsChunk *__restrict__ chunks, int nchunk, int nc,
                                                       int ntile, int nlast, double *__restrict__ part,
                                                       const double *__restrict__ uu,
                                                       const int *__restrict__ bstart, int NB,
                                                       int *__restrict__ no_last, double *__restrict__ us_last)
{
    __shared__ MapsSmem L;
    const int id = blockIdx.x;
    if (id < nlast) {
        if (id < NB) maps_last(L, s, order, uu, bstart, nc, id, no_last, us_last);
        return;
    }
    const int w = id - nlast, xcd = w % NXCD, slot = w / NXCD;
    const int npair = ntile * (ntile + 1) / 2;
    const int chunk = xcd + NXCD * (slot / npair);
    if (chunk >= nchunk) return;
    int p = slot % npair, ti = 0;
    while (p >= ntile - ti) {
        p -= ntile - ti;
        ++ti;
    }
    maps_tile(L, v, mask, nslab, order, chunks[chunk], chunk, ti, ti + p, nc, part);
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

/* block (32 lags, bin): thread (lag dl, segment g) sums n1 in
 * [g S, (g+1) S) sequentially, S = ceil(nc / 8); the 8 partials are then
 * added in segment order */
__global__ __launch_bounds__(256) void k_maps_lag(const double *__restrict__ part, const int *__restrict__ nop,
                                                  const int *__restrict__ bchunk, int nc, double *__restrict__ cl)
{
    __shared__ double ps[8][32];
    const int dl = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int d = blockIdx.x * 32 + dl, b = blockIdx.y;
    const int c0 = bchunk[b], c1 = bchunk[b + 1];
    const int S = (nc + 7) / 8;
    double acc = 0.0;
    if (d < nc && c1 > c0) {
        const double *P = part + (size_t)c0 * nc * nc;
        const int *no = nop + (size_t)c0 * nc;
        const int e = min((g + 1) * S, nc - d);
#pragma unroll 8
        for (int n1 = g * S; n1 < e; ++n1) {
            double v = P[(size_t)n1 * nc + n1 + d];
            const int k = no[n1];
            if (k != 0) v /= (1. * k);
            acc += v;
        }
    }
    ps[g][dl] = acc;
    __syncthreads();
    if (g == 0 && d < nc) {
        double t = ps[0][dl];
        for (int i = 1; i < 8; ++i) t += ps[i][dl];
        cl[(size_t)b * nc + d] = t;
    }
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
    if (N < 2 || nc < 1 || nc > (1 << 20) || !(Du > 0) || NB < 1 || NB > 65535 || !(umin > 0) || !(umax > umin))
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
    if ((double)(nchunk + NXCD) * npair + NB + NXCD > 2e9)
        return fail(c, AMB_EARG, "MAPS: %d cell chunks x %d tile pairs exceed the launch grid", nchunk, npair);

    /* one device block: cube, sampling, tables, partials, results */
    const size_t nv = (size_t)npix * nc;
    size_t off = 0;
    auto place = [&](size_t bytes) { size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
    const size_t o_v = place(sizeof(double2) * nv), o_s = place(sizeof(double) * nv);
    const int nslab = (int)((nc + 63) / 64);
    const size_t o_mk = place(sizeof(unsigned long long) * (size_t)npix * nslab);
    const size_t o_ord = place(sizeof(int) * order.size()), o_uu = place(sizeof(double) * ouu.size());
    const size_t o_ch = place(sizeof(MapsChunk) * std::max(nchunk, 1));
    const size_t o_bs = place(sizeof(int) * (NB + 1)), o_bc = place(sizeof(int) * (NB + 1));
    const size_t o_part = place(sizeof(double) * (size_t)std::max(nchunk, 1) * nc * nc);
    const size_t o_nop = place(sizeof(int) * (size_t)std::max(nchunk, 1) * nc);
    const size_t o_cl = place(sizeof(double) * NB * nc);
    const size_t o_nl = place(sizeof(int) * NB), o_ul = place(sizeof(double) * NB);
    const size_t o_bad = place(sizeof(unsigned));
    void *base = nullptr;
    int rc = dev_alloc(c, &base, off, "MAPS cube and partials");
    if (rc) return rc;
    char *d = (char *)base;
    double2 *d_v = (double2 *)(d + o_v);
    unsigned long long *d_mk = (unsigned long long *)(d + o_mk);
    double *d_s = (double *)(d + o_s), *d_uu = (double *)(d + o_uu), *d_part = (double *)(d + o_part);
    double *d_cl = (double *)(d + o_cl), *d_ul = (double *)(d + o_ul);
    int *d_ord = (int *)(d + o_ord), *d_bs = (int *)(d + o_bs), *d_bc = (int *)(d + o_bc);
    int *d_nop = (int *)(d + o_nop), *d_nl = (int *)(d + o_nl);
    MapsChunk *d_ch = (MapsChunk *)(d + o_ch);
    unsigned *d_bad = (unsigned *)(d + o_bad);

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
    if (e == hipSuccess) e = amb_launch_fill_u32(st, d_bad, 1, 0u);
    if (e == hipSuccess) {
        timer_start(c, &c->st.ms_maps_total);
        if (nchunk)
            hipLaunchKernelGGL(k_maps_prep, dim3((unsigned)nslab, nchunk), dim3(256), 0, st, d_v, d_s, d_mk, nslab,
                               d_ord, d_ch, (int)nc, d_nop, d_bad);
        const int nlast = (NB + NXCD - 1) / NXCD * NXCD;
        const long ntb = (long)(nchunk + NXCD - 1) / NXCD * NXCD * npair;
        timer_start(c, &c->st.ms_maps);
        hipLaunchKernelGGL(k_maps_pairs, dim3((unsigned)(nlast + ntb)), dim3(256), 0, st, d_v, d_mk, nslab, d_s,
                           d_ord, d_ch, nchunk, (int)nc, ntile, nlast, d_part, d_uu, d_bs, NB, d_nl, d_ul);
        timer_stop(c);
        if (nchunk)
            hipLaunchKernelGGL(k_maps_sum, dim3((unsigned)(((size_t)nc * nc + 255) / 256), NB), dim3(256), 0, st,
                               d_part, d_nop, d_bc, (int)nc);
        hipLaunchKernelGGL(k_maps_lag, dim3((unsigned)((nc + 31) / 32), NB), dim3(256), 0, st, d_part, d_nop, d_bc,
                           (int)nc, d_cl);
        timer_stop(c);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(cl_h.data(), d_cl, sizeof(double) * NB * nc, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(nl.data(), d_nl, sizeof(int) * NB, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ul.data(), d_ul, sizeof(double) * NB, hipMemcpyDeviceToHost, st);
    unsigned bad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(unsigned), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    timers_collect(c);
    dev_free(c, base, off);
    if (e != hipSuccess) return fail(c, AMB_EHIP, "MAPS: %s", hipGetErrorString(e));
    if (bad) return fail(c, AMB_EARG, "MAPS: negative sampling count in the cube");
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

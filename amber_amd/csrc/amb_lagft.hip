/*
 * Lag-Fourier form of the triple-product accumulation (amber.c:233-283).
 *
 * The reference sums, for every lag cell (d2, d3),
 *   B[d2][d3] = sum_{p1} sum_x D1[p1](x) D2[p1 + d2](x) D3[p1 + d3](x)
 * over the triplets of its channel set (amber_funcs.c:2207-2310).  For a
 * fixed pixel x that is a triple correlation along the channel axis, so
 * with f, g, h the channel sequences of D1, D2, D3 at x (zero on channels
 * outside the set) and their length-L DFTs F, G, H (L >= span + the lag
 * window, so no lag inside the window wraps onto another):
 *   sum_{d2,d3} B[d2][d3] e^{-2 pi i (k2 d2 + k3 d3)/L}
 *       = S(k2, k3) = sum_x conj(F(k2 + k3)) G(k2) H(k3),
 * and B is the 2-D inverse DFT of S divided by L^2.  S costs L^2/2 complex
 * triple products per pixel (half the plane: B is real) instead of the
 * nc^3 real ones of the direct sum: 48x fewer at 768 channels.
 *
 * Stages (FP32 products, FP64 sums and transforms after them):
 *   k_chan_spectra  per pixel pair, the real channel sequences packed as
 *                   one complex sequence, forward DFT of length L in LDS;
 *                   X[x][k], k = 0 .. L/2 (a plane set kind of its own);
 *   k_lagft_tiles   S over 128 x 128 (k2, k3) tiles, k3 < L/2, for one
 *                   segment of pixels per work item: FP32 accumulators in
 *                   registers, written as the item's partial;
 *   k_lagft_nyq     the k3 = L/2 column;
 *   k_lagft_reduce  the partials summed over segments in FP64, in segment
 *                   order (deterministic), transposed to S^T[k3][k2];
 *   k_lagft_cols    FP64 inverse DFT over k2 of every row k3;
 *   k_lagft_rows    FP64 inverse DFT over k3 (half spectrum, two real
 *                   rows per complex transform) of every window row d2,
 *                   times 2^(E1+E2+E3-21)/L^2; cells without triplets are
 *                   written as exact zeros, as the direct sum leaves them.
 * Counts (amber.c:268) do not depend on the triangle: one integer kernel
 * per channel set (k_lag_counts), kept with the cube.
 */
#include "amb_ctx.h"
#include "amb_fftcore.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

using namespace amb;

namespace {

constexpr int LT = 128;             /* tile side in k2 and in k3 */
constexpr int KPX = 4;              /* pixels per LDS chunk */
constexpr int BLKB = 80;            /* LDS bytes per block of 8 complex (64 + 16 pad) */
constexpr int PXB = 64 * BLKB;      /* one pixel: G 16 blocks, H 16, conj F 32 */
constexpr int SEG = 2048;           /* pixels per work item */
constexpr int MAXL = 2048;          /* longest channel transform: 2 L FP64 complex in 64 KiB of LDS */
constexpr size_t ALIGN = 256;
constexpr size_t PART_CAP = (size_t)8 << 30; /* partial sums per wave */
constexpr size_t OUT_CAP = (size_t)1 << 30;  /* result bytes per wave */
constexpr int MAX_JOBS = 512;

size_t al(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct LagJob {
    const float2 *F, *G, *H; /* channel spectra [P][lhp] of r1, r2, r3 */
    const int *ex1, *ex2, *ex3; /* exponents E of the three sets */
    float2 *part;            /* [nseg][ntile][LT*LT] tile partials */
    float2 *nyq;             /* [nseg][L] partials of the k3 = L/2 column */
    long P;
    long item0, nitem0;      /* first tile item, first nyquist item in the launch */
    int lhp, L, nseg, seg;
};

__host__ __device__ inline int tiles3(int L) { return (L / 2 + LT - 1) / LT; }
__host__ __device__ inline int tiles2(int L) { return (L + LT - 1) / LT; }

/* blocks b = x (mod 8) run on XCD x: XCD x takes items [x per, (x+1) per) */
__device__ __forceinline__ long xcd_item(long b, long nb)
{
    const long per = nb / 8;
    return (b % 8) * per + b / 8;
}

__device__ __forceinline__ int find_job(const LagJob *__restrict__ jobs, int njob, long it, bool nyq)
{
    int lo = 0, hi = njob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const long v = nyq ? jobs[mid].nitem0 : jobs[mid].item0;
        if (v <= it) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* X(k) of a pixel row for k in [0, L): stored for k <= L/2, else conj X(L - k) */
__device__ __forceinline__ float2 spec_at(const float2 *__restrict__ X, long row, int k, int L)
{
    if (2 * k <= L) return X[row + k];
    const float2 v = X[row + (L - k)];
    return make_float2(v.x, -v.y);
}

/* ---------------------------------------------------------------- spectra */
/* grid: multiple of 8 blocks of nteam x 64 threads, each team ppt pixel
   pairs; dynamic LDS: tw[L] + nteam * 2 * L complex.  Plane i of the
   normalised set carries (-1)^i, removed here; channel ch = act0 + p sits
   at position p of the sequence. */
__global__ __launch_bounds__(256) void k_chan_spectra(const float *__restrict__ D, long pitch, const int *__restrict__ inv,
                                                      int nc, int act0, long P, FftPlan pl,
                                                      const float2 *__restrict__ twg, int lhp, int ppt,
                                                      float2 *__restrict__ X, const unsigned *__restrict__ slot_in,
                                                      unsigned *__restrict__ slot_out)
{
    extern __shared__ __align__(16) float2 lsm[];
    const int L = pl.n, h = L / 2;
    const int nteam = blockDim.x >> 6, team = threadIdx.x >> 6, t = threadIdx.x & 63;
    float2 *tw = lsm, *bx = lsm + L + (size_t)team * 2 * L, *by = bx + L;
    for (int i = threadIdx.x; i < L; i += blockDim.x) tw[i] = twg[i];
    if (blockIdx.x == 0 && threadIdx.x < 2) slot_out[threadIdx.x] = slot_in[threadIdx.x];
    const long npair = (P + 1) / 2;
    const long pair0 = (xcd_item(blockIdx.x, gridDim.x) * nteam + team) * ppt;
    __syncthreads();
    for (int q = 0; q < ppt; ++q) {
        const long pr = pair0 + q;
        const bool live = pr < npair;
        const long x0 = 2 * pr;
        for (int p = t; p < L; p += 64) {
            float2 v = make_float2(0.f, 0.f);
            const int ch = act0 + p;
            if (live && ch < nc) {
                const int i = inv[ch];
                if (i >= 0) {
                    /* x0 even and pitch a multiple of 64: an aligned pair; for
                       odd P the second value is the zeroed pitch tail */
                    const float2 d = *(const float2 *)(D + (size_t)i * pitch + x0);
                    v = (i & 1) ? make_float2(-d.x, -d.y) : d;
                }
            }
            bx[p] = v;
        }
        __syncthreads();
        const float2 *Z = fftc::transform<-1, 5>(bx, by, tw, pl, t, 64);
        if (live) {
            float2 *xa = X + (size_t)x0 * lhp, *xb = xa + lhp;
            const bool two = x0 + 1 < P;
            for (int k = t; k <= h; k += 64) {
                const float2 a = Z[k], b = Z[k == 0 ? 0 : L - k];
                /* A = (Z(k) + conj Z(L-k)) / 2, B = (Z(k) - conj Z(L-k)) / 2i */
                xa[k] = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
                if (two) xb[k] = make_float2(0.5f * (a.y + b.y), -0.5f * (a.x - b.x));
            }
        }
        __syncthreads(); /* the next pair overwrites both buffers */
    }
}

/* ------------------------------------------------------------------ tiles */
__global__ __launch_bounds__(256, 2) void k_lagft_tiles(const LagJob *__restrict__ jobs, int njob, long total)
{
    __shared__ __align__(16) char lds[2][KPX * PXB];
    const long it = xcd_item(blockIdx.x, gridDim.x);
    if (it >= total) return;
    const int jb = find_job(jobs, njob, it, false);
    const float2 *__restrict__ F = jobs[jb].F;
    const long P = jobs[jb].P;
    const int lhp = jobs[jb].lhp, L = jobs[jb].L, seg = jobs[jb].seg;
    const int nt3 = tiles3(L), ntile = tiles2(L) * nt3;
    const long local = it - jobs[jb].item0;
    const int sg = (int)(local / ntile), tile = (int)(local - (long)sg * ntile);
    const int a = (tile / nt3) * LT, b = (tile % nt3) * LT;
    const long xb = (long)sg * seg, xe = xb + seg < P ? xb + seg : P;
    const int tid = threadIdx.x;

    /* the two values this thread stages per pixel: slot tid (G for
       tid < 128, else H) and slot 256 + tid (conj F) */
    const float2 *__restrict__ src0;
    int k0;
    bool ok0, cj0;
    if (tid < LT) {
        const int k2 = a + tid;
        ok0 = k2 < L;
        cj0 = 2 * k2 > L;
        k0 = !ok0 ? 0 : cj0 ? L - k2 : k2;
        src0 = jobs[jb].G;
    } else {
        const int k3 = b + tid - LT;
        ok0 = 2 * k3 < L;
        cj0 = false;
        k0 = ok0 ? k3 : 0;
        src0 = jobs[jb].H;
    }
    const int s1 = (a + b + tid) % L;
    const bool m1 = 2 * s1 > L;          /* F(s) = conj X(L - s) */
    const int k1 = m1 ? L - s1 : s1;     /* conj F(s): X(L - s), or conj X(s) */
    const int off0 = (tid >> 3) * BLKB + (tid & 7) * 8;
    const int off1 = (32 + (tid >> 3)) * BLKB + (tid & 7) * 8;

    float2 r0[KPX], r1[KPX];
    auto fetch = [&](long x0) {
#pragma unroll
        for (int pp = 0; pp < KPX; ++pp) {
            const long x = x0 + pp;
            float2 v0 = make_float2(0.f, 0.f), v1 = make_float2(0.f, 0.f);
            if (x < xe) {
                const size_t row = (size_t)x * lhp;
                if (ok0) {
                    v0 = src0[row + k0];
                    if (cj0) v0.y = -v0.y;
                }
                v1 = F[row + k1];
                if (!m1) v1.y = -v1.y;
            }
            r0[pp] = v0;
            r1[pp] = v1;
        }
    };
    auto stash = [&](char *st) {
#pragma unroll
        for (int pp = 0; pp < KPX; ++pp) {
            *(float2 *)(st + pp * PXB + off0) = r0[pp];
            *(float2 *)(st + pp * PXB + off1) = r1[pp];
        }
    };

    const int li = tid >> 4, lj = tid & 15;
    float acr[8][8], aci[8][8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int v = 0; v < 8; ++v) acr[u][v] = aci[u][v] = 0.f;

    const long nch = (xe - xb + KPX - 1) / KPX;
    fetch(xb);
    stash(lds[0]);
    __syncthreads();
    for (long ch = 0; ch < nch; ++ch) {
        if (ch + 1 < nch) fetch(xb + (ch + 1) * KPX);
        const char *st = lds[ch & 1];
#pragma unroll 1
        for (int pp = 0; pp < KPX; ++pp) {
            const char *px = st + pp * PXB;
            float gr[8], gi[8], hr[8], hi[8], fr[16], fi[16];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 g4 = *(const float4 *)(px + li * BLKB + 16 * q);
                const float4 h4 = *(const float4 *)(px + (16 + lj) * BLKB + 16 * q);
                const float4 f4 = *(const float4 *)(px + (32 + li + lj) * BLKB + 16 * q);
                const float4 e4 = *(const float4 *)(px + (33 + li + lj) * BLKB + 16 * q);
                gr[2 * q] = g4.x; gi[2 * q] = g4.y; gr[2 * q + 1] = g4.z; gi[2 * q + 1] = g4.w;
                hr[2 * q] = h4.x; hi[2 * q] = h4.y; hr[2 * q + 1] = h4.z; hi[2 * q + 1] = h4.w;
                fr[2 * q] = f4.x; fi[2 * q] = f4.y; fr[2 * q + 1] = f4.z; fi[2 * q + 1] = f4.w;
                fr[8 + 2 * q] = e4.x; fi[8 + 2 * q] = e4.y; fr[9 + 2 * q] = e4.z; fi[9 + 2 * q] = e4.w;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    /* t = G H, acc += conj(F) t */
                    const float tr = fmaf(gr[u], hr[v], -gi[u] * hi[v]);
                    const float ti = fmaf(gr[u], hi[v], gi[u] * hr[v]);
                    acr[u][v] = fmaf(fr[u + v], tr, fmaf(-fi[u + v], ti, acr[u][v]));
                    aci[u][v] = fmaf(fr[u + v], ti, fmaf(fi[u + v], tr, aci[u][v]));
                }
        }
        if (ch + 1 < nch) stash(lds[(ch + 1) & 1]);
        __syncthreads();
    }
    float2 *out = jobs[jb].part + ((size_t)sg * ntile + tile) * (LT * LT);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        float4 *o = (float4 *)(out + (size_t)(8 * li + u) * LT + 8 * lj);
#pragma unroll
        for (int v = 0; v < 8; v += 2) o[v / 2] = make_float4(acr[u][v], aci[u][v], acr[u][v + 1], aci[u][v + 1]);
    }
}

/* --------------------------------------------------------------- nyquist */
/* S(k2, L/2) over one segment per block; thread t owns k2 = t + 256 q */
__global__ __launch_bounds__(256) void k_lagft_nyq(const LagJob *__restrict__ jobs, int njob, long total)
{
    const long it = xcd_item(blockIdx.x, gridDim.x);
    if (it >= total) return;
    const int jb = find_job(jobs, njob, it, true);
    const float2 *__restrict__ F = jobs[jb].F, *__restrict__ G = jobs[jb].G, *__restrict__ H = jobs[jb].H;
    const long P = jobs[jb].P;
    const int lhp = jobs[jb].lhp, L = jobs[jb].L, seg = jobs[jb].seg, h = L / 2;
    const int sg = (int)(it - jobs[jb].nitem0);
    const long xb = (long)sg * seg, xe = xb + seg < P ? xb + seg : P;
    constexpr int NQ = MAXL / 256;
    float acr[NQ], aci[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acr[q] = aci[q] = 0.f;
    for (long x = xb; x < xe; ++x) {
        const size_t row = (size_t)x * lhp;
        const float2 hv = H[row + h];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int k2 = threadIdx.x + 256 * q;
            if (256 * q < L && k2 < L) {
                const float2 g = spec_at(G, row, k2, L);
                int s = k2 + h;
                if (s >= L) s -= L;
                const float2 f = spec_at(F, row, s, L);
                const float tr = fmaf(g.x, hv.x, -g.y * hv.y), ti = fmaf(g.x, hv.y, g.y * hv.x);
                acr[q] = fmaf(f.x, tr, fmaf(f.y, ti, acr[q]));
                aci[q] = fmaf(f.x, ti, fmaf(-f.y, tr, aci[q]));
            }
        }
    }
    float2 *o = jobs[jb].nyq + (size_t)sg * L;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int k2 = threadIdx.x + 256 * q;
        if (256 * q < L && k2 < L) o[k2] = make_float2(acr[q], aci[q]);
    }
}

/* ----------------------------------------------------------------- reduce */
/* St[k3][k2] (k3 < L/2) = sum over segments of the tile partials, FP64 in
   segment order; 32 x 32 (k2, k3) blocks transposed through LDS */
__global__ __launch_bounds__(256) void k_lagft_reduce(const float2 *__restrict__ part, int nseg, int L,
                                                      double2 *__restrict__ St)
{
    __shared__ double2 tr[32][33];
    const int nt3 = tiles3(L);
    const size_t tsz = (size_t)tiles2(L) * nt3 * LT * LT;
    const int X = blockIdx.x * 32, Y = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int k3 = Y + tx;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k2 = X + ty + 8 * q;
        double sx = 0.0, sy = 0.0;
        if (k2 < L && 2 * k3 < L) {
            const size_t off = (size_t)((k2 / LT) * nt3 + k3 / LT) * (LT * LT) + (k2 % LT) * LT + (k3 % LT);
            for (int sg = 0; sg < nseg; ++sg) {
                const float2 v = part[sg * tsz + off];
                sx += (double)v.x;
                sy += (double)v.y;
            }
        }
        tr[ty + 8 * q][tx] = make_double2(sx, sy);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k3w = Y + ty + 8 * q, k2w = X + tx;
        if (2 * k3w < L && k2w < L) St[(size_t)k3w * L + k2w] = tr[tx][ty + 8 * q];
    }
}

/* St[L/2][k2] = sum over segments of the nyquist partials (FP64, in order) */
__global__ __launch_bounds__(256) void k_lagft_reduce_nyq(const float2 *__restrict__ nyq, int nseg, int L,
                                                          double2 *__restrict__ St)
{
    const int k2 = blockIdx.x * 256 + threadIdx.x;
    if (k2 >= L) return;
    double sx = 0.0, sy = 0.0;
    for (int sg = 0; sg < nseg; ++sg) {
        const float2 v = nyq[(size_t)sg * L + k2];
        sx += (double)v.x;
        sy += (double)v.y;
    }
    St[(size_t)(L / 2) * L + k2] = make_double2(sx, sy);
}

/* ---------------------------------------------------------- 2-D inverse */
/* one 64-thread team per row k3 of St: inverse DFT over k2, in place;
   dynamic LDS 2 L complex double, twiddles read from device memory */
__global__ __launch_bounds__(64) void k_lagft_cols(double2 *__restrict__ St, FftPlan pl, const double2 *__restrict__ tw)
{
    extern __shared__ __align__(16) double2 dsm[];
    const int L = pl.n;
    double2 *bx = dsm, *by = dsm + L;
    double2 *row = St + (size_t)blockIdx.x * L;
    for (int t = threadIdx.x; t < L; t += 64) bx[t] = row[t];
    __syncthreads();
    const double2 *res = fftc::transform<+1, 5>(bx, by, tw, pl, threadIdx.x, 64);
    for (int t = threadIdx.x; t < L; t += 64) row[t] = res[t];
}

/* window rows wa = 2 blockIdx.x and wb = wa + 1 (d2 = w - nc10): the
   half spectra St[k3][d2], k3 <= L/2, extended Hermitian, as one complex
   sequence A + i B; its inverse DFT carries row wa in the real part and
   row wb in the imaginary part (DC and Nyquist terms: real parts only) */
__global__ __launch_bounds__(64) void k_lagft_rows(const double2 *__restrict__ St, FftPlan pl,
                                                   const double2 *__restrict__ tw, int nc10,
                                                   const int *__restrict__ cnt, const int *__restrict__ ex1,
                                                   const int *__restrict__ ex2, const int *__restrict__ ex3,
                                                   float *__restrict__ bl, double *__restrict__ bl64)
{
    extern __shared__ __align__(16) double2 dsm[];
    const int L = pl.n, W = 2 * nc10 + 1;
    double2 *bx = dsm, *by = dsm + L;
    const int wa = 2 * blockIdx.x, wb = wa + 1 < W ? wa + 1 : wa;
    const int da = ((wa - nc10) % L + L) % L, db = ((wb - nc10) % L + L) % L;
    for (int k = threadIdx.x; k < L; k += 64) {
        double2 y;
        if (k == 0 || 2 * k == L) {
            y = make_double2(St[(size_t)k * L + da].x, St[(size_t)k * L + db].x);
        } else if (2 * k < L) {
            const double2 A = St[(size_t)k * L + da], B = St[(size_t)k * L + db];
            y = make_double2(A.x - B.y, A.y + B.x);
        } else {
            const int kk = L - k;
            const double2 A = St[(size_t)kk * L + da], B = St[(size_t)kk * L + db];
            y = make_double2(A.x + B.y, B.x - A.y);
        }
        bx[k] = y;
    }
    __syncthreads();
    const double2 *res = fftc::transform<+1, 5>(bx, by, tw, pl, threadIdx.x, 64);
    const double scale = ldexp(1.0, *ex1 + *ex2 + *ex3 - 3 * AMB_SET_EXP) / ((double)L * (double)L);
    for (int w3 = threadIdx.x; w3 < W; w3 += 64) {
        const int idx = ((w3 - nc10) % L + L) % L;
        const double2 z = res[idx];
        const size_t ca = (size_t)wa * W + w3, cb = (size_t)wb * W + w3;
        const double va = cnt[ca] ? z.x * scale : 0.0;
        bl[ca] = (float)va;
        if (bl64) bl64[ca] = va;
        if (wb != wa) {
            const double vb = cnt[cb] ? z.y * scale : 0.0;
            bl[cb] = (float)vb;
            if (bl64) bl64[cb] = vb;
        }
    }
}

/* --------------------------------------------------------------- counts */
/* cnt[d2][d3] = #{p1 in C : p1 + d2 in C, p1 + d3 in C}, |d| <= nc10 */
__global__ __launch_bounds__(256) void k_lag_counts(const int *__restrict__ act, const int *__restrict__ inv, int nc,
                                                    int nact, int nc10, int *__restrict__ cnt)
{
    extern __shared__ unsigned char on[];
    for (int p = threadIdx.x; p < nc; p += blockDim.x) on[p] = inv[p] >= 0;
    __syncthreads();
    const long W = 2 * nc10 + 1;
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= W * W) return;
    const int d2 = (int)(cell / W) - nc10, d3 = (int)(cell % W) - nc10;
    int n = 0;
    for (int i = 0; i < nact; ++i) {
        const int p = act[i], p2 = p + d2, p3 = p + d3;
        n += (p2 >= 0 && p2 < nc && p3 >= 0 && p3 < nc && on[p2] && on[p3]) ? 1 : 0;
    }
    cnt[cell] = n;
}

/* ------------------------------------------------------------------ host */

int nice_length(int lmin)
{
    for (int L = std::max(lmin, 2);; ++L) {
        if (L & 1) continue;
        int m = L;
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        if (m == 1) return L;
    }
}

int spectra_teams(int L)
{
    int t = (8192 / L - 1) / 2; /* tw + 2 L per team in 64 KiB */
    return std::min(4, t);
}

} // namespace

/* channel-axis transform length, row pitch and counts of a cube's channel set */
int amb::lag_setup(amb_ctx *c, Cube &cb)
{
    if (cb.lag_L) return AMB_OK;
    const int na = cb.nact();
    if (na < 1) return fail(c, AMB_ESTATE, "empty channel set");
    const int span = cb.act.back() - cb.act.front() + 1;
    const int L = nice_length(span + std::min(cb.nc10, span - 1));
    if (L > MAXL) return fail(c, AMB_EARG, "lag-Fourier form: channel transform %d exceeds %d", L, MAXL);
    if (spectra_teams(L) < 1) return fail(c, AMB_EARG, "lag-Fourier form: transform %d exceeds LDS", L);
    const size_t W = 2 * (size_t)cb.nc10 + 1, cells = W * W;
    void *d = nullptr;
    int rc = dev_alloc(c, &d, sizeof(int) * cells, "lag counts");
    if (rc) return rc;
    hipLaunchKernelGGL(k_lag_counts, dim3((unsigned)((cells + 255) / 256)), dim3(256), (size_t)cb.nc, c->stream,
                       cb.d_act, cb.d_inv, (int)cb.nc, na, cb.nc10, (int *)d);
    std::vector<int32_t> h(cells);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, sizeof(int) * cells, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        dev_free(c, d, sizeof(int) * cells);
        return fail(c, AMB_EHIP, "lag counts: %s", hipGetErrorString(e));
    }
    int64_t ev = 0, pr = 0;
    for (size_t i = 0; i < cells; ++i) {
        ev += h[i] > 0;
        pr += h[i];
    }
    cb.lag_L = L;
    cb.lag_lhp = (L / 2 + 1 + 3) / 4 * 4;
    cb.d_lagcnt = (int *)d;
    cb.h_lagcnt = std::move(h);
    cb.lag_evals = ev;
    cb.lag_pairs = pr;
    return AMB_OK;
}

void amb::lag_release(amb_ctx *c, Cube &cb)
{
    if (cb.d_lagcnt) dev_free(c, cb.d_lagcnt, sizeof(int) * cb.h_lagcnt.size());
    cb.d_lagcnt = nullptr;
    cb.h_lagcnt.clear();
    cb.lag_L = cb.lag_lhp = 0;
}

size_t amb::lag_spectra_bytes(const Cube &cb, long ng)
{
    return sizeof(float2) * (size_t)ng * ng * cb.lag_lhp + AMB_SET_SLOT_BYTES;
}

hipError_t amb::lag_launch_spectra(amb_ctx *c, hipStream_t s, const BuildOp &op)
{
    const Cube &cb = c->cubes[op.cube];
    const long ng = op.g.ng, P = ng * ng, pitch = plane_pitch(ng);
    const int L = cb.lag_L;
    auto itw = c->tw.find(L);
    if (itw == c->tw.end()) return hipErrorInvalidValue;
    const FftPlan &pl = c->plans[L];
    const int nteam = spectra_teams(L), ppt = 4;
    const long npair = (P + 1) / 2, per_blk = (long)nteam * ppt;
    const long nblk = ((npair + per_blk - 1) / per_blk + 7) / 8 * 8;
    const size_t lds = sizeof(float2) * (size_t)L * (1 + 2 * nteam);
    const size_t setw = (size_t)pitch * cb.nact();
    const unsigned *slot_in = (const unsigned *)((const float *)op.src + setw);
    unsigned *slot_out = (unsigned *)((float2 *)op.dst + (size_t)P * cb.lag_lhp);
    hipLaunchKernelGGL(k_chan_spectra, dim3((unsigned)nblk), dim3(64 * nteam), lds, s, (const float *)op.src, pitch,
                       cb.d_inv, (int)cb.nc, cb.act.front(), P, pl, itw->second, cb.lag_lhp, ppt, (float2 *)op.dst,
                       slot_in, slot_out);
    return hipGetLastError();
}

/* triangles of the lag-Fourier form in waves; bl64[i] (may be null, or
   hold nulls): FP64 lag arrays of request i */
int amb::lagft_run(amb_ctx *c, const amb_tri_req *req, int n, double t0, double *done_ms, double *const *bl64)
{
    hipSetDevice(c->device);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return req[x].cube_id != req[y].cube_id ? req[x].cube_id < req[y].cube_id : req[x].r1 < req[y].r1;
    });
    const PlaneSrc src{&c->planes};
    struct J {
        int q;
        Cube *cb;
        long ng, P;
        int nseg;
        void *X[3];
        size_t part, nyq, out, out64;
    };
    std::vector<J> w;
    size_t part_tot = 0, out_tot = 0;
    std::vector<int> norm_redo;

    auto flush = [&]() -> int {
        if (w.empty()) return AMB_OK;
        /* buffers */
        size_t st_need = 0;
        for (auto &j : w) st_need = std::max(st_need, sizeof(double2) * (size_t)(j.cb->lag_L / 2 + 1) * j.cb->lag_L);
        void *p;
        int rc = AMB_OK;
        p = c->lag_part;
        if (!rc) rc = grow(c, &p, &c->lag_part_bytes, part_tot, "lag-Fourier partials");
        c->lag_part = p;
        p = c->lag_st;
        if (!rc) rc = grow(c, &p, &c->lag_st_bytes, st_need, "lag-Fourier spectrum");
        c->lag_st = p;
        p = c->lag_out;
        if (!rc) rc = grow(c, &p, &c->lag_out_bytes, out_tot, "lag arrays");
        c->lag_out = p;
        if (!rc) rc = grow(c, &c->wdesc, &c->wdesc_bytes, sizeof(LagJob) * w.size(), "wave descriptors");
        if (!rc) rc = grow_host(c, &c->hstage[0], &c->hstage_bytes[0], out_tot);
        if (rc) { unpin_all(c); w.clear(); return rc; }
        std::vector<LagJob> d(w.size());
        long items = 0, nitems = 0;
        for (size_t i = 0; i < w.size(); ++i) {
            const J &j = w[i];
            const int L = j.cb->lag_L;
            LagJob &g = d[i];
            g.F = (const float2 *)j.X[0];
            g.G = (const float2 *)j.X[1];
            g.H = (const float2 *)j.X[2];
            for (int k = 0; k < 3; ++k) {
                const int *ex = (const int *)((const float2 *)j.X[k] + (size_t)j.P * j.cb->lag_lhp) + 1;
                (k == 0 ? g.ex1 : k == 1 ? g.ex2 : g.ex3) = ex;
            }
            g.part = (float2 *)((char *)c->lag_part + j.part);
            g.nyq = (float2 *)((char *)c->lag_part + j.nyq);
            g.P = j.P;
            g.lhp = j.cb->lag_lhp;
            g.L = L;
            g.nseg = j.nseg;
            g.seg = SEG;
            g.item0 = items;
            g.nitem0 = nitems;
            items += (long)j.nseg * tiles2(L) * tiles3(L);
            nitems += j.nseg;
        }
        hipError_t e = hipMemcpyAsync(c->wdesc, d.data(), sizeof(LagJob) * d.size(), hipMemcpyHostToDevice, c->stream);
        const LagJob *dj = (const LagJob *)c->wdesc;
        timer_start(c, &c->st.ms_contract);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_lagft_tiles, dim3((unsigned)((items + 7) / 8 * 8)), dim3(256), 0, c->stream, dj,
                               (int)d.size(), items);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_lagft_nyq, dim3((unsigned)((nitems + 7) / 8 * 8)), dim3(256), 0, c->stream, dj,
                               (int)d.size(), nitems);
            e = hipGetLastError();
        }
        timer_stop(c);
        c->st.n_contract_launch++;
        timer_start(c, &c->st.ms_reduce);
        for (size_t i = 0; i < w.size() && e == hipSuccess; ++i) {
            const J &j = w[i];
            const Cube &cb = *j.cb;
            const int L = cb.lag_L;
            double2 *St = (double2 *)c->lag_st;
            hipLaunchKernelGGL(k_lagft_reduce, dim3((unsigned)((L + 31) / 32), (unsigned)((L / 2 + 31) / 32)), dim3(256),
                               0, c->stream, d[i].part, j.nseg, L, St);
            hipLaunchKernelGGL(k_lagft_reduce_nyq, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, c->stream, d[i].nyq,
                               j.nseg, L, St);
            const FftPlan &pl = c->plans[L];
            const double2 *tw = c->tw64[L];
            const size_t lds = sizeof(double2) * 2 * (size_t)L;
            hipLaunchKernelGGL(k_lagft_cols, dim3((unsigned)(L / 2 + 1)), dim3(64), lds, c->stream, St, pl, tw);
            const int W = 2 * cb.nc10 + 1;
            float *o = (float *)((char *)c->lag_out + j.out);
            double *o64 = j.out64 != (size_t)-1 ? (double *)((char *)c->lag_out + j.out64) : nullptr;
            hipLaunchKernelGGL(k_lagft_rows, dim3((unsigned)((W + 1) / 2)), dim3(64), lds, c->stream, St, pl, tw, cb.nc10,
                               cb.d_lagcnt, d[i].ex1, d[i].ex2, d[i].ex3, o, o64);
            e = hipGetLastError();
        }
        timer_stop(c);
        if (e == hipSuccess) e = hipMemcpyAsync(c->hstage[0], c->lag_out, out_tot, hipMemcpyDeviceToHost, c->stream);
        const hipError_t es = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = es;
        timers_collect(c);
        unpin_all(c);
        if (e != hipSuccess) {
            w.clear();
            return fail(c, AMB_EHIP, "lag-Fourier wave of %zu triangles: %s", d.size(), hipGetErrorString(e));
        }
        const double tdone = now_ms() - t0;
        for (auto &j : w) {
            const amb_tri_req &q = req[j.q];
            const Cube &cb = *j.cb;
            const size_t cells = cb.h_lagcnt.size();
            memcpy(q.bl, (const char *)c->hstage[0] + j.out, sizeof(float) * cells);
            if (bl64 && bl64[j.q]) memcpy(bl64[j.q], (const char *)c->hstage[0] + j.out64, sizeof(double) * cells);
            if (q.cnt) memcpy(q.cnt, cb.h_lagcnt.data(), sizeof(int32_t) * cells);
            const int L = cb.lag_L;
            c->st.evals += cb.lag_evals;
            c->st.flop_contract += 2.0 * (double)cb.lag_pairs * (double)j.P;
            c->st.flop_contract_distinct += 2.0 * (double)cb.lag_pairs * (double)j.P;
            /* 14 flops per (k2, k3, pixel) of the half plane k3 <= L/2 */
            c->st.flop_lagft += 14.0 * (double)L * (double)(L / 2 + 1) * (double)j.P;
            c->st.flop_contract_exec += 14.0 * (double)tiles2(L) * LT * ((double)tiles3(L) * LT + 1) * (double)j.P;
            if (done_ms) done_ms[j.q] = tdone;
            if (q.norm) norm_redo.push_back(j.q);
        }
        w.clear();
        part_tot = out_tot = 0;
        return AMB_OK;
    };

    for (int k = 0; k < n; ++k) {
        const int qi = order[k];
        const amb_tri_req &q = req[qi];
        if (!q.bl) return fail(c, AMB_EARG, "request %d has no bl buffer", qi);
        Cube *cb;
        int rc = find_cube(c, q.cube_id, &cb);
        if (rc) { unpin_all(c); return rc; }
        if (!cb->chans) { unpin_all(c); return fail(c, AMB_ESTATE, "channels not set for cube %d", q.cube_id); }
        const int nr = (int)c->ulow.size();
        if (nr == 0) { unpin_all(c); return fail(c, AMB_ESTATE, "rings not set"); }
        if (q.r1 < 0 || q.r1 >= nr || q.r2 < 0 || q.r2 >= nr || q.r3 < 0 || q.r3 >= nr) {
            unpin_all(c);
            return fail(c, AMB_EARG, "triangle (%d,%d,%d) outside [0,%d)", q.r1, q.r2, q.r3, nr);
        }
        if ((rc = lag_setup(c, *cb))) { unpin_all(c); return rc; }
        const long ng = amb_ng(c->uhigh[q.r1], cb->Du);
        if (ng <= 0 || ng > 2730) { unpin_all(c); return fail(c, AMB_EARG, "Ng=%ld out of range", ng); }
        const int L = cb->lag_L;
        {   /* FP64 twiddles and plan of L for the 2-D inverse */
            float2 *tw;
            FftPlan pl;
            if ((rc = twiddles(c, L, &tw, &pl)) || (rc = twiddles64(c, L, nullptr))) { unpin_all(c); return rc; }
        }
        J j{};
        j.q = qi;
        j.cb = cb;
        j.ng = ng;
        j.P = ng * ng;
        j.nseg = (int)((j.P + SEG - 1) / SEG);
        const size_t pb = sizeof(float2) * (size_t)j.nseg * ((size_t)tiles2(L) * tiles3(L) * LT * LT + L);
        const size_t cells = cb->h_lagcnt.size();
        const bool want64 = bl64 && bl64[qi];
        const size_t ob = al(sizeof(float) * cells) + (want64 ? al(sizeof(double) * cells) : 0);
        if (!w.empty() && (part_tot + al(pb) > PART_CAP || out_tot + ob > OUT_CAP || (int)w.size() >= MAX_JOBS))
            if ((rc = flush())) return rc;
        const int rr[3] = {q.r1, q.r2, q.r3};
        for (int t = 0; t < 3 && !rc; ++t) {
            rc = get_planes(c, src, q.cube_id, 9, rr[t], (int)ng, &j.X[t]);
            if (rc == AMB_EBUDGET && !w.empty()) {
                /* the wave's planes hold the memory: run it, then resolve again */
                if ((rc = flush())) return rc;
                t = -1;
                rc = AMB_OK;
            }
        }
        if (rc) { unpin_all(c); return rc; }
        j.part = part_tot;
        j.nyq = part_tot + sizeof(float2) * (size_t)j.nseg * tiles2(L) * tiles3(L) * LT * LT;
        part_tot += al(pb);
        j.out = out_tot;
        j.out64 = want64 ? out_tot + al(sizeof(float) * cells) : (size_t)-1;
        out_tot += ob;
        w.push_back(j);
    }
    int rc = flush();
    if (rc) return rc;
    /* v2 normalisation of this form: the exact FP64 path */
    for (int i : norm_redo) {
        const amb_tri_req &q = req[i];
        if ((rc = norm_exact(c, q.cube_id, q.r1, q.r2, q.r3, q.norm, nullptr, nullptr))) return rc;
        if (done_ms) done_ms[i] = now_ms() - t0;
    }
    return AMB_OK;
}

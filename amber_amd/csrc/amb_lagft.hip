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
 * Stages (data: FP32 products, FP64 sums and transforms after them; the v2
 * normalisation runs the same stages in FP64 throughout, its sums rounded
 * to their 1/8 grid of Ng^2 with a certificate):
 *   k_chan_spectra  per pixel pair, the real channel sequences packed as
 *                   one complex sequence, forward DFT of length L in LDS;
 *                   X[x][k], k = 0 .. L/2 (a plane set kind of its own);
 *   k_lagft_tiles   S over 128 x 128 (k2, k3) tiles, k3 < L/2, for one
 *                   segment of pixels per work item: FP32 accumulators in
 *                   registers, written as the item's partial;
 *                   the k3 = L/2 column rides along in the last k3 tile
 *                   (which stages every conj F it needs), or is a column
 *                   of the last, partial tile when TB does not divide L/2;
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

constexpr int BLKB = 80;            /* LDS bytes per unit: 64 B of values + 16 B pad */
constexpr int SEG = 1024;           /* pixels per tile work item: a launch's last round leaves CUs
                                       idle, and shorter items shorten that tail (2048 was 2.5%
                                       slower per pass, with fewer partials to sum) */
constexpr int MAXL = 2048;          /* longest channel transform: 2 L FP64 complex in 64 KiB of LDS */
constexpr size_t ALIGN = 256;
constexpr size_t PART_CAP = (size_t)8 << 30; /* partial sums per wave */
/* result bytes per wave: a wave's copies to host memory (PCIe, then into
   the caller's arrays) must fit behind the next wave's device work */
constexpr size_t OUT_CAP = (size_t)384 << 20;
constexpr int MAX_JOBS = 512;

size_t al(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

/* Per scalar type: the complex type, each lane's block of BU k2 x BV k3
 * accumulators (16 x 16 lanes: tiles of TA = 16 BU by TB = 16 BV), pixels
 * per LDS chunk, waves per SIMD the kernel is built for, whether the k3 =
 * L/2 column rides in the last tile (FOLD).  FP32: 8 x 8
 * (128 x 128 tiles); FP64: 4 x 8 (64 x 128), whose 32 complex double
 * accumulators take as many registers as the FP32 form's 64. */
template <typename R>
struct LagT;
template <>
struct LagT<float> {
    typedef float2 C;
    static constexpr int BU = 8, BV = 8, KPX = 4, WPS = 2;
    static constexpr bool FOLD = true;
};
template <>
struct LagT<double> {
    typedef double2 C;
    /* the k3 = L/2 column in a kernel of its own: folding it into the tiles
       would take the registers of a second wave per SIMD */
    static constexpr int BU = 4, BV = 8, KPX = 2, WPS = 2;
    static constexpr bool FOLD = false;
};

template <typename R>
struct LagJob {
    typedef typename LagT<R>::C C;
    const C *F, *G, *H;         /* channel spectra [P][lhp] of r1, r2, r3 */
    const int *ex1, *ex2, *ex3; /* data sets: exponents E of the three sets (null: unscaled) */
    C *part;                    /* [nseg][ntile][TA*TB] tile partials */
    C *nyq;                     /* [nseg][L] partials of the k3 = L/2 column, when a
                                   kernel of its own sums it (nyq_sep) */
    const int2 *tlist;          /* mode 1: (I, J) of each tile */
    const int *nyql;            /* mode 1: the k2 of the k3 = L/2 column that no listed
                                   tile holds (null: every k2) */
    int nnyq;
    long P;
    long item0, nitem0;         /* first tile item, first nyquist item in the launch */
    long ibase;                 /* the triangle's item that the launch's item0 is (a
                                   launch may cover a sub-range of a triangle's items:
                                   each item owns its partial block) */
    int lhp, L, nseg, seg;      /* tile segments and their pixels */
    int mode;                   /* 0: (k2, k3) tiles; 1: r1 == r2, (i, j) tiles, i <= j;
                                   2: r2 == r3, listed (k2, k3) tiles (tiles below the
                                   diagonal of [0, L/2)^2 dropped: S is symmetric there);
                                   3: r1 == r2, sheared (t, k3) tiles; 4: r2 == r3, mode 2's
                                   tiles without those past the anti-diagonal for k2 >= L/2;
                                   5: r1 == r2 == r3, the listed sheared tiles of the
                                   principal domain */
    int ntile;                  /* tiles per segment */
    int nyq_sep;                /* the k3 = L/2 column is summed by k_lagft_nyq */
};

/* most k2 of the k3 = L/2 column a pair-tile job leaves to k_lagft_nyq */
constexpr int NYQ_LIST_MAX = 32;

template <typename R>
__host__ __device__ inline int tile_a() { return 16 * LagT<R>::BU; }
template <typename R>
__host__ __device__ inline int tile_b() { return 16 * LagT<R>::BV; }
template <typename R>
__host__ __device__ inline int tiles2(int L) { return (L + tile_a<R>() - 1) / tile_a<R>(); }
template <typename R>
__host__ __device__ inline int tiles3(int L) { return (L / 2 + tile_b<R>() - 1) / tile_b<R>(); }

/* blocks b = x (mod 8) run on XCD x: XCD x takes items [x per, (x+1) per) */
__device__ __forceinline__ long xcd_item(long b, long nb)
{
    const long per = nb / 8;
    return (b % 8) * per + b / 8;
}

template <typename J>
__device__ __forceinline__ int find_job(const J *__restrict__ jobs, int njob, long it, bool nyq = false)
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
template <typename C>
__device__ __forceinline__ C spec_at(const C *__restrict__ X, long row, int k, int L)
{
    if (2 * k <= L) return X[row + k];
    C v = X[row + (L - k)];
    v.y = -v.y;
    return v;
}

/* the 64 value bytes of one LDS unit as complex values, in 16-byte reads */
__device__ __forceinline__ void load_unit(const char *p, float2 *v)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 w = *(const float4 *)(p + 16 * q);
        v[2 * q] = make_float2(w.x, w.y);
        v[2 * q + 1] = make_float2(w.z, w.w);
    }
}
__device__ __forceinline__ void load_unit(const char *p, double2 *v)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = *(const double2 *)(p + 16 * q);
}

/* ---------------------------------------------------------------- spectra */
/* grid: multiple of 8 blocks of nteam x 64 threads, each team ppt
   consecutive pixel pairs; dynamic LDS: [tw L if TWL] + nteam * 2 * L
   complex.  SCALED: a data set as computed (kind 11), read times 2^(7 - E),
   E = amb_set_exp of its max word, which goes with E after the spectra.
   Channel ch = act0 + p sits at position p of the sequence.
   The next pair's plane values are loaded into registers before the
   current pair's transform, so their latency hides behind it; a
   workgroup's pairs cover whole 128-byte lines of every plane. */
template <typename R, bool SCALED, bool TWL>
__global__ __launch_bounds__(256) void k_chan_spectra(const R *__restrict__ D, long pitch, const int *__restrict__ inv,
                                                      int nc, int act0, long P, FftPlan pl,
                                                      const typename LagT<R>::C *__restrict__ twg, int lhp, int ppt,
                                                      typename LagT<R>::C *__restrict__ X,
                                                      const unsigned *__restrict__ slot_in, unsigned *__restrict__ slot_out)
{
    typedef typename LagT<R>::C C;
    constexpr int NPF = MAXL / 64; /* sequence positions per thread */
    extern __shared__ __align__(16) char lraw[];
    C *lsm = (C *)lraw;
    const int L = pl.n, h = L / 2;
    const int team = threadIdx.x >> 6, t = threadIdx.x & 63, nteam = blockDim.x >> 6;
    C *bx = lsm + (TWL ? L : 0) + (size_t)team * 2 * L, *by = bx + L;
    const C *tw = TWL ? lsm : twg;
    if (TWL)
        for (int i = threadIdx.x; i < L; i += blockDim.x) lsm[i] = twg[i];
    const int E = SCALED ? amb_set_exp(__uint_as_float(slot_in[0])) : 0;
    const R sc = SCALED ? (R)ldexpf(1.f, AMB_SET_EXP - E) : (R)1;
    if (SCALED && blockIdx.x == 0 && threadIdx.x == 0) {
        slot_out[0] = slot_in[0];
        slot_out[1] = (unsigned)E;
    }
    const long npair = (P + 1) / 2;
    const long pair0 = (xcd_item(blockIdx.x, gridDim.x) * nteam + team) * ppt;
    const int pmax = min(L, nc - act0); /* positions that can hold a channel */
    /* FP64 (the normalisation's spectra) loads straight into LDS: its
       prefetch registers would double the kernel's VGPRs */
    constexpr bool PF = sizeof(R) == 4;
    C pf[PF ? NPF : 1];
    auto fetch = [&](long pr) {
        const bool live = pr < npair;
        const long x0 = 2 * pr;
#pragma unroll
        for (int u = 0; u < (PF ? NPF : (L + 63) / 64); ++u) {
            const int p = t + 64 * u;
            C v;
            v.x = v.y = (R)0;
            if (live && p < pmax) {
                const int i = inv[act0 + p];
                if (i >= 0) {
                    /* x0 even and the pitch a multiple of 64: an aligned pair;
                       for odd P the second value is the zeroed pitch tail */
                    const R *d = D + (size_t)i * pitch + x0;
                    v.x = d[0];
                    v.y = d[1];
                    if (SCALED) { v.x *= sc; v.y *= sc; } /* a power of two: exact */
                }
            }
            if constexpr (PF) pf[u] = v;
            else if (p < L) bx[p] = v;
        }
    };
    if (PF) fetch(pair0);
    __syncthreads();
    for (int q = 0; q < ppt; ++q) {
        const long pr = pair0 + q;
        const bool live = pr < npair;
        const long x0 = 2 * pr;
        if constexpr (PF) {
#pragma unroll
            for (int u = 0; u < NPF; ++u) {
                const int p = t + 64 * u;
                if (p < L) bx[p] = pf[u];
            }
            if (q + 1 < ppt) fetch(pr + 1);
        } else {
            fetch(pr);
        }
        __syncthreads();
        const C *Z = fftc::transform<-1, 5>(bx, by, tw, pl, t, 64);
        if (live) {
            C *xa = X + (size_t)x0 * lhp, *xb = xa + lhp;
            const bool two = x0 + 1 < P;
            for (int k = t; k <= h; k += 64) {
                const C a = Z[k], b = Z[k == 0 ? 0 : L - k];
                /* A = (Z(k) + conj Z(L-k)) / 2, B = (Z(k) - conj Z(L-k)) / 2i */
                C A, B;
                A.x = (R)0.5 * (a.x + b.x);
                A.y = (R)0.5 * (a.y - b.y);
                B.x = (R)0.5 * (a.y + b.y);
                B.y = (R)-0.5 * (a.x - b.x);
                xa[k] = A;
                if (two) xb[k] = B;
            }
        }
        __syncthreads(); /* the next pair overwrites both buffers */
    }
}

/* Channel spectra of a block of 32 pixels (16 pairs) per workgroup, FP32
 * sets of at most SPB_PLANES planes.  The block's values are loaded once
 * with whole-line reads (a wave reads 128 contiguous bytes of 4 planes per
 * instruction) into registers: thread tid holds pixel pair tid % 16 of
 * planes tid / 16 + 16 m.  Four rounds then each transform 4 pairs, one
 * per 64-thread team: the holders of a round's pairs write their values at
 * the digit-reversed positions pos[i] of a zeroed team buffer, and the
 * in-place transform leaves the spectrum in natural order.  The team
 * buffers are L + 8 values apart and addressed through fftc::IdxXor32 (a
 * permutation inside each aligned block of 32): the 16 writers of a
 * scatter (4 buffers x 4 planes, whose digit-reversed positions are L/8
 * apart) and the first two radix-8 stages then spread over the banks
 * instead of landing 4 to 16 on one.  Dynamic LDS: [tw L if TWL] +
 * 4 (L + 8) complex + nact positions.  The source is a data set as
 * computed (kind 11), read times 2^(7 - E), E = amb_set_exp of its max
 * word; the max word and E follow the spectra. */
constexpr int SPB_PLANES = 768;
template <bool TWL>
__global__ __launch_bounds__(256, 2) void k_chan_spectra_blk(const float *__restrict__ D, long pitch,
                                                             const int *__restrict__ posg, int nact, long P,
                                                             FftPlan pl, const float2 *__restrict__ twg, int lhp,
                                                             float2 *__restrict__ X,
                                                             const unsigned *__restrict__ slot_in,
                                                             unsigned *__restrict__ slot_out)
{
    constexpr int NQ = SPB_PLANES / 16;
    extern __shared__ __align__(16) char lraw[];
    const int L = pl.n, h = L / 2;
    /* TWL: the twiddles in LDS; otherwise read through the cache, and
       three workgroups fit a CU instead of two */
    const int LB = L + 8;
    const fftc::IdxXor32 sw(L);
    float2 *twl = (float2 *)lraw, *buf = twl + (TWL ? L : 0);
    const float2 *tw = TWL ? twl : twg;
    int *pos = (int *)(buf + 4 * LB);
    const int tid = threadIdx.x, team = tid >> 6, t = tid & 63;
    if (TWL)
        for (int i = tid; i < L; i += 256) twl[i] = twg[i];
    for (int i = tid; i < nact; i += 256) pos[i] = sw(posg[i]);
    const int E = amb_set_exp(__uint_as_float(slot_in[0]));
    const float sc = ldexpf(1.f, AMB_SET_EXP - E);
    if (blockIdx.x == 0 && tid == 0) {
        slot_out[0] = slot_in[0];
        slot_out[1] = (unsigned)E;
    }
    const long x0 = 32 * xcd_item(blockIdx.x, gridDim.x);
    if (x0 >= P) return; /* whole workgroups: no barrier is skipped by part of one */
    const int j = tid & 15, i0 = tid >> 4;
    /* x0 + 2 j + 1 < x0 + 32 <= pitch (the pitch is a multiple of 64, its tail zero) */
    float2 v[NQ];
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
        const int i = i0 + 16 * m;
        float2 w = make_float2(0.f, 0.f);
        if (i < nact) {
            w = *(const float2 *)(D + (size_t)i * pitch + x0 + 2 * j);
            w.x *= sc; /* a power of two: exact */
            w.y *= sc;
        }
        v[m] = w;
    }
    float2 *bx = buf + (size_t)team * LB;
    for (int r = 0; r < 4; ++r) {
        for (int p = t; p < L; p += 64) bx[p] = make_float2(0.f, 0.f);
        __syncthreads();
        if ((j >> 2) == r) {
            float2 *bt = buf + (size_t)(j & 3) * LB;
#pragma unroll
            for (int m = 0; m < NQ; ++m) {
                const int i = i0 + 16 * m;
                if (i < nact) bt[pos[i]] = v[m];
            }
        }
        __syncthreads();
        fftc::transform_ip<-1>(bx, tw, pl, t, 64, sw);
        const long xa = x0 + 2 * (4 * r + team);
        if (xa < P) {
            float2 *oa = X + (size_t)xa * lhp, *ob = oa + lhp;
            const bool two = xa + 1 < P;
            for (int k = t; k <= h; k += 64) {
                const float2 a = bx[sw(k)], b = bx[sw(k == 0 ? 0 : L - k)];
                /* A = (Z(k) + conj Z(L-k)) / 2, B = (Z(k) - conj Z(L-k)) / 2i */
                oa[k] = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
                if (two) ob[k] = make_float2(0.5f * (a.y + b.y), -0.5f * (a.x - b.x));
            }
        }
        __syncthreads(); /* the next round zeroes the buffers */
    }
}

/* ------------------------------------------------------------------ tiles */
/* One work item: a TA x TB tile over one segment of pixels, as
 * acc(p, q) += C(p + q) A(p) B(q) per pixel, the three operands staged in
 * padded 80-byte LDS units (16 lanes reading 16 units at a stride of 20
 * banks touch all 64 banks once); lane (li, lj) accumulates p = pa + BU li
 * + u, q = pb + BV lj + v.
 *   mode 0: p = k2, q = k3 (k3 < L/2): A = G(k2), B = H(k3), C = conj F(k2+k3);
 *   mode 1 (r1 == r2, F = G): p = i = -k2, q = j = k2 + k3: A = conj F(i),
 *           B = conj F(j), C = H(i + j), symmetric in (i, j), so only tiles
 *           I <= J of the band (i + j) mod L in [0, L/2) are summed (the
 *           job's tile list);
 *   mode 2 (r2 == r3, G = H): as mode 0 over the job's tile list, which
 *           leaves out the tiles of [0, L/2)^2 below its diagonal (S is
 *           symmetric in (k2, k3) there).
 * FOLD (FP32): in mode 0 with TB | L/2 the last k3 tile also sums the k3 =
 * L/2 column, S(k2, L/2) = conj F(k2 + L/2) G(k2) H(L/2): every conj F it
 * needs is staged (at k2 - a + TB), and H(L/2) takes one more slot.  When
 * TB does not divide L/2 the column is a column of the last, partial tile;
 * otherwise (FP64, mode 1) k_lagft_nyq sums it. */
/* ---------------------------------------------------------- sheared tiles */
/* Mode 3 (FP32, r1 == r2, F = G, 256 | L).  S(k2, k3) = S(-k2 - k3, k3):
 * for a row k3 the map k2 -> -k2 - k3 (mod L) is a reflection, and in the
 * coordinate t = k2 + floor(k3/2) it is t -> -1 - t (k3 odd) or t -> -t
 * (k3 even, fixed points t = 0 and t = L/2).  So t in [0, L/2) holds every
 * distinct sum of a row once, plus t = L/2 for even k3: the half plane's
 * distinct sums are the rectangle (t, k3) in [0, L/2)^2, the fixed-point
 * column t = L/2 (even k3 < L/2) and the row k3 = L/2 (t in [0, L/2]).
 * A 128 x 128 tile (t in [a, a+128), k3 in [b, b+128)) stages per pixel
 *   G(k2) for k2 = a - b/2 - 64 + r, r in [0, 193)      (units 0-24)
 *   conj F(s) for s = a + b/2 + r,   r in [0, 193)      (units 25-49)
 *   H(k3) for k3 = b + r,            r in [0, 129)      (units 50-66)
 * Lane (li, lj) = (tid % 16, tid / 16) owns t = a + 8 li + u, k3 = b + 8 lj
 * + v: G at r = 64 + 8 li - 4 lj + u - floor(v/2) (a 12-value window from
 * 8 li - 4 lj + 60), conj F at r = 8 li + 4 lj + u + ceil(v/2) (12 values
 * from 8 li + 4 lj), H at 8 lj + v.  Lanes that an LDS cycle serves
 * together differ in li, so their windows start 8 values (one 80-byte
 * unit, 20 banks) apart: conflict-free.  Per pixel that is 6 + 6 + 4
 * 16-byte reads for 64 products, as the (k2, k3) tiles.
 * The edge sums ride in the tiles that stage their operands, one per
 * thread, into the job's [segment][L] column block at index k2:
 *   k3 = L/2 (tiles with b = L/2 - 128): thread t < 128, t = a + tid,
 *     k2 = t - L/4: G at r = tid, conj F at tid + 64, H(L/2) at 128;
 *   t = L/2, even k3 (tiles with a = L/2 - 128): thread 128 + m, k3 = b +
 *     2m, k2 = L/2 - k3/2: G at 192 - m, conj F at 128 + m, H at 2m;
 *   (t, k3) = (L/2, L/2) (both): thread 192, k2 = L/4: G at 128, conj F at
 *     192, H at 128. */
constexpr int SH_UG = 25, SH_UF = 25, SH_UH = 17;
constexpr int SH_PXB = (SH_UG + SH_UF + SH_UH) * BLKB;
constexpr int SH_NG = 193, SH_NF = 193, SH_NH = 129;

__device__ __forceinline__ int sh_off(int unit0, int r) { return (unit0 + (r >> 3)) * BLKB + (r & 7) * 8; }

template <int KPX, int STRIDE>
__device__ __forceinline__ void shear_tile(const LagJob<float> &J, int sg, int tile, char *lds0, char *lds1)
{
    typedef float2 C;
    constexpr int SPT = (SH_NG + SH_NF + SH_NH + 255) / 256;
    const long P = J.P;
    const int lhp = J.lhp, L = J.L, seg = J.seg, half = L / 2;
    const int nta = L / 256;
    int a = (tile / nta) * 128, b = (tile % nta) * 128;
    if (J.mode == 5) { /* r1 == r2 == r3: the listed tiles of the principal domain */
        const int2 ij = J.tlist[tile];
        a = ij.x * 128;
        b = ij.y * 128;
    }
    const long xb = (long)sg * seg, xe = xb + seg < P ? xb + seg : P;
    const int tid = threadIdx.x;

    const C *src[SPT];
    int kk[SPT], off[SPT];
    bool ok[SPT], cj[SPT];
#pragma unroll
    for (int u = 0; u < SPT; ++u) {
        const int sl = tid + 256 * u;
        src[u] = J.G;
        kk[u] = 0;
        ok[u] = cj[u] = false;
        off[u] = -1;
        int k;
        bool conj_of_value;
        const C *x;
        if (sl < SH_NG) {
            k = a - b / 2 - 64 + sl;
            x = J.G;
            conj_of_value = false;
            off[u] = sh_off(0, sl);
        } else if (sl < SH_NG + SH_NF) {
            k = a + b / 2 + (sl - SH_NG);
            x = J.F;
            conj_of_value = true;
            off[u] = sh_off(SH_UG, sl - SH_NG);
        } else if (sl < SH_NG + SH_NF + SH_NH) {
            k = b + (sl - SH_NG - SH_NF); /* <= L/2 */
            x = J.H;
            conj_of_value = false;
            off[u] = sh_off(SH_UG + SH_UF, sl - SH_NG - SH_NF);
        } else {
            continue;
        }
        k %= L;
        if (k < 0) k += L;
        const bool m = 2 * k > L;
        src[u] = x;
        kk[u] = m ? L - k : k;
        cj[u] = m != conj_of_value;
        ok[u] = true;
    }

    C rg[KPX][SPT];
    auto fetch = [&](long x0) {
#pragma unroll
        for (int pp = 0; pp < KPX; ++pp) {
            const long x = x0 + pp;
            const size_t row = (size_t)x * lhp;
#pragma unroll
            for (int u = 0; u < SPT; ++u) {
                C v = make_float2(0.f, 0.f);
                if (x < xe && ok[u]) {
                    v = src[u][row + kk[u]];
                    if (cj[u]) v.y = -v.y;
                }
                rg[pp][u] = v;
            }
        }
    };
    auto stash = [&](char *st) {
#pragma unroll
        for (int pp = 0; pp < KPX; ++pp)
#pragma unroll
            for (int u = 0; u < SPT; ++u)
                if (off[u] >= 0) *(C *)(st + pp * STRIDE + off[u]) = rg[pp][u];
    };

    const int li = tid & 15, lj = tid >> 4;
    /* window starts (multiples of 4 values): reads 0, 1 (and 4, 5) sit at
       +0, +16 (+80, +96) bytes of the window's first unit; reads 2, 3 at
       +32, +48, or in the next unit at +80, +96 when the window starts
       mid-unit */
    const int gw = 8 * li - 4 * lj + 60, fw = 8 * li + 4 * lj;
    const int gbase = sh_off(0, gw), fbase = sh_off(SH_UG, fw);
    const int gmid = (gw & 7) ? 16 : 0, fmid = (fw & 7) ? 16 : 0;
    const int hbase = (SH_UG + SH_UF + lj) * BLKB;

    float acr[8][8], aci[8][8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int v = 0; v < 8; ++v) acr[u][v] = aci[u][v] = 0.f;

    /* the edge sum of this thread, if any */
    int eg = -1, ef = 0, eh = 0, ek2 = 0;
    if (b == half - 128 && tid < 128) {
        eg = sh_off(0, tid);
        ef = sh_off(SH_UG, tid + 64);
        eh = sh_off(SH_UG + SH_UF, 128);
        ek2 = (a + tid - L / 4 + L) % L;
    } else if (a == half - 128 && tid >= 128 && tid < 192) {
        const int m = tid - 128;
        eg = sh_off(0, 192 - m);
        ef = sh_off(SH_UG, 128 + m);
        eh = sh_off(SH_UG + SH_UF, 2 * m);
        ek2 = half - b / 2 - m;
    } else if (a == half - 128 && b == half - 128 && tid == 192) {
        eg = sh_off(0, 128);
        ef = sh_off(SH_UG, 192);
        eh = sh_off(SH_UG + SH_UF, 128);
        ek2 = L / 4;
    }
    const bool ex = eg >= 0;
    float enr = 0.f, eni = 0.f;

    const long nch = (xe - xb + KPX - 1) / KPX;
    fetch(xb);
    stash(lds0);
    __syncthreads();
    for (long ch = 0; ch < nch; ++ch) {
        if (ch + 1 < nch) fetch(xb + (ch + 1) * KPX);
        const char *st = (ch & 1) ? lds1 : lds0;
#pragma unroll 1
        for (int pp = 0; pp < KPX; ++pp) {
            const char *px = st + pp * STRIDE;
            C g[12], f[12], hv[8];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const int o = (q < 2 ? 16 * q : q < 4 ? 16 * q + gmid : 16 * q + 16);
                const float4 w = *(const float4 *)(px + gbase + o);
                g[2 * q] = make_float2(w.x, w.y);
                g[2 * q + 1] = make_float2(w.z, w.w);
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const int o = (q < 2 ? 16 * q : q < 4 ? 16 * q + fmid : 16 * q + 16);
                const float4 w = *(const float4 *)(px + fbase + o);
                f[2 * q] = make_float2(w.x, w.y);
                f[2 * q + 1] = make_float2(w.z, w.w);
            }
            load_unit(px + hbase, hv);
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    const C gg = g[4 + u - (v >> 1)], ff = f[u + ((v + 1) >> 1)];
                    const float tr = fmaf(gg.x, hv[v].x, -gg.y * hv[v].y);
                    const float ti = fmaf(gg.x, hv[v].y, gg.y * hv[v].x);
                    acr[u][v] = fmaf(ff.x, tr, fmaf(-ff.y, ti, acr[u][v]));
                    aci[u][v] = fmaf(ff.x, ti, fmaf(ff.y, tr, aci[u][v]));
                }
            if (ex) {
                const C gg = *(const C *)(px + eg), ff = *(const C *)(px + ef), hh = *(const C *)(px + eh);
                const float tr = fmaf(gg.x, hh.x, -gg.y * hh.y), ti = fmaf(gg.x, hh.y, gg.y * hh.x);
                enr = fmaf(ff.x, tr, fmaf(-ff.y, ti, enr));
                eni = fmaf(ff.x, ti, fmaf(ff.y, tr, eni));
            }
        }
        if (ch + 1 < nch) stash((ch & 1) ? lds0 : lds1);
        __syncthreads();
    }
    if (ex) J.nyq[(size_t)sg * L + ek2] = make_float2(enr, eni);
    C *out = J.part + ((size_t)sg * J.ntile + tile) * (128 * 128);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        C *o = out + (size_t)(8 * li + u) * 128 + 8 * lj;
#pragma unroll
        for (int v = 0; v < 8; ++v) o[v] = make_float2(acr[u][v], aci[u][v]);
    }
}

/* Mode 3 in FP64 (the exact v2 normalisation's sums): the same sheared
 * (t, k3) tiling with the FP64 kernel's 64 x 128 tiles, lane (li, lj) =
 * (tid % 16, tid / 16) owning t = a + 4 li + u (u < 4), k3 = b + 8 lj + v.
 * LDS units of 4 complex doubles (64 B + 16 B pad): per pixel
 *   G(k2) for k2 = a - b/2 - 64 + r, r in [0, 129)      (units 0-32)
 *   conj F(s) for s = a + b/2 + r,   r in [0, 129)      (units 33-65)
 *   H(k3) for k3 = b + r,            r in [0, 129)      (units 66-98)
 * G window of 8 values from 60 + 4 li - 4 lj (index 4 + u - floor(v/2)),
 * conj F window of 8 from 4 li + 4 lj (index u + ceil(v/2)), H 8 from
 * 8 lj: whole units, and lanes an LDS cycle serves one unit apart.  Edge
 * sums, one per thread, into the column block at index k2:
 *   k3 = L/2 (tiles with b = L/2 - 128): thread t < 64, t = a + tid,
 *     k2 = t - L/4: G at r = tid, conj F at tid + 64, H(L/2) at 128;
 *   t = L/2, even k3 (tiles with a = L/2 - 64): thread 64 + m, k3 = b + 2m,
 *     k2 = L/2 - k3/2: G at 128 - m, conj F at 64 + m, H at 2m;
 *   (L/2, L/2) (both): thread 128, k2 = L/4: G at 64, conj F at 128, H at 128. */
constexpr int SH64_UG = 0, SH64_UF = 33, SH64_UH = 66;
constexpr int SH64_PXB = 99 * BLKB;
constexpr int SH64_N = 129;

__device__ __forceinline__ int sh64_off(int unit0, int r) { return (unit0 + (r >> 2)) * BLKB + (r & 3) * 16; }

template <int KPX, int STRIDE>
__device__ __forceinline__ void shear_tile64(const LagJob<double> &J, int sg, int tile, char *lds0, char *lds1)
{
    typedef double2 C;
    constexpr int SPT = (3 * SH64_N + 255) / 256;
    const long P = J.P;
    const int lhp = J.lhp, L = J.L, seg = J.seg, half = L / 2;
    const int ntb = L / 256;
    int a = (tile / ntb) * 64, b = (tile % ntb) * 128;
    if (J.mode == 5) { /* r1 == r2 == r3: the listed tiles of the principal domain */
        const int2 ij = J.tlist[tile];
        a = ij.x * 64;
        b = ij.y * 128;
    }
    const long xb = (long)sg * seg, xe = xb + seg < P ? xb + seg : P;
    const int tid = threadIdx.x;

    const C *src[SPT];
    int kk[SPT], off[SPT];
    bool ok[SPT], cj[SPT];
#pragma unroll
    for (int u = 0; u < SPT; ++u) {
        const int sl = tid + 256 * u;
        src[u] = J.G;
        kk[u] = 0;
        ok[u] = cj[u] = false;
        off[u] = -1;
        int k;
        bool conj_of_value;
        const C *x;
        if (sl < SH64_N) {
            k = a - b / 2 - 64 + sl;
            x = J.G;
            conj_of_value = false;
            off[u] = sh64_off(SH64_UG, sl);
        } else if (sl < 2 * SH64_N) {
            k = a + b / 2 + (sl - SH64_N);
            x = J.F;
            conj_of_value = true;
            off[u] = sh64_off(SH64_UF, sl - SH64_N);
        } else if (sl < 3 * SH64_N) {
            k = b + (sl - 2 * SH64_N); /* <= L/2 */
            x = J.H;
            conj_of_value = false;
            off[u] = sh64_off(SH64_UH, sl - 2 * SH64_N);
        } else {
            continue;
        }
        k %= L;
        if (k < 0) k += L;
        const bool m = 2 * k > L;
        src[u] = x;
        kk[u] = m ? L - k : k;
        cj[u] = m != conj_of_value;
        ok[u] = true;
    }

    C rg[KPX][SPT];
    auto fetch = [&](long x0) {
#pragma unroll
        for (int pp = 0; pp < KPX; ++pp) {
            const long x = x0 + pp;
            const size_t row = (size_t)x * lhp;
#pragma unroll
            for (int u = 0; u < SPT; ++u) {
                C v = make_double2(0.0, 0.0);
                if (x < xe && ok[u]) {
                    v = src[u][row + kk[u]];
                    if (cj[u]) v.y = -v.y;
                }
                rg[pp][u] = v;
            }
        }
    };
    auto stash = [&](char *st) {
#pragma unroll
        for (int pp = 0; pp < KPX; ++pp)
#pragma unroll
            for (int u = 0; u < SPT; ++u)
                if (off[u] >= 0) *(C *)(st + pp * STRIDE + off[u]) = rg[pp][u];
    };

    const int li = tid & 15, lj = tid >> 4;
    const int gbase = sh64_off(SH64_UG, 60 + 4 * li - 4 * lj), fbase = sh64_off(SH64_UF, 4 * li + 4 * lj);
    const int hbase = sh64_off(SH64_UH, 8 * lj);

    double acr[4][8], aci[4][8];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 8; ++v) acr[u][v] = aci[u][v] = 0.0;

    int eg = -1, ef = 0, eh = 0, ek2 = 0;
    if (b == half - 128 && tid < 64) {
        eg = sh64_off(SH64_UG, tid);
        ef = sh64_off(SH64_UF, tid + 64);
        eh = sh64_off(SH64_UH, 128);
        ek2 = (a + tid - L / 4 + L) % L;
    } else if (a == half - 64 && tid >= 64 && tid < 128) {
        const int m = tid - 64;
        eg = sh64_off(SH64_UG, 128 - m);
        ef = sh64_off(SH64_UF, 64 + m);
        eh = sh64_off(SH64_UH, 2 * m);
        ek2 = half - b / 2 - m;
    } else if (a == half - 64 && b == half - 128 && tid == 128) {
        eg = sh64_off(SH64_UG, 64);
        ef = sh64_off(SH64_UF, 128);
        eh = sh64_off(SH64_UH, 128);
        ek2 = L / 4;
    }
    const bool ex = eg >= 0;
    double enr = 0.0, eni = 0.0;

    const long nch = (xe - xb + KPX - 1) / KPX;
    fetch(xb);
    stash(lds0);
    __syncthreads();
    for (long ch = 0; ch < nch; ++ch) {
        if (ch + 1 < nch) fetch(xb + (ch + 1) * KPX);
        const char *st = (ch & 1) ? lds1 : lds0;
#pragma unroll 1
        for (int pp = 0; pp < KPX; ++pp) {
            const char *px = st + pp * STRIDE;
            /* two halves of v, each with only the window values it reads
               (a complex double is one 16-byte LDS read): v < 4 reads G
               3..7, conj F 0..5, H 0..3; v >= 4 reads G 1..5, conj F
               2..7, H 4..7 */
#pragma unroll
            for (int hv2 = 0; hv2 < 2; ++hv2) {
                const int g0 = hv2 ? 1 : 3, f0 = hv2 ? 2 : 0;
                C g[5], f[6], hv[4];
#pragma unroll
                for (int i = 0; i < 5; ++i) g[i] = *(const C *)(px + gbase + (g0 + i) / 4 * BLKB + (g0 + i) % 4 * 16);
#pragma unroll
                for (int i = 0; i < 6; ++i) f[i] = *(const C *)(px + fbase + (f0 + i) / 4 * BLKB + (f0 + i) % 4 * 16);
                load_unit(px + hbase + hv2 * BLKB, hv);
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int vv = 0; vv < 4; ++vv) {
                        const int v = 4 * hv2 + vv;
                        const C gg = g[4 + u - (v >> 1) - g0], ff = f[u + ((v + 1) >> 1) - f0];
                        const double tr = fma(gg.x, hv[vv].x, -gg.y * hv[vv].y);
                        const double ti = fma(gg.x, hv[vv].y, gg.y * hv[vv].x);
                        acr[u][v] = fma(ff.x, tr, fma(-ff.y, ti, acr[u][v]));
                        aci[u][v] = fma(ff.x, ti, fma(ff.y, tr, aci[u][v]));
                    }
            }
            if (ex) {
                const C gg = *(const C *)(px + eg), ff = *(const C *)(px + ef), hh = *(const C *)(px + eh);
                const double tr = fma(gg.x, hh.x, -gg.y * hh.y), ti = fma(gg.x, hh.y, gg.y * hh.x);
                enr = fma(ff.x, tr, fma(-ff.y, ti, enr));
                eni = fma(ff.x, ti, fma(ff.y, tr, eni));
            }
        }
        if (ch + 1 < nch) stash((ch & 1) ? lds0 : lds1);
        __syncthreads();
    }
    if (ex) J.nyq[(size_t)sg * L + ek2] = make_double2(enr, eni);
    C *out = J.part + ((size_t)sg * J.ntile + tile) * (64 * 128);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        C *o = out + (size_t)(4 * li + u) * 128 + 8 * lj;
#pragma unroll
        for (int v = 0; v < 8; ++v) o[v] = make_double2(acr[u][v], aci[u][v]);
    }
}

template <typename R>
__global__ __launch_bounds__(256, LagT<R>::WPS) void k_lagft_tiles(const LagJob<R> *__restrict__ jobs, int njob, long total)
{
    typedef typename LagT<R>::C C;
    constexpr int BU = LagT<R>::BU, BV = LagT<R>::BV, KPX = LagT<R>::KPX;
    constexpr bool FOLD = LagT<R>::FOLD;
    constexpr int TA = 16 * BU, TB = 16 * BV;
    constexpr int VPU = 64 / (int)sizeof(C);           /* values per LDS unit */
    constexpr int UG = TA / VPU, UH = TB / VPU, UF = (TA + TB) / VPU;
    constexpr int PXB = (UG + UH + UF + (FOLD ? 1 : 0)) * BLKB; /* one pixel (+ a unit for H(L/2)) */
    constexpr int NSLOT = 2 * (TA + TB);               /* A, B, then TA + TB of C */
    constexpr int SPT = (NSLOT + (FOLD ? 1 : 0) + 255) / 256; /* slots per thread */
    constexpr int NF = BU + BV - 1, NFU = (NF + VPU - 1) / VPU;
    constexpr int PXBM = sizeof(R) == 4 ? (SH_PXB > PXB ? SH_PXB : PXB) : (SH64_PXB > PXB ? SH64_PXB : PXB);
    __shared__ __align__(16) char lds[2][KPX * PXBM];
    const long it = xcd_item(blockIdx.x, gridDim.x);
    if (it >= total) return;
    const int jb = find_job(jobs, njob, it);
    const long P = jobs[jb].P;
    const int lhp = jobs[jb].lhp, L = jobs[jb].L, seg = jobs[jb].seg, mode = jobs[jb].mode;
    const int nt3 = tiles3<R>(L), ntile = jobs[jb].ntile;
    const long local = it - jobs[jb].item0 + jobs[jb].ibase;
    const int sg = (int)(local / ntile), tile = (int)(local - (long)sg * ntile);
    if constexpr (sizeof(R) == 4) {
        if (mode == 3 || mode == 5) {
            shear_tile<KPX, SH_PXB>(jobs[jb], sg, tile, lds[0], lds[1]);
            return;
        }
    } else {
        if (mode == 3 || mode == 5) {
            shear_tile64<1, SH64_PXB>(jobs[jb], sg, tile, lds[0], lds[1]);
            return;
        }
    }
    int a, b;
    if (mode == 0) {
        a = (tile / nt3) * TA;
        b = (tile % nt3) * TB;
    } else {
        const int2 ij = jobs[jb].tlist[tile];
        a = ij.x * TA;
        b = ij.y * TB;
    }
    const long xb = (long)sg * seg, xe = xb + seg < P ? xb + seg : P;
    const int tid = threadIdx.x;
    const int half = L / 2;
    const bool nyq_col = half % TB != 0;
    const bool nyqx = FOLD && mode != 1 && !nyq_col && b + TB == half;
    /* mode 4: the tile (I - 1, J) left of an unlisted corner (I, J), I + J =
       L / 128, also sums the corner's one needed cell S(a + TA, b), on the
       anti-diagonal k2 + k3 = L (G(a + TA) takes one more staging slot) */
    const bool cornerx = FOLD && mode == 4 && a >= half && a / TA + b / TB == 2 * nt3 - 1 && b >= TB;

    /* the values this thread stages per pixel: the spectrum, its index
       (stored half: k <= L/2, else the conjugate at L - k) and whether the
       staged value is the conjugate of what is read */
    const C *src[SPT];
    int kk[SPT], off[SPT];
    bool ok[SPT], cj[SPT];
    auto at = [&](int u, const C *x, int k, bool conj_of_value) {
        /* value x(k) (or its conjugate) for k in [0, L) */
        const bool m = 2 * k > L;
        src[u] = x;
        kk[u] = m ? L - k : k;
        cj[u] = m != conj_of_value;
        ok[u] = true;
    };
#pragma unroll
    for (int u = 0; u < SPT; ++u) {
        const int sl = tid + 256 * u;
        src[u] = jobs[jb].G;
        kk[u] = 0;
        ok[u] = cj[u] = false;
        off[u] = -1;
        int r, base;
        if (sl < TA) {
            const int p = a + sl;
            if (p < L) at(u, mode == 1 ? jobs[jb].F : jobs[jb].G, p, mode == 1);
            r = sl;
            base = 0;
        } else if (sl < TA + TB) {
            const int q = b + sl - TA;
            if (mode == 1) at(u, jobs[jb].F, q, true);
            else if (2 * q < L || (nyq_col && q == half)) at(u, jobs[jb].H, q, false);
            r = sl - TA;
            base = UG;
        } else if (sl < NSLOT) {
            const int s1 = (a + b + sl - TA - TB) % L;
            if (mode == 1) at(u, jobs[jb].H, s1, false);
            else at(u, jobs[jb].F, s1, true);
            r = sl - TA - TB;
            base = UG + UH;
        } else if (FOLD && sl == NSLOT && nyqx) {
            at(u, jobs[jb].H, half, false); /* H(L/2), in the extra unit */
            r = 0;
            base = UG + UH + UF;
        } else if (FOLD && sl == NSLOT + 1 && cornerx) {
            at(u, jobs[jb].G, a + TA, false); /* G(a + TA), beside it */
            r = 1;
            base = UG + UH + UF;
        } else {
            continue;
        }
        off[u] = (base + r / VPU) * BLKB + (r % VPU) * (int)sizeof(C);
    }

    C rg[KPX][SPT];
    auto fetch = [&](long x0) {
#pragma unroll
        for (int pp = 0; pp < KPX; ++pp) {
            const long x = x0 + pp;
            const size_t row = (size_t)x * lhp;
#pragma unroll
            for (int u = 0; u < SPT; ++u) {
                C v;
                v.x = v.y = (R)0;
                if (x < xe && ok[u]) {
                    v = src[u][row + kk[u]];
                    if (cj[u]) v.y = -v.y;
                }
                rg[pp][u] = v;
            }
        }
    };
    auto stash = [&](char *st) {
#pragma unroll
        for (int pp = 0; pp < KPX; ++pp)
#pragma unroll
            for (int u = 0; u < SPT; ++u)
                if (off[u] >= 0) *(C *)(st + pp * PXB + off[u]) = rg[pp][u];
    };

    const int li = tid >> 4, lj = tid & 15;
    const int ug = (BU * li) / VPU, uh = UG + (BV * lj) / VPU, uf = UG + UH + (BU * li + BV * lj) / VPU;
    R acr[BU][BV], aci[BU][BV];
#pragma unroll
    for (int u = 0; u < BU; ++u)
#pragma unroll
        for (int v = 0; v < BV; ++v) acr[u][v] = aci[u][v] = (R)0;
    /* k3 = L/2 column of k2 = a + tid (tid < TA, whole waves); thread TA
       the mode 4 corner cell: G(a + TA), H(b), conj F(a + b + TA) */
    const bool nx = FOLD && nyqx && tid < TA;
    const bool cx4 = FOLD && cornerx && tid == TA;
    const int offg = cx4 ? (UG + UH + UF) * BLKB + (int)sizeof(C) : (tid / VPU) * BLKB + (tid % VPU) * (int)sizeof(C);
    const int offf = cx4 ? (UG + UH + TA / VPU) * BLKB + (TA % VPU) * (int)sizeof(C)
                         : (UG + UH + (tid + TB) / VPU) * BLKB + ((tid + TB) % VPU) * (int)sizeof(C);
    const int offh = cx4 ? UG * BLKB : (UG + UH + UF) * BLKB;
    const bool ex = nx || cx4;
    R anr = (R)0, ani = (R)0;

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
            C g[(BU + VPU - 1) / VPU * VPU], hv[(BV + VPU - 1) / VPU * VPU], f[NFU * VPU];
#pragma unroll
            for (int q = 0; q < (BU + VPU - 1) / VPU; ++q) load_unit(px + (ug + q) * BLKB, g + q * VPU);
#pragma unroll
            for (int q = 0; q < (BV + VPU - 1) / VPU; ++q) load_unit(px + (uh + q) * BLKB, hv + q * VPU);
#pragma unroll
            for (int q = 0; q < NFU; ++q) load_unit(px + (uf + q) * BLKB, f + q * VPU);
#pragma unroll
            for (int u = 0; u < BU; ++u)
#pragma unroll
                for (int v = 0; v < BV; ++v) {
                    /* t = G H, acc += conj(F) t */
                    const R tr = fma(g[u].x, hv[v].x, -g[u].y * hv[v].y);
                    const R ti = fma(g[u].x, hv[v].y, g[u].y * hv[v].x);
                    acr[u][v] = fma(f[u + v].x, tr, fma(-f[u + v].y, ti, acr[u][v]));
                    aci[u][v] = fma(f[u + v].x, ti, fma(f[u + v].y, tr, aci[u][v]));
                }
            if (FOLD && ex) {
                const C gg = *(const C *)(px + offg), ff = *(const C *)(px + offf);
                const C hh = *(const C *)(px + offh);
                const R tr = fma(gg.x, hh.x, -gg.y * hh.y), ti = fma(gg.x, hh.y, gg.y * hh.x);
                anr = fma(ff.x, tr, fma(-ff.y, ti, anr));
                ani = fma(ff.x, ti, fma(ff.y, tr, ani));
            }
        }
        if (ch + 1 < nch) stash(lds[(ch + 1) & 1]);
        __syncthreads();
    }
    if (FOLD && ex && (cx4 || a + tid < L)) {
        C w;
        w.x = anr;
        w.y = ani;
        jobs[jb].nyq[(size_t)sg * L + (cx4 ? a + TA : a + tid)] = w;
    }
    C *out = jobs[jb].part + ((size_t)sg * ntile + tile) * (TA * TB);
#pragma unroll
    for (int u = 0; u < BU; ++u) {
        C *o = out + (size_t)(BU * li + u) * TB + BV * lj;
#pragma unroll
        for (int v = 0; v < BV; ++v) {
            C w;
            w.x = acr[u][v];
            w.y = aci[u][v];
            o[v] = w;
        }
    }
}

/* --------------------------------------------------------------- nyquist */
/* S(k2, L/2) = sum_x conj F(k2 + L/2) G(k2) H(L/2) over one segment per
   block, for jobs with nyq_sep.  Every k2: thread t owns k2 = t + 256 q.
   A k2 list (pair-tile jobs, whose listed tiles hold the rest of the
   column): thread t sums pixels xb + t + 256 m for every listed k2, then a
   fixed-order tree over the block. */
template <typename R>
__global__ __launch_bounds__(256) void k_lagft_nyq(const LagJob<R> *__restrict__ jobs, int njob, long total)
{
    typedef typename LagT<R>::C C;
    const long it = xcd_item(blockIdx.x, gridDim.x);
    if (it >= total) return;
    const int jb = find_job(jobs, njob, it, true);
    const C *__restrict__ F = jobs[jb].F, *__restrict__ G = jobs[jb].G, *__restrict__ H = jobs[jb].H;
    const long P = jobs[jb].P;
    const int lhp = jobs[jb].lhp, L = jobs[jb].L, seg = jobs[jb].seg, h = L / 2;
    const int sg = (int)(it - jobs[jb].nitem0);
    const long xb = (long)sg * seg, xe = xb + seg < P ? xb + seg : P;
    if constexpr (sizeof(R) == 4) if (jobs[jb].nyql) {
        const int *__restrict__ kl = jobs[jb].nyql;
        const int nk = jobs[jb].nnyq;
        R lr[NYQ_LIST_MAX], li[NYQ_LIST_MAX];
#pragma unroll
        for (int q = 0; q < NYQ_LIST_MAX; ++q) lr[q] = li[q] = (R)0;
        for (long x = xb + threadIdx.x; x < xe; x += 256) {
            const size_t row = (size_t)x * lhp;
            const C hv = H[row + h];
#pragma unroll
            for (int q = 0; q < NYQ_LIST_MAX; ++q) {
                if (q < nk) {
                    const int k2 = kl[q];
                    const C g = spec_at(G, row, k2, L);
                    int s2 = k2 + h;
                    if (s2 >= L) s2 -= L;
                    const C f = spec_at(F, row, s2, L);
                    const R tr = fma(g.x, hv.x, -g.y * hv.y), ti = fma(g.x, hv.y, g.y * hv.x);
                    lr[q] = fma(f.x, tr, fma(f.y, ti, lr[q]));
                    li[q] = fma(f.x, ti, fma(-f.y, tr, li[q]));
                }
            }
        }
        __shared__ R red[2][256];
        for (int q = 0; q < nk; ++q) {
            R vr = (R)0, vi = (R)0;
#pragma unroll
            for (int u = 0; u < NYQ_LIST_MAX; ++u)
                if (u == q) { vr = lr[u]; vi = li[u]; }
            red[0][threadIdx.x] = vr;
            red[1][threadIdx.x] = vi;
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if ((int)threadIdx.x < w) {
                    red[0][threadIdx.x] += red[0][threadIdx.x + w];
                    red[1][threadIdx.x] += red[1][threadIdx.x + w];
                }
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                C w;
                w.x = red[0][0];
                w.y = red[1][0];
                jobs[jb].nyq[(size_t)sg * L + kl[q]] = w;
            }
            __syncthreads();
        }
        return;
    }
    constexpr int NQ = MAXL / 256;
    R acr[NQ], aci[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acr[q] = aci[q] = (R)0;
    for (long x = xb; x < xe; ++x) {
        const size_t row = (size_t)x * lhp;
        const C hv = H[row + h];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int k2 = threadIdx.x + 256 * q;
            if (256 * q < L && k2 < L) {
                const C g = spec_at(G, row, k2, L);
                int s = k2 + h;
                if (s >= L) s -= L;
                const C f = spec_at(F, row, s, L);
                const R tr = fma(g.x, hv.x, -g.y * hv.y), ti = fma(g.x, hv.y, g.y * hv.x);
                acr[q] = fma(f.x, tr, fma(f.y, ti, acr[q]));
                aci[q] = fma(f.x, ti, fma(-f.y, tr, aci[q]));
            }
        }
    }
    C *o = jobs[jb].nyq + (size_t)sg * L;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int k2 = threadIdx.x + 256 * q;
        if (256 * q < L && k2 < L) {
            C w;
            w.x = acr[q];
            w.y = aci[q];
            o[k2] = w;
        }
    }
}

/* ----------------------------------------------------------------- reduce */
/* St[k3][k2] = sum over segments of the tile partials, FP64 in segment
   order, for k3 < L/2 (and k3 = L/2 when it is a column of the last,
   partial tile); 32 x 32 (k2, k3) blocks transposed through LDS.
   MODE 1: the value sits at (i, j) = (-k2, k2 + k3) mod L, in listed
   tile tmap[I][J] of (i, j) or, when I > J, of (j, i) (symmetric).
   MODE 2: listed tile tmap[a][b] of (k2, k3), or of (k3, k2) for a tile of
   [0, L/2)^2 below its diagonal (TA = TB).
   MODE 4: as MODE 2 on [0, L/2)^2; for k2 >= L/2 the cells past the
   anti-diagonal (k2 + k3 > L) from their conjugate mirror (L - k3, L - k2),
   the anti-diagonal corners of unlisted tiles and the k3 = L/2 column from
   the column block.
   MODE 5 (r1 == r2 == r3): S is the same for every permutation of (k1, k2,
   k3), k1 = -k2 - k3, and conjugate under negation; the cell's orbit meets
   the principal domain {0 <= k3 <= k2, 2 k2 + k3 <= L} (the bispectrum's
   inner and outer triangles), whose (t, k3) lie in the listed sheared tiles
   or on the fixed-point column t = L/2.
   MODE 3: the sheared tiles' (t, k3), t = k2 + floor(k3/2), or its mirror
   (-1 - t for odd k3, -t for even) when t is past L/2; the fixed points t =
   L/2 and the row k3 = L/2 from the column block after the partials. */
template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_lagft_reduce(const typename LagT<R>::C *__restrict__ part, int nseg, int ntile,
                                                      int L, const int *__restrict__ tmap, double2 *__restrict__ St)
{
    typedef typename LagT<R>::C C;
    constexpr int TA = 16 * LagT<R>::BU, TB = 16 * LagT<R>::BV;
    __shared__ double2 tr[32][33];
    const int nt3 = tiles3<R>(L);
    const size_t tsz = (size_t)ntile * TA * TB;
    const int X = blockIdx.x * 32, Y = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int k3 = Y + tx, half = L / 2;
    /* k3 = L/2 is a tile column (mode 0, TB does not divide L/2), or in mode
       1 a diagonal of the listed tiles (k_lagft_reduce_nyq then writes the
       k2 that no listed tile holds) */
    const bool col = MODE == 1 || MODE >= 3 || half % TB != 0;
    const C *nyq = part + (size_t)nseg * tsz;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k2 = X + ty + 8 * q;
        double sx = 0.0, sy = 0.0;
        if (k2 < L && (k3 < half || (col && k3 == half))) {
            long off = 0;
            int ni = -1;
            bool cj = false;
            if (MODE == 5) {
                const int k1 = (2 * L - k2 - k3) % L;
                const int cs[6][2] = {{k2, k3}, {k3, k2}, {k1, k3}, {k3, k1}, {k1, k2}, {k2, k1}};
                int p = -1, q = -1;
                for (int ng = 0; ng < 2 && p < 0; ++ng)
                    for (int c = 0; c < 6 && p < 0; ++c) {
                        int x = cs[c][0], y = cs[c][1];
                        if (ng) {
                            x = x ? L - x : 0;
                            y = y ? L - y : 0;
                        }
                        if (y <= x && y <= L - 2 * x) {
                            p = x;
                            q = y;
                            cj = ng;
                        }
                    }
                off = -1;
                if (p >= 0) {
                    const int t = p + (q >> 1);
                    if (t == half) {
                        ni = half - (q >> 1);
                    } else {
                        const int tl = tmap[(t / TA) * (half / TB) + q / TB];
                        if (tl >= 0) off = (long)tl * (TA * TB) + (t % TA) * TB + (q % TB);
                    }
                }
            } else if (MODE == 4) {
                int p2 = k2, p3 = k3;
                if (p2 >= half && p2 + p3 > L) {
                    const int q2 = L - p3;
                    p3 = L - p2;
                    p2 = q2;
                    cj = true;
                }
                off = -1;
                if (p3 == half) {
                    ni = p2;
                } else {
                    int a = p2 / TA, b = p3 / TB;
                    if (p2 < half && a > b) {
                        const int t = p2; p2 = p3; p3 = t;
                        a = p2 / TA;
                        b = p3 / TB;
                    }
                    const int tl = tmap[a * nt3 + b];
                    if (tl < 0) ni = p2; /* a corner: p2 + p3 = L */
                    else off = (long)tl * (TA * TB) + (p2 - a * TA) * TB + (p3 - b * TB);
                }
            } else if (MODE == 3) {
                const int fl = k3 >> 1;
                int t = (k2 + fl) % L;
                off = -1;
                if (k3 == half) {
                    ni = t <= half ? k2 : (3 * (L / 4) - t + L) % L;
                } else {
                    if (t > half || (t == half && (k3 & 1))) t = (k3 & 1) ? L - 1 - t : L - t;
                    if (t == half) ni = half - fl;
                    else off = (long)((t / TA) * (L / 256) + k3 / TB) * (TA * TB) + (t % TA) * TB + (k3 % TB);
                }
            } else if (MODE == 2) {
                int p2 = k2, p3 = k3, a = k2 / TA, b = k3 / TB;
                if (a > b && (a + 1) * TA <= half) {
                    p2 = k3;
                    p3 = k2;
                    a = p2 / TA;
                    b = p3 / TB;
                }
                const int tl = tmap[a * nt3 + b];
                off = tl < 0 ? -1 : (long)tl * (TA * TB) + (p2 - a * TA) * TB + (p3 - b * TB);
            } else if (MODE == 1) {
                int i = k2 == 0 ? 0 : L - k2, j = k2 + k3;
                if (j >= L) j -= L;
                int I = i / TA, J = j / TB;
                if (I > J) {
                    const int t = i; i = j; j = t;
                    const int T = I; I = J; J = T;
                }
                const int tl = tmap[I * (L / TA) + J];
                off = tl < 0 ? -1 : (long)tl * (TA * TB) + (i - I * TA) * TB + (j - J * TB);
            } else {
                off = (long)((k2 / TA) * nt3 + k3 / TB) * (TA * TB) + (k2 % TA) * TB + (k3 % TB);
            }
            for (int sg = 0; sg < nseg && off >= 0; ++sg) {
                const C v = part[sg * tsz + off];
                sx += (double)v.x;
                sy += (double)v.y;
            }
            for (int sg = 0; sg < nseg && ni >= 0; ++sg) {
                const C v = nyq[(size_t)sg * L + ni];
                sx += (double)v.x;
                sy += (double)v.y;
            }
            if (cj) sy = -sy;
        }
        tr[ty + 8 * q][tx] = make_double2(sx, sy);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k3w = Y + ty + 8 * q, k2w = X + tx;
        if ((k3w < half || (col && k3w == half)) && k2w < L) St[(size_t)k3w * L + k2w] = tr[tx][ty + 8 * q];
    }
}

/* St[L/2][k2] = sum over segments of the k3 = L/2 partials (FP64, in
   order), for every k2 or the nk listed ones (kl non-null) */
template <typename R>
__global__ __launch_bounds__(256) void k_lagft_reduce_nyq(const typename LagT<R>::C *__restrict__ nyq, int nseg, int L,
                                                          const int *__restrict__ kl, int nk, double2 *__restrict__ St)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (kl ? idx >= nk : idx >= L) return;
    const int k2 = kl ? kl[idx] : idx;
    double sx = 0.0, sy = 0.0;
    for (int sg = 0; sg < nseg; ++sg) {
        const typename LagT<R>::C v = nyq[(size_t)sg * L + k2];
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
   row wb in the imaginary part (DC and Nyquist terms: real parts only).
   Data (snapP == 0): times gscale 2^(E1+E2+E3-21)/L^2 (exponents null:
   gscale/L^2), into out32 and out64 (normalisation: gscale = 1); gscale = (Ng/n)^2 carries a sum over
   the n^2 pixels of a compute grid to the reference's Ng^2.  Normalisation
   (snapP = n^2, outP = Ng^2): the sum divided by L^2 is rounded to the
   nearest 1/8 of n^2, and that many eighths of Ng^2 go to out64 (and
   out32); the largest distance to the grid, in units of n^2, goes to
   margin (float bits).  Cells without triplets are exact zeros. */
__global__ __launch_bounds__(64) void k_lagft_rows(const double2 *__restrict__ St, FftPlan pl,
                                                   const double2 *__restrict__ tw, int nc10,
                                                   const int *__restrict__ cnt, const int *__restrict__ ex1,
                                                   const int *__restrict__ ex2, const int *__restrict__ ex3,
                                                   double snapP, double outP, double gscale,
                                                   float *__restrict__ out32, double *__restrict__ out64,
                                                   unsigned *__restrict__ margin)
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
    const double scale = (ex1 ? ldexp(gscale, *ex1 + *ex2 + *ex3 - 3 * AMB_SET_EXP) : gscale) / ((double)L * (double)L);
    float mg = 0.f;
    auto put = [&](size_t cell, double v) {
        if (!cnt[cell]) v = 0.0;
        if (snapP > 0.0) {
            const double q = rint(8.0 * v / snapP) / 8.0;
            mg = fmaxf(mg, (float)fabs(v / snapP - q));
            v = outP * q;
        }
        out32[cell] = (float)v;
        if (out64) out64[cell] = v;
    };
    for (int w3 = threadIdx.x; w3 < W; w3 += 64) {
        const int idx = ((w3 - nc10) % L + L) % L;
        const double2 z = res[idx];
        put((size_t)wa * W + w3, z.x * scale);
        if (wb != wa) put((size_t)wb * W + w3, z.y * scale);
    }
    if (margin) {
        for (int o = 32; o > 0; o >>= 1) mg = fmaxf(mg, __shfl_xor(mg, o));
        if (threadIdx.x == 0) atomicMax(margin, __float_as_uint(mg));
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
    /* digit-reversed sequence positions of the active planes */
    FftPlan pl;
    fftc::make_plan(L, &pl);
    std::vector<int> hp(na);
    for (int i = 0; i < na; ++i) hp[i] = fftc::digit_reversed(cb.act[i] - cb.act.front(), pl);
    void *dp = nullptr;
    if ((rc = dev_alloc(c, &dp, sizeof(int) * na, "lag positions"))) {
        dev_free(c, d, sizeof(int) * cells);
        return rc;
    }
    e = hipMemcpy(dp, hp.data(), sizeof(int) * na, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        dev_free(c, d, sizeof(int) * cells);
        dev_free(c, dp, sizeof(int) * na);
        return fail(c, AMB_EHIP, "lag positions: %s", hipGetErrorString(e));
    }
    cb.d_lagpos = (int *)dp;
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
    if (cb.d_lagpos) dev_free(c, cb.d_lagpos, sizeof(int) * cb.act.size());
    cb.d_lagcnt = nullptr;
    cb.d_lagpos = nullptr;
    cb.h_lagcnt.clear();
    cb.lag_L = cb.lag_lhp = 0;
}

size_t amb::lag_spectra_bytes(const Cube &cb, long ng, bool fp64)
{
    return (fp64 ? sizeof(double2) : sizeof(float2)) * (size_t)ng * ng * cb.lag_lhp + AMB_SET_SLOT_BYTES;
}

/* kind 9: FP32 spectra of the data set op.src (kind 11) times 2^(7 - E),
   its [max, E] words after the spectra; kind 10: FP64 spectra of the FP64
   sampling images op.src */
hipError_t amb::lag_launch_spectra(amb_ctx *c, hipStream_t s, const BuildOp &op)
{
    const Cube &cb = c->cubes[op.cube];
    const long ng = op.g.ng, P = ng * ng, pitch = plane_pitch(ng);
    const int L = cb.lag_L;
    const long npair = (P + 1) / 2;
    /* pairs per team: up to 64 pixels of every plane per workgroup (FP32),
       fewer while that leaves under 2048 workgroups; FP64: 4 */
    const int nt9 = spectra_teams(L);
    const int ppt = op.kind == 9 ? (int)std::max(1L, std::min<long>(32 / nt9, npair / (2048L * nt9))) : 4;
    const FftPlan &pl = c->plans[L];
    if (op.kind == 9) {
        auto itw = c->tw.find(L);
        if (itw == c->tw.end()) return hipErrorInvalidValue;
        const size_t setw = (size_t)pitch * cb.nact();
        const unsigned *slot_in = (const unsigned *)((const float *)op.src + setw);
        unsigned *slot_out = (unsigned *)((float2 *)op.dst + (size_t)P * cb.lag_lhp);
        /* the in-place transform's radices: 2, 3, 4, 5 and 8 */
        bool smooth5 = true;
        for (int s2 = 0; s2 < pl.nst; ++s2) smooth5 = smooth5 && (pl.rad[s2] <= 5 || pl.rad[s2] == 8);
        /* four team buffers of L + 8 values (the bank swizzle's padding) */
        const size_t lds_blk = sizeof(float2) * 4 * (size_t)(L + 8) + sizeof(int) * (size_t)cb.nact();
        if (cb.nact() <= SPB_PLANES && smooth5 && lds_blk <= 65536) {
            const long nblk = ((P + 31) / 32 + 7) / 8 * 8;
            hipLaunchKernelGGL(k_chan_spectra_blk<false>, dim3((unsigned)nblk), dim3(256), lds_blk, s,
                               (const float *)op.src, pitch, (const int *)cb.d_lagpos, cb.nact(), P, pl, itw->second,
                               cb.lag_lhp, (float2 *)op.dst, slot_in, slot_out);
            return hipGetLastError();
        }
        const int nteam = spectra_teams(L);
        const long per_blk = (long)nteam * ppt;
        const long nblk = ((npair + per_blk - 1) / per_blk + 7) / 8 * 8;
        const size_t lds = sizeof(float2) * (size_t)L * (1 + 2 * nteam);
        hipLaunchKernelGGL((k_chan_spectra<float, true, true>), dim3((unsigned)nblk), dim3(64 * nteam), lds, s,
                           (const float *)op.src, pitch, cb.d_inv, (int)cb.nc, cb.act.front(), P, pl, itw->second,
                           cb.lag_lhp, ppt, (float2 *)op.dst, slot_in, slot_out);
        return hipGetLastError();
    }
    auto itw = c->tw64.find(L);
    if (itw == c->tw64.end()) return hipErrorInvalidValue;
    const long nblk = ((npair + ppt - 1) / ppt + 7) / 8 * 8;
    const size_t lds = sizeof(double2) * 2 * (size_t)L;
    hipLaunchKernelGGL((k_chan_spectra<double, false, false>), dim3((unsigned)nblk), dim3(64), lds, s,
                       (const double *)op.src, pitch, cb.d_inv, (int)cb.nc, cb.act.front(), P, pl, itw->second,
                       cb.lag_lhp, ppt, (double2 *)op.dst, nullptr, nullptr);
    return hipGetLastError();
}

namespace {

/* one triangle's stage of a wave: its data sums (FP32) or its
   normalisation (FP64) */
struct LJ {
    int q;          /* request */
    bool norm;      /* FP64 normalisation job */
    Cube *cb;
    long ng, P;     /* Ng(r1); pixels of the compute grid */
    long ngc, Pref; /* the compute grid; Ng(r1)^2 */
    int nseg;
    int mode, ntile, nyq_sep;
    const int2 *tlist; /* mode 1 */
    const int *tmap;   /* mode 1 */
    const int *nyql;   /* mode 1: k2 of the k3 = L/2 column left to k_lagft_nyq (null: all) */
    int nnyq;
    void *X[3];     /* spectra of r1, r2, r3 */
    size_t part, nyq, out, out64;
};

/* mode 1 tiles of a transform length L (128 | L): (I, J), I <= J, on the
   tile diagonals m = (I + J) mod nt that hold cells with (i + j) mod L in
   [0, L/2): m in [0, (64 nt - 1) / 128] and m = nt - 1 (a cell of tile
   diagonal m has i + j in [128 m, 128 m + 254]) */
std::vector<int2> pair_tiles(int L)
{
    const int nt = L / 128, mmax = (64 * nt - 1) / 128;
    std::vector<int2> t;
    for (int I = 0; I < nt; ++I)
        for (int J = I; J < nt; ++J) {
            const int m = (I + J) % nt;
            if (m <= mmax || m == nt - 1) t.push_back(make_int2(I, J));
        }
    return t;
}

template <typename R>
size_t job_part_bytes(const LJ &j, int L)
{
    const size_t cb = sizeof(typename LagT<R>::C);
    return cb * (size_t)j.nseg * ((size_t)j.ntile * tile_a<R>() * tile_b<R>() + L);
}

template <typename R>
void fill_job(LagJob<R> &g, const LJ &j, char *partbase, long *items, long *nitems)
{
    typedef typename LagT<R>::C C;
    const int L = j.cb->lag_L;
    g.F = (const C *)j.X[0];
    g.G = (const C *)j.X[1];
    g.H = (const C *)j.X[2];
    g.ex1 = g.ex2 = g.ex3 = nullptr;
    if (!j.norm) {
        const int *ex[3];
        for (int k = 0; k < 3; ++k) ex[k] = (const int *)((const C *)j.X[k] + (size_t)j.P * j.cb->lag_lhp) + 1;
        g.ex1 = ex[0];
        g.ex2 = ex[1];
        g.ex3 = ex[2];
    }
    g.part = (C *)(partbase + j.part);
    g.nyq = (C *)(partbase + j.nyq);
    g.tlist = j.tlist;
    g.nyql = j.nyql;
    g.nnyq = j.nnyq;
    g.P = j.P;
    g.lhp = j.cb->lag_lhp;
    g.L = L;
    g.nseg = j.nseg;
    g.seg = SEG;
    g.mode = j.mode;
    g.ntile = j.ntile;
    g.nyq_sep = j.nyq_sep;
    g.item0 = *items;
    g.nitem0 = *nitems;
    g.ibase = 0;
    *items += (long)j.nseg * j.ntile;
    if (j.nyq_sep) *nitems += j.nseg;
}

template <typename R>
hipError_t launch_products(hipStream_t s, const LagJob<R> *d, int n, long items, long nitems)
{
    hipLaunchKernelGGL(k_lagft_tiles<R>, dim3((unsigned)((items + 7) / 8 * 8)), dim3(256), 0, s, d, n, items);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nitems == 0) return e;
    hipLaunchKernelGGL(k_lagft_nyq<R>, dim3((unsigned)((nitems + 7) / 8 * 8)), dim3(256), 0, s, d, n, nitems);
    return hipGetLastError();
}

/* share of the half spectrum's (k2, k3) whose sums are distinct under the
   triangle's symmetries */
double lag_distinct_r(int r1, int r2, int r3)
{
    return r1 == r2 && r2 == r3 ? 1.0 / 3.0 : (r1 == r2 || r2 == r3) ? 0.5 : 1.0;
}
double lag_distinct(const amb_tri_req &q) { return lag_distinct_r(q.r1, q.r2, q.r3); }

/* sums outside whole tiles: the k3 = L/2 column (or the k2 of it that a
   pair-tile job leaves to k_lagft_nyq); sheared tiles: the row k3 = L/2
   (t <= L/2) and the fixed points t = L/2 of the even rows */
double edge_cells(int mode, int L, const int *nyql, int nnyq)
{
    if (mode == 3) return (double)(L / 2 + 1 + L / 4);
    if (mode == 5) return (double)nnyq; /* the fixed points the listed last-t tiles sum */
    if (mode == 4) return (double)((L / 256) * 128 + 128 + (L / 256 - 1)); /* the folded column of the listed last-k3 tiles, corners */
    return (double)(nyql ? nnyq : L);
}



/* tiling of a triangle: (k2, k3) tiles, or for FP32 jobs the symmetric
   (i, j) tiles of r1 == r2 or the listed (k2, k3) tiles of r2 == r3 where
   they are fewer (tables built once per L); whether k_lagft_nyq sums the
   k3 = L/2 column (pair tiles, and FP64 where the column is not a tile
   column) */
int choose_tiles(amb_ctx *c, int L, bool fp64, bool r12, bool r23, int *mode, int *ntile, int *nyq_sep,
                 const int2 **tlist, const int **tmap, const int **nyql, int *nnyq)
{
    *mode = 0;
    *ntile = fp64 ? tiles2<double>(L) * tiles3<double>(L) : tiles2<float>(L) * tiles3<float>(L);
    *tlist = nullptr;
    *tmap = nullptr;
    *nyql = nullptr;
    *nnyq = 0;
    if (r12 && r23 && !c->no_sym && !c->no_shear && L % 256 == 0) {
        /* r1 == r2 == r3: the sheared tiles (t width 128 in FP32, 64 in
           FP64; k3 width 128) that hold a cell of the principal domain
           {0 <= k3 <= k2, 2 k2 + k3 <= L}, and the last-t tiles whose fixed
           points t = L/2 lie in it; listed once per L and precision */
        const int key = fp64 ? -L : L;
        auto it = c->lag_equi_tab.find(key);
        if (it == c->lag_equi_tab.end()) {
            const int half = L / 2, ta = fp64 ? 64 : 128, nta = half / ta, ntb = half / 128;
            auto indom = [&](int p, int q) { return q >= 0 && q <= p && q <= L - 2 * p; };
            std::vector<int> tm((size_t)nta * ntb, -1);
            for (int k3 = 0; k3 < half; ++k3)
                for (int k2 = 0; k2 <= half; ++k2) {
                    const int t = k2 + (k3 >> 1);
                    if (t < half && indom(k2, k3)) tm[(size_t)(t / ta) * ntb + k3 / 128] = 1;
                }
            for (int k3 = 0; k3 < half; k3 += 2)
                if (indom(half - k3 / 2, k3)) tm[(size_t)(nta - 1) * ntb + k3 / 128] = 1;
            std::vector<int2> tl;
            int fixed = 0;
            for (int A = 0; A < nta; ++A)
                for (int B = 0; B < ntb; ++B)
                    if (tm[(size_t)A * ntb + B] > 0) {
                        tm[(size_t)A * ntb + B] = (int)tl.size();
                        tl.push_back(make_int2(A, B));
                        if (A == nta - 1) fixed += 64;
                    }
            void *d = nullptr;
            const size_t tb = sizeof(int2) * tl.size(), mb = sizeof(int) * tm.size();
            const int rc = dev_alloc(c, &d, tb + mb, "lag-Fourier tile tables");
            if (rc) return rc;
            HIPCHK(c, hipMemcpy(d, tl.data(), tb, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy((char *)d + tb, tm.data(), mb, hipMemcpyHostToDevice));
            it = c->lag_equi_tab.emplace(key, std::make_tuple(d, (int)tl.size(), tb + mb, fixed)).first;
        }
        *mode = 5;
        *ntile = std::get<1>(it->second);
        *tlist = (const int2 *)std::get<0>(it->second);
        *tmap = (const int *)((const char *)std::get<0>(it->second) + sizeof(int2) * *ntile);
        *nnyq = std::get<3>(it->second); /* fixed-point cells, for the issued-work count */
        *nyq_sep = 0;
        return AMB_OK;
    }
    if (r12 && !c->no_sym && !c->no_shear && L % 256 == 0) {
        /* sheared (t, k3) tiles: exactly the distinct sums, edges folded
           (FP32 128 x 128, FP64 64 x 128 tiles) */
        *mode = 3;
        *ntile = (L / 2 / tile_a<float>()) * (L / 256);
        if (fp64) *ntile = (L / 2 / tile_a<double>()) * (L / 256);
        *nyq_sep = 0;
        return AMB_OK;
    }
    if (!fp64 && r12 && !c->no_sym && L % 128 == 0) {
        auto it = c->lag_pairs_tab.find(L);
        if (it == c->lag_pairs_tab.end()) {
            /* tiles, the tile map, and the k2 whose k3 = L/2 cell, (i, j) =
               (-k2, k2 + L/2), lies in no listed tile */
            const std::vector<int2> tl = pair_tiles(L);
            const int nt = L / 128;
            std::vector<int> tm((size_t)nt * nt, -1), miss;
            for (size_t t = 0; t < tl.size(); ++t) tm[(size_t)tl[t].x * nt + tl[t].y] = (int)t;
            for (int k2 = 0; k2 < L; ++k2) {
                const int i = (L - k2) % L, j = (k2 + L / 2) % L;
                const int I = std::min(i, j) / 128, J = std::max(i, j) / 128;
                if (tm[(size_t)I * nt + J] < 0) miss.push_back(k2);
            }
            const int nm = (int)miss.size() <= NYQ_LIST_MAX ? (int)miss.size() : -1; /* -1: every k2 */
            void *d = nullptr;
            const size_t tb = sizeof(int2) * tl.size(), mb = sizeof(int) * tm.size();
            const size_t lb = sizeof(int) * std::max(nm, 0);
            const int rc = dev_alloc(c, &d, tb + mb + lb, "lag-Fourier tile tables");
            if (rc) return rc;
            HIPCHK(c, hipMemcpy(d, tl.data(), tb, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy((char *)d + tb, tm.data(), mb, hipMemcpyHostToDevice));
            if (nm > 0) HIPCHK(c, hipMemcpy((char *)d + tb + mb, miss.data(), lb, hipMemcpyHostToDevice));
            it = c->lag_pairs_tab.emplace(L, std::make_tuple(d, (int)tl.size(), tb + mb + lb, nm)).first;
        }
        const int n1 = std::get<1>(it->second);
        if (n1 < *ntile) {
            const int nt = L / 128;
            *mode = 1;
            *ntile = n1;
            *tlist = (const int2 *)std::get<0>(it->second);
            *tmap = (const int *)((const char *)std::get<0>(it->second) + sizeof(int2) * n1);
            const int nm = std::get<3>(it->second);
            if (nm >= 0) {
                *nyql = *tmap + (size_t)nt * nt;
                *nnyq = nm;
            }
        }
    }
    if (!fp64 && !r12 && r23 && !c->no_sym && L % 256 == 0) {
        /* (k2, k3) tiles: on [0, L/2)^2 those on and above its diagonal, for
           k2 >= L/2 those with a cell on or before the anti-diagonal k2 + k3
           = L other than a corner (I + J < L/128); tables keyed by -L */
        auto it = c->lag_r23_tab.find(-L);
        if (it == c->lag_r23_tab.end()) {
            const int n2 = tiles2<float>(L), n3 = tiles3<float>(L);
            std::vector<int2> tl;
            std::vector<int> tm((size_t)n2 * n3, -1);
            for (int a = 0; a < n2; ++a)
                for (int b = 0; b < n3; ++b)
                    if (a < n3 ? a <= b : a + b <= 2 * n3 - 1) {
                        tm[(size_t)a * n3 + b] = (int)tl.size();
                        tl.push_back(make_int2(a, b));
                    }
            void *d = nullptr;
            const size_t tb = sizeof(int2) * tl.size(), mb = sizeof(int) * tm.size();
            const int rc = dev_alloc(c, &d, tb + mb, "lag-Fourier tile tables");
            if (rc) return rc;
            HIPCHK(c, hipMemcpy(d, tl.data(), tb, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy((char *)d + tb, tm.data(), mb, hipMemcpyHostToDevice));
            it = c->lag_r23_tab.emplace(-L, std::make_tuple(d, (int)tl.size(), tb + mb)).first;
        }
        *mode = 4;
        *ntile = std::get<1>(it->second);
        *tlist = (const int2 *)std::get<0>(it->second);
        *tmap = (const int *)((const char *)std::get<0>(it->second) + sizeof(int2) * *ntile);
        *nyq_sep = 0;
        return AMB_OK;
    }
    if (!fp64 && !r12 && r23 && !c->no_sym) {
        /* (k2, k3) tiles without those of [0, L/2)^2 below its diagonal */
        auto it = c->lag_r23_tab.find(L);
        if (it == c->lag_r23_tab.end()) {
            const int n2 = tiles2<float>(L), n3 = tiles3<float>(L), T = tile_a<float>();
            std::vector<int2> tl;
            std::vector<int> tm((size_t)n2 * n3, -1);
            for (int a = 0; a < n2; ++a)
                for (int b = 0; b < n3; ++b)
                    if (!(a > b && (a + 1) * T <= L / 2)) {
                        tm[(size_t)a * n3 + b] = (int)tl.size();
                        tl.push_back(make_int2(a, b));
                    }
            void *d = nullptr;
            const size_t tb = sizeof(int2) * tl.size(), mb = sizeof(int) * tm.size();
            const int rc = dev_alloc(c, &d, tb + mb, "lag-Fourier tile tables");
            if (rc) return rc;
            HIPCHK(c, hipMemcpy(d, tl.data(), tb, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy((char *)d + tb, tm.data(), mb, hipMemcpyHostToDevice));
            it = c->lag_r23_tab.emplace(L, std::make_tuple(d, (int)tl.size(), tb + mb)).first;
        }
        const int n1 = std::get<1>(it->second);
        if (n1 < *ntile) {
            *mode = 2;
            *ntile = n1;
            *tlist = (const int2 *)std::get<0>(it->second);
            *tmap = (const int *)((const char *)std::get<0>(it->second) + sizeof(int2) * n1);
        }
    }
    /* pair tiles whose listed tiles hold the whole k3 = L/2 column need no
       kernel for it */
    *nyq_sep = ((*mode == 1 && !(*nyql && *nnyq == 0)) || (fp64 && (L / 2) % tile_b<double>() == 0)) ? 1 : 0;
    return AMB_OK;
}

} // namespace

/* triangles of the lag-Fourier form in waves; bl64[i] (may be null, or
   hold nulls): FP64 lag arrays of request i; norm64[i] likewise for the
   normalisation of requests that ask for one */
int amb::lagft_run(amb_ctx *c, const amb_tri_req *req, int n, double t0, double *done_ms, double *const *bl64,
                   double *const *norm64)
{
    hipSetDevice(c->device);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return req[x].cube_id != req[y].cube_id ? req[x].cube_id < req[y].cube_id : req[x].r1 < req[y].r1;
    });
    const PlaneSrc src{&c->planes};
    std::vector<LJ> w;
    size_t part_tot = 0, out_tot = 0;
    c->lag_margin = 0.f;
    /* a wave's lag arrays land in device buffer hb and are copied to pinned
       staging hb on the copy stream (ev_cp[hb]); the previous wave's host
       copies (prev, buffer hb ^ 1) run while this wave's kernels run */
    int hb = 0;
    std::vector<LJ> prev;
    int rc0 = copy_stream(c);
    if (rc0) return rc0;

    auto finish_prev = [&]() -> int {
        if (prev.empty()) return AMB_OK;
        const hipError_t e = hipEventSynchronize(c->ev_cp[hb ^ 1]);
        if (e != hipSuccess) {
            prev.clear();
            return fail(c, AMB_EHIP, "copy of a lag-Fourier wave's results: %s", hipGetErrorString(e));
        }
        const char *h = (const char *)c->hstage[hb ^ 1];
        std::vector<std::tuple<void *, const void *, size_t>> cp;
        for (auto &j : prev) {
            const amb_tri_req &q = req[j.q];
            const Cube &cb = *j.cb;
            const size_t cells = cb.h_lagcnt.size();
            if (j.norm) {
                cp.emplace_back(q.norm, h + j.out, sizeof(float) * cells);
                if (norm64 && norm64[j.q]) cp.emplace_back(norm64[j.q], h + j.out64, sizeof(double) * cells);
                continue;
            }
            cp.emplace_back(q.bl, h + j.out, sizeof(float) * cells);
            if (bl64 && bl64[j.q]) cp.emplace_back(bl64[j.q], h + j.out64, sizeof(double) * cells);
            if (q.cnt) cp.emplace_back(q.cnt, cb.h_lagcnt.data(), sizeof(int32_t) * cells);
        }
        par_copy(cp);
        const double tdone = now_ms() - t0;
        if (done_ms)
            for (auto &j : prev) done_ms[j.q] = tdone;
        prev.clear();
        return AMB_OK;
    };

    auto flush = [&]() -> int {
        if (w.empty()) return AMB_OK;
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
        /* buffer hb's last copy finished with finish_prev of the wave before */
        void **outp = hb ? &c->lag_out2 : &c->lag_out;
        p = *outp;
        if (!rc) rc = grow(c, &p, hb ? &c->lag_out2_bytes : &c->lag_out_bytes, out_tot, "lag arrays");
        *outp = p;
        const size_t dbytes = (sizeof(LagJob<float>) + sizeof(LagJob<double>)) * w.size() + 256;
        if (!rc) rc = grow(c, &c->wdesc, &c->wdesc_bytes, dbytes, "wave descriptors");
        if (!rc) rc = grow_host(c, &c->hstage[hb], &c->hstage_bytes[hb], out_tot);
        if (rc) { unpin_all(c); w.clear(); return rc; }
        char *dout = (char *)*outp;
        std::vector<LagJob<float>> d32;
        std::vector<LagJob<double>> d64;
        std::vector<int> slot(w.size()); /* index of each job in its descriptor array */
        long it32 = 0, it64 = 0, nit32 = 0, nit64 = 0;
        for (size_t i = 0; i < w.size(); ++i) {
            if (w[i].norm) {
                slot[i] = (int)d64.size();
                d64.emplace_back();
                fill_job(d64.back(), w[i], (char *)c->lag_part, &it64, &nit64);
            } else {
                slot[i] = (int)d32.size();
                d32.emplace_back();
                fill_job(d32.back(), w[i], (char *)c->lag_part, &it32, &nit32);
            }
        }
        char *dd = (char *)c->wdesc;
        LagJob<float> *g32 = (LagJob<float> *)dd;
        LagJob<double> *g64 = (LagJob<double> *)(dd + (sizeof(LagJob<float>) * d32.size() + 255) / 256 * 256);
        hipError_t e = hipSuccess;
        if (!d32.empty()) e = hipMemcpyAsync(g32, d32.data(), sizeof(LagJob<float>) * d32.size(), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && !d64.empty())
            e = hipMemcpyAsync(g64, d64.data(), sizeof(LagJob<double>) * d64.size(), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && !d32.empty()) {
            timer_start(c, &c->st.ms_contract);
            e = launch_products<float>(c->stream, g32, (int)d32.size(), it32, nit32);
            timer_stop(c);
            c->st.n_contract_launch++;
        }
        const bool any_norm = !d64.empty();
        if (e == hipSuccess && any_norm) {
            e = amb_launch_fill_u32(c->stream, c->margin + 2, 1, 0u);
            timer_start(c, &c->st.ms_norm);
            if (e == hipSuccess) e = launch_products<double>(c->stream, g64, (int)d64.size(), it64, nit64);
            timer_stop(c);
        }
        timer_start(c, &c->st.ms_reduce);
        for (size_t i = 0; i < w.size() && e == hipSuccess; ++i) {
            const LJ &j = w[i];
            const Cube &cb = *j.cb;
            const int L = cb.lag_L;
            double2 *St = (double2 *)c->lag_st;
            const dim3 gr((unsigned)((L + 31) / 32), (unsigned)((L / 2 + 1 + 31) / 32));
            if (j.norm) {
                const LagJob<double> &g = d64[slot[i]];
                if (j.mode == 3)
                    hipLaunchKernelGGL((k_lagft_reduce<double, 3>), gr, dim3(256), 0, c->stream, g.part, j.nseg,
                                       j.ntile, L, nullptr, St);
                else if (j.mode == 5)
                    hipLaunchKernelGGL((k_lagft_reduce<double, 5>), gr, dim3(256), 0, c->stream, g.part, j.nseg,
                                       j.ntile, L, j.tmap, St);
                else
                    hipLaunchKernelGGL((k_lagft_reduce<double, 0>), gr, dim3(256), 0, c->stream, g.part, j.nseg,
                                       j.ntile, L, nullptr, St);
                if (j.nyq_sep)
                    hipLaunchKernelGGL(k_lagft_reduce_nyq<double>, dim3((unsigned)((L + 255) / 256)), dim3(256), 0,
                                       c->stream, g.nyq, j.nseg, L, nullptr, 0, St);
            } else {
                const LagJob<float> &g = d32[slot[i]];
                if (j.mode == 1)
                    hipLaunchKernelGGL((k_lagft_reduce<float, 1>), gr, dim3(256), 0, c->stream, g.part, j.nseg,
                                       j.ntile, L, j.tmap, St);
                else if (j.mode == 2)
                    hipLaunchKernelGGL((k_lagft_reduce<float, 2>), gr, dim3(256), 0, c->stream, g.part, j.nseg,
                                       j.ntile, L, j.tmap, St);
                else if (j.mode == 3)
                    hipLaunchKernelGGL((k_lagft_reduce<float, 3>), gr, dim3(256), 0, c->stream, g.part, j.nseg,
                                       j.ntile, L, nullptr, St);
                else if (j.mode == 4)
                    hipLaunchKernelGGL((k_lagft_reduce<float, 4>), gr, dim3(256), 0, c->stream, g.part, j.nseg,
                                       j.ntile, L, j.tmap, St);
                else if (j.mode == 5)
                    hipLaunchKernelGGL((k_lagft_reduce<float, 5>), gr, dim3(256), 0, c->stream, g.part, j.nseg,
                                       j.ntile, L, j.tmap, St);
                else
                    hipLaunchKernelGGL((k_lagft_reduce<float, 0>), gr, dim3(256), 0, c->stream, g.part, j.nseg,
                                       j.ntile, L, nullptr, St);
                /* the column rode in the last tile (modes 0 and 2, TB | L/2),
                   was a tile column (TB does not divide L/2), or had its own
                   kernel */
                const bool folded = (j.mode == 0 || j.mode == 2) && (L / 2) % tile_b<float>() == 0;
                if (j.nyq_sep || folded)
                    hipLaunchKernelGGL(k_lagft_reduce_nyq<float>, dim3((unsigned)((L + 255) / 256)), dim3(256), 0,
                                       c->stream, g.nyq, j.nseg, L, j.nyql, j.nnyq, St);
            }
            const FftPlan &pl = c->plans[L];
            const double2 *tw = c->tw64[L];
            const size_t lds = sizeof(double2) * 2 * (size_t)L;
            hipLaunchKernelGGL(k_lagft_cols, dim3((unsigned)(L / 2 + 1)), dim3(64), lds, c->stream, St, pl, tw);
            const int W = 2 * cb.nc10 + 1;
            float *o = (float *)(dout + j.out);
            double *o64 = j.out64 != (size_t)-1 ? (double *)(dout + j.out64) : nullptr;
            const int *e1 = j.norm ? nullptr : d32[slot[i]].ex1, *e2 = j.norm ? nullptr : d32[slot[i]].ex2,
                      *e3 = j.norm ? nullptr : d32[slot[i]].ex3;
            hipLaunchKernelGGL(k_lagft_rows, dim3((unsigned)((W + 1) / 2)), dim3(64), lds, c->stream, St, pl, tw, cb.nc10,
                               cb.d_lagcnt, e1, e2, e3, j.norm ? (double)j.P : 0.0, (double)j.Pref,
                               j.norm ? 1.0 : (double)j.Pref / (double)j.P, o, o64,
                               j.norm ? c->margin + 2 : nullptr);
            e = hipGetLastError();
        }
        timer_stop(c);
        unsigned mbits = 0;
        if (e == hipSuccess) e = hipEventRecord(c->ev_red, c->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(c->cstream, c->ev_red, 0);
        if (e == hipSuccess) e = hipMemcpyAsync(c->hstage[hb], dout, out_tot, hipMemcpyDeviceToHost, c->cstream);
        if (e == hipSuccess) e = hipEventRecord(c->ev_cp[hb], c->cstream);
        if (e == hipSuccess && any_norm)
            e = hipMemcpyAsync(&mbits, c->margin + 2, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream);
        /* the previous wave's host copies while this wave runs */
        const int prc = finish_prev();
        const hipError_t es = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = es;
        timers_collect(c);
        unpin_all(c);
        if (e != hipSuccess) {
            hipStreamSynchronize(c->cstream);
            w.clear();
            return fail(c, AMB_EHIP, "lag-Fourier wave of %zu jobs: %s", slot.size(), hipGetErrorString(e));
        }
        if (prc) { hipStreamSynchronize(c->cstream); w.clear(); return prc; }
        if (any_norm) {
            float m;
            memcpy(&m, &mbits, sizeof m);
            if (m > c->st.cert_margin_max) c->st.cert_margin_max = m;
            if (m > c->lag_margin) c->lag_margin = m;
            if (!(m <= 1.0f / 32.0f)) {
                hipStreamSynchronize(c->cstream);
                w.clear();
                return fail(c, AMB_ECERT, "lag-Fourier normalisation: FP64 sums off their 1/8 grid by %g", (double)m);
            }
        }
        for (auto &j : w) {
            const Cube &cb = *j.cb;
            if (j.norm) {
                ++c->st.n_norm_exact;
                continue;
            }
            const int L = cb.lag_L;
            c->st.evals += cb.lag_evals;
            /* the direct sum's work on the reference grid */
            c->st.flop_contract += 2.0 * (double)cb.lag_pairs * (double)j.Pref;
            c->st.flop_contract_distinct += 2.0 * (double)cb.lag_pairs * (double)j.Pref;
            /* 14 flops per distinct (k2, k3, pixel of the compute grid) of the
               half spectrum k3 <= L/2: half of them where r1 == r2 or r2 == r3 pair equal
               sums, a third where r1 == r2 == r3 */
            c->st.flop_lagft += 14.0 * (double)L * (double)(L / 2 + 1) * (double)j.P * lag_distinct(req[j.q]);
            /* issued: whole tiles, plus the k3 = L/2 column (or its k2 left
               to k_lagft_nyq by pair tiles) */
            c->st.flop_contract_exec +=
                14.0 * ((double)j.ntile * tile_a<float>() * tile_b<float>() + edge_cells(j.mode, L, j.nyql, j.nnyq)) *
                (double)j.P;
        }
        prev = std::move(w);
        w.clear();
        hb ^= 1;
        part_tot = out_tot = 0;
        return AMB_OK;
    };

    /* Wave cuts: each wave's results go to host memory (PCIe, then into the
       caller's arrays) while the next wave computes, and the last wave's
       copies end the call.  From the end backwards: the last wave is the
       last triangle, and each earlier wave takes as many triangles as the
       following wave's estimated device time covers with its copies
       (device time ~ 14 flop per (k2, k3, pixel) at 60 TFLOP/s, halved for
       pair tiles, FP64 normalisation 8x; copies ~ 15 GB/s). */
    std::vector<char> cut(n + 1, 0);
    if (n >= 2) {
        std::vector<double> tdev(n), tcp(n);
        for (int k = 0; k < n; ++k) {
            const amb_tri_req &q = req[order[k]];
            auto it = c->cubes.find(q.cube_id);
            const int nr = (int)c->ulow.size();
            if (it == c->cubes.end() || q.r1 < 0 || q.r1 >= nr || !it->second.chans || lag_setup(c, it->second)) {
                tdev[k] = tcp[k] = 0;
                continue;
            }
            const Cube &cb = it->second;
            const double ng = (double)compute_grid(c, cb, q.r1, q.r2, q.r3), L = cb.lag_L;
            const double cells = (double)cb.h_lagcnt.size();
            tdev[k] = 14.0 * L * (L / 2 + 1) * ng * ng * (q.r1 == q.r2 ? 0.55 : 1.0) * (q.norm ? 9.0 : 1.0) / 60e12;
            const double bytes = cells * (4.0 + (q.cnt ? 4.0 : 0.0) + (q.norm ? 4.0 : 0.0) +
                                          (bl64 && bl64[order[k]] ? 8.0 : 0.0) + (norm64 && norm64[order[k]] ? 8.0 : 0.0));
            tcp[k] = bytes / 15e9;
        }
        /* each wave at least 1.5x the one after it (where copies outweigh
           device time more waves would not hide them) */
        int end = n - 1; /* first triangle of the wave sized last; the last wave is [n-1] */
        int after = 1;   /* its size */
        cut[n - 1] = 1;
        double cover = tdev[n - 1];
        while (end > 0) {
            int start = end - 1;
            double cp = tcp[start], dv = tdev[start];
            const int minw = (3 * after + 1) / 2;
            while (start > 0 && (cp + tcp[start - 1] <= cover || end - start < minw)) {
                --start;
                cp += tcp[start];
                dv += tdev[start];
            }
            cut[start] = 1;
            cover = dv;
            after = end - start;
            end = start;
        }
    }

    auto run_all = [&]() -> int {
        for (int k = 0; k < n; ++k) {
            const int qi = order[k];
            const amb_tri_req &q = req[qi];
            if (!q.bl) return fail(c, AMB_EARG, "request %d has no bl buffer", qi);
            if (cut[k] && !w.empty()) {
                int rc = flush();
                if (rc) return rc;
            }
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
            const long ng = amb_ng(c->uhigh[q.r1], cb->Du), ngc = compute_grid(c, *cb, q.r1, q.r2, q.r3);
            if (ng <= 0 || ng > 2730) { unpin_all(c); return fail(c, AMB_EARG, "Ng=%ld out of range", ng); }
            const int L = cb->lag_L;
            {   /* twiddles and plan of L: FP32 for the spectra, FP64 for the 2-D inverse */
                float2 *tw;
                FftPlan pl;
                if ((rc = twiddles(c, L, &tw, &pl)) || (rc = twiddles64(c, L, nullptr))) { unpin_all(c); return rc; }
            }
            const size_t cells = cb->h_lagcnt.size();
            const int rr[3] = {q.r1, q.r2, q.r3};
            for (int pass = 0; pass < (q.norm ? 2 : 1); ++pass) {
                const bool nm = pass == 1;
                LJ j{};
                j.q = qi;
                j.norm = nm;
                j.cb = cb;
                j.ng = ng;
                j.ngc = ngc;
                j.P = ngc * ngc;
                j.Pref = ng * ng;
                j.nseg = (int)((j.P + SEG - 1) / SEG);
                if ((rc = choose_tiles(c, L, nm, q.r1 == q.r2, q.r2 == q.r3, &j.mode, &j.ntile, &j.nyq_sep, &j.tlist,
                                       &j.tmap, &j.nyql, &j.nnyq))) {
                    unpin_all(c);
                    return rc;
                }
                const size_t pb = nm ? job_part_bytes<double>(j, L) : job_part_bytes<float>(j, L);
                const bool want64 = nm ? (norm64 && norm64[qi]) : (bl64 && bl64[qi]);
                const size_t ob = al(sizeof(float) * cells) + (want64 ? al(sizeof(double) * cells) : 0);
                if (!w.empty() && (part_tot + al(pb) > PART_CAP || out_tot + ob > OUT_CAP || (int)w.size() >= MAX_JOBS))
                    if ((rc = flush())) return rc;
                for (int t = 0; t < 3 && !rc; ++t) {
                    rc = get_planes(c, src, q.cube_id, nm ? 10 : 9, rr[t], (int)ngc, &j.X[t], (int)ng);
                    if (rc == AMB_EBUDGET && !w.empty()) {
                        /* the wave's planes hold the memory: run it, then resolve again */
                        if ((rc = flush())) return rc;
                        t = -1;
                        rc = AMB_OK;
                    }
                }
                if (rc) { unpin_all(c); return rc; }
                j.part = part_tot;
                j.nyq = part_tot + (nm ? job_part_bytes<double>(j, L) - sizeof(double2) * (size_t)j.nseg * L
                                       : job_part_bytes<float>(j, L) - sizeof(float2) * (size_t)j.nseg * L);
                part_tot += al(pb);
                j.out = out_tot;
                j.out64 = want64 ? out_tot + al(sizeof(float) * cells) : (size_t)-1;
                out_tot += ob;
                w.push_back(j);
            }
        }
        int rc = flush();
        return rc ? rc : finish_prev();
    };
    const int rc = run_all();
    hipStreamSynchronize(c->cstream); /* no copy left running, also after an error */
    return rc;
}

/* ------------------------------------------------ the serving queue's view */

int amb::lag_prepare(amb_ctx *c, int cube, int r1, int r2, int r3, LagTri *t)
{
    Cube *cb;
    int rc = find_cube(c, cube, &cb);
    if (rc) return rc;
    if (!cb->chans) return fail(c, AMB_ESTATE, "channels not set for cube %d", cube);
    const int nr = (int)c->ulow.size();
    if (nr == 0) return fail(c, AMB_ESTATE, "rings not set");
    if (r1 < 0 || r1 >= nr || r2 < 0 || r2 >= nr || r3 < 0 || r3 >= nr)
        return fail(c, AMB_EARG, "triangle (%d,%d,%d) outside [0,%d)", r1, r2, r3, nr);
    if ((rc = lag_setup(c, *cb))) return rc;
    const long ng = amb_ng(c->uhigh[r1], cb->Du), ngc = compute_grid(c, *cb, r1, r2, r3);
    if (ng <= 0 || ng > 2730) return fail(c, AMB_EARG, "Ng=%ld out of range", ng);
    const int L = cb->lag_L;
    float2 *tw;
    FftPlan pl;
    if ((rc = twiddles(c, L, &tw, &pl)) || (rc = twiddles64(c, L, nullptr))) return rc;
    *t = LagTri{};
    t->cube = cube; t->r1 = r1; t->r2 = r2; t->r3 = r3;
    t->ng = ng;
    t->ngc = ngc;
    t->P = ngc * ngc;
    t->L = L;
    t->nseg = (int)((t->P + SEG - 1) / SEG);
    t->cells = cb->h_lagcnt.size();
    if ((rc = choose_tiles(c, L, false, r1 == r2, r2 == r3, &t->mode, &t->ntile, &t->nyq_sep, &t->tlist, &t->tmap,
                           &t->nyql, &t->nnyq)))
        return rc;
    const PlaneSrc src{&c->planes};
    const int rr[3] = {r1, r2, r3};
    for (int k = 0; k < 3; ++k)
        if ((rc = get_planes(c, src, cube, 9, rr[k], (int)ngc, &t->X[k], (int)ng))) return rc;
    return AMB_OK;
}

size_t amb::lag_part_bytes(const LagTri &t)
{
    LJ j{};
    j.nseg = t.nseg;
    j.ntile = t.ntile;
    return job_part_bytes<float>(j, t.L);
}

long amb::lag_items(const LagTri &t) { return (long)t.nseg * t.ntile; }

long amb::lag_item_px() { return SEG; }

size_t amb::lag_desc_bytes() { return sizeof(LagJob<float>); }

size_t amb::lag_st_cap() { return sizeof(double2) * (size_t)(MAXL / 2 + 1) * MAXL; }

/* the descriptor of items [i0, i0 + n) of t, numbered from item0 in its launch */
void amb::lag_fill_range(amb_ctx *c, const LagTri &t, void *part, long i0, long item0, void *h)
{
    LJ j{};
    j.norm = false;
    j.cb = &c->cubes[t.cube];
    j.ng = t.ng;
    j.ngc = t.ngc;
    j.P = t.P;
    j.Pref = t.ng * t.ng;
    j.nseg = t.nseg;
    j.mode = t.mode;
    j.ntile = t.ntile;
    j.nyq_sep = t.nyq_sep;
    j.tlist = t.tlist;
    j.tmap = t.tmap;
    j.nyql = t.nyql;
    j.nnyq = t.nnyq;
    for (int k = 0; k < 3; ++k) j.X[k] = t.X[k];
    j.part = 0;
    j.nyq = sizeof(float2) * (size_t)j.nseg * j.ntile * tile_a<float>() * tile_b<float>();
    LagJob<float> g;
    long items = item0, nitems = 0;
    fill_job(g, j, (char *)part, &items, &nitems);
    g.ibase = i0;
    memcpy(h, &g, sizeof g);
}

hipError_t amb::lag_launch_tiles(hipStream_t s, const void *d, int n, long total)
{
    if (n <= 0 || total <= 0) return hipSuccess;
    return launch_products<float>(s, (const LagJob<float> *)d, n, total, 0);
}

/* after the last of t's items: the k3 = L/2 column where it has a kernel of
   its own (d_whole: t's descriptor with item0 = ibase = 0), the segment sums,
   the 2-D inverse and the lag array bl (f32, W x W) */
hipError_t amb::lag_launch_finish(amb_ctx *c, hipStream_t s, const LagTri &t, const void *d_whole, void *part,
                                  void *St, float *bl)
{
    const Cube &cb = c->cubes[t.cube];
    const int L = t.L;
    const LagJob<float> *g = (const LagJob<float> *)d_whole;
    const float2 *pt = (const float2 *)part;
    const float2 *nyq = pt + (size_t)t.nseg * t.ntile * tile_a<float>() * tile_b<float>();
    if (t.nyq_sep) {
        hipLaunchKernelGGL(k_lagft_nyq<float>, dim3((unsigned)((t.nseg + 7) / 8 * 8)), dim3(256), 0, s, g, 1,
                           (long)t.nseg);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    double2 *st = (double2 *)St;
    const dim3 gr((unsigned)((L + 31) / 32), (unsigned)((L / 2 + 1 + 31) / 32));
    if (t.mode == 1)
        hipLaunchKernelGGL((k_lagft_reduce<float, 1>), gr, dim3(256), 0, s, pt, t.nseg, t.ntile, L, t.tmap, st);
    else if (t.mode == 2)
        hipLaunchKernelGGL((k_lagft_reduce<float, 2>), gr, dim3(256), 0, s, pt, t.nseg, t.ntile, L, t.tmap, st);
    else if (t.mode == 3)
        hipLaunchKernelGGL((k_lagft_reduce<float, 3>), gr, dim3(256), 0, s, pt, t.nseg, t.ntile, L, nullptr, st);
    else if (t.mode == 4)
        hipLaunchKernelGGL((k_lagft_reduce<float, 4>), gr, dim3(256), 0, s, pt, t.nseg, t.ntile, L, t.tmap, st);
    else if (t.mode == 5)
        hipLaunchKernelGGL((k_lagft_reduce<float, 5>), gr, dim3(256), 0, s, pt, t.nseg, t.ntile, L, t.tmap, st);
    else
        hipLaunchKernelGGL((k_lagft_reduce<float, 0>), gr, dim3(256), 0, s, pt, t.nseg, t.ntile, L, nullptr, st);
    const bool folded = (t.mode == 0 || t.mode == 2) && (L / 2) % tile_b<float>() == 0;
    if (t.nyq_sep || folded)
        hipLaunchKernelGGL(k_lagft_reduce_nyq<float>, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, s, nyq, t.nseg, L,
                           t.nyql, t.nnyq, st);
    const FftPlan &pl = c->plans[L];
    const double2 *tw = c->tw64[L];
    const size_t lds = sizeof(double2) * 2 * (size_t)L;
    hipLaunchKernelGGL(k_lagft_cols, dim3((unsigned)(L / 2 + 1)), dim3(64), lds, s, st, pl, tw);
    const int W = 2 * cb.nc10 + 1;
    const int *ex[3];
    for (int k = 0; k < 3; ++k) ex[k] = (const int *)((const float2 *)t.X[k] + (size_t)t.P * cb.lag_lhp) + 1;
    hipLaunchKernelGGL(k_lagft_rows, dim3((unsigned)((W + 1) / 2)), dim3(64), lds, s, st, pl, tw, cb.nc10,
                       cb.d_lagcnt, ex[0], ex[1], ex[2], 0.0, (double)(t.ng * t.ng),
                       (double)(t.ng * t.ng) / (double)t.P, bl, nullptr, nullptr);
    return hipGetLastError();
}

/* counts, evals and work of one finished triangle */
void amb::lag_account(amb_ctx *c, const LagTri &t, int32_t *cnt)
{
    const Cube &cb = c->cubes[t.cube];
    if (cnt) memcpy(cnt, cb.h_lagcnt.data(), sizeof(int32_t) * cb.h_lagcnt.size());
    const int L = t.L;
    c->st.evals += cb.lag_evals;
    c->st.flop_contract += 2.0 * (double)cb.lag_pairs * (double)(t.ng * t.ng);
    c->st.flop_contract_distinct += 2.0 * (double)cb.lag_pairs * (double)(t.ng * t.ng);
    c->st.flop_lagft += 14.0 * (double)L * (double)(L / 2 + 1) * (double)t.P * lag_distinct_r(t.r1, t.r2, t.r3);
    c->st.flop_contract_exec +=
        14.0 * ((double)t.ntile * tile_a<float>() * tile_b<float>() + edge_cells(t.mode, L, t.nyql, t.nnyq)) *
        (double)t.P;
}

/*
 * Ring gather + c2r for gfx950: the filtered image planes of Dfg/Ifg
 * (amber_funcs.c:573-621) as used by fill_s2d_float / fill_s2I
 * (amber_funcs.c:967-1010, 1109-1131), in FP32.
 *
 * The c2r of an Ng x (Ng/2+1) half plane whose nonzeros are the cells of one
 * annulus runs in two passes, each transform owned by one 64-thread team of
 * the LDS Stockham core (amb_fftcore.h: register codelets for 2, 3, 4 and
 * the primes up to 47, a direct stage for larger primes, so every even Ng,
 * 2*109, 2*181, 2*349 included, runs without padding):
 *   cols: for every column j < jcnt (the only columns a ring can touch,
 *         jcnt = min(Ng/3, ceil(UHigh))) gather the ring cells straight from
 *         the plane-major cube, inverse-DFT along the first axis and write
 *         G[plane][j][x] (x contiguous, coalesced).
 *   rows: for every pair of output rows (xa, xb) build the Hermitian
 *         extension Z(k) = A(k) + i B(k), Z(Ng-k) = conj A(k) + i conj B(k),
 *         one complex inverse DFT gives row xa in Re and row xb in Im; the
 *         row is written once as Ng floats (coalesced).
 * The j=0 column enters the row pass through its real part, which is the
 * c2r of the Hermitian-symmetrised column (the FFTW/pocketfft behaviour).
 * Ring membership is an integer test on ia^2 + j^2 against thresholds the
 * host derives with the reference's double sqrt.  A plane's workgroups run
 * as one contiguous range on one XCD, so its G rows stay in one L2.
 */
#include "amb_internal.h"
#include "amb_fftcore.h"
#include <math.h>
#include <stdio.h>

#include <algorithm>

#define FFT_THREADS 256

void amb_make_plan(int n, FftPlan *pl) { fftc::make_plan(n, pl); }

namespace {

using fftc::cx;
template <typename T>
using Cx = typename fftc::CT<T>::c;

/* blocks b = x (mod 8) run on XCD x: XCD x takes the contiguous work range
   [x per, (x+1) per), so the blocks of one plane share an L2 */
__device__ __forceinline__ long xcd_item(long b, long nb)
{
    const long per = nb / 8;
    return (b % 8) * per + b / 8;
}

/* largest codelet prime per scalar type (larger codelets spill: FP32 past 23,
   FP64 past 19); larger prime factors take the direct stage */
template <typename R>
struct MaxP { static constexpr int v = 23; };
template <>
struct MaxP<double> { static constexpr int v = 19; };

} // namespace

/* Column pass: one 64-thread team per (plane, column j < jcnt).  The ring
 * cells of column j (s_lo <= ia^2 + j^2 < s_hi, ia in (-Ng/2, Ng/2]) are
 * gathered from the plane-major cube (unit stride in ia), inverse DFT along
 * the first axis, written to G[plane][j][x].
 * 1-D grid of nblk x nteam teams, a multiple of 8 blocks; dynamic LDS:
 * [tw n if TWL] + nteam * 2 n complex. */
template <typename R, bool TWL>
__global__ __launch_bounds__(FFT_THREADS) void k_ring_cols(int kind, const float2 *__restrict__ vis_t,
                                                           const float *__restrict__ samp_t,
                                                           const int *__restrict__ planes, int npl, int N, int Nh,
                                                           FftPlan pl, const Cx<R> *__restrict__ twg, long s_lo,
                                                           long s_hi, int jcnt, int nbx, Cx<R> *__restrict__ G)
{
    typedef Cx<R> C;
    extern __shared__ __align__(16) char smem_raw[];
    C *sm = (C *)smem_raw;
    const int n = pl.n;
    const int nteam = blockDim.x >> 6, team = threadIdx.x >> 6, t = threadIdx.x & 63;
    C *bx = sm + (TWL ? n : 0) + (size_t)team * 2 * n, *by = bx + n;
    const C *tw = TWL ? sm : twg;
    if (TWL)
        for (int i = threadIdx.x; i < n; i += blockDim.x) sm[i] = twg[i];
    const long item = xcd_item(blockIdx.x, gridDim.x);
    const int pli = (int)(item / nbx);
    const int j = (int)(item - (long)pli * nbx) * nteam + team;
    const bool live = pli < npl && j < jcnt;
    const int ch = live ? planes[pli] : 0;
    for (int ing = t; ing < n; ing += 64) {
        const int ia = ing <= n / 2 ? ing : ing - n;
        const long s2 = (long)ia * ia + (long)j * j;
        C v = cx((R)0, (R)0);
        if (live && s2 >= s_lo && s2 < s_hi) {
            const int in = ia < 0 ? N + ia : ia;
            const size_t src = ((size_t)ch * Nh + j) * N + in;
            if (kind == 0) v = cx((R)vis_t[src].x, (R)vis_t[src].y);
            else v = cx((R)samp_t[src], (R)0);
        }
        bx[ing] = v;
    }
    __syncthreads();
    const C *res = fftc::transform<+1, MaxP<R>::v>(bx, by, tw, pl, t, 64);
    if (live) {
        C *g = G + ((size_t)pli * jcnt + j) * n;
        for (int x = t; x < n; x += 64) g[x] = res[x];
    }
}

/* Row pass: one team per (plane, pair of output rows xa, xb), a workgroup
 * holding the 2 nteam consecutive rows x0 .. x0 + 2 nteam - 1.  The
 * Hermitian extension of two rows' half spectra as one complex sequence
 * Z(k) = A(k) + i B(k), Z(Ng - k) = conj A(k) + i conj B(k); its inverse DFT
 * holds row xa in the real part and row xb in the imaginary part (the j = 0
 * column enters through its real part, the c2r of the Hermitian-symmetrised
 * column).  The workgroup reads G[k][x0 ..] for 32 columns k per pass, each
 * lane one (k, x) value (2 nteam x 8 contiguous bytes per k), and the lane
 * of row xa takes row xb's value from its neighbour lane.  mx (may be
 * null): max |value| over the set, as the bits of a non-negative float
 * (one atomic per workgroup). */
template <typename R, bool TWL>
__global__ __launch_bounds__(FFT_THREADS) void k_ring_rows(const Cx<R> *__restrict__ G, int jcnt, int npl, FftPlan pl,
                                                           const Cx<R> *__restrict__ twg, int nbx,
                                                           R *__restrict__ out, long pitch, unsigned *mx)
{
    typedef Cx<R> C;
    extern __shared__ __align__(16) char smem_raw[];
    C *sm = (C *)smem_raw;
    const int n = pl.n;
    const int nteam = blockDim.x >> 6, team = threadIdx.x >> 6, t = threadIdx.x & 63;
    C *bx = sm + (TWL ? n : 0) + (size_t)team * 2 * n, *by = bx + n;
    const C *tw = TWL ? sm : twg;
    if (TWL)
        for (int i = threadIdx.x; i < n; i += blockDim.x) sm[i] = twg[i];
    const long item = xcd_item(blockIdx.x, gridDim.x);
    const int pli = (int)(item / nbx), bxi = (int)(item - (long)pli * nbx);
    const int W = 2 * nteam, x0 = W * bxi;
    const int xa = x0 + 2 * team, xb = xa + 1;
    const bool live = pli < npl && xb < n;
    /* the columns no ring cell reaches: jcnt <= n / 2 */
    for (int k = jcnt + t; k <= n - jcnt; k += 64) bx[k] = cx((R)0, (R)0);
    if (pli < npl) {
        const C *Gp = G + (size_t)pli * jcnt * n;
        const int xx = threadIdx.x % W, kq = threadIdx.x / W; /* 32 columns per pass */
        for (int k0 = 0; k0 < jcnt; k0 += 32) {
            const int k = k0 + kq;
            C v = cx((R)0, (R)0);
            if (k < jcnt && x0 + xx < n) v = Gp[(size_t)k * n + x0 + xx];
            /* W is even: lanes 2m, 2m + 1 hold rows xa, xb of one pair at one k */
            C o;
            o.x = __shfl_xor(v.x, 1);
            o.y = __shfl_xor(v.y, 1);
            if (k < jcnt && !(xx & 1)) {
                C *bt = sm + (TWL ? n : 0) + (size_t)(xx >> 1) * 2 * n;
                if (k == 0) {
                    bt[0] = cx(v.x, o.x);
                } else {
                    bt[k] = cx(v.x - o.y, v.y + o.x);
                    bt[n - k] = cx(v.x + o.y, o.x - v.y);
                }
            }
        }
    }
    __syncthreads();
    const C *res = fftc::transform<+1, MaxP<R>::v>(bx, by, tw, pl, t, 64);
    float m = 0.f;
    if (live) {
        R *op = out + (size_t)pli * pitch;
        if (bxi == 0 && team == 0)
            for (long q = (long)n * n + t; q < pitch; q += 64) op[q] = (R)0;
        for (int y = t; y < n; y += 64) {
            const C z = res[y];
            op[(size_t)xa * n + y] = z.x;
            op[(size_t)xb * n + y] = z.y;
            m = fmaxf(m, fmaxf(fabsf((float)z.x), fabsf((float)z.y)));
        }
    }
    if (mx) {
        __shared__ float red[FFT_THREADS / 64];
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if (t == 0) red[team] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < nteam; ++w) m = fmaxf(m, red[w]);
            atomicMax(mx, __float_as_uint(m));
        }
    }
}

namespace {

/* teams per workgroup and whether the twiddles fit LDS beside them */
struct FftLaunch {
    int nteam = 0;
    bool twl = true;
    size_t lds = 0;
};
FftLaunch fft_launch(int n, size_t cbytes)
{
    FftLaunch f;
    const size_t cap = 65536 / cbytes; /* complex values in 64 KiB */
    f.nteam = (int)std::min<size_t>(4, cap > (size_t)n ? (cap - n) / (2 * (size_t)n) : 0);
    if (f.nteam < 1) {
        f.twl = false;
        f.nteam = (int)std::min<size_t>(4, cap / (2 * (size_t)n));
    }
    f.lds = cbytes * ((f.twl ? n : 0) + 2 * (size_t)f.nteam * n);
    return f;
}

template <typename R>
hipError_t ring_passes(hipStream_t s, int kind, const float2 *vis_t, const float *samp_t, const int *d_planes, int npl,
                       int N, int Nh, const FftPlan &pl, const Cx<R> *d_tw, const RingGeom &g, Cx<R> *G, R *out,
                       long pitch, unsigned *mx)
{
    const int n = pl.n;
    const FftLaunch f = fft_launch(n, sizeof(Cx<R>));
    if (f.nteam < 1) return hipErrorInvalidValue;
    const int nbc = (g.jcnt + f.nteam - 1) / f.nteam;
    const long gc = ((long)npl * nbc + 7) / 8 * 8;
    if (f.twl)
        hipLaunchKernelGGL((k_ring_cols<R, true>), dim3((unsigned)gc), dim3(64 * f.nteam), f.lds, s, kind, vis_t, samp_t,
                           d_planes, npl, N, Nh, pl, d_tw, g.s_lo, g.s_hi, g.jcnt, nbc, G);
    else
        hipLaunchKernelGGL((k_ring_cols<R, false>), dim3((unsigned)gc), dim3(64 * f.nteam), f.lds, s, kind, vis_t,
                           samp_t, d_planes, npl, N, Nh, pl, d_tw, g.s_lo, g.s_hi, g.jcnt, nbc, G);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nbr = (n / 2 + f.nteam - 1) / f.nteam;
    const long gr = ((long)npl * nbr + 7) / 8 * 8;
    if (f.twl)
        hipLaunchKernelGGL((k_ring_rows<R, true>), dim3((unsigned)gr), dim3(64 * f.nteam), f.lds, s, G, g.jcnt, npl, pl,
                           d_tw, nbr, out, pitch, mx);
    else
        hipLaunchKernelGGL((k_ring_rows<R, false>), dim3((unsigned)gr), dim3(64 * f.nteam), f.lds, s, G, g.jcnt, npl,
                           pl, d_tw, nbr, out, pitch, mx);
    return hipGetLastError();
}

} // namespace

/* Normalises a data plane set in place: plane i (compact channel index)
 * times (-1)^i 2^(7 - E), E = amb_set_exp(max |x|) from slot[0], so every
 * stored value lies in [-2^7, 2^7]; slot[1] gets E.  Both factors are
 * exact and undone exactly by whoever reads the set.  The alternating sign
 * makes the products of a lag cell's triplets (p1, p1 + d2, p1 + d3) carry
 * (-1)^(p1 + d2 + d3): every other reference plane is accumulated negated,
 * so a rounding bias of the matrix cores that does not follow the sign of
 * its operands cancels in the sum over p1 instead of adding up. */
__global__ void k_normalise_set(float *__restrict__ D, long pitch4, unsigned *slot)
{
    const int e = amb_set_exp(__uint_as_float(slot[0]));
    const float f = ldexpf(1.f, AMB_SET_EXP - e);
    const float g = (blockIdx.y & 1) ? -f : f;
    float4 *p = (float4 *)D + (size_t)blockIdx.y * pitch4;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < pitch4; i += (long)gridDim.x * blockDim.x) {
        float4 v = p[i];
        v.x *= g; v.y *= g; v.z *= g; v.w *= g;
        p[i] = v;
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) slot[1] = (unsigned)e;
}

__global__ void k_zero64(double *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_zero(float *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = 0.f;
}

hipError_t amb_launch_ring_c2r(hipStream_t s, int kind, const float2 *vis_t, const float *samp_t,
                               const int *d_planes, int npl, int N, int Nh, const FftPlan &pl,
                               const float2 *d_tw, RingGeom g, float2 *G, float *out, long pitch, unsigned *slot,
                               bool raw)
{
    const int n = pl.n;
    if (npl <= 0) return hipSuccess;
    if (pitch < (long)n * n || pitch % 4 != 0 || (n & 1)) return hipErrorInvalidValue;
    if (slot) {
        /* max and exponent words start at zero: an empty set keeps E = 0 */
        hipError_t e = amb_launch_fill_u32(s, slot, 2, 0u);
        if (e != hipSuccess) return e;
    }
    if (g.jcnt <= 0) {
        size_t tot = (size_t)npl * pitch;
        hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, out, tot);
        return hipGetLastError();
    }
    hipError_t e = ring_passes<float>(s, kind, vis_t, samp_t, d_planes, npl, N, Nh, pl, d_tw, g, G, out, pitch, slot);
    if (e != hipSuccess || !slot || raw) return e;
    const long p4 = pitch / 4;
    const int bx = (int)std::min<long>((p4 + 255) / 256, 64);
    hipLaunchKernelGGL(k_normalise_set, dim3((unsigned)bx, (unsigned)npl), dim3(256), 0, s, out, p4, slot);
    return hipGetLastError();
}

/* FP64 c2r of sampling planes' ring (the calCR images, and the exact v2
 * normalisation's, whose triple sums must resolve a 1/8 grid of Ng^2): the
 * same two passes in double. */
hipError_t amb_launch_ring_c2r64(hipStream_t s, const float *samp_t, const int *d_planes, int npl, int N, int Nh,
                                 const FftPlan &pl, const double2 *d_tw, RingGeom g, double2 *G, double *out,
                                 long pitch)
{
    const int n = pl.n;
    if (npl <= 0) return hipSuccess;
    if (pitch < (long)n * n || (n & 1)) return hipErrorInvalidValue;
    if (g.jcnt <= 0) {
        hipLaunchKernelGGL(k_zero64, dim3(1024), dim3(256), 0, s, out, (size_t)npl * pitch);
        return hipGetLastError();
    }
    return ring_passes<double>(s, 1, nullptr, samp_t, d_planes, npl, N, Nh, pl, d_tw, g, G, out, pitch, nullptr);
}

/* tw[m] = exp(+2 pi i m / n) in FP64 */
__global__ void k_twiddles64(double2 *tw, int n)
{
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) {
        double sn, cs;
        sincospi(2.0 * (double)m / (double)n, &sn, &cs);
        tw[m] = make_double2(cs, sn);
    }
}

hipError_t amb_launch_twiddles64(hipStream_t s, double2 *tw, int n)
{
    hipLaunchKernelGGL(k_twiddles64, dim3((n + 255) / 256), dim3(256), 0, s, tw, n);
    return hipGetLastError();
}

/* Upload: file layout [rows][Nh][nc]{re,im} doubles -> plane-major
 * vis_t[nc][Nh][N] complex64 and samp_t[nc][Nh][N] float, with
 * normalize_data (amber_funcs.c:1946-1978) applied in FP64 first.  For each
 * uv column j this is a transpose of the (row, channel) block: a workgroup
 * reads a 64-row x 64-channel tile (whole 1 KB runs of one row's channels),
 * normalises it into LDS, and writes it back as 64 channels x 64 rows
 * (512-byte runs of one plane's rows), so both sides move whole lines. */
__global__ __launch_bounds__(256) void k_upload(const double2 *__restrict__ gv, const double *__restrict__ samp,
                                                int i0, int rows, int N, int Nh, int nc,
                                                float2 *__restrict__ vis_t, float *__restrict__ samp_t)
{
    __shared__ float2 tv[64][65];
    __shared__ float ts[64][65];
    const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64, j = blockIdx.z;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int rr = ty; rr < 64; rr += 4) {
        const int r = r0 + rr, c = c0 + tx;
        if (r < rows && c < nc) {
            const size_t idx = ((size_t)r * Nh + j) * nc + c;
            double sv = samp[idx];
            const double2 g = gv[idx];
            double re = g.x, im = g.y;
            if (sv > 0) { re /= sv; im /= sv; sv = 1.0; }
            tv[rr][tx] = make_float2((float)re, (float)im);
            ts[rr][tx] = (float)sv;
        }
    }
    __syncthreads();
    for (int cc = ty; cc < 64; cc += 4) {
        const int c = c0 + cc, r = r0 + tx;
        if (r < rows && c < nc) {
            const size_t dst = ((size_t)c * Nh + j) * N + (i0 + r);
            vis_t[dst] = tv[tx][cc];
            samp_t[dst] = ts[tx][cc];
        }
    }
}

hipError_t amb_launch_upload(hipStream_t s, const double *gv, const double *samp, int i0, int rows,
                             int N, int Nh, int nc, float2 *vis_t, float *samp_t)
{
    if (rows < 1 || nc < 1 || Nh < 1 || Nh > 65535 || i0 < 0 || i0 + rows > N) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((nc + 63) / 64), (unsigned)((rows + 63) / 64), (unsigned)Nh);
    hipLaunchKernelGGL(k_upload, grid, dim3(256), 0, s, (const double2 *)gv, samp, i0, rows, N, Nh, nc, vis_t,
                       samp_t);
    return hipGetLastError();
}

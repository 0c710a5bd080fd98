/*
 * Ring gather + c2r for gfx950: the filtered image planes of Dfg/Ifg
 * (amber_funcs.c:573-621) as used by fill_s2d_float / fill_s2I
 * (amber_funcs.c:967-1010, 1109-1131), in FP32.
 *
 * The c2r of an Ng x (Ng/2+1) half plane whose nonzeros are the cells of one
 * annulus runs in two LDS-staged Stockham passes:
 *   cols: for every column j < jcnt (the only columns a ring can touch,
 *         jcnt = min(Ng/3, ceil(UHigh))) gather the ring cells straight from
 *         the plane-major cube, inverse-DFT along the first axis in LDS and
 *         write G[plane][j][x] (x contiguous, coalesced).
 *   rows: for every pair of output rows (xa, xb) build the Hermitian
 *         extension Z(k) = A(k) + i B(k), Z(Ng-k) = conj A(k) + i conj B(k),
 *         one complex inverse DFT gives row xa in Re and row xb in Im; the
 *         row is written once as Ng floats (coalesced).
 * The j=0 column enters the row pass through its real part, which is the
 * c2r of the Hermitian-symmetrised column (the FFTW/pocketfft behaviour).
 * Radices 2, 3, 4 have register butterflies; any other prime factor runs a
 * direct radix-p stage parallelised over its outputs, so every even Ng
 * (2*109, 2*181, 2*349, ...) is handled without padding.
 */
#include "amb_internal.h"
#include <math.h>
#include <stdio.h>

#include <algorithm>

#define FFT_THREADS 256

void amb_make_plan(int n, FftPlan *pl)
{
    pl->n = n;
    pl->nst = 0;
    int m = n;
    while (m % 4 == 0 && pl->nst < AMB_MAX_RADIX_STAGES) { pl->rad[pl->nst++] = 4; m /= 4; }
    while (m % 2 == 0 && pl->nst < AMB_MAX_RADIX_STAGES) { pl->rad[pl->nst++] = 2; m /= 2; }
    while (m % 3 == 0 && pl->nst < AMB_MAX_RADIX_STAGES) { pl->rad[pl->nst++] = 3; m /= 3; }
    for (int p = 5; m > 1 && pl->nst < AMB_MAX_RADIX_STAGES; p += 2)
        while (m % p == 0) { pl->rad[pl->nst++] = p; m /= p; }
}

/* complex type of a real scalar type: FP32 for the filtered planes, FP64
 * for the calCR images */
template <typename R>
struct CT;
template <>
struct CT<float> { typedef float2 c; };
template <>
struct CT<double> { typedef double2 c; };

__device__ __forceinline__ float2 cmk(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ double2 cmk(double x, double y) { return make_double2(x, y); }
template <typename C>
__device__ __forceinline__ C cmul(C a, C b)
{
    return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { return cmk(a.x + b.x, a.y + b.y); }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { return cmk(a.x - b.x, a.y - b.y); }

/* nb transforms of length n in x[b*n + i]; inverse direction exp(+2 pi i/n).
 * tw[m] = exp(+2 pi i m / n).  Returns the buffer holding the result. */
template <typename T>
__device__ typename CT<T>::c *lds_fft_batch(typename CT<T>::c *x, typename CT<T>::c *y,
                                            const typename CT<T>::c *tw, const FftPlan &pl, int nb)
{
    typedef typename CT<T>::c float2; /* the complex type of this instance */
    const int n = pl.n;
    const int tid = threadIdx.x, nth = blockDim.x;
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.rad[s];
        const int nR = n / R;
        const int tstep = n / (Ns * R);
        if (R == 4) {
            for (int t = tid; t < nb * nR; t += nth) {
                int b = t / nR, j = t - b * nR, k = j % Ns;
                const float2 *xi = x + b * n;
                float2 v0 = xi[j];
                float2 v1 = cmul(xi[j + nR], tw[k * tstep]);
                float2 v2 = cmul(xi[j + 2 * nR], tw[2 * k * tstep]);
                float2 v3 = cmul(xi[j + 3 * nR], tw[3 * k * tstep]);
                float2 s02 = cadd(v0, v2), d02 = csub(v0, v2);
                float2 s13 = cadd(v1, v3), d13 = csub(v1, v3);
                float2 id13 = cmk(-d13.y, d13.x); /* i*(v1-v3) */
                float2 *yo = y + b * n + (j / Ns) * Ns * 4 + k;
                yo[0] = cadd(s02, s13);
                yo[Ns] = cadd(d02, id13);
                yo[2 * Ns] = csub(s02, s13);
                yo[3 * Ns] = csub(d02, id13);
            }
        } else if (R == 2) {
            for (int t = tid; t < nb * nR; t += nth) {
                int b = t / nR, j = t - b * nR, k = j % Ns;
                const float2 *xi = x + b * n;
                float2 v0 = xi[j];
                float2 v1 = cmul(xi[j + nR], tw[k * tstep]);
                float2 *yo = y + b * n + (j / Ns) * Ns * 2 + k;
                yo[0] = cadd(v0, v1);
                yo[Ns] = csub(v0, v1);
            }
        } else if (R == 3) {
            const T h = (T)0.86602540378443864676; /* sqrt(3)/2 */
            for (int t = tid; t < nb * nR; t += nth) {
                int b = t / nR, j = t - b * nR, k = j % Ns;
                const float2 *xi = x + b * n;
                float2 v0 = xi[j];
                float2 v1 = cmul(xi[j + nR], tw[k * tstep]);
                float2 v2 = cmul(xi[j + 2 * nR], tw[2 * k * tstep]);
                float2 sm = cadd(v1, v2), df = csub(v1, v2);
                float2 m = cmk(v0.x - (T)0.5 * sm.x, v0.y - (T)0.5 * sm.y);
                float2 r = cmk(-h * df.y, h * df.x); /* i*h*(v1-v2) */
                float2 *yo = y + b * n + (j / Ns) * Ns * 3 + k;
                yo[0] = cadd(v0, sm);
                yo[Ns] = cadd(m, r);
                yo[2 * Ns] = csub(m, r);
            }
        } else {
            /* direct radix-R stage, one thread per output:
               out[(j/Ns)Ns R + k + q Ns] = sum_r x[j + r nR] w_{Ns R}^{r (k + q Ns)} */
            const int NsR = Ns * R;
            for (int t = tid; t < nb * n; t += nth) {
                int b = t / n, o = t - b * n;
                int g = o / NsR, rem = o - g * NsR;
                int q = rem / Ns, k = rem - q * Ns;
                int j = g * Ns + k;
                int e = k + q * Ns; /* exponent step in units of w_{NsR} */
                const float2 *xi = x + b * n + j;
                float2 acc = cmk((T)0, (T)0);
                int ex = 0;
                for (int r = 0; r < R; ++r) {
                    acc = cadd(acc, cmul(xi[r * nR], tw[ex * tstep]));
                    ex += e;
                    if (ex >= NsR) ex -= NsR;
                }
                y[b * n + o] = acc;
            }
        }
        __syncthreads();
        float2 *tmp = x; x = y; y = tmp;
        Ns *= R;
    }
    return x;
}

/* grid (ceil(jcnt/nb), npl); dynamic LDS: tw[n] + 2 * nb * n complex */
template <typename R>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_cols(
    int kind, const float2 *__restrict__ vis_t, const float *__restrict__ samp_t,
    const int *__restrict__ planes, int N, int Nh, FftPlan pl, const typename CT<R>::c *__restrict__ twg,
    double lo, double hi, int jcnt, int nb, typename CT<R>::c *__restrict__ G)
{
    typedef typename CT<R>::c C;
    extern __shared__ __align__(16) char smem_raw[];
    C *smem = (C *)smem_raw;
    const int n = pl.n;
    C *tw = smem, *bx = smem + n, *by = bx + nb * n;
    const int pli = blockIdx.y;
    const int ch = planes[pli];
    const int j0 = blockIdx.x * nb;
    for (int t = threadIdx.x; t < n; t += blockDim.x) tw[t] = twg[t];
    const int jlim = n / 3; /* column cut j < Ng/3 (amber_funcs.c:582) */
    for (int t = threadIdx.x; t < nb * n; t += blockDim.x) {
        int c = t / n, ing = t - c * n;
        int j = j0 + c;
        int ia = ing <= n / 2 ? ing : ing - n; /* ia in (-Ng/2, Ng/2] */
        C v = cmk((R)0, (R)0);
        if (j < jcnt && j < jlim) {
            double uu = sqrt(1. * (double)(ia * ia + j * j));
            if (uu >= lo && uu < hi && uu > 0.0) {
                int in = ia < 0 ? N + ia : ia;
                size_t src = ((size_t)ch * Nh + j) * N + in;
                if (kind == 0) v = cmk((R)vis_t[src].x, (R)vis_t[src].y);
                else v = cmk((R)samp_t[src], (R)0);
            }
        }
        bx[t] = v;
    }
    __syncthreads();
    C *res = lds_fft_batch<R>(bx, by, tw, pl, nb);
    for (int t = threadIdx.x; t < nb * n; t += blockDim.x) {
        int c = t / n, x = t - c * n;
        int j = j0 + c;
        if (j < jcnt) G[((size_t)pli * jcnt + j) * n + x] = res[t];
    }
}

/* grid (ceil(n/(2 nb)), npl); each block nb transforms = 2 nb output rows.
 * mx (may be null): max |value| over the whole set, as the bit pattern of
 * a non-negative float (one atomic per workgroup). */
template <typename R>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_rows(
    const typename CT<R>::c *__restrict__ G, int jcnt, FftPlan pl, const typename CT<R>::c *__restrict__ twg,
    int nb, R *__restrict__ out, long pitch, unsigned *mx)
{
    typedef typename CT<R>::c float2; /* the complex type of this instance */
    extern __shared__ __align__(16) char smem_raw[];
    float2 *smem = (float2 *)smem_raw;
    const int n = pl.n;
    float2 *tw = smem, *bx = smem + n, *by = bx + nb * n;
    const int pli = blockIdx.y;
    const int x0 = blockIdx.x * 2 * nb;
    for (int t = threadIdx.x; t < n; t += blockDim.x) tw[t] = twg[t];
    const float2 *Gp = G + (size_t)pli * jcnt * n;
    /* b fastest so a warp reads 2*nb consecutive x of one column */
    for (int t = threadIdx.x; t < nb * n; t += blockDim.x) {
        int k = t / nb, b = t - k * nb;
        int xa = x0 + 2 * b, xb = xa + 1;
        float2 z = cmk((R)0, (R)0);
        if (xb < n) {
            if (k == 0) {
                float2 A = Gp[xa], B = Gp[xb];
                z = cmk(A.x, B.x);
            } else if (k < jcnt) {
                float2 A = Gp[(size_t)k * n + xa], B = Gp[(size_t)k * n + xb];
                z = cmk(A.x - B.y, A.y + B.x);
            } else if (n - k < jcnt) {
                int kk = n - k;
                float2 A = Gp[(size_t)kk * n + xa], B = Gp[(size_t)kk * n + xb];
                z = cmk(A.x + B.y, B.x - A.y);
            }
        }
        bx[b * n + k] = z;
    }
    __syncthreads();
    float2 *res = lds_fft_batch<R>(bx, by, tw, pl, nb);
    R *op = out + (size_t)pli * pitch;
    if (blockIdx.x == 0)
        for (long t = (long)n * n + threadIdx.x; t < pitch; t += blockDim.x) op[t] = (R)0;
    float m = 0.f;
    for (int t = threadIdx.x; t < nb * n; t += blockDim.x) {
        int b = t / n, yv = t - b * n;
        int xa = x0 + 2 * b;
        if (xa + 1 < n) {
            float2 z = res[t];
            op[(size_t)xa * n + yv] = z.x;
            op[(size_t)(xa + 1) * n + yv] = z.y;
            m = fmaxf(m, fmaxf(fabsf((float)z.x), fabsf((float)z.y)));
        }
    }
    if (mx) {
        __shared__ float red[FFT_THREADS / 64];
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < FFT_THREADS / 64; ++w) m = fmaxf(m, red[w]);
            atomicMax(mx, __float_as_uint(m));
        }
    }
}

/* Smallest E with m <= 2^E (0 for m == 0 or not finite), clamped to
 * [-60, 60]. */
__device__ __forceinline__ int set_exp(float m)
{
    if (!(m > 0.f) || !(m <= 3.0e38f)) return 0;
    int e;
    const float f = frexpf(m, &e); /* m = f 2^e, f in [0.5, 1) */
    if (f == 0.5f) --e;
    return e < -60 ? -60 : e > 60 ? 60 : e;
}

/* Normalises a data plane set in place: plane i (compact channel index)
 * times (-1)^i 2^(7 - E), E = set_exp(max |x|) from slot[0], so every
 * stored value lies in [-2^7, 2^7]; slot[1] gets E.  Both factors are
 * exact and undone exactly by whoever reads the set.  The alternating sign
 * makes the products of a lag cell's triplets (p1, p1 + d2, p1 + d3) carry
 * (-1)^(p1 + d2 + d3): every other reference plane is accumulated negated,
 * so a rounding bias of the matrix cores that does not follow the sign of
 * its operands cancels in the sum over p1 instead of adding up. */
__global__ void k_normalise_set(float *__restrict__ D, size_t n4, long pitch4, unsigned *slot)
{
    const int e = set_exp(__uint_as_float(slot[0]));
    const float f = ldexpf(1.f, AMB_SET_EXP - e);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        const float g = ((i / (size_t)pitch4) & 1) ? -f : f;
        float4 v = ((float4 *)D)[i];
        v.x *= g; v.y *= g; v.z *= g; v.w *= g;
        ((float4 *)D)[i] = v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) slot[1] = (unsigned)e;
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

static int lds_batch(int n, int cbytes = 8)
{   /* transforms per block with tw + 2 buffers inside 64 KiB (cbytes per complex) */
    int nb = (65536 / (cbytes * n) - 1) / 2;
    if (nb > 8) nb = 8;
    if (nb < 1) nb = 1;
    return nb;
}

hipError_t amb_launch_ring_c2r(hipStream_t s, int kind, const float2 *vis_t, const float *samp_t,
                               const int *d_planes, int npl, int N, int Nh, const FftPlan &pl,
                               const float2 *d_tw, RingGeom g, float2 *G, float *out, long pitch, unsigned *slot)
{
    const int n = pl.n;
    if (npl <= 0) return hipSuccess;
    if (pitch < (long)n * n || pitch % 4 != 0) return hipErrorInvalidValue;
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
    int nb = lds_batch(n);
    size_t lds = sizeof(float2) * (size_t)(n + 2 * nb * n);
    dim3 gc((g.jcnt + nb - 1) / nb, npl);
    hipLaunchKernelGGL(k_fft_cols<float>, gc, dim3(FFT_THREADS), lds, s, kind, vis_t, samp_t, d_planes, N, Nh,
                       pl, d_tw, g.lo, g.hi, g.jcnt, nb, G);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 gr((n + 2 * nb - 1) / (2 * nb), npl);
    hipLaunchKernelGGL(k_fft_rows<float>, gr, dim3(FFT_THREADS), lds, s, G, g.jcnt, pl, d_tw, nb, out, pitch, slot);
    e = hipGetLastError();
    if (e != hipSuccess || !slot) return e;
    const size_t n4 = (size_t)npl * pitch / 4;
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>((n4 + 255) / 256, 8192));
    hipLaunchKernelGGL(k_normalise_set, dim3(blocks), dim3(256), 0, s, out, n4, pitch / 4, slot);
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
    if (pitch < (long)n * n) return hipErrorInvalidValue;
    if (g.jcnt <= 0) {
        hipLaunchKernelGGL(k_zero64, dim3(1024), dim3(256), 0, s, out, (size_t)npl * pitch);
        return hipGetLastError();
    }
    const int nb = lds_batch(n, 16);
    const size_t lds = sizeof(double2) * (size_t)(n + 2 * nb * n);
    if (lds > 65536) return hipErrorInvalidValue;
    dim3 gc((g.jcnt + nb - 1) / nb, npl);
    hipLaunchKernelGGL(k_fft_cols<double>, gc, dim3(FFT_THREADS), lds, s, 1, nullptr, samp_t, d_planes, N, Nh, pl,
                       d_tw, g.lo, g.hi, g.jcnt, nb, G);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 gr((n + 2 * nb - 1) / (2 * nb), npl);
    hipLaunchKernelGGL(k_fft_rows<double>, gr, dim3(FFT_THREADS), lds, s, G, g.jcnt, pl, d_tw, nb, out, pitch,
                       nullptr);
    return hipGetLastError();
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

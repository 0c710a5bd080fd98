/*
 * Lag-space post-processing of bispectrum lag arrays, the consumer side of
 * the result files: get_cyl_BS (mabs_funcs.py:375-388) and the windowed
 * cylbs_win (mfunc_winds.py:113-127).  Per triangle:
 *   fold      F[d2][d3] = B[d2][d3] + B[-d2][-d3] for d2 >= 1
 *             (fold_quadrants, mabs_funcs.py:210-222), float32 adds as the
 *             reference's arrays do, the same for the denominator;
 *   normalise T = fl32(F / fl32(mabs_norm)) / Dfold (float64 for int32
 *             counts, float32 for v2 norms), cut n rows and columns
 *             (cut_nan, :259-260), optional Blackman-Nuttall window
 *             (mfunc_winds.py:16-30, 94-101);
 *   transform fftshift(irfft2(ifftshift(X, axes=1), axes=(1,0), s=(N,N)))
 *             * rc^4 L^2, L = N dnu rc' (Bdnu2Bk_fft, :325-334), in FP64 by
 *             direct sums over a twiddle table: a complex inverse DFT along
 *             the columns, then the half-spectrum inverse along the rows.
 * N = W - 2n is odd (W = 2 nc10 + 1), so the row transform has no Nyquist
 * term; like numpy's c2r it uses only the real part of the m = 0 row.
 * Cost is about N^3 FP64 multiply-adds per triangle (N = 1519: 3.5e9),
 * under a millisecond; the arrays are small enough that HBM is no concern.
 */
#include "amb_ctx.h"

#include <cmath>
#include <vector>

namespace {

/* fold + normalise + cut (+ window), written column-ifftshifted:
 * X[m][k] = T[m][(k + N/2) % N] for m < M = nc10 + 1 - n, k < N */
template <bool INT_DEN>
__global__ void k_cyl_prep(const float *__restrict__ bl, const void *__restrict__ den, int W, int nc10, int ncut,
                           float mn, const double *__restrict__ win, double *__restrict__ X)
{
    const int N = W - 2 * ncut, M = nc10 + 1 - ncut;
    const int k = blockIdx.x * blockDim.x + threadIdx.x, m = blockIdx.y;
    if (k >= N || m >= M) return;
    const int j = (k + N / 2) % N;     /* column of the cut grid */
    const int r = nc10 + m, c = j + ncut; /* row d2 = m, column d3 = c - nc10 */
    float f = bl[(size_t)r * W + c];
    if (m >= 1) f += bl[(size_t)(nc10 - m) * W + (W - 1 - c)];
    const float t = f / mn;
    double v;
    if constexpr (INT_DEN) {
        const int *d = (const int *)den;
        int dv = d[(size_t)r * W + c];
        if (m >= 1) dv += d[(size_t)(nc10 - m) * W + (W - 1 - c)];
        v = (double)t / (double)dv;
    } else {
        const float *d = (const float *)den;
        float dv = d[(size_t)r * W + c];
        if (m >= 1) dv += d[(size_t)(nc10 - m) * W + (W - 1 - c)];
        v = (double)(t / dv);
    }
    if (win) v *= win[(size_t)m * N + j];
    X[(size_t)m * N + k] = v;
}

/* Z[m][k] = (1/N) sum_j X[m][j] e^{+2 pi i j k / N}; a block stages one row
 * and the twiddles in LDS */
__global__ void k_cyl_cols(const double *__restrict__ X, const double2 *__restrict__ tw, int N, int M,
                           double2 *__restrict__ Z)
{
    extern __shared__ double sm[];
    double *row = sm;
    double2 *t = (double2 *)(sm + N);
    const int m = blockIdx.y;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        row[i] = X[(size_t)m * N + i];
        t[i] = tw[i];
    }
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    double re = 0.0, im = 0.0;
    int idx = 0; /* j k mod N */
    for (int j = 0; j < N; ++j) {
        re = fma(row[j], t[idx].x, re);
        im = fma(row[j], t[idx].y, im);
        idx += k;
        if (idx >= N) idx -= N;
    }
    Z[(size_t)m * N + k] = make_double2(re / N, im / N);
}

/* y[tt][k] = (1/N)(Re Z[0][k] + 2 sum_{m>=1} Re(Z[m][k] e^{+2 pi i m tt / N})),
 * written fftshifted on both axes and scaled */
__global__ void k_cyl_rows(const double2 *__restrict__ Z, const double2 *__restrict__ tw, int N, int M, double scale,
                           double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x, tt = blockIdx.y;
    if (k >= N) return;
    double acc = Z[k].x;
    double s2 = 0.0;
    int idx = 0; /* m tt mod N */
    for (int m = 1; m < M; ++m) {
        idx += tt;
        if (idx >= N) idx -= N;
        const double2 z = Z[(size_t)m * N + k], w = tw[idx];
        s2 = fma(z.x, w.x, s2);
        s2 = fma(-z.y, w.y, s2);
    }
    acc = (acc + 2.0 * s2) / N;
    const int a = (tt + N / 2) % N, b = (k + N / 2) % N; /* fftshift: out[a] = y[a - N/2] */
    out[(size_t)a * N + b] = acc * scale;
}

} // namespace

hipError_t amb_launch_cyl(hipStream_t s, const float *bl, const void *den, int den_is_int, int W, int nc10, int ncut,
                          float mabs_norm, const double *win, const double2 *tw, double scale, double *X, double2 *Z,
                          double *out)
{
    const int N = W - 2 * ncut, M = nc10 + 1 - ncut;
    if (N < 1 || M < 1 || W != 2 * nc10 + 1) return hipErrorInvalidValue;
    dim3 g1((N + 255) / 256, M);
    if (den_is_int) hipLaunchKernelGGL(k_cyl_prep<true>, g1, dim3(256), 0, s, bl, den, W, nc10, ncut, mabs_norm, win, X);
    else hipLaunchKernelGGL(k_cyl_prep<false>, g1, dim3(256), 0, s, bl, den, W, nc10, ncut, mabs_norm, win, X);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t lds = sizeof(double) * N + sizeof(double2) * N;
    hipLaunchKernelGGL(k_cyl_cols, g1, dim3(256), lds, s, X, tw, N, M, Z);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    dim3 g2((N + 255) / 256, N);
    hipLaunchKernelGGL(k_cyl_rows, g2, dim3(256), 0, s, Z, tw, N, M, scale, out);
    return hipGetLastError();
}

using namespace amb;

extern "C" int amb_cyl_bispec(amb_ctx *c, int n, int W, const float *bl, const void *den, int den_is_int,
                              const amb_cyl_params *p, double *out)
{
    if (!c || n < 0 || !bl || !den || !p || !out) return AMB_EARG;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    hipSetDevice(c->device);
    if (W < 3 || W % 2 == 0) return fail(c, AMB_EARG, "W=%d must be odd (2 nc10 + 1)", W);
    const int nc10 = (W - 1) / 2, ncut = p->ncut;
    const int N = W - 2 * ncut, M = nc10 + 1 - ncut;
    if (ncut < 0 || M < 1) return fail(c, AMB_EARG, "cut %d leaves no lags of W=%d", ncut, W);
    if (24L * N > 65536) return fail(c, AMB_EARG, "N=%d exceeds the transform's LDS row (2730)", N);
    if (!(p->mabs_norm != 0.0)) return fail(c, AMB_EARG, "mabs_norm is zero");
    const size_t WW = (size_t)W * W, MN = (size_t)M * N, NN = (size_t)N * N;
    /* device scratch: bl, den, X, Z, out, window, twiddles */
    const size_t bytes = sizeof(float) * WW * 2 + sizeof(double) * MN + sizeof(double2) * MN + sizeof(double) * NN +
                         sizeof(double) * MN + sizeof(double2) * N + 8 * 256;
    void *d = nullptr;
    int rc = dev_alloc(c, &d, bytes, "post-processing scratch");
    if (rc) return rc;
    char *q = (char *)d;
    auto take = [&](size_t b) { char *r = q; q += (b + 255) / 256 * 256; return r; };
    float *d_bl = (float *)take(sizeof(float) * WW);
    void *d_den = take(sizeof(float) * WW);
    double *d_X = (double *)take(sizeof(double) * MN);
    double2 *d_Z = (double2 *)take(sizeof(double2) * MN);
    double *d_out = (double *)take(sizeof(double) * NN);
    double *d_win = (double *)take(sizeof(double) * MN);
    double2 *d_tw = (double2 *)take(sizeof(double2) * N);
    std::vector<double2> tw(N);
    for (int t = 0; t < N; ++t) tw[t] = make_double2(cos(2.0 * M_PI * t / N), sin(2.0 * M_PI * t / N));
    std::vector<double> win;
    if (p->window) { /* blackman_nuttall (mfunc_winds.py:16-30), folded half of the outer product */
        std::vector<double> w(N);
        const double a0 = 0.3635819, a1 = 0.4891775, a2 = 0.1365995, a3 = 0.0106411;
        for (int i = 0; i < N; ++i) {
            const double x = (double)i / (N - 1);
            w[i] = a0 - a1 * cos(2 * M_PI * x) + a2 * cos(4 * M_PI * x) - a3 * cos(6 * M_PI * x);
        }
        win.resize(MN);
        for (int m = 0; m < M; ++m)
            for (int j = 0; j < N; ++j) win[(size_t)m * N + j] = w[N / 2 + m] * w[j];
    }
    const double L = N * p->dnu * p->rc_prime;
    const double scale = pow(p->rc, 4) * L * L;
    hipError_t e = hipMemcpyAsync(d_tw, tw.data(), sizeof(double2) * N, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && p->window)
        e = hipMemcpyAsync(d_win, win.data(), sizeof(double) * MN, hipMemcpyHostToDevice, c->stream);
    for (int t = 0; t < n && e == hipSuccess; ++t) {
        e = hipMemcpyAsync(d_bl, bl + WW * t, sizeof(float) * WW, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_den, (const char *)den + sizeof(float) * WW * t, sizeof(float) * WW,
                               hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = amb_launch_cyl(c->stream, d_bl, d_den, den_is_int, W, nc10, ncut, (float)p->mabs_norm,
                               p->window ? d_win : nullptr, d_tw, scale, d_X, d_Z, d_out);
        if (e == hipSuccess)
            e = hipMemcpyAsync(out + NN * t, d_out, sizeof(double) * NN, hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free(c, d, bytes);
    if (e != hipSuccess) return fail(c, AMB_EHIP, "lag-space post-processing: %s", hipGetErrorString(e));
    return AMB_OK;
}

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
 *             * rc^4 L^2, L = N dnu rc' (Bdnu2Bk_fft, :325-334), in FP64 as
 *             two direct-DFT matrix products (rocBLAS DGEMM on the FP64
 *             matrix cores):
 *             [Zr | Zi] = X [cos | sin] / N (the complex inverse DFT along the
 *             columns, e^{+2 pi i j k / N}), then
 *             y = [w cos | -w sin] [Zr ; Zi] / N (the half-spectrum inverse
 *             along the rows, w_0 = 1, w_m = 2).
 * N = W - 2n is odd (W = 2 nc10 + 1), so the row transform has no Nyquist
 * term; like numpy's c2r it uses only the real part of the m = 0 row.  The
 * twiddle matrices are built once per call from one table of e^{2 pi i t/N}
 * indexed by (j k) mod N.  Work: 8 M N^2 flop per triangle (N = 1519,
 * M = 760: 1.4e10).
 */
#include "amb_ctx.h"

#include <rocblas/rocblas.h>

#include <algorithm>
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

/* B1 = [cos | sin](2 pi (j k mod N) / N), N x 2N; A2 = [w cos | -w sin](2 pi
 * (m t mod N) / N), N x 2M (row t, column m), w_0 = 1, w_m = 2 */
__global__ void k_cyl_tables(const double2 *__restrict__ tw, int N, int M, double *__restrict__ B1,
                             double *__restrict__ A2)
{
    const int r = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) {
        const double2 t = tw[(int)(((long)r * k) % N)];
        B1[(size_t)r * 2 * N + k] = t.x;
        B1[(size_t)r * 2 * N + N + k] = t.y;
    }
    if (k < M) {
        const double2 t = tw[(int)(((long)k * r) % N)];
        const double w = k == 0 ? 1.0 : 2.0;
        A2[(size_t)r * 2 * M + k] = w * t.x;
        A2[(size_t)r * 2 * M + M + k] = -(w * t.y);
    }
}

/* [Zr ; Zi] (2M x N) from C1 = X [cos | sin] (M x 2N), divided by N */
__global__ void k_cyl_stack(const double *__restrict__ C1, int N, int M, double *__restrict__ Z2)
{
    const int m = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    Z2[(size_t)m * N + k] = C1[(size_t)m * 2 * N + k] / N;
    Z2[(size_t)(M + m) * N + k] = C1[(size_t)m * 2 * N + N + k] / N;
}

/* y = C2 / N, written fftshifted on both axes and scaled */
__global__ void k_cyl_out(const double *__restrict__ C2, int N, double scale, double *__restrict__ out)
{
    const int t = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int a = (t + N / 2) % N, b = (k + N / 2) % N; /* fftshift: out[a] = y[a - N/2] */
    out[(size_t)a * N + b] = C2[(size_t)t * N + k] / N * scale;
}

} // namespace

namespace amb {

/* C[M][N] = A[M][K] B[K][N], row-major with leading dimensions: two plain
 * FP64 matrix products, on rocBLAS's gfx950 DGEMM (FP64 matrix cores).  In
 * rocBLAS's column-major terms the product is C^T = B^T A^T, so the call
 * passes B first with m = N and n = M.  The handle lives with the context
 * and runs on its stream.  A rocBLAS failure is returned as its status
 * name in *why. */
static hipError_t dgemm(amb_ctx *c, hipStream_t s, const double *A, int lda, const double *B, int ldb, double *C,
                        int ldc, int M, int N, int K, const char **why)
{
    rocblas_status r = rocblas_status_success;
    if (!c->blas) {
        rocblas_handle h = nullptr;
        r = rocblas_create_handle(&h);
        if (r == rocblas_status_success) c->blas = h;
    }
    if (r == rocblas_status_success) r = rocblas_set_stream((rocblas_handle)c->blas, s);
    const double one = 1.0, zero = 0.0;
    if (r == rocblas_status_success)
        r = rocblas_dgemm((rocblas_handle)c->blas, rocblas_operation_none, rocblas_operation_none, N, M, K, &one, B,
                          ldb, A, lda, &zero, C, ldc);
    if (r == rocblas_status_success) return hipGetLastError();
    *why = rocblas_status_to_string(r);
    return hipErrorUnknown;
}

void post_release(amb_ctx *c)
{
    if (c->blas) rocblas_destroy_handle((rocblas_handle)c->blas);
    c->blas = nullptr;
}

} // namespace amb

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
    if (N > 32768) return fail(c, AMB_EARG, "N=%d exceeds the transform tables", N);
    if (!(p->mabs_norm != 0.0)) return fail(c, AMB_EARG, "mabs_norm is zero");
    const size_t WW = (size_t)W * W, MN = (size_t)M * N, NN = (size_t)N * N;
    /* device scratch: bl, den, X, C1 = X [cos|sin], Z2 = [Zr;Zi], C2, out,
       window, the two twiddle matrices, the twiddle table */
    size_t bytes = 0;
    auto sz = [&](size_t b) { bytes += (b + 255) / 256 * 256; };
    sz(sizeof(float) * WW); sz(sizeof(float) * WW); sz(sizeof(double) * MN); sz(sizeof(double) * 2 * MN);
    sz(sizeof(double) * 2 * MN); sz(sizeof(double) * NN); sz(sizeof(double) * 2 * NN); sz(sizeof(double) * MN);
    sz(sizeof(double) * 2 * NN); sz(sizeof(double) * 2 * MN); sz(sizeof(double2) * N);
    void *d = nullptr;
    int rc = dev_alloc(c, &d, bytes, "post-processing scratch");
    if (rc) return rc;
    char *q = (char *)d;
    auto take = [&](size_t b) { char *r = q; q += (b + 255) / 256 * 256; return r; };
    float *d_bl = (float *)take(sizeof(float) * WW);
    void *d_den = take(sizeof(float) * WW);
    double *d_X = (double *)take(sizeof(double) * MN);
    double *d_C1 = (double *)take(sizeof(double) * 2 * MN);
    double *d_Z2 = (double *)take(sizeof(double) * 2 * MN);
    double *d_C2 = (double *)take(sizeof(double) * NN);
    double *d_out = (double *)take(sizeof(double) * 2 * NN); /* two slots */
    double *d_win = (double *)take(sizeof(double) * MN);
    double *d_B1 = (double *)take(sizeof(double) * 2 * NN);
    double *d_A2 = (double *)take(sizeof(double) * 2 * MN);
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
    hipStream_t st = c->stream;
    const char *why = nullptr; /* a rocBLAS status, if that is what failed */
    hipError_t e = hipMemcpyAsync(d_tw, tw.data(), sizeof(double2) * N, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && p->window)
        e = hipMemcpyAsync(d_win, win.data(), sizeof(double) * MN, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cyl_tables, dim3((std::max(N, M) + 255) / 256, N), dim3(256), 0, st, d_tw, N, M, d_B1,
                           d_A2);
        e = hipGetLastError();
    }
    /* Triangle t uses slot t & 1: a pinned host block [bl | den | result] and
       a device result array. Its inputs are staged by host threads while
       the GPU runs triangle t - 1; its result goes to the host on the copy
       stream while the GPU runs triangle t + 1; slot reuse (t + 2) waits for
       that copy (ev_cp), which also orders the reuse of the input block. */
    const size_t in_b = sizeof(float) * WW, slot_b = 2 * in_b + sizeof(double) * NN;
    for (int k = 0; k < 2 && !rc; ++k) rc = grow_host(c, &c->hstage[k], &c->hstage_bytes[k], slot_b);
    if (!rc) rc = copy_stream(c);
    if (rc) {
        hipStreamSynchronize(st);
        dev_free(c, d, bytes);
        return rc;
    }
    auto drain = [&](int t) { /* result of triangle t from its slot to out */
        char *h = (char *)c->hstage[t & 1];
        par_copy({{out + NN * t, h + 2 * in_b, sizeof(double) * NN}});
    };
    const dim3 g1((N + 255) / 256, M), g2((N + 255) / 256, N);
    for (int t = 0; t < n && e == hipSuccess; ++t) {
        const int sl = t & 1;
        char *h = (char *)c->hstage[sl];
        double *d_res = d_out + NN * sl;
        std::vector<std::tuple<void *, const void *, size_t>> cp{
            {h, bl + WW * t, in_b}, {h + in_b, (const char *)den + in_b * t, in_b}};
        if (t >= 2) {
            e = hipEventSynchronize(c->ev_cp[sl]);
            if (e != hipSuccess) break;
            cp.emplace_back(out + NN * (t - 2), h + 2 * in_b, sizeof(double) * NN);
        }
        par_copy(cp);
        e = hipMemcpyAsync(d_bl, h, in_b, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d_den, h + in_b, in_b, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) break;
        const double *win_p = p->window ? d_win : nullptr;
        if (den_is_int)
            hipLaunchKernelGGL(k_cyl_prep<true>, g1, dim3(256), 0, st, d_bl, d_den, W, nc10, ncut,
                               (float)p->mabs_norm, win_p, d_X);
        else
            hipLaunchKernelGGL(k_cyl_prep<false>, g1, dim3(256), 0, st, d_bl, d_den, W, nc10, ncut,
                               (float)p->mabs_norm, win_p, d_X);
        timer_start(c, &c->st.ms_post);
        e = dgemm(c, st, d_X, N, d_B1, 2 * N, d_C1, 2 * N, M, 2 * N, N, &why);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_cyl_stack, g1, dim3(256), 0, st, d_C1, N, M, d_Z2);
            e = dgemm(c, st, d_A2, 2 * M, d_Z2, N, d_C2, N, N, N, 2 * M, &why);
        }
        timer_stop(c);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_cyl_out, g2, dim3(256), 0, st, d_C2, N, scale, d_res);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(c->ev_red, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(c->cstream, c->ev_red, 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h + 2 * in_b, d_res, sizeof(double) * NN, hipMemcpyDeviceToHost, c->cstream);
        if (e == hipSuccess) e = hipEventRecord(c->ev_cp[sl], c->cstream);
    }
    hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es;
    es = hipStreamSynchronize(c->cstream);
    if (e == hipSuccess) e = es;
    if (e == hipSuccess)
        for (int t = std::max(0, n - 2); t < n; ++t) drain(t);
    timers_collect(c);
    c->st.flop_post += 8.0 * M * (double)N * N * n;
    dev_free(c, d, bytes);
    if (e != hipSuccess)
        return fail(c, AMB_EHIP, "lag-space post-processing: %s", why ? why : hipGetErrorString(e));
    return AMB_OK;
}

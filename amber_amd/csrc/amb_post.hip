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
 *             * rc^4 L^2, L = N dnu rc' (Bdnu2Bk_fft, :325-334), in FP64 on
 *             the LDS Stockham core (amb_fftcore.h):
 *             Z[m][j] = sum_k X[m][k] e^{+2 pi i j k / N} (the complex inverse
 *             DFT along the columns, two real rows per complex transform),
 *             then y[t][j] = sum_m w_m Re(Z[m][j] e^{+2 pi i m t / N}) (the
 *             half-spectrum inverse along the rows, w_0 = 1, w_m = 2, as a
 *             length-N transform of the Hermitian extension, two columns per
 *             complex transform), both divided by N.
 * N = W - 2n is odd (W = 2 nc10 + 1, N = 2 M - 1), so the row transform has
 * no Nyquist term; like numpy's c2r it uses only the real part of the
 * m = 0 row.  At N = 1519 = 7^2 31 the radix-7 stages are register codelets
 * and the radix-31 stage a direct one.  Nominal work 5 N log2 N flop per
 * transform, ceil(M/2) + ceil(N/2) transforms per triangle.
 */
#include "amb_ctx.h"
#include "amb_fftcore.h"

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

/* Z[m][j] = sum_k X[m][k] e^{+2 pi i j k / N} for rows m1 = 2 blockIdx.x
   and m2 = m1 + 1 at once: z = X[m1] + i X[m2], T its inverse DFT, then
   Z1(j) = (T(j) + conj T(N-j)) / 2, Z2(j) = (T(j) - conj T(N-j)) / 2i.
   One 64-thread team; dynamic LDS 2 N complex double. */
__global__ __launch_bounds__(64) void k_cyl_rows(const double *__restrict__ X, int M, FftPlan pl,
                                                 const double2 *__restrict__ tw, double2 *__restrict__ Z)
{
    extern __shared__ __align__(16) double2 dsm[];
    const int N = pl.n;
    double2 *bx = dsm, *by = dsm + N;
    const int m1 = 2 * blockIdx.x, m2 = m1 + 1;
    const bool two = m2 < M;
    for (int k = threadIdx.x; k < N; k += 64)
        bx[k] = make_double2(X[(size_t)m1 * N + k], two ? X[(size_t)m2 * N + k] : 0.0);
    __syncthreads();
    const double2 *T = fftc::transform<+1, 19>(bx, by, tw, pl, threadIdx.x, 64);
    for (int j = threadIdx.x; j < N; j += 64) {
        const double2 a = T[j], b = T[j == 0 ? 0 : N - j];
        Z[(size_t)m1 * N + j] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
        if (two) Z[(size_t)m2 * N + j] = make_double2(0.5 * (a.y + b.y), -0.5 * (a.x - b.x));
    }
}

/* y[t][j] for columns ja = 2 blockIdx.x and jb = ja + 1: the Hermitian
   extension Zh(m) = Z[m][j] (m < M; real part only at m = 0),
   conj Z[N - m][j] (m >= M) of both columns as A + i B, one inverse DFT of
   length N; written fftshifted on both axes, times scale / N^2 */
__global__ __launch_bounds__(64) void k_cyl_cols(const double2 *__restrict__ Z, int M, FftPlan pl,
                                                 const double2 *__restrict__ tw, double scale, double *__restrict__ out)
{
    extern __shared__ __align__(16) double2 dsm[];
    const int N = pl.n;
    double2 *bx = dsm, *by = dsm + N;
    const int ja = 2 * blockIdx.x, jb = ja + 1 < N ? ja + 1 : ja;
    for (int m = threadIdx.x; m < N; m += 64) {
        double2 y;
        if (m == 0) {
            y = make_double2(Z[ja].x, Z[jb].x);
        } else if (m < M) {
            const double2 A = Z[(size_t)m * N + ja], B = Z[(size_t)m * N + jb];
            y = make_double2(A.x - B.y, A.y + B.x);
        } else {
            const int mm = N - m;
            const double2 A = Z[(size_t)mm * N + ja], B = Z[(size_t)mm * N + jb];
            y = make_double2(A.x + B.y, B.x - A.y);
        }
        bx[m] = y;
    }
    __syncthreads();
    const double2 *yv = fftc::transform<+1, 19>(bx, by, tw, pl, threadIdx.x, 64);
    const double f = scale / ((double)N * (double)N);
    const int ba = (ja + N / 2) % N, bb = (jb + N / 2) % N;
    for (int t = threadIdx.x; t < N; t += 64) {
        const size_t row = (size_t)((t + N / 2) % N) * N; /* fftshift: out[a] = y[a - N/2] */
        out[row + ba] = yv[t].x * f;
        if (jb != ja) out[row + bb] = yv[t].y * f;
    }
}

} // namespace

namespace amb {

void post_release(amb_ctx *) {}

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
    if (N > 2048) return fail(c, AMB_EARG, "N=%d exceeds the LDS transform size (2048)", N);
    if (!(p->mabs_norm != 0.0)) return fail(c, AMB_EARG, "mabs_norm is zero");
    const size_t WW = (size_t)W * W, MN = (size_t)M * N, NN = (size_t)N * N;
    /* device scratch: bl, den, X, Z (complex, M x N), out (two slots), window */
    size_t bytes = 0;
    auto sz = [&](size_t b) { bytes += (b + 255) / 256 * 256; };
    sz(sizeof(float) * WW); sz(sizeof(float) * WW); sz(sizeof(double) * MN); sz(sizeof(double2) * MN);
    sz(sizeof(double) * 2 * NN); sz(sizeof(double) * MN);
    FftPlan pl;
    fftc::make_plan(N, &pl);
    double2 *d_tw = nullptr;
    int rc = twiddles64(c, N, &d_tw);
    if (rc) return rc;
    void *d = nullptr;
    if ((rc = dev_alloc(c, &d, bytes, "post-processing scratch"))) return rc;
    char *q = (char *)d;
    auto take = [&](size_t b) { char *r = q; q += (b + 255) / 256 * 256; return r; };
    float *d_bl = (float *)take(sizeof(float) * WW);
    void *d_den = take(sizeof(float) * WW);
    double *d_X = (double *)take(sizeof(double) * MN);
    double2 *d_Z = (double2 *)take(sizeof(double2) * MN);
    double *d_out = (double *)take(sizeof(double) * 2 * NN); /* two slots */
    double *d_win = (double *)take(sizeof(double) * MN);
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
    hipError_t e = hipSuccess;
    if (p->window) e = hipMemcpyAsync(d_win, win.data(), sizeof(double) * MN, hipMemcpyHostToDevice, st);
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
    const dim3 g1((N + 255) / 256, M);
    const size_t lds = sizeof(double2) * 2 * (size_t)N;
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
        hipLaunchKernelGGL(k_cyl_rows, dim3((unsigned)((M + 1) / 2)), dim3(64), lds, st, d_X, M, pl, d_tw, d_Z);
        hipLaunchKernelGGL(k_cyl_cols, dim3((unsigned)((N + 1) / 2)), dim3(64), lds, st, d_Z, M, pl, d_tw, scale,
                           d_res);
        e = hipGetLastError();
        timer_stop(c);
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
    /* nominal 5 N log2 N flop per length-N transform */
    c->st.flop_post += 5.0 * N * log2((double)N) * (double)((M + 1) / 2 + (N + 1) / 2) * n;
    dev_free(c, d, bytes);
    if (e != hipSuccess) return fail(c, AMB_EHIP, "lag-space post-processing: %s", hipGetErrorString(e));
    return AMB_OK;
}

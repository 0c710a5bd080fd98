/*
 * Triple-product contraction for gfx950 (the hot loop of amber.c:233-283 and
 * amber_v2.c:229-288), recast as one dense product per reference plane p1:
 *
 *   C[p1][p2][p3] = sum_x (D_r1[p1](x) * D_r2[p2](x)) * D_r3[p3](x)
 *
 * i.e. (D_r1[p1] o D_r2) . D_r3^T over pixels, for p2, p3 in the lag window
 * of p1.  A workgroup owns a 128x128 (p2, p3) tile of one p1; its 4 waves
 * each hold 64x64 of the tile in MFMA accumulators (2x2 blocks of 32x32).
 * The Hadamard product with D_r1[p1] is formed while staging the D_r2 rows
 * into LDS, so the A operand is never materialised in HBM.  Workgroups are
 * placed so that the ones resident on an XCD share their (p2, p3) tile and
 * differ only in p1: the D_r2/D_r3 K-chunks they stream are then L2 hits.
 *
 * Precisions (the accumulate is FP32 in every case):
 *   FP32   v_mfma_f32_32x32x2_f32 on FP32 operands
 *   BF16x3 operands split a = hi + lo (hi = bf16(a), lo = bf16(a - hi));
 *          3 passes lo*hi + hi*lo + hi*hi of v_mfma_f32_32x32x16_bf16
 *   FP8    e4m3 with per-plane absmax scales, v_mfma_scale_f32_32x32x64_f8f6f4
 *
 * The sums over p1 (amber.c:265, private_bldnu[idx] += Bispec) happen in the
 * reduction kernel below, in ascending p1 order with compensated FP32
 * summation, so a result never depends on scheduling or batch composition.
 * The sampling contraction (v2 numdnu, amber_v2.c:264,268) reuses the GEMM
 * with an epilogue that rounds S/P to the nearest 1/8 (the exact value is a
 * count of closed sampled triangles, a multiple of 1/8 even when a j=0 column
 * is not Hermitian) and records the worst rounding distance as a certificate.
 */
#include "amb_internal.h"
#include "amber_mi355x.h"

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

#define TM 128
#define TN 128
#define KT 32

template <int MODE, int NORM>
__global__ __launch_bounds__(256) void k_contract(ContractArgs a, int nt3, int per_xcd, int nwork)
{
    const int b = blockIdx.x;
    const int s = b >> 3;
    const int w = (b & 7) * per_xcd + s;
    if (s >= per_xcd || w >= nwork) return;
    const int nact = a.nact;
    const int t = w / nact, i1 = w - t * nact;
    const int t2 = t / nt3, t3 = t - t2 * nt3;
    const int r0 = t2 * TM, c0 = t3 * TN;
    const int lo = a.wlo[i1], hi = a.whi[i1];
    if (r0 >= hi || r0 + TM <= lo || c0 >= hi || c0 + TN <= lo) return;

    const long P = a.P;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int srow = tid >> 3, kq = (tid & 7) * 4;

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

    const float *d1p = a.D1 + (size_t)i1 * P;

    if constexpr (MODE == AMB_PREC_FP32) {
        __shared__ float As[TM][KT + 4];
        __shared__ float Bs[TN][KT + 4];
        for (long k0 = 0; k0 < P; k0 += KT) {
            const long kk = k0 + kq;
            const bool kin = kk < P;
            float4 d1 = kin ? *(const float4 *)(d1p + kk) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int pass = 0; pass < 4; ++pass) {
                const int row = pass * 32 + srow;
                const int ga = r0 + row, gb = c0 + row;
                float4 av = make_float4(0.f, 0.f, 0.f, 0.f), bv = av;
                if (kin && ga < nact) av = *(const float4 *)(a.D2 + (size_t)ga * P + kk);
                if (kin && gb < nact) bv = *(const float4 *)(a.D3 + (size_t)gb * P + kk);
                As[row][kq] = av.x * d1.x; As[row][kq + 1] = av.y * d1.y;
                As[row][kq + 2] = av.z * d1.z; As[row][kq + 3] = av.w * d1.w;
                Bs[row][kq] = bv.x; Bs[row][kq + 1] = bv.y; Bs[row][kq + 2] = bv.z; Bs[row][kq + 3] = bv.w;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < KT; k += 2) {
                float af[2], bf[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) af[i] = As[wm * 64 + i * 32 + (lane & 31)][k + (lane >> 5)];
#pragma unroll
                for (int j = 0; j < 2; ++j) bf[j] = Bs[wn * 64 + j * 32 + (lane & 31)][k + (lane >> 5)];
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[i], bf[j], acc[i][j], 0, 0, 0);
            }
            __syncthreads();
        }
    } else {
        /* BF16x3: hi/lo planes of A = D1 o D2 and of B = D3 */
        __shared__ __attribute__((aligned(16))) __bf16 Ah[TM][KT + 8];
        __shared__ __attribute__((aligned(16))) __bf16 Al[TM][KT + 8];
        __shared__ __attribute__((aligned(16))) __bf16 Bh[TN][KT + 8];
        __shared__ __attribute__((aligned(16))) __bf16 Bl[TN][KT + 8];
        for (long k0 = 0; k0 < P; k0 += KT) {
            const long kk = k0 + kq;
            const bool kin = kk < P;
            float4 d1 = kin ? *(const float4 *)(d1p + kk) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int pass = 0; pass < 4; ++pass) {
                const int row = pass * 32 + srow;
                const int ga = r0 + row, gb = c0 + row;
                float4 av = make_float4(0.f, 0.f, 0.f, 0.f), bv = av;
                if (kin && ga < nact) av = *(const float4 *)(a.D2 + (size_t)ga * P + kk);
                if (kin && gb < nact) bv = *(const float4 *)(a.D3 + (size_t)gb * P + kk);
                float x[4] = {av.x * d1.x, av.y * d1.y, av.z * d1.z, av.w * d1.w};
                float y[4] = {bv.x, bv.y, bv.z, bv.w};
                bf16x4 xh, xl, yh, yl;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    __bf16 h = (__bf16)x[e];
                    xh[e] = h;
                    xl[e] = (__bf16)(x[e] - (float)h);
                    __bf16 g = (__bf16)y[e];
                    yh[e] = g;
                    yl[e] = (__bf16)(y[e] - (float)g);
                }
                *(bf16x4 *)&Ah[row][kq] = xh;
                *(bf16x4 *)&Al[row][kq] = xl;
                *(bf16x4 *)&Bh[row][kq] = yh;
                *(bf16x4 *)&Bl[row][kq] = yl;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < KT; k += 16) {
                bf16x8 ah[2], al[2], bh[2], bl[2];
                const int kc = k + (lane >> 5) * 8;
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const int r = wm * 64 + i * 32 + (lane & 31);
                    ah[i] = *(const bf16x8 *)&Ah[r][kc];
                    al[i] = *(const bf16x8 *)&Al[r][kc];
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int r = wn * 64 + j * 32 + (lane & 31);
                    bh[j] = *(const bf16x8 *)&Bh[r][kc];
                    bl[j] = *(const bf16x8 *)&Bl[r][kc];
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                    }
            }
            __syncthreads();
        }
    }

    /* epilogue: C layout of a 32x32 MFMA block, lane l register v holds
       row (v/4)*8 + (l/32)*4 + v%4, column l%32 */
    const double invP = 1.0 / (double)P;
    float worst = 0.f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int gi = r0 + wm * 64 + i * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                const int gj = c0 + wn * 64 + j * 32 + (lane & 31);
                if (gi < lo || gi >= hi || gj < lo || gj >= hi) continue;
                const size_t idx = ((size_t)i1 * nact + gi) * nact + gj;
                if constexpr (NORM) {
                    double val = (double)acc[i][j][v] * invP;
                    double q8 = rint(val * 8.0);
                    float d = (float)fabs(val - q8 * 0.125);
                    worst = fmaxf(worst, d);
                    a.Cq[idx] = (int)q8;
                } else {
                    a.C[idx] = acc[i][j][v];
                }
            }
    if constexpr (NORM) {
        /* non-negative floats order like their bit patterns */
        if (worst > 0.f) atomicMax(a.margin, __float_as_uint(worst));
    }
}

hipError_t amb_launch_contract(hipStream_t s, int prec, int norm, const ContractArgs &a)
{
    const int nt2 = (a.nact + TM - 1) / TM, nt3 = (a.nact + TN - 1) / TN;
    const int nwork = nt2 * nt3 * a.nact;
    const int per_xcd = (nwork + 7) / 8;
    dim3 grid(8 * per_xcd), block(256);
    if (a.P % 4 != 0) return hipErrorInvalidValue;
    if (prec == AMB_PREC_FP32) {
        if (norm) hipLaunchKernelGGL((k_contract<AMB_PREC_FP32, 1>), grid, block, 0, s, a, nt3, per_xcd, nwork);
        else hipLaunchKernelGGL((k_contract<AMB_PREC_FP32, 0>), grid, block, 0, s, a, nt3, per_xcd, nwork);
    } else {
        if (norm) hipLaunchKernelGGL((k_contract<AMB_PREC_BF16X3, 1>), grid, block, 0, s, a, nt3, per_xcd, nwork);
        else hipLaunchKernelGGL((k_contract<AMB_PREC_BF16X3, 0>), grid, block, 0, s, a, nt3, per_xcd, nwork);
    }
    return hipGetLastError();
}

/* Lag reduction: cell (d2, d3) sums C[p1][p1+d2][p1+d3] over valid p1 in
 * ascending order with Neumaier-compensated FP32 (data) and exact int64
 * (normalisation, in units of 1/8), and counts valid p1 (v1 numdnu,
 * amber.c:266). */
__global__ void k_reduce(ReduceArgs a, const int *__restrict__ first_ge)
{
    const int W = 2 * a.nc10 + 1;
    const long cells = (long)W * W;
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    const int d2 = (int)(cell / W) - a.nc10, d3 = (int)(cell % W) - a.nc10;
    int plo = 0;
    if (-d2 > plo) plo = -d2;
    if (-d3 > plo) plo = -d3;
    int phi = a.nc;
    if (a.nc - d2 < phi) phi = a.nc - d2;
    if (a.nc - d3 < phi) phi = a.nc - d3;
    float sum = 0.f, comp = 0.f;
    long long q = 0;
    int n = 0;
    const size_t na = (size_t)a.nact;
    if (plo < phi) {
        for (int i1 = first_ge[plo]; i1 < a.nact; ++i1) {
            const int p1 = a.act[i1];
            if (p1 >= phi) break;
            const int j2 = a.inv[p1 + d2], j3 = a.inv[p1 + d3];
            if (j2 < 0 || j3 < 0) continue;
            const size_t idx = ((size_t)i1 * na + j2) * na + j3;
            if (a.C) {
                const float x = a.C[idx];
                const float tsum = sum + x;
                if (fabsf(sum) >= fabsf(x)) comp += (sum - tsum) + x;
                else comp += (x - tsum) + sum;
                sum = tsum;
            }
            if (a.Cq) q += a.Cq[idx];
            ++n;
        }
    }
    if (a.bl) a.bl[cell] = sum + comp;
    if (a.bl64) a.bl64[cell] = (double)sum + (double)comp;
    if (a.cnt) a.cnt[cell] = n;
    if (a.Cq) {
        const double v = (double)a.P * ((double)q * 0.125);
        if (a.norm) a.norm[cell] = (float)v;
        if (a.norm64) a.norm64[cell] = v;
    }
}

hipError_t amb_launch_reduce(hipStream_t s, const ReduceArgs &a)
{
    /* first_ge is appended after inv by the caller: inv[nc] then first_ge[nc+1] */
    const int W = 2 * a.nc10 + 1;
    const long cells = (long)W * W;
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, a, a.inv + a.nc);
    return hipGetLastError();
}

/* calCR triple sums (amber_funcs.c:451-456): one block per triangle, FP64
 * accumulation in a fixed thread/tree order. */
__global__ void k_triple_sum(const float *__restrict__ II, long P, const int *__restrict__ tri, double *out)
{
    __shared__ double red[256];
    const int t = blockIdx.x;
    const float *a = II + (size_t)tri[3 * t] * P, *b = II + (size_t)tri[3 * t + 1] * P,
                *c = II + (size_t)tri[3 * t + 2] * P;
    double s = 0.0;
    for (long x = threadIdx.x; x < P; x += blockDim.x) s += (double)a[x] * (double)b[x] * (double)c[x];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[t] = red[0];
}

hipError_t amb_launch_triple_sum(hipStream_t s, const float *II, long P, const int *tri, int ntri, double *out)
{
    if (ntri <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_triple_sum, dim3(ntri), dim3(256), 0, s, II, P, tri, out);
    return hipGetLastError();
}

__global__ void k_absmax(const float *__restrict__ D, long P, float *out)
{
    __shared__ float red[256];
    const float *p = D + (size_t)blockIdx.x * P;
    float m = 0.f;
    for (long x = threadIdx.x; x < P; x += blockDim.x) m = fmaxf(m, fabsf(p[x]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

hipError_t amb_launch_absmax(hipStream_t s, const float *D, long P, int npl, float *out)
{
    if (npl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_absmax, dim3(npl), dim3(256), 0, s, D, P, out);
    return hipGetLastError();
}

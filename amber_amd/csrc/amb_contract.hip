/*
 * Triple-product contraction for gfx950 (the hot loop of amber.c:233-283 and
 * amber_v2.c:229-288), recast as one dense product per reference plane p1:
 *
 *   C[p1][p2][p3] = sum_x (D_r1[p1](x) * D_r2[p2](x)) * D_r3[p3](x)
 *
 * i.e. (D_r1[p1] o D_r2) . D_r3^T over pixels, for p2, p3 in the lag window
 * of p1 (compact channel indices).
 *
 * Work decomposition.  A workgroup owns one 128x128 (p2, p3) tile and 4
 * reference planes p1; each plane's 128x128 block lives in the MFMA
 * accumulators of one wave (4x4 blocks of 32x32) or of two waves (2x4 each,
 * the 8-wave variant).  Per 32-pixel chunk the
 * workgroup stages the D_r2 rows (FP32), the D_r3 rows (pre-split hi/lo BF16
 * or FP32) and its 4 D_r1 rows into LDS once; every wave then forms its own
 * A operand D_r1[p1] o D_r2 in registers (multiply, split) and runs the MFMAs
 * against the shared B operand, so the staged bytes feed 4x the MFMA work of
 * a one-plane tile.  LDS is double-buffered and the next chunk's global loads
 * are in flight while the current chunk is multiplied.  Workgroups are
 * numbered tile-major and dealt to the 8 XCDs in contiguous runs, so the
 * workgroups resident on one XCD share their (p2, p3) tile and stream the
 * same D_r2 / D_r3 chunks through that XCD's L2.  The BF16x3 data and
 * sampling contractions run the pipelined form further down by default
 * (64-pixel chunks loaded straight into LDS, k-steps streamed across chunk
 * boundaries), with the same arithmetic in the same order.
 *
 * Precisions (the accumulate is FP32 in every case):
 *   FP32   v_mfma_f32_32x32x2_f32 on FP32 operands
 *   BF16x3 a = hi + lo (hi = bf16(a), lo = bf16(a - hi)); three passes
 *          lo*hi + hi*lo + hi*hi of v_mfma_f32_32x32x16_bf16
 *   FP8    e4m3 operands with per-plane power-of-two absmax scales: the A
 *          operand D_r1[p1] o D_r2[p2] is quantised as
 *          (D_r1 s1[p1]) (D_r2 s2[p2]) with s = 2^floor(log2(sqrt(448)/max|D|)),
 *          the B operand as D_r3 t[p3] with t = 2^floor(log2(448/max|D|)),
 *          one v_mfma_f32_32x32x16_fp8_fp8 per 16 pixels, and the epilogue
 *          divides by s1 s2 t exactly
 *
 * The sums over p1 (amber.c:265, private_bldnu[idx] += Bispec) happen in the
 * reduction kernel below, in ascending p1 order with compensated FP32
 * summation, so a result never depends on scheduling or batch composition.
 * The sampling contraction (v2 numdnu, amber_v2.c:264,268) reuses the GEMM
 * with an epilogue that rounds S/P to the nearest 1/8 (the exact value is a
 * count of closed sampled triangles, a multiple of 1/8 even when a j=0 column
 * is not Hermitian) and records the worst rounding distance as a certificate.
 *
 * Planes are stored with a pitch that is a multiple of 32 elements and a
 * zero tail, so every chunk load is aligned and in bounds.
 */
#include "amb_internal.h"
#include "amber_mi355x.h"

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

/* hi/lo split forms and their 16-bit element types */
template <int MODE>
constexpr bool is_split() { return MODE == AMB_PREC_BF16X3 || MODE == AMB_PREC_FP16X3; }
/* forms whose operands carry per-plane power-of-two scales */
template <int MODE>
constexpr bool is_scaled() { return MODE == AMB_PREC_FP8; }
/* Unscale factor of a data triangle: the stored planes of r1, r2, r3 are
 * (-1)^i D 2^(AMB_SET_EXP - E) (exactly), and the FP16x3 form's B operand
 * carries another 2^AMB_F16_B_EXP, so accumulators times this power of two,
 * and the sign of the element's three planes, are the sums of the planes
 * as computed, bit for bit. */
template <int MODE>
__device__ __forceinline__ float data_unscale(const ContractArgs &a)
{
    const int e = *a.ex1 + *a.ex2 + *a.ex3 - 3 * AMB_SET_EXP - (MODE == AMB_PREC_FP16X3 ? AMB_F16_B_EXP : 0);
    return ldexpf(1.f, e);
}

#define QP 4     /* reference planes per workgroup */
#define TT 128   /* p2 rows = p3 columns per workgroup */
#define KC 32    /* pixels per LDS chunk */
#define SEGC 256 /* chunks per accumulator segment (8192 pixels) */
#define PADF 4   /* FP32 row padding (floats) */
#define PADH 8   /* BF16 row padding (elements) */

/* Geometry: NWR = 2 (the launched form) runs 8 waves, wave w holding
 * plane w%4 and rows 64*(w/4)..+64 (2x4 blocks of 32x32, 128 accumulators
 * per lane).  NWR = 1 (4 waves of 4x4 blocks) needs all 256 AGPRs for
 * accumulators and spills once the segment flush is in the loop, so it
 * is not launched.
 * Variants (var):
 *   0: PIPE = 0, a k-step forms all its fragments, then multiplies
 *   1: PIPE = 1, (BF16x3) a k-step forms the A fragment of row block i+1
 *      while the 12 MFMAs of block i issue; scheduling hints interleave
 *      MFMA with the LDS reads and VALU work (other forms: as var 0)
 *   2: the split forms' k_contract_v2 / k_contract_grouped_v2 below (direct
 *      global->LDS loads, pipelined k-steps)
 * All three sum the same products in the same order.  The launch
 * configuration table of the context (amb_tune.hip) picks the variant per
 * shape bucket. */
template <int NWR>
struct Geo {
    static constexpr int NT = 256 * NWR; /* threads */
    static constexpr int RB = 4 / NWR;   /* 32-row blocks per wave */
    static constexpr int NA = 1024 / NT; /* float4 of a 128x32 FP32 tile per thread */
    static constexpr int NH = 512 / NT;  /* uint4 of a 128x32 BF16 tile per thread */
};

#define A16 __attribute__((aligned(16)))

template <int MODE>
struct Stage;

template <>
struct A16 Stage<AMB_PREC_BF16X3> {
    A16 float a[TT][KC + PADF];         /* D_r2 rows */
    A16 __bf16 bh[TT][KC + PADH];       /* D_r3 hi */
    A16 __bf16 bl[TT][KC + PADH];       /* D_r3 lo */
    A16 float d1[QP][KC];               /* D_r1 rows of the 4 planes */
};

template <>
struct A16 Stage<AMB_PREC_FP16X3> {
    A16 float a[TT][KC + PADF];         /* D_r2 rows */
    A16 _Float16 bh[TT][KC + PADH];     /* D_r3 hi (scaled) */
    A16 _Float16 bl[TT][KC + PADH];     /* D_r3 lo (scaled) */
    A16 float d1[QP][KC];               /* D_r1 rows of the 4 planes */
};

template <>
struct A16 Stage<AMB_PREC_FP32> {
    A16 float a[TT][KC + PADF];
    A16 float b[TT][KC + PADF];
    A16 float d1[QP][KC];
};

template <>
struct A16 Stage<AMB_PREC_FP8> {
    A16 float a[TT][KC + PADF];         /* D_r2 rows times s2 */
    A16 uint8_t b[TT][KC + 16];         /* D_r3 rows times t3, e4m3 */
    A16 float d1[QP][KC];               /* D_r1 rows times s1 */
};

/* global -> register prefetch of one chunk (per thread) */
template <int MODE, int NWR>
struct Pref {
    float4 a[Geo<NWR>::NA];
    float4 b[Geo<NWR>::NA];    /* FP32 form */
    uint4 bh[Geo<NWR>::NH];    /* BF16x3 form */
    uint4 bl[Geo<NWR>::NH];
    uint4 q;                   /* FP8 form */
    float4 d1;
};

__device__ __forceinline__ float4 ldg4(const float *p) { return *(const float4 *)p; }
__device__ __forceinline__ uint4 ldg4u(const void *p) { return *(const uint4 *)p; }

/* thread -> staged element maps (NT threads):
 *   FP32 rows (D_r2; D_r3 in the FP32 form): 128 rows x 8 float4,
 *     row = p*(NT/8) + tid/8, float4 slot tid%8;
 *   BF16 rows: 128 rows x 4 uint4 per half, row = p*(NT/4) + tid/4, slot tid%4;
 *   e4m3 rows: 128 rows x 2 uint4, threads 0..255, row = tid/2, slot tid%2;
 *   D_r1: 4 rows x 8 float4, threads 0..31. */
template <int MODE, int NWR>
__device__ __forceinline__ void prefetch(Pref<MODE, NWR> &r, const ContractArgs &a, long k0, int r0, int c0,
                                         int g0, int tid)
{
    using G = Geo<NWR>;
    const long pitch = a.pitch;
    const int nact = a.nact;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    const uint4 zu = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int p = 0; p < G::NA; ++p) {
        const int ga = r0 + p * (G::NT / 8) + (tid >> 3);
        r.a[p] = ga < nact ? ldg4(a.D2 + (size_t)ga * pitch + k0 + (tid & 7) * 4) : z;
    }
    if constexpr (is_split<MODE>()) {
#pragma unroll
        for (int p = 0; p < G::NH; ++p) {
            const int gb = c0 + p * (G::NT / 4) + (tid >> 2);
            const size_t off = (size_t)gb * pitch + k0 + (tid & 3) * 8;
            r.bh[p] = gb < nact ? ldg4u(a.H3 + off) : zu;
            r.bl[p] = gb < nact ? ldg4u(a.L3 + off) : zu;
        }
    } else if constexpr (MODE == AMB_PREC_FP32) {
#pragma unroll
        for (int p = 0; p < G::NA; ++p) {
            const int gb = c0 + p * (G::NT / 8) + (tid >> 3);
            r.b[p] = gb < nact ? ldg4(a.D3 + (size_t)gb * pitch + k0 + (tid & 7) * 4) : z;
        }
    } else {
        const int gb = c0 + (tid >> 1);
        r.q = (tid < 256 && gb < nact) ? ldg4u(a.Q3 + (size_t)gb * pitch + k0 + (tid & 1) * 16) : zu;
    }
    if (tid < QP * KC / 4) {
        const int g = g0 + (tid >> 3);
        r.d1 = g < nact ? ldg4(a.D1 + (size_t)g * pitch + k0 + (tid & 7) * 4) : z;
    }
}

template <int MODE, int NWR>
__device__ __forceinline__ void stage_store(Stage<MODE> &s, const Pref<MODE, NWR> &r, int tid, const float *rsc,
                                            const float *s1v)
{
    using G = Geo<NWR>;
#pragma unroll
    for (int p = 0; p < G::NA; ++p) {
        const int row = p * (G::NT / 8) + (tid >> 3);
        float4 v = r.a[p];
        if constexpr (is_scaled<MODE>()) { /* fold the plane scale s2 in */
            const float f = rsc[row];
            v.x *= f; v.y *= f; v.z *= f; v.w *= f;
        }
        *(float4 *)&s.a[row][(tid & 7) * 4] = v;
    }
    if constexpr (is_split<MODE>()) {
#pragma unroll
        for (int p = 0; p < G::NH; ++p) {
            const int row = p * (G::NT / 4) + (tid >> 2);
            *(uint4 *)&s.bh[row][(tid & 3) * 8] = r.bh[p];
            *(uint4 *)&s.bl[row][(tid & 3) * 8] = r.bl[p];
        }
    } else if constexpr (MODE == AMB_PREC_FP32) {
#pragma unroll
        for (int p = 0; p < G::NA; ++p) *(float4 *)&s.b[p * (G::NT / 8) + (tid >> 3)][(tid & 7) * 4] = r.b[p];
    } else {
        if (tid < 256) *(uint4 *)&s.b[tid >> 1][(tid & 1) * 16] = r.q;
    }
    if (tid < QP * KC / 4) {
        float4 v = r.d1;
        if constexpr (is_scaled<MODE>()) { /* and s1 of the reference plane */
            const float f = s1v[tid >> 3];
            v.x *= f; v.y *= f; v.z *= f; v.w *= f;
        }
        *(float4 *)&s.d1[tid >> 3][(tid & 7) * 4] = v;
    }
}

/* Operand fragments of one k-step (16 pixels; 8 for FP32 which runs four
 * k=2 MFMAs per 8 pixels and lane half) for RB row blocks and 4 column
 * blocks. */
template <int MODE, int RB>
struct Frag;
template <int RB>
struct Frag<AMB_PREC_BF16X3, RB> {
    bf16x8 ah[RB], al[RB], bh[4], bl[4];
};
template <int RB>
struct Frag<AMB_PREC_FP16X3, RB> {
    f16x8 ah[RB], al[RB], bh[4], bl[4];
};
template <int RB>
struct Frag<AMB_PREC_FP8, RB> {
    long aq[RB], bq[4];
};
template <int RB>
struct Frag<AMB_PREC_FP32, RB> {
    float av[RB][4], bv[4][4];
};

/* k-steps per chunk */
template <int MODE>
constexpr int ksteps() { return MODE == AMB_PREC_FP32 ? 4 : 2; }

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

/* x*d split into hi + lo BF16 for 8 elements, two at a time: hi rounds to
 * nearest even, lo is the BF16 of the exact residual x*d - hi (one FMA).
 * Converting pairs needs one v_cvt_pk_bf16_f32 per two values, and the
 * FP32 value of each hi comes back from the packed word by a shift or a
 * mask. */
__device__ __forceinline__ void split8(const float (&x)[8], const float (&d)[8], bf16x8 &hi, bf16x8 &lo)
{
    u32x4 hp, lp;
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        const f32x2 v = {x[e] * d[e], x[e + 1] * d[e + 1]};
        const unsigned hb = __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
        const float h0 = __uint_as_float(hb << 16), h1 = __uint_as_float(hb & 0xffff0000u);
        const f32x2 r = {fmaf(x[e], d[e], -h0), fmaf(x[e + 1], d[e + 1], -h1)};
        hp[e / 2] = hb;
        lp[e / 2] = __builtin_bit_cast(unsigned, __builtin_convertvector(r, bf16x2));
    }
    hi = __builtin_bit_cast(bf16x8, hp);
    lo = __builtin_bit_cast(bf16x8, lp);
}

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

/* 8-element 16-bit operand vector of a split form */
template <int MODE>
struct HalfT;
template <>
struct HalfT<AMB_PREC_BF16X3> { typedef bf16x8 v8; };
template <>
struct HalfT<AMB_PREC_FP16X3> { typedef f16x8 v8; };

/* x*d split into hi + lo FP16 for 8 elements, two at a time: the FP32
 * product v rounds to nearest into hi (v_cvt_pk_f16_f32); v - hi is exact
 * (v_fma_mix_f32 reads hi as FP16) and rounds to nearest into lo.  The
 * stored planes are normalised to |x|, |d| <= 2^7, so |x*d| <= 2^14 lies
 * inside FP16's range. */
__device__ __forceinline__ void split8(const float (&x)[8], const float (&d)[8], f16x8 &hi, f16x8 &lo)
{
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        const f32x2 v = {x[e] * d[e], x[e + 1] * d[e + 1]};
        const f16x2 h = __builtin_convertvector(v, f16x2);
        const f32x2 r = {v[0] - (float)h[0], v[1] - (float)h[1]};
        const f16x2 l = __builtin_convertvector(r, f16x2);
        hi[e] = h[0]; hi[e + 1] = h[1];
        lo[e] = l[0]; lo[e + 1] = l[1];
    }
}

template <int MODE, int RB>
__device__ __forceinline__ void load_frag(const Stage<MODE> &s, int step, int q, int rb, int lane, Frag<MODE, RB> &f)
{
    const int rl = lane & 31, hl = lane >> 5;
    if constexpr (is_split<MODE>()) {
        using V = typename HalfT<MODE>::v8;
        const int kc = step * 16 + hl * 8;
        float d1v[8];
        *(float4 *)&d1v[0] = *(const float4 *)&s.d1[q][kc];
        *(float4 *)&d1v[4] = *(const float4 *)&s.d1[q][kc + 4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            f.bh[j] = *(const V *)&s.bh[j * 32 + rl][kc];
            f.bl[j] = *(const V *)&s.bl[j * 32 + rl][kc];
        }
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            float x[8];
            *(float4 *)&x[0] = *(const float4 *)&s.a[rb + i * 32 + rl][kc];
            *(float4 *)&x[4] = *(const float4 *)&s.a[rb + i * 32 + rl][kc + 4];
            split8(x, d1v, f.ah[i], f.al[i]);
        }
    } else if constexpr (MODE == AMB_PREC_FP8) {
        const int kc = step * 16 + hl * 8;
        float d1v[8];
        *(float4 *)&d1v[0] = *(const float4 *)&s.d1[q][kc];
        *(float4 *)&d1v[4] = *(const float4 *)&s.d1[q][kc + 4];
#pragma unroll
        for (int j = 0; j < 4; ++j) f.bq[j] = *(const long *)&s.b[j * 32 + rl][kc];
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const float4 x0 = *(const float4 *)&s.a[rb + i * 32 + rl][kc];
            const float4 x1 = *(const float4 *)&s.a[rb + i * 32 + rl][kc + 4];
            int lo4 = __builtin_amdgcn_cvt_pk_fp8_f32(x0.x * d1v[0], x0.y * d1v[1], 0, false);
            lo4 = __builtin_amdgcn_cvt_pk_fp8_f32(x0.z * d1v[2], x0.w * d1v[3], lo4, true);
            int hi4 = __builtin_amdgcn_cvt_pk_fp8_f32(x1.x * d1v[4], x1.y * d1v[5], 0, false);
            hi4 = __builtin_amdgcn_cvt_pk_fp8_f32(x1.z * d1v[6], x1.w * d1v[7], hi4, true);
            f.aq[i] = (long)(((unsigned long)(unsigned)hi4 << 32) | (unsigned)lo4);
        }
    } else {
        /* k-permuted: lane half hl takes pixels hl*16 + 4*step .. +4, so a
           lane reads 4 consecutive floats per row with one b128 load */
        const int kc = hl * 16 + step * 4;
        const float4 d1 = *(const float4 *)&s.d1[q][kc];
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const float4 t = *(const float4 *)&s.a[rb + i * 32 + rl][kc];
            f.av[i][0] = t.x * d1.x; f.av[i][1] = t.y * d1.y; f.av[i][2] = t.z * d1.z; f.av[i][3] = t.w * d1.w;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float4 t = *(const float4 *)&s.b[j * 32 + rl][kc];
            f.bv[j][0] = t.x; f.bv[j][1] = t.y; f.bv[j][2] = t.z; f.bv[j][3] = t.w;
        }
    }
}

template <int MODE, int RB>
__device__ __forceinline__ void mfma_frag(const Frag<MODE, RB> &f, f32x16 (&acc)[RB][4])
{
    if constexpr (MODE == AMB_PREC_BF16X3) {
        /* small terms first; the three passes of one block are 4*RB MFMAs apart */
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.al[i], f.bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.ah[i], f.bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.ah[i], f.bh[j], acc[i][j], 0, 0, 0);
    } else if constexpr (MODE == AMB_PREC_FP16X3) {
        /* the order of the BF16x3 passes, on the FP16 instruction */
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.al[i], f.bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.ah[i], f.bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(f.ah[i], f.bh[j], acc[i][j], 0, 0, 0);
    } else if constexpr (MODE == AMB_PREC_FP8) {
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(f.aq[i], f.bq[j], acc[i][j], 0, 0, 0);
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(f.av[i][e], f.bv[j][e], acc[i][j], 0, 0, 0);
    }
}

/* LDS of one workgroup */
template <int MODE, int NST>
struct A16 Smem {
    Stage<MODE> sm[NST];
    float rsc[TT], csc[TT]; /* scaled forms: s2 of the tile rows, 1/t of its columns */
    float s1v[QP];          /* scaled forms: s1 of the 4 reference planes */
};

template <int PIPE>
constexpr int nstages() { return 2; }

/* BF16x3 k-step with the A formation of block i+1 overlapping the MFMAs of
 * block i (one extra A fragment live). */
template <int RB>
__device__ __forceinline__ void bf16_step_ilv(const Stage<AMB_PREC_BF16X3> &s, int step, int q, int rb, int lane,
                                              f32x16 (&acc)[RB][4])
{
    const int rl = lane & 31, hl = lane >> 5;
    const int kc = step * 16 + hl * 8;
    float d1v[8];
    *(float4 *)&d1v[0] = *(const float4 *)&s.d1[q][kc];
    *(float4 *)&d1v[4] = *(const float4 *)&s.d1[q][kc + 4];
    bf16x8 bh[4], bl[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        bh[j] = *(const bf16x8 *)&s.bh[j * 32 + rl][kc];
        bl[j] = *(const bf16x8 *)&s.bl[j * 32 + rl][kc];
    }
    auto form = [&](int i, bf16x8 &ah, bf16x8 &al) {
        float x[8];
        *(float4 *)&x[0] = *(const float4 *)&s.a[rb + i * 32 + rl][kc];
        *(float4 *)&x[4] = *(const float4 *)&s.a[rb + i * 32 + rl][kc + 4];
        split8(x, d1v, ah, al);
    };
    bf16x8 ah0, al0, ah1, al1;
    form(0, ah0, al0);
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        bf16x8 &ah = (i & 1) ? ah1 : ah0, &al = (i & 1) ? al1 : al0;
        if (i + 1 < RB) {
            if (i & 1) form(i + 1, ah0, al0);
            else form(i + 1, ah1, al1);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
        if (i + 1 < RB) {
            /* 12 MFMAs, each followed by a share of the next block's work:
               2 LDS reads and 40 VALU instructions spread over them */
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); /* MFMA */
            __builtin_amdgcn_sched_group_barrier(0x100, 2, 0); /* DS read */
#pragma unroll
            for (int k = 0; k < 11; ++k) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); /* MFMA */
                __builtin_amdgcn_sched_group_barrier(0x002, 4, 0); /* VALU */
            }
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        }
    }
}

/* Adds (first: stores) the data-form accumulators of one wave into C and
 * clears them.  s1: the FP8 form's reference-plane scale.  A 32x32 MFMA block leaves row (v/4)*8 + (l/32)*4 + v%4,
 * column l%32 in register v of lane l.  Flushing every segment keeps each
 * register sum short: the rounding of a running sum grows with its length
 * (for the 16-bit forms all the more: their MFMAs add each exact sum of 16
 * products into the accumulator truncated toward zero, so the error of
 * every add is biased), and the pixel sum becomes a sum of short sums. */
template <int MODE, int RB>
__device__ __forceinline__ void flush_acc(const ContractArgs &a, f32x16 (&acc)[RB][4], bool first, int i1,
                                          int r0, int c0, int rb, int lane, float s1, const float *rsc,
                                          const float *csc)
{
    /* opaque bases: otherwise the 128 element addresses are loop invariants
       of the chunk loop, get hoisted out of it and spill */
    int nact = a.nact, lrow = rb + (lane >> 5) * 4, lcol = lane & 31;
    asm volatile("" : "+s"(nact));
    asm volatile("" : "+v"(lrow), "+v"(lcol));
    const int lo = a.wlo[i1], hi = a.whi[i1];
    /* recomputed here (flushes are rare) rather than held in a register
       across the chunk loop */
    const float us = data_unscale<MODE>(a);
    const float usl = ((i1 ^ lane) & 1) ? -us : us;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int li = lrow + i * 32 + (v >> 2) * 8 + (v & 3);
                const int lj = lcol + j * 32;
                const int gi = r0 + li, gj = c0 + lj;
                float x = acc[i][j][v];
                acc[i][j][v] = 0.f;
                if (gi < lo || gi >= hi || gj < lo || gj >= hi) continue;
                if constexpr (is_scaled<MODE>()) x = x * (1.f / (s1 * rsc[li])) * csc[lj];
                /* the stored planes' signs (-1)^(i1 + gi + gj), with the scale:
                   r0, c0 and the wave's row base are even, so gi is odd with
                   v and gj with the lane */
                x = x * ((v & 1) ? -usl : usl);
                const size_t idx = ((size_t)i1 * nact + gi) * nact + gj;
                /* one lane owns each element, so the adds land in a fixed
                   order; an add needs no register for the old value */
                if (first) a.C[idx] = x;
                else __hip_atomic_fetch_add(a.C + idx, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
    /* the first segment's stores complete before any later add reaches them */
    if (first) __builtin_amdgcn_s_waitcnt(0x0F70); /* vmcnt(0) */
}

/* Normalisation epilogue: the whole-pixel-range sum of reference plane i1
 * (never segmented: it must be whole to be certified) rounded to the
 * nearest 1/8 and stored as an integer count of eighths; the worst distance
 * from that grid goes to the margin word. */
template <int RB>
__device__ __forceinline__ void norm_store(const ContractArgs &a, f32x16 (&acc)[RB][4], int i1, int r0, int c0,
                                           int rb, int lane)
{
    const int nact = a.nact, lo = a.wlo[i1], hi = a.whi[i1];
    const double invP = 1.0 / (double)a.P;
    float worst = 0.f;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int li = rb + i * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                const int lj = j * 32 + (lane & 31);
                const int gi = r0 + li, gj = c0 + lj;
                if (gi < lo || gi >= hi || gj < lo || gj >= hi) continue;
                const size_t idx = ((size_t)i1 * nact + gi) * nact + gj;
                double val = (double)acc[i][j][v] * invP;
                double q8 = rint(val * 8.0);
                if (fabs(q8) < 2147483647.0) {
                    worst = fmaxf(worst, (float)fabs(val - q8 * 0.125));
                    a.Cq[idx] = (int)q8;
                } else { /* not representable: flagged, stored clamped */
                    worst = AMB_CERT_RANGE;
                    a.Cq[idx] = q8 > 0 ? 2147483647 : -2147483647;
                }
            }
    /* non-negative floats order like their bit patterns */
    if (worst > 0.f) atomicMax(a.margin, __float_as_uint(worst));
}

/* One work item: the (tile, group) product of one triangle. */
template <int MODE, int NORM, int NWR, int PIPE>
__device__ __forceinline__ void contract_item(const ContractArgs &a, int tile, int grp,
                                              Smem<MODE, nstages<PIPE>()> &L)
{
    using G = Geo<NWR>;
    constexpr int RB = G::RB;
    constexpr int NS = ksteps<MODE>();
    const int nact = a.nact;
    const int nt3 = (nact + TT - 1) / TT;
    const int t2 = tile / nt3, t3 = tile - t2 * nt3;
    const int r0 = t2 * TT, c0 = t3 * TT, g0 = grp * QP;
    /* skip a tile outside every window of the group */
    int lo = a.wlo[g0], hi = a.whi[min(g0 + QP, nact) - 1];
    if (r0 >= hi || r0 + TT <= lo || c0 >= hi || c0 + TT <= lo) return;
    /* r1 == r2: C[p1][p2][.] = C[p2][p1][.] bit for bit (the A operand is a
       product of the same two floats), so rows below every reference plane
       of the group are read transposed and never computed */
    if (a.sym12 && r0 + TT <= g0) return;
    if (a.sym23 && r0 >= c0 + TT) return;

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int q = wv & 3, rb = (wv >> 2) * (TT / NWR);
    const bool live = g0 + q < nact;
    float s1 = 1.f;
    if constexpr (is_scaled<MODE>()) {
        if (tid < TT) {
            L.rsc[tid] = r0 + tid < nact ? a.s2[r0 + tid] : 1.f;
            L.csc[tid] = c0 + tid < nact ? 1.f / a.t3[c0 + tid] : 1.f;
        }
        s1 = live ? a.s1[g0 + q] : 1.f;
        if (tid < QP) L.s1v[tid] = g0 + tid < nact ? a.s1[g0 + tid] : 1.f;
        __syncthreads();
    }

    f32x16 acc[RB][4];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

    const long nchunk = a.pitch / KC;
    Pref<MODE, NWR> pf;
    if constexpr (PIPE == 0) {
        prefetch<MODE, NWR>(pf, a, 0, r0, c0, g0, tid);
        stage_store<MODE, NWR>(L.sm[0], pf, tid, L.rsc, L.s1v);
        __syncthreads();
        for (long c = 0; c < nchunk; ++c) {
            const int cur = (int)(c & 1);
            if (c + 1 < nchunk) prefetch<MODE, NWR>(pf, a, (c + 1) * KC, r0, c0, g0, tid);
            if (live) {
#pragma unroll 1
                for (int st = 0; st < NS; ++st) {
                    Frag<MODE, RB> f;
                    load_frag<MODE, RB>(L.sm[cur], st, q, rb, lane, f);
                    mfma_frag<MODE, RB>(f, acc);
                }
                if (!NORM && (c + 1) % SEGC == 0 && c + 1 < nchunk)
                    flush_acc<MODE, RB>(a, acc, c + 1 == SEGC, g0 + q, r0, c0, rb, lane, s1, L.rsc, L.csc);
            }
            if (c + 1 < nchunk) stage_store<MODE, NWR>(L.sm[cur ^ 1], pf, tid, L.rsc, L.s1v);
            __syncthreads();
        }
    } else {
        prefetch<MODE, NWR>(pf, a, 0, r0, c0, g0, tid);
        stage_store<MODE, NWR>(L.sm[0], pf, tid, L.rsc, L.s1v);
        __syncthreads();
        for (long c = 0; c < nchunk; ++c) {
            const int cur = (int)(c & 1);
            if (c + 1 < nchunk) prefetch<MODE, NWR>(pf, a, (c + 1) * KC, r0, c0, g0, tid);
            if (live) {
#pragma unroll 1
                for (int st = 0; st < NS; ++st) {
                    if constexpr (MODE == AMB_PREC_BF16X3) {
                        bf16_step_ilv<RB>(L.sm[cur], st, q, rb, lane, acc);
                    } else {
                        Frag<MODE, RB> f;
                        load_frag<MODE, RB>(L.sm[cur], st, q, rb, lane, f);
                        mfma_frag<MODE, RB>(f, acc);
                    }
                }
                if (!NORM && (c + 1) % SEGC == 0 && c + 1 < nchunk)
                    flush_acc<MODE, RB>(a, acc, c + 1 == SEGC, g0 + q, r0, c0, rb, lane, s1, L.rsc, L.csc);
            }
            if (c + 1 < nchunk) stage_store<MODE, NWR>(L.sm[cur ^ 1], pf, tid, L.rsc, L.s1v);
            __syncthreads();
        }
    }
    if (!live) return;

    const int i1 = g0 + q;
    if constexpr (!NORM) flush_acc<MODE, RB>(a, acc, nchunk <= SEGC, i1, r0, c0, rb, lane, s1, L.rsc, L.csc);
    else norm_store<RB>(a, acc, i1, r0, c0, rb, lane);
}

/* ---------------------------------------------------------------------------
 * Pipelined BF16x3 data contraction (variant 2).  Same work item and the
 * same arithmetic in the same order as the form above, so its results are
 * identical bit for bit; only the data movement differs:
 *   - 64-pixel chunks in a two-stage LDS ring filled by direct
 *     global -> LDS loads (global_load_lds_dwordx4: lane i's 16 bytes land
 *     at the instruction's LDS base + 16 i), one chunk ahead, so staging
 *     costs neither registers nor LDS store instructions;
 *   - rows stored without padding; a row's 16-byte slots are XOR-permuted so
 *     that the 16 lanes an LDS cycle serves for a 128-bit read hit 16
 *     different 4-bank groups: A rows (16 slots, 256 bytes) by the low 4
 *     bits of the row, B rows (8 slots, 128 bytes, so neighbouring rows sit
 *     in opposite bank halves) by bits 1..3 of the row;
 *   - one barrier per chunk; the LDS reads and A formation of k-step s+1
 *     are interleaved with the MFMAs of k-step s, across chunk boundaries
 *     too (the barrier sits before a chunk's last k-step).
 * Rows and columns past nact and planes past nact read clamped (valid)
 * rows; their results fall outside every window and are never stored.
 * ------------------------------------------------------------------------- */
#define KC2 64
typedef __attribute__((address_space(3))) void lds_void_t;

/* One LDS stage: views of separate __shared__ arrays (see the kernels) */
struct V2Stage {
    float *a;      /* D_r2 row r, float4 slot c at r*64 + 4*(c ^ (r & 15)) */
    uint16_t *bh;  /* D_r3 hi col j, 16-byte slot c at j*64 + 8*(c ^ ((j >> 1) & 7)) */
    uint16_t *bl;  /* D_r3 lo, same layout */
    float *d1;     /* D_r1 of the 4 planes, plain rows */
};

/* FORM: AMB_PREC_BF16X3 or AMB_PREC_FP16X3 (16-bit elements of the split) */
template <int FORM>
struct F2 {
    typedef typename HalfT<FORM>::v8 V;
    V ah[2], al[2], bh[4], bl[4];
};

__device__ __forceinline__ void lds_load16(const void *g, void *l)
{
    __builtin_amdgcn_global_load_lds(g, (lds_void_t *)l, 16, 0, 0);
}

/* One chunk into stage s; wave w issues A rows 16w..16w+15 (4 loads), hi and
 * lo columns 16w..16w+15 (2 + 2 loads), wave 0 also the 4 D_r1 rows.  Each
 * array holds nact rows of pitch elements and the launchers run this form
 * only when that is under 4 GiB (v2_fits), so a load is the array base plus
 * the chunk offset (scalar registers) plus a 32-bit lane offset, formed
 * with 24-bit multiplies (row < 2^24, pitch < 2^24). */
__device__ __forceinline__ void v2_issue(const V2Stage &s, const ContractArgs &a, long k0, int r0, int c0, int g0,
                                         int wv, int lane)
{
    const char *pa = (const char *)a.D2 + k0 * 4, *pd = (const char *)a.D1 + k0 * 4;
    const char *ph = (const char *)a.H3 + k0 * 2, *pl = (const char *)a.L3 + k0 * 2;
    /* opaque scalar bases and lane: otherwise the chunk loop's strength
       reduction or hoisting keeps per-load 64-bit addresses in vector
       registers across the loop, and they spill */
    asm volatile("" : "+s"(pa), "+s"(pd), "+s"(ph), "+s"(pl));
    asm volatile("" : "+v"(lane));
    const unsigned pitch = (unsigned)a.pitch;
    const int last = a.nact - 1;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int t = wv * 4 + u;                  /* 1 KB block of the A region */
        const int r = t * 4 + (lane >> 4), c = (lane & 15) ^ (r & 15);
        const unsigned off = (__umul24((unsigned)min(r0 + r, last), pitch) + (unsigned)c * 4u) * 4u;
        lds_load16(pa + off, &s.a[t * 256]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int t = wv * 2 + u;
        const int j = t * 8 + (lane >> 3), c = (lane & 7) ^ ((j >> 1) & 7);
        const unsigned off = (__umul24((unsigned)min(c0 + j, last), pitch) + (unsigned)c * 8u) * 2u;
        lds_load16(ph + off, &s.bh[t * 512]);
        lds_load16(pl + off, &s.bl[t * 512]);
    }
    if (wv == 0) {
        const unsigned row = (unsigned)min(g0 + (lane >> 4), last);
        const unsigned off = (__umul24(row, pitch) + (unsigned)(lane & 15) * 4u) * 4u;
        lds_load16(pd + off, &s.d1[0]);
    }
}

/* fragments of k-step st (16 pixels) of a chunk */
template <int FORM>
__device__ __forceinline__ void v2_frag(const V2Stage &s, int st, int q, int rb, int lane, F2<FORM> &f)
{
    typedef typename F2<FORM>::V V;
    const int rl = lane & 31, hl = lane >> 5;
    const int kc = st * 16 + hl * 8;
    float d1v[8];
    *(float4 *)&d1v[0] = *(const float4 *)&s.d1[q * KC2 + kc];
    *(float4 *)&d1v[4] = *(const float4 *)&s.d1[q * KC2 + kc + 4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = j * 32 + rl;
        const int ph = (kc >> 3) ^ ((col >> 1) & 7);
        f.bh[j] = *(const V *)&s.bh[col * KC2 + ph * 8];
        f.bl[j] = *(const V *)&s.bl[col * KC2 + ph * 8];
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int r = rb + i * 32 + rl;
        const int p0 = (kc >> 2) ^ (r & 15);
        float x[8];
        *(float4 *)&x[0] = *(const float4 *)&s.a[r * KC2 + p0 * 4];
        *(float4 *)&x[4] = *(const float4 *)&s.a[r * KC2 + (p0 ^ 1) * 4];
        split8(x, d1v, f.ah[i], f.al[i]);
    }
}

/* one 32x32x16 MFMA of the split form */
__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma16(f16x8 a, f16x8 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

/* the three passes of one k-step, in the order of mfma_frag */
template <int FORM>
__device__ __forceinline__ void v2_mfma(const F2<FORM> &f, f32x16 (&acc)[2][4])
{
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mfma16(f.al[i], f.bh[j], acc[i][j]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mfma16(f.ah[i], f.bl[j], acc[i][j]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mfma16(f.ah[i], f.bh[j], acc[i][j]);
}

/* A-side operands of a k-step as read from LDS: the D_r2 rows still to be
 * multiplied by D_r1 and split. */
struct R2 {
    float x[2][8];
    float d1v[8];
};

__device__ __forceinline__ void v2_read_a(const V2Stage &s, int st, int q, int rb, int lane, R2 &r)
{
    const int rl = lane & 31, hl = lane >> 5;
    const int kc = st * 16 + hl * 8;
    *(float4 *)&r.d1v[0] = *(const float4 *)&s.d1[q * KC2 + kc];
    *(float4 *)&r.d1v[4] = *(const float4 *)&s.d1[q * KC2 + kc + 4];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int row = rb + i * 32 + rl;
        const int p0 = (kc >> 2) ^ (row & 15);
        *(float4 *)&r.x[i][0] = *(const float4 *)&s.a[row * KC2 + p0 * 4];
        *(float4 *)&r.x[i][4] = *(const float4 *)&s.a[row * KC2 + (p0 ^ 1) * 4];
    }
}

template <int FORM>
__device__ __forceinline__ void v2_read_b(const V2Stage &s, int st, int lane, F2<FORM> &f)
{
    typedef typename F2<FORM>::V V;
    const int rl = lane & 31, hl = lane >> 5;
    const int kc = st * 16 + hl * 8;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = j * 32 + rl;
        const int ph = (kc >> 3) ^ ((col >> 1) & 7);
        f.bh[j] = *(const V *)&s.bh[col * KC2 + ph * 8];
        f.bl[j] = *(const V *)&s.bl[col * KC2 + ph * 8];
    }
}

/* pair e (0..7) of the hi/lo split of a k-step's A operand: row block e/4,
 * elements 2(e%4), 2(e%4)+1; the same arithmetic as split8 */
template <int FORM>
__device__ __forceinline__ void v2_split_pair(const R2 &r, int e, u32x4 (&hp)[2], u32x4 (&lp)[2])
{
    const int i = e >> 2, k = (e & 3) * 2;
    const f32x2 v = {r.x[i][k] * r.d1v[k], r.x[i][k + 1] * r.d1v[k + 1]};
    if constexpr (FORM == AMB_PREC_BF16X3) {
        const unsigned hb = __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
        const float h0 = __uint_as_float(hb << 16), h1 = __uint_as_float(hb & 0xffff0000u);
        const f32x2 res = {fmaf(r.x[i][k], r.d1v[k], -h0), fmaf(r.x[i][k + 1], r.d1v[k + 1], -h1)};
        hp[i][e & 3] = hb;
        lp[i][e & 3] = __builtin_bit_cast(unsigned, __builtin_convertvector(res, bf16x2));
    } else {
        /* the residual of the rounded product (exact: v and hi are within a
           factor 2 of each other) */
        const f16x2 h = __builtin_convertvector(v, f16x2);
        const f32x2 res = {v[0] - (float)h[0], v[1] - (float)h[1]};
        hp[i][e & 3] = __builtin_bit_cast(unsigned, h);
        lp[i][e & 3] = __builtin_bit_cast(unsigned, __builtin_convertvector(res, f16x2));
    }
}

/* MFMA m (0..23) of a k-step in the order of mfma_frag: 8 lo.hi, 8 hi.lo,
 * 8 hi.hi, each over (row block, column block) */
template <int FORM>
__device__ __forceinline__ void v2_mfma_one(const F2<FORM> &f, f32x16 (&acc)[2][4], int m)
{
    const int pass = m >> 3, i = (m >> 2) & 1, j = m & 3;
    if (pass == 0) acc[i][j] = mfma16(f.al[i], f.bh[j], acc[i][j]);
    else if (pass == 1) acc[i][j] = mfma16(f.ah[i], f.bl[j], acc[i][j]);
    else acc[i][j] = mfma16(f.ah[i], f.bh[j], acc[i][j]);
}

/* The 24 MFMAs of one k-step with the next k-step's work placed between
 * them in a fixed order: after MFMA 0 the 6 LDS reads of the next A side,
 * after MFMAs 8..15 one pair each of its hi/lo split (the reads have had 7
 * MFMAs to land), after MFMA 16 (the current lo halves are dead by then)
 * the 8 reads of the next B side.  sched_barrier keeps the compiler from
 * clustering the MFMAs again, so a wave's VALU and LDS work runs beside its
 * own MFMAs.  The empty asm consumes the split here: otherwise, where the
 * next k-step lies beyond a branch (the segment flush), the split is sunk
 * past it and runs with no MFMA beside it. */
template <int FORM>
__device__ __forceinline__ void v2_step(const F2<FORM> &cur, f32x16 (&acc)[2][4], const V2Stage &s, int st_next,
                                        int q, int rb, int lane, F2<FORM> &nxt)
{
    typedef typename F2<FORM>::V V;
    R2 r;
    u32x4 hp[2], lp[2];
#pragma unroll
    for (int m = 0; m < 24; ++m) {
        v2_mfma_one<FORM>(cur, acc, m);
        if (m == 0) v2_read_a(s, st_next, q, rb, lane, r);
        if (m >= 8 && m < 16) v2_split_pair<FORM>(r, m - 8, hp, lp);
        if (m == 16) v2_read_b<FORM>(s, st_next, lane, nxt);
        __builtin_amdgcn_sched_barrier(0);
    }
    asm volatile("" : "+v"(hp[0]), "+v"(hp[1]), "+v"(lp[0]), "+v"(lp[1]));
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        nxt.ah[i] = __builtin_bit_cast(V, hp[i]);
        nxt.al[i] = __builtin_bit_cast(V, lp[i]);
    }
}

/* Chunk c of a work item, not its last: k-steps 1..3 from Sc, the barrier,
 * chunk c + 2 into Sc, then k-step 3's MFMAs with k-step 0 of chunk c + 1
 * from Sn. */
template <int FORM, int NORM>
__device__ __forceinline__ void v2_chunk(const ContractArgs &a, long c, long nchunk, const V2Stage &Sc,
                                         const V2Stage &Sn, F2<FORM> &f0, F2<FORM> &f1, f32x16 (&acc)[2][4], int r0,
                                         int c0, int g0, int wv, int q, int rb, int lane)
{
    constexpr int SEG2 = SEGC * KC / KC2; /* the same 8192-pixel segments */
    v2_step<FORM>(f0, acc, Sc, 1, q, rb, lane, f1);
    v2_step<FORM>(f1, acc, Sc, 2, q, rb, lane, f0);
    v2_step<FORM>(f0, acc, Sc, 3, q, rb, lane, f1);
    /* this wave's reads of Sc and loads of chunk c + 1 are complete; after
       the barrier, every wave's are */
    __builtin_amdgcn_s_waitcnt(0x0070); /* vmcnt(0) lgkmcnt(0) */
    __syncthreads();
    if (c + 2 < nchunk) v2_issue(Sc, a, (c + 2) * KC2, r0, c0, g0, wv, lane);
    v2_step<FORM>(f1, acc, Sn, 0, q, rb, lane, f0);
    if (!NORM && (c + 1) % SEG2 == 0)
        flush_acc<FORM, 2>(a, acc, c + 1 == SEG2, g0 + q, r0, c0, rb, lane, 1.f, nullptr, nullptr);
}

/* The last chunk from S, and the final flush (or the normalisation epilogue). */
template <int FORM, int NORM>
__device__ __forceinline__ void v2_last(const ContractArgs &a, long nchunk, const V2Stage &S, F2<FORM> &f0,
                                        F2<FORM> &f1, f32x16 (&acc)[2][4], int r0, int c0, int g0, int q, int rb,
                                        int lane)
{
    constexpr int SEG2 = SEGC * KC / KC2;
    v2_step<FORM>(f0, acc, S, 1, q, rb, lane, f1);
    v2_step<FORM>(f1, acc, S, 2, q, rb, lane, f0);
    v2_step<FORM>(f0, acc, S, 3, q, rb, lane, f1);
    v2_mfma<FORM>(f1, acc);
    if constexpr (NORM) norm_store<2>(a, acc, g0 + q, r0, c0, rb, lane);
    else flush_acc<FORM, 2>(a, acc, nchunk <= SEG2, g0 + q, r0, c0, rb, lane, 1.f, nullptr, nullptr);
}

/* The two LDS stages are separate objects and every access names one of
 * them (the chunk loop is unrolled by two): the compiler can then tell that
 * the direct loads into one stage do not alias reads of the other and does
 * not wait for the loads before those reads.  With one array indexed by
 * chunk parity it waits for all loads before the first read after them, so
 * a chunk's prefetch would land before any further MFMA issues. */
template <int FORM, int NORM>
__device__ __forceinline__ void contract_item_v2(const ContractArgs &a, int tile, int grp, const V2Stage &S0,
                                                 const V2Stage &S1)
{
    const int nact = a.nact;
    const int nt3 = (nact + TT - 1) / TT;
    const int t2 = tile / nt3, t3 = tile - t2 * nt3;
    const int r0 = t2 * TT, c0 = t3 * TT, g0 = grp * QP;
    int lo = a.wlo[g0], hi = a.whi[min(g0 + QP, nact) - 1];
    if (r0 >= hi || r0 + TT <= lo || c0 >= hi || c0 + TT <= lo) return;
    if (a.sym12 && r0 + TT <= g0) return;
    if (a.sym23 && r0 >= c0 + TT) return;

    /* the wave index is uniform across the wave: branches on it are scalar */
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int q = wv & 3, rb = (wv >> 2) * 64;
    const bool live = g0 + q < nact;
    f32x16 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

    const long nchunk = a.pitch / KC2;
    /* The k-steps run as one stream across chunks: the MFMAs of a chunk's
       last k-step carry the LDS reads and A formation of the next chunk's
       first, so no k-step starts without operands.  The barrier therefore
       sits before that last k-step: by then every wave has read the
       chunk's LDS data (step 3 operands are in flight to registers) and
       waited for its loads of the next chunk, which was issued one
       barrier earlier; after it the stage just read takes chunk c + 2.
       Chunk c lives in stage c & 1. */
    v2_issue(S0, a, 0, r0, c0, g0, wv, lane);
    __builtin_amdgcn_s_waitcnt(0x0F70); /* vmcnt(0) */
    __syncthreads();
    if (nchunk > 1) v2_issue(S1, a, KC2, r0, c0, g0, wv, lane);
    if (!live) {
        /* no plane for this wave: it loads its share and keeps the barriers */
        for (long c = 0; c + 1 < nchunk; ++c) {
            __builtin_amdgcn_s_waitcnt(0x0070); /* vmcnt(0) lgkmcnt(0) */
            __syncthreads();
            if (c + 2 < nchunk) v2_issue((c & 1) ? S1 : S0, a, (c + 2) * KC2, r0, c0, g0, wv, lane);
        }
        return;
    }
    /* Straight-line loop body with one carried fragment set (f0) and the
       last chunk peeled: conditions on the wave or on the last chunk inside
       the loop made the compiler spill the accumulators. */
    F2<FORM> f0, f1;
    v2_frag<FORM>(S0, 0, q, rb, lane, f0);
    long c = 0;
    for (; c + 2 < nchunk; c += 2) {
        v2_chunk<FORM, NORM>(a, c, nchunk, S0, S1, f0, f1, acc, r0, c0, g0, wv, q, rb, lane);
        v2_chunk<FORM, NORM>(a, c + 1, nchunk, S1, S0, f0, f1, acc, r0, c0, g0, wv, q, rb, lane);
    }
    if (c + 1 < nchunk) v2_chunk<FORM, NORM>(a, c, nchunk, S0, S1, f0, f1, acc, r0, c0, g0, wv, q, rb, lane);
    v2_last<FORM, NORM>(a, nchunk, ((nchunk - 1) & 1) ? S1 : S0, f0, f1, acc, r0, c0, g0, q, rb, lane);
}

/* The two stages: A rows, hi then lo B rows, D_r1 rows, each its own array */
#define V2_STAGES                                                                                                  \
    __shared__ A16 float sA0[TT * KC2], sA1[TT * KC2];                                                             \
    __shared__ A16 uint16_t sB0[2 * TT * KC2], sB1[2 * TT * KC2];                                                  \
    __shared__ A16 float sD0[QP * KC2], sD1[QP * KC2];                                                             \
    const V2Stage S0{sA0, sB0, sB0 + TT * KC2, sD0}, S1{sA1, sB1, sB1 + TT * KC2, sD1};

template <int FORM, int NORM>
__global__ __launch_bounds__(512, 1) void k_contract_v2(ContractArgs a, int ngroups, int per_xcd, int nwork)
{
    V2_STAGES
    const int b = blockIdx.x;
    const int slot = b >> 3;
    const int w = (b & 7) * per_xcd + slot;
    if (slot >= per_xcd || w >= nwork) return;
    contract_item_v2<FORM, NORM>(a, w / ngroups, w % ngroups, S0, S1);
}

/* XCD queues of the grouped contraction: runs of 2^xl consecutive items (one
 * tile's p1 groups, which stream the same D2 / D3 chunks) are dealt
 * round-robin to the 8 XCDs; workgroup b sits on XCD b mod 8 and takes items
 * from that XCD's counter (xq starts at b mod 8), then from the others once
 * it runs dry.  Within a queue the item index only grows, so a queue that
 * yields an index past the end is empty for good: each caller makes at most
 * 8 such failed takes (left), then gets -1.  The run length is a tuning
 * parameter (32 by default); the items each triangle computes do not
 * depend on it. */
__device__ __forceinline__ long xcd_take(unsigned long long *counter, long total, int &xq, int &left, int xl)
{
    const unsigned long long msk = (1ull << xl) - 1;
    while (left > 0) {
        const unsigned long long j = atomicAdd(counter + xq, 1ull);
        const long cand = ((long)(j >> xl) * 8 + xq) * (1l << xl) + (long)(j & msk);
        if (cand < total) return cand;
        xq = (xq + 1) & 7;
        --left;
    }
    return -1;
}

template <int FORM, int NORM>
__global__ __launch_bounds__(512, 1) void k_contract_grouped_v2(const ContractArgs *__restrict__ descs,
                                                                const long *__restrict__ begin, int ntri,
                                                                unsigned long long *counter, int xl)
{
    V2_STAGES
    __shared__ long item;
    const long total = begin[ntri];
    int xq = blockIdx.x & 7, left = 8;
    for (;;) {
        if (threadIdx.x == 0) item = xcd_take(counter, total, xq, left, xl);
        __syncthreads();
        const long w = item;
        __syncthreads();
        if (w < 0) return;
        int lo = 0, hi = ntri - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (begin[mid] <= w) lo = mid;
            else hi = mid - 1;
        }
        const ContractArgs &a = descs[lo];
        const int ngroups = (a.nact + QP - 1) / QP;
        const long r = w - begin[lo] + a.item0;
        contract_item_v2<FORM, NORM>(a, (int)(r / ngroups), (int)(r % ngroups), S0, S1);
        __syncthreads();
    }
}

/* One triangle: workgroups numbered tile-major and dealt to the 8 XCDs in
 * contiguous runs (hardware places workgroup b on XCD b mod 8). */
template <int MODE, int NORM, int NWR, int PIPE>
__global__ __launch_bounds__(256 * NWR, 1) void k_contract(ContractArgs a, int ngroups, int per_xcd, int nwork)
{
    __shared__ Smem<MODE, nstages<PIPE>()> L;
    const int b = blockIdx.x;
    const int slot = b >> 3;
    const int w = (b & 7) * per_xcd + slot;
    if (slot >= per_xcd || w >= nwork) return;
    contract_item<MODE, NORM, NWR, PIPE>(a, w / ngroups, w % ngroups, L);
}

/* Many triangles in one persistent launch: each workgroup takes items from
 * the XCD queues (xcd_take) until they are empty.  Items are ordered
 * by the host (largest triangles first, tile-major within a triangle).
 * begin[t] is the first item of triangle t, begin[ntri] the total. */
template <int MODE, int NORM, int NWR, int PIPE>
__global__ __launch_bounds__(256 * NWR, 1) void k_contract_grouped(const ContractArgs *__restrict__ descs,
                                                                   const long *__restrict__ begin, int ntri,
                                                                   unsigned long long *counter, int xl)
{
    __shared__ Smem<MODE, nstages<PIPE>()> L;
    __shared__ long item;
    const long total = begin[ntri];
    int xq = blockIdx.x & 7, left = 8;
    for (;;) {
        if (threadIdx.x == 0) item = xcd_take(counter, total, xq, left, xl);
        __syncthreads();
        const long w = item;
        __syncthreads();
        if (w < 0) return;
        int lo = 0, hi = ntri - 1;
        while (lo < hi) { /* last t with begin[t] <= w */
            const int mid = (lo + hi + 1) >> 1;
            if (begin[mid] <= w) lo = mid;
            else hi = mid - 1;
        }
        const ContractArgs &a = descs[lo]; /* uniform: scalar loads */
        const int ngroups = (a.nact + QP - 1) / QP;
        const long r = w - begin[lo] + a.item0;
        contract_item<MODE, NORM, NWR, PIPE>(a, (int)(r / ngroups), (int)(r % ngroups), L);
        __syncthreads();
    }
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

__global__ void k_fill_u32(unsigned *p, long n, unsigned v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t amb_launch_fill_u32(hipStream_t s, unsigned *p, long n, unsigned v)
{
    if (n <= 0) return hipSuccess;
    const int blocks = (int)std::min<long>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

static bool args_ok(int prec, int norm, const ContractArgs &a)
{
    if (a.pitch % KC != 0 || a.nact <= 0) return false;
    if ((prec == AMB_PREC_BF16X3 || prec == AMB_PREC_FP16X3) && (!a.H3 || !a.L3)) return false;
    /* data forms read the exponents of the normalised plane sets; the FP16x3
       form carries no normalisation epilogue */
    if (!norm && (!a.ex1 || !a.ex2 || !a.ex3)) return false;
    if (prec == AMB_PREC_FP16X3 && norm) return false;
    if (prec == AMB_PREC_FP32 && !a.D3) return false;
    if (prec == AMB_PREC_FP8 && !a.Q3) return false;
    /* scaled forms carry no normalisation epilogue */
    if (prec == AMB_PREC_FP8 && (norm || !a.s1 || !a.s2 || !a.t3)) return false;
    return true;
}

int amb_contract_grouped_blocks(int dev)
{
    int cus = 0, occ = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    /* every launched form shares the 8-wave geometry; the BF16x3 data form
       has the largest LDS footprint */
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_contract_grouped<AMB_PREC_BF16X3, 0, 2, 0>, 512, 0) !=
            hipSuccess || occ < 1)
        occ = 1;
    (void)hipGetLastError();
    return cus * occ;
}

long amb_contract_items(int nact)
{
    const long nt = (nact + TT - 1) / TT;
    return nt * nt * ((nact + QP - 1) / QP);
}

long amb_contract_live_items(int nact, const int *wlo, const int *whi, int sym12, int sym23)
{   /* the skip tests at the top of contract_item */
    const int nt = (nact + TT - 1) / TT, ng = (nact + QP - 1) / QP;
    long n = 0;
    for (int g = 0; g < ng; ++g) {
        const int g0 = g * QP, lo = wlo[g0], hi = whi[std::min(g0 + QP, nact) - 1];
        for (int t2 = 0; t2 < nt; ++t2)
            for (int t3 = 0; t3 < nt; ++t3) {
                const int r0 = t2 * TT, c0 = t3 * TT;
                if (r0 >= hi || r0 + TT <= lo || c0 >= hi || c0 + TT <= lo) continue;
                if (sym12 && r0 + TT <= g0) continue;
                if (sym23 && r0 >= c0 + TT) continue;
                ++n;
            }
    }
    return n;
}

double amb_contract_item_flops(int prec, long pitch)
{
    return 2.0 * QP * TT * TT * (double)pitch * ((prec == AMB_PREC_BF16X3 || prec == AMB_PREC_FP16X3) ? 3.0 : 1.0);
}

template <int MODE, int NORM>
static void launch_grouped_var(int var, dim3 grid, hipStream_t s, const ContractArgs *d, const long *b, int n,
                               unsigned long long *ctr, int xl)
{
    if constexpr (is_split<MODE>()) {
        if (var == 2) {
            hipLaunchKernelGGL((k_contract_grouped_v2<MODE, NORM>), grid, dim3(512), 0, s, d, b, n, ctr, xl);
            return;
        }
    }
    if (var == 1) hipLaunchKernelGGL((k_contract_grouped<MODE, NORM, 2, 1>), grid, dim3(512), 0, s, d, b, n, ctr, xl);
    else hipLaunchKernelGGL((k_contract_grouped<MODE, NORM, 2, 0>), grid, dim3(512), 0, s, d, b, n, ctr, xl);
}

bool amb_contract_v2_fits(int nact, long pitch)
{   /* the 32-bit lane offsets of v2_issue: a plane array of nact rows of
       pitch FP32 values lies within 4 GiB, and rows and pitch fit the
       24-bit multiplies */
    return nact > 0 && pitch > 0 && pitch < (1l << 24) &&
           (unsigned long long)nact * (unsigned long long)pitch * 4ull <= 0xFFFFFFFFull;
}

template <int MODE, int NORM>
static void launch_var(int var, dim3 grid, hipStream_t s, const ContractArgs &a, int ngroups, int per_xcd, int nwork)
{
    if (var == 2 && !amb_contract_v2_fits(a.nact, a.pitch)) var = 0;
    if constexpr (is_split<MODE>()) {
        if (var == 2) {
            hipLaunchKernelGGL((k_contract_v2<MODE, NORM>), grid, dim3(512), 0, s, a, ngroups, per_xcd, nwork);
            return;
        }
    }
    if (var == 1) hipLaunchKernelGGL((k_contract<MODE, NORM, 2, 1>), grid, dim3(512), 0, s, a, ngroups, per_xcd, nwork);
    else hipLaunchKernelGGL((k_contract<MODE, NORM, 2, 0>), grid, dim3(512), 0, s, a, ngroups, per_xcd, nwork);
}

hipError_t amb_launch_contract_grouped(hipStream_t s, int prec, int norm, int var, const ContractArgs *d_descs,
                                       const long *d_begin, int ntri, long total, unsigned long long *d_counter,
                                       int nblocks, int xrun_log2)
{
    if (xrun_log2 < 0 || xrun_log2 > 12) return hipErrorInvalidValue;
    if (ntri <= 0 || total <= 0) return hipSuccess;
    if ((prec == AMB_PREC_FP8 || prec == AMB_PREC_FP16X3) && norm) return hipErrorInvalidValue;
    /* a kernel fill, not a memset, so the reset can live in a captured graph */
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, d_counter, 0ull, AMB_GROUPED_COUNTERS);
    if ((long)nblocks > total) nblocks = (int)total;
    dim3 grid(nblocks);
    if (prec == AMB_PREC_FP8) launch_grouped_var<AMB_PREC_FP8, 0>(var, grid, s, d_descs, d_begin, ntri, d_counter, xrun_log2);
    else if (prec == AMB_PREC_FP16X3)
        launch_grouped_var<AMB_PREC_FP16X3, 0>(var, grid, s, d_descs, d_begin, ntri, d_counter, xrun_log2);
    else if (prec == AMB_PREC_FP32) {
        if (norm) launch_grouped_var<AMB_PREC_FP32, 1>(var, grid, s, d_descs, d_begin, ntri, d_counter, xrun_log2);
        else launch_grouped_var<AMB_PREC_FP32, 0>(var, grid, s, d_descs, d_begin, ntri, d_counter, xrun_log2);
    } else {
        if (norm) launch_grouped_var<AMB_PREC_BF16X3, 1>(var, grid, s, d_descs, d_begin, ntri, d_counter, xrun_log2);
        else launch_grouped_var<AMB_PREC_BF16X3, 0>(var, grid, s, d_descs, d_begin, ntri, d_counter, xrun_log2);
    }
    return hipGetLastError();
}

bool amb_contract_args_ok(int prec, int norm, const ContractArgs &a) { return args_ok(prec, norm, a); }

hipError_t amb_launch_contract(hipStream_t s, int prec, int norm, int var, const ContractArgs &a)
{
    if (!args_ok(prec, norm, a)) return hipErrorInvalidValue;
    const int nt = (a.nact + TT - 1) / TT;
    const int ngroups = (a.nact + QP - 1) / QP;
    const int nwork = nt * nt * ngroups;
    const int per_xcd = (nwork + 7) / 8;
    dim3 grid(8 * per_xcd);
    if (prec == AMB_PREC_FP8) launch_var<AMB_PREC_FP8, 0>(var, grid, s, a, ngroups, per_xcd, nwork);
    else if (prec == AMB_PREC_FP16X3) launch_var<AMB_PREC_FP16X3, 0>(var, grid, s, a, ngroups, per_xcd, nwork);
    else if (prec == AMB_PREC_FP32) {
        if (norm) launch_var<AMB_PREC_FP32, 1>(var, grid, s, a, ngroups, per_xcd, nwork);
        else launch_var<AMB_PREC_FP32, 0>(var, grid, s, a, ngroups, per_xcd, nwork);
    } else {
        if (norm) launch_var<AMB_PREC_BF16X3, 1>(var, grid, s, a, ngroups, per_xcd, nwork);
        else launch_var<AMB_PREC_BF16X3, 0>(var, grid, s, a, ngroups, per_xcd, nwork);
    }
    return hipGetLastError();
}

/* hi/lo BF16 split of FP32 planes (the B operand of the BF16x3 form) */
__global__ void k_split(const float *__restrict__ D, size_t n, __bf16 *__restrict__ H, __bf16 *__restrict__ L)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float x = D[i];
        const __bf16 h = (__bf16)x;
        H[i] = h;
        L[i] = (__bf16)(x - (float)h);
    }
}

hipError_t amb_launch_split(hipStream_t s, const float *D, size_t n, __bf16 *H, __bf16 *L)
{
    int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_split, dim3(blocks), dim3(256), 0, s, D, n, H, L);
    return hipGetLastError();
}

/* FP16 hi / lo split of the B operand: stored x times 2^AMB_F16_B_EXP
 * (|x| <= 2^AMB_SET_EXP, so the product stays within 2^15) */
__global__ void k_split_f16(const float *__restrict__ D, size_t n4, _Float16 *__restrict__ H, _Float16 *__restrict__ L)
{
    const float sb = (float)(1 << AMB_F16_B_EXP);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        const float4 v = ((const float4 *)D)[i];
        const f32x2 a = {v.x * sb, v.y * sb}, b = {v.z * sb, v.w * sb};
        const f16x2 ha = __builtin_convertvector(a, f16x2), hb = __builtin_convertvector(b, f16x2);
        const f32x2 ra = {a[0] - (float)ha[0], a[1] - (float)ha[1]}, rb = {b[0] - (float)hb[0], b[1] - (float)hb[1]};
        const f16x2 la = __builtin_convertvector(ra, f16x2), lb = __builtin_convertvector(rb, f16x2);
        ((uint2 *)H)[i] = make_uint2(__builtin_bit_cast(unsigned, ha), __builtin_bit_cast(unsigned, hb));
        ((uint2 *)L)[i] = make_uint2(__builtin_bit_cast(unsigned, la), __builtin_bit_cast(unsigned, lb));
    }
}

hipError_t amb_launch_split_f16(hipStream_t s, const float *D, size_t n, _Float16 *H, _Float16 *L)
{
    if (n % 4 != 0) return hipErrorInvalidValue;
    const size_t n4 = n / 4;
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>((n4 + 255) / 256, 65536));
    hipLaunchKernelGGL(k_split_f16, dim3(blocks), dim3(256), 0, s, D, n4, H, L);
    return hipGetLastError();
}

/* Per-plane power-of-two scales for the FP8 form from max|D| over each
 * plane: sc[i] = 2^floor(log2(sqrt(448)/m)) (A-side factors) and
 * sc[n+i] = 2^floor(log2(448/m)) (B side); 1 for an all-zero plane. */
__global__ void k_fp8_scales(const float *__restrict__ D, long pitch, float *__restrict__ sc, int n, float ta,
                             float tb)
{
    __shared__ float red[256];
    const float *p = D + (size_t)blockIdx.x * pitch;
    float m = 0.f;
    for (long x = threadIdx.x; x < pitch; x += blockDim.x) m = fmaxf(m, fabsf(p[x]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        m = red[0];
        float a = 1.f, b = 1.f;
        if (m > 0.f) {
            a = exp2f(floorf(log2f(ta / m)));
            b = exp2f(floorf(log2f(tb / m)));
            /* log2f may round up across an integer: keep |x| s within range */
            while (m * a > ta) a *= 0.5f;
            while (m * b > tb) b *= 0.5f;
        }
        sc[blockIdx.x] = a;
        sc[n + blockIdx.x] = b;
    }
}

__global__ void k_fp8_quant(const float *__restrict__ D, long pitch, int n, const float *__restrict__ t,
                            uint8_t *__restrict__ Q)
{
    const size_t tot4 = (size_t)n * pitch / 4;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot4; i += (size_t)gridDim.x * blockDim.x) {
        const size_t e = 4 * i;
        const float f = t[e / pitch];
        const float4 v = *(const float4 *)(D + e);
        int w = __builtin_amdgcn_cvt_pk_fp8_f32(v.x * f, v.y * f, 0, false);
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v.z * f, v.w * f, w, true);
        *(int *)(Q + e) = w;
    }
}

hipError_t amb_launch_fp8_scales(hipStream_t s, const float *D, long pitch, int n, float *sc)
{
    /* e4m3 max 448: A-side factors to sqrt(448) each, B side to 448 */
    hipLaunchKernelGGL(k_fp8_scales, dim3(n), dim3(256), 0, s, D, pitch, sc, n, 21.166010488516726f, 448.f);
    return hipGetLastError();
}

hipError_t amb_launch_fp8_quant(hipStream_t s, const float *D, long pitch, int n, const float *t, uint8_t *Q)
{
    if (pitch % 4 != 0) return hipErrorInvalidValue;
    size_t tot4 = (size_t)n * pitch / 4;
    int blocks = (int)std::min<size_t>((tot4 + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fp8_quant, dim3(blocks), dim3(256), 0, s, D, pitch, n, t, Q);
    return hipGetLastError();
}

/* Lag reduction: cell (d2, d3) sums C[p1][p1+d2][p1+d3] over valid p1 in
 * ascending order with Neumaier-compensated FP32 (data) and exact int64
 * (normalisation, in units of 1/8), and counts valid p1 (v1 numdnu,
 * amber.c:266). */
__device__ __forceinline__ void reduce_cell(const ReduceArgs &a, long cell)
{
    const int W = 2 * a.nc10 + 1;
    const int *__restrict__ first_ge = a.inv + a.nc; /* inv[nc] then first_ge[nc+1] */
    const int d2 = (int)(cell / W) - a.nc10, d3 = (int)(cell % W) - a.nc10;
    /* r2 == r3: cell (d2, d3) with d2 > d3 sums the same elements in the same
       order as (d3, d2), which writes both */
    if (a.sym23 && d2 > d3) return;
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
    /* eight p1 at a time: the index loads and the C loads of a batch are
       independent, the compensated sums still run in ascending p1 order */
    constexpr int U = 8;
    int i1 = plo < phi ? first_ge[plo] : a.nact;
    while (i1 < a.nact) {
        size_t idx[U];
        bool ok[U];
        bool past = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i1 + u;
            const int p1 = i < a.nact ? a.act[i] : phi;
            past |= p1 >= phi;
            ok[u] = p1 < phi;
            const int j2 = ok[u] ? a.inv[p1 + d2] : -1, j3 = ok[u] ? a.inv[p1 + d3] : -1;
            ok[u] = ok[u] && j2 >= 0 && j3 >= 0;
            /* r1 == r2: rows below i are read from their mirror; r2 == r3
               needs none, d2 <= d3 here and compact indices ascend with
               the channel, so j2 <= j3 */
            const bool sw = a.sym12 && j2 < i;
            idx[u] = ok[u] ? ((size_t)(sw ? j2 : i) * na + (sw ? i : j2)) * na + j3 : 0;
        }
        float x[U];
        int qx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            x[u] = (a.C && ok[u]) ? a.C[idx[u]] : 0.f;
            qx[u] = (a.Cq && ok[u]) ? a.Cq[idx[u]] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!ok[u]) continue;
            const float tsum = sum + x[u];
            if (fabsf(sum) >= fabsf(x[u])) comp += (sum - tsum) + x[u];
            else comp += (x[u] - tsum) + sum;
            sum = tsum;
            q += qx[u];
            ++n;
        }
        if (past) break;
        i1 += U;
    }
    const long mirror = (long)(d3 + a.nc10) * W + (d2 + a.nc10);
    const int nw = (a.sym23 && d2 != d3) ? 2 : 1;
    /* sums over a compute grid scaled to the reference grid in FP64 */
    const bool scaled = a.gscale > 0.0 && a.gscale != 1.0;
    const double sd = ((double)sum + (double)comp) * (scaled ? a.gscale : 1.0);
    for (int w = 0; w < nw; ++w) {
        const long o = w ? mirror : cell;
        if (a.bl) a.bl[o] = scaled ? (float)sd : sum + comp;
        if (a.bl64) a.bl64[o] = sd;
        if (a.cnt) a.cnt[o] = n;
        if (a.Cq) {
            const double v = (a.Pout > 0.0 ? a.Pout : (double)a.P) * ((double)q * 0.125);
            if (a.norm) a.norm[o] = (float)v;
            if (a.norm64) a.norm64[o] = v;
        }
    }
}

__global__ void k_reduce(ReduceArgs a)
{
    const int W = 2 * a.nc10 + 1;
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell < (long)W * W) reduce_cell(a, cell);
}

/* many triangles: cbegin[t] is the first global cell of triangle t */
__global__ void k_reduce_grouped(const ReduceArgs *__restrict__ ra, const long *__restrict__ cbegin, int ntri)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= cbegin[ntri]) return;
    int lo = 0, hi = ntri - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cbegin[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    reduce_cell(ra[lo], g - cbegin[lo]);
}

hipError_t amb_launch_reduce_grouped(hipStream_t s, const ReduceArgs *d_ra, const long *d_cbegin, int ntri,
                                     long total_cells)
{
    if (ntri <= 0 || total_cells <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_grouped, dim3((unsigned)((total_cells + 255) / 256)), dim3(256), 0, s, d_ra,
                       d_cbegin, ntri);
    return hipGetLastError();
}

hipError_t amb_launch_reduce(hipStream_t s, const ReduceArgs &a)
{
    const int W = 2 * a.nc10 + 1;
    const long cells = (long)W * W;
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

/* calCR triple sums (amber_funcs.c:451-456): one block per triangle, FP64
 * accumulation in a fixed thread/tree order. */
__global__ void k_triple_sum(const double *__restrict__ II, long P, const int *__restrict__ tri, double *out)
{
    __shared__ double red[256];
    const int t = blockIdx.x;
    const double *a = II + (size_t)tri[3 * t] * P, *b = II + (size_t)tri[3 * t + 1] * P,
                 *c = II + (size_t)tri[3 * t + 2] * P;
    double s = 0.0;
    for (long x = threadIdx.x; x < P; x += blockDim.x) s += a[x] * b[x] * c[x];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[t] = red[0];
}

hipError_t amb_launch_triple_sum(hipStream_t s, const double *II, long P, const int *tri, int ntri, double *out)
{
    if (ntri <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_triple_sum, dim3(ntri), dim3(256), 0, s, II, P, tri, out);
    return hipGetLastError();
}

/* ---------------------------------------------------------------------------
 * The exact v2 normalisation (amber_v2.c:262-268): for the triangles whose
 * split or FP32 sampling sums are not within 1/32 of their 1/8 grid, the
 * same products on FP64 images with FP64 MFMA and whole-range FP64 sums.
 * An item is a 64 x 64 (p2, p3) tile and 4 reference planes; wave q holds
 * plane g0 + q as 4 x 4 blocks of v_mfma_f64_16x16x4f64 (A[m][k] = lane
 * m + 16 k, B[k][n] = lane n + 16 k, lane l's four results are rows
 * 4 r + l / 16 of column l % 16).  Chunks of 16 pixels are double-buffered
 * through LDS (rows padded by one double so a read's 16 rows hit distinct
 * banks); the next chunk's loads are in flight during the current one's
 * products.  Items are taken in a grid-stride loop.
 * ------------------------------------------------------------------------- */
#define T64 64
#define KC64 16

struct Pref64 {
    double2 a[2], b[2]; /* D2 and D3: rows tid/4, pixels 4 (tid%4) .. +4 */
    double d1;          /* D1: plane tid/16, pixel tid%16 (threads 0..63) */
};

__device__ __forceinline__ void pref64(Pref64 &r, const ContractArgs64 &a, long k0, int r0, int c0, int g0, int tid)
{
    const int row = tid >> 2, k = (tid & 3) * 4;
    const int last = a.nact - 1;
    /* rows past nact read a valid row; their results are never stored */
    const double *pa = a.D2 + (size_t)min(r0 + row, last) * a.pitch + k0 + k;
    const double *pb = a.D3 + (size_t)min(c0 + row, last) * a.pitch + k0 + k;
    r.a[0] = *(const double2 *)pa; r.a[1] = *(const double2 *)(pa + 2);
    r.b[0] = *(const double2 *)pb; r.b[1] = *(const double2 *)(pb + 2);
    if (tid < QP * KC64) r.d1 = a.D1[(size_t)min(g0 + (tid >> 4), last) * a.pitch + k0 + (tid & 15)];
}

struct Smem64 {
    double A[2][T64][KC64 + 1];
    double B[2][T64][KC64 + 1];
    double D[2][QP][KC64];
};

__device__ __forceinline__ void stage64(Smem64 &L, int st, const Pref64 &r, int tid)
{
    const int row = tid >> 2, k = (tid & 3) * 4;
    L.A[st][row][k] = r.a[0].x; L.A[st][row][k + 1] = r.a[0].y;
    L.A[st][row][k + 2] = r.a[1].x; L.A[st][row][k + 3] = r.a[1].y;
    L.B[st][row][k] = r.b[0].x; L.B[st][row][k + 1] = r.b[0].y;
    L.B[st][row][k + 2] = r.b[1].x; L.B[st][row][k + 3] = r.b[1].y;
    if (tid < QP * KC64) L.D[st][tid >> 4][tid & 15] = r.d1;
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_norm64(ContractArgs64 a, int ngroups, int nt, long nwork)
{
    __shared__ Smem64 L;
    const int tid = threadIdx.x, lane = tid & 63;
    const int q = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int fm = lane & 15, fk = lane >> 4;
    const long nchunk = a.pitch / KC64;
    const double invP = 1.0 / (double)a.P;
    for (long w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int grp = (int)(w % ngroups), tile = (int)(w / ngroups);
        const int t2 = tile / nt, t3 = tile - t2 * nt;
        const int r0 = t2 * T64, c0 = t3 * T64, g0 = grp * QP;
        const int lo = a.wlo[g0], hi = a.whi[min(g0 + QP, a.nact) - 1];
        if (r0 >= hi || r0 + T64 <= lo || c0 >= hi || c0 + T64 <= lo) continue;
        if (a.sym12 && r0 + T64 <= g0) continue;
        if (a.sym23 && r0 >= c0 + T64) continue;
        const bool live = g0 + q < a.nact;
        f64x4 acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = (f64x4){0.0, 0.0, 0.0, 0.0};
        Pref64 pf;
        pref64(pf, a, 0, r0, c0, g0, tid);
        __syncthreads(); /* the previous item's reads of L are done */
        stage64(L, 0, pf, tid);
        __syncthreads();
        for (long c = 0; c < nchunk; ++c) {
            const int cur = (int)(c & 1);
            if (c + 1 < nchunk) pref64(pf, a, (c + 1) * KC64, r0, c0, g0, tid);
            if (live) {
#pragma unroll
                for (int s = 0; s < KC64 / 4; ++s) {
                    const int k = 4 * s + fk;
                    const double d1 = L.D[cur][q][k];
                    double fa[4], fb[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) fa[i] = L.A[cur][16 * i + fm][k] * d1;
#pragma unroll
                    for (int j = 0; j < 4; ++j) fb[j] = L.B[cur][16 * j + fm][k];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
                }
            }
            if (c + 1 < nchunk) stage64(L, cur ^ 1, pf, tid);
            __syncthreads();
        }
        if (!live) continue;
        const int i1 = g0 + q, wlo = a.wlo[i1], whi = a.whi[i1];
        float worst = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int gi = r0 + 16 * i + 4 * r + fk, gj = c0 + 16 * j + fm;
                    if (gi < wlo || gi >= whi || gj < wlo || gj >= whi) continue;
                    const double val = acc[i][j][r] * invP;
                    const double q8 = rint(val * 8.0);
                    const size_t idx = ((size_t)i1 * a.nact + gi) * a.nact + gj;
                    if (fabs(q8) < 2147483647.0) {
                        worst = fmaxf(worst, (float)fabs(val - q8 * 0.125));
                        a.Cq[idx] = (int)q8;
                    } else {
                        worst = AMB_CERT_RANGE;
                        a.Cq[idx] = q8 > 0 ? 2147483647 : -2147483647;
                    }
                }
        if (worst > 0.f) atomicMax(a.margin, __float_as_uint(worst));
    }
}

hipError_t amb_launch_norm64(hipStream_t s, const ContractArgs64 &a, int nblocks)
{
    if (a.nact <= 0 || a.pitch % KC64 != 0 || !a.D1 || !a.D2 || !a.D3 || !a.Cq || !a.margin) return hipErrorInvalidValue;
    const int nt = (a.nact + T64 - 1) / T64, ngroups = (a.nact + QP - 1) / QP;
    const long nwork = (long)nt * nt * ngroups;
    const int blocks = (int)std::max<long>(1, std::min<long>(nwork, nblocks));
    hipLaunchKernelGGL(k_norm64, dim3(blocks), dim3(256), 0, s, a, ngroups, nt, nwork);
    return hipGetLastError();
}

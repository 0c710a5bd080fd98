/* Internal declarations shared by the HIP translation units of amber_mi355x. */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AMB_MAX_RADIX_STAGES 24
/* bytes of the static device scratch of the allocation-free entry points */
#define AMB_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* Stockham plan for one transform length n: radices in the order applied. */
struct FftPlan {
    int n;
    int nst;
    int rad[AMB_MAX_RADIX_STAGES];
};

/* Ring geometry of one filtered-plane set. */
struct RingGeom {
    double lo, hi;  /* [ULow, UHigh) in cells */
    int ng;         /* reduced grid side */
    int jcnt;       /* columns that can hold ring cells: min(ng/3, ceil(hi)) */
    long s_lo, s_hi; /* a cell (ia, j) is in the ring iff s_lo <= ia^2 + j^2 < s_hi:
                        the least s >= 1 with sqrt(s) >= ULow and the least s with
                        sqrt(s) >= UHigh, sqrt in double as the reference's test
                        (amber_funcs.c:588) evaluates it */
};

/* ---- FFT stage (amb_fft.hip) ---- */
/* Data plane sets are stored normalised: D 2^(AMB_SET_EXP - E) with
 * E = ceil(log2 max |D|) over the set, so |stored| <= 2^AMB_SET_EXP; two
 * words after the set hold max |D| (float bits) and E.  Products of three
 * stored planes carry 2^(3 AMB_SET_EXP - E1 - E2 - E3), removed exactly. */
#define AMB_SET_EXP 7
#define AMB_SET_SLOT_BYTES 256
/* Smallest E with m <= 2^E (0 for m == 0 or not finite), clamped to
 * [-60, 60]: the exponent of a set whose max |D| is m. */
__device__ __forceinline__ int amb_set_exp(float m)
{
    if (!(m > 0.f) || !(m <= 3.0e38f)) return 0;
    int e;
    const float f = frexpf(m, &e); /* m = f 2^e, f in [0.5, 1) */
    if (f == 0.5f) --e;
    return e < -60 ? -60 : e > 60 ? 60 : e;
}
/* kind 0: data (complex vis), kind 1: sampling (0/1 real).
 * G scratch: [npl][jcnt][n] complex64.  out: npl planes of n*n floats at
 * `pitch` floats apart; the pitch - n*n tail of each plane is zeroed.
 * slot (data sets; null: stored as computed): the set's max word (and E
 * after it), the set normalised in place unless raw (then the max word
 * only, the planes as computed). */
hipError_t amb_launch_ring_c2r(hipStream_t s, int kind, const float2 *vis_t, const float *samp_t,
                               const int *d_planes, int npl, int N, int Nh, const FftPlan &pl,
                               const float2 *d_tw, RingGeom g, float2 *G, float *out, long pitch,
                               unsigned *slot, bool raw = false);
void amb_make_plan(int n, FftPlan *pl);
/* FP64 c2r of npl sampling planes' ring (d_planes: their channel ids on
 * the device) into out[npl][pitch] (tails zeroed); G: [npl][jcnt][n]
 * complex128 scratch; d_tw: n twiddles from amb_launch_twiddles64 */
hipError_t amb_launch_ring_c2r64(hipStream_t s, const float *samp_t, const int *d_planes, int npl, int N, int Nh,
                                 const FftPlan &pl, const double2 *d_tw, RingGeom g, double2 *G, double *out,
                                 long pitch);
hipError_t amb_launch_twiddles64(hipStream_t s, double2 *tw, int n);

/* ---- cube upload (amb_fft.hip) ---- */
hipError_t amb_launch_upload(hipStream_t s, const double *gv, const double *samp, int i0, int rows,
                             int N, int Nh, int nc, float2 *vis_t, float *samp_t);

/* ---- contraction (amb_contract.hip) ---- */
struct ContractArgs {
    const float *D1, *D2;      /* [nact][pitch] FP32 */
    const float *D3;           /* FP32 form: [nact][pitch] */
    const uint16_t *H3, *L3;   /* split forms: hi / lo halves of D3 (BF16 bits,
                                  or FP16 bits of the stored D3 times 2^8) */
    const uint8_t *Q3;         /* FP8 form: e4m3 D3 times t3 */
    const float *s1, *s2, *t3; /* FP8 form: per-plane scales of D1, D2, D3 */
    const int *ex1, *ex2, *ex3; /* data forms: exponents E of the normalised
                                   plane sets of r1, r2, r3 */
    long P, pitch;             /* pixels, plane pitch (multiple of 32) */
    int nact;
    int sym12;                 /* r1 == r2 and full windows: rows below the
                                  reference planes are never needed */
    int sym23;                 /* r2 == r3 (r1 != r2): tiles below the
                                  (p2, p3) diagonal are never needed */
    const int *wlo, *whi;      /* per-i1 window in compact channel indices */
    float *C;                  /* data: [nact][nact][nact] f32 */
    int *Cq;                   /* norm: [nact][nact][nact] int32 of 8*count */
    unsigned *margin;          /* norm: max |S/P - q| as float bits */
    long item0;                /* grouped launches: the triangle's item that the
                                  descriptor's first item is (a launch may run a
                                  sub-range of a triangle's items; each item owns
                                  its C block, so ranges run in any order) */
};
/* var: kernel variant (0: plain k-steps, 1: BF16x3 k-steps interleaving
 * MFMA with the next block's A formation, 2: BF16x3 data and normalisation
 * forms with direct global->LDS loads and k-step pipelining; other forms
 * run variant 0) */
#define AMB_NVAR 3
/* margin word value of a normalisation sum beyond the int32 range of eighths */
#define AMB_CERT_RANGE 1e30f
/* variant 2 addresses each plane array with 32-bit byte offsets: true when
 * nact rows of pitch FP32 values fit in 4 GiB; otherwise run variant 0 */
bool amb_contract_v2_fits(int nact, long pitch);
hipError_t amb_launch_contract(hipStream_t s, int prec, int norm, int var, const ContractArgs &a);
hipError_t amb_launch_split(hipStream_t s, const float *D, size_t n, __bf16 *H, __bf16 *L);
/* FP16 hi / lo split of n stored (normalised) floats times 2^8, so
 * |x 2^8| <= 2^15: H = fp16(x 2^8), L = fp16(x 2^8 - H), both rounded to
 * nearest */
#define AMB_F16_B_EXP 8
hipError_t amb_launch_split_f16(hipStream_t s, const float *D, size_t n, _Float16 *H, _Float16 *L);
/* FP8 preparation of n planes: sc[0..n) A-side scales, sc[n..2n) B-side
 * scales; quant writes the e4m3 planes D*t[i] */
hipError_t amb_launch_fp8_scales(hipStream_t s, const float *D, long pitch, int n, float *sc);
hipError_t amb_launch_fp8_quant(hipStream_t s, const float *D, long pitch, int n, const float *t, uint8_t *Q);

/* The exact v2 normalisation: FP64 sampling images, FP64 MFMA products,
 * whole-pixel-range FP64 sums rounded to their 1/8 grid (Cq, margin). */
struct ContractArgs64 {
    const double *D1, *D2, *D3; /* [nact][pitch] FP64 sampling images of r1, r2, r3 */
    long P, pitch;
    int nact, sym12, sym23;
    const int *wlo, *whi;
    int *Cq;                    /* [nact][nact][nact] int32 of 8 * count */
    unsigned *margin;           /* max |S/P - q| as float bits */
};
hipError_t amb_launch_norm64(hipStream_t s, const ContractArgs64 &a, int nblocks);

struct ReduceArgs {
    const float *C;     /* data partials or NULL */
    const int *Cq;      /* norm partials or NULL */
    const int *act, *inv;
    int nc, nact, nc10;
    int sym12;          /* read C[p2][p1][p3] when p2 < p1 */
    int sym23;          /* r2 == r3: sum cells d2 <= d3, write their mirrors */
    long P;
    double gscale;      /* data sums times (Ng/n)^2, the compute grid's sums as the reference
                           grid's (0 or 1: as summed) */
    double Pout;        /* norm: written as eighths of this many pixels, Ng^2 (0: P) */
    float *bl;
    double *bl64;
    int *cnt;
    float *norm;
    double *norm64;
};
hipError_t amb_launch_reduce(hipStream_t s, const ReduceArgs &a);

/* ---- many triangles per launch (batches, serving) ---- */
long amb_contract_items(int nact); /* work items of one triangle */
/* items the kernel computes (not skipped by its window and r1 == r2 tests),
 * from host copies of the per-plane windows */
long amb_contract_live_items(int nact, const int *wlo, const int *whi, int sym12, int sym23);
/* MFMA flops one computed item issues at pixel pitch `pitch` */
double amb_contract_item_flops(int prec, long pitch);
bool amb_contract_args_ok(int prec, int norm, const ContractArgs &a);
/* persistent contraction over ntri triangles: descs/begin in device memory,
 * begin[t] = first item of triangle t (largest first), begin[ntri] = total;
 * d_counter (AMB_GROUPED_COUNTERS words, one work queue per XCD) is reset
 * by a kernel inside the launch sequence */
#define AMB_GROUPED_COUNTERS 8
hipError_t amb_launch_contract_grouped(hipStream_t s, int prec, int norm, int var, const ContractArgs *d_descs,
                                       const long *d_begin, int ntri, long total, unsigned long long *d_counter,
                                       int nblocks, int xrun_log2);
hipError_t amb_launch_reduce_grouped(hipStream_t s, const ReduceArgs *d_ra, const long *d_cbegin, int ntri,
                                     long total_cells);
/* workgroups of the persistent grouped contraction that are resident at
 * once on device `dev` (CUs x occupancy of the launched form) */
int amb_contract_grouped_blocks(int dev);
/* n words set to v by a kernel (graph-capturable, unlike a memset) */
hipError_t amb_launch_fill_u32(hipStream_t s, unsigned *p, long n, unsigned v);


/* calCR triple sums: one FP64 sum per triangle over P pixels of FP64 images. */
hipError_t amb_launch_triple_sum(hipStream_t s, const double *II, long P, const int *tri, int ntri,
                                 double *out);

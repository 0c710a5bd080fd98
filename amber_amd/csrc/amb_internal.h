/* Internal declarations shared by the HIP translation units of amber_mi355x. */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AMB_MAX_RADIX_STAGES 24

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
};

/* ---- FFT stage (amb_fft.hip) ---- */
/* kind 0: data (complex vis), kind 1: sampling (0/1 real).
 * G scratch: [npl][jcnt][n] complex64.  out: [npl][n][n] float. */
hipError_t amb_launch_ring_c2r(hipStream_t s, int kind, const float2 *vis_t, const float *samp_t,
                               const int *d_planes, int npl, int N, int Nh, const FftPlan &pl,
                               const float2 *d_tw, RingGeom g, float2 *G, float *out);
void amb_make_plan(int n, FftPlan *pl);

/* ---- cube upload (amb_fft.hip) ---- */
hipError_t amb_launch_upload(hipStream_t s, const double *gv, const double *samp, int i0, int rows,
                             int N, int Nh, int nc, float2 *vis_t, float *samp_t);

/* ---- contraction (amb_contract.hip) ---- */
struct ContractArgs {
    const float *D1, *D2, *D3; /* [nact][P] */
    long P;
    int nact;
    const int *wlo, *whi;      /* per-i1 window in compact channel indices */
    float *C;                  /* data: [nact][nact][nact] f32 */
    int *Cq;                   /* norm: [nact][nact][nact] int32 of 8*count */
    unsigned *margin;          /* norm: max |S/P - q| as float bits */
    const float *scale;        /* fp8: per-plane scale, [3][nact] */
};
hipError_t amb_launch_contract(hipStream_t s, int prec, int norm, const ContractArgs &a);

struct ReduceArgs {
    const float *C;     /* data partials or NULL */
    const int *Cq;      /* norm partials or NULL */
    const int *act, *inv;
    int nc, nact, nc10;
    long P;
    float *bl;
    double *bl64;
    int *cnt;
    float *norm;
    double *norm64;
};
hipError_t amb_launch_reduce(hipStream_t s, const ReduceArgs &a);

/* calCR triple sums: one FP64 sum per triangle over P pixels. */
hipError_t amb_launch_triple_sum(hipStream_t s, const float *II, long P, const int *tri, int ntri,
                                 double *out);

/* per-plane absmax for FP8 scaling: out[i] = max |x| over plane i */
hipError_t amb_launch_absmax(hipStream_t s, const float *D, long P, int npl, float *out);

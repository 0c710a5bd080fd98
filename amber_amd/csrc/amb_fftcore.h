/*
 * Mixed-radix Stockham FFT in LDS for gfx950, shared by the ring c2r, the
 * channel-axis transforms of the lag-Fourier form and the lag-space
 * post-processing.
 *
 * A transform of length n is owned by a team of threads (one or more
 * 64-wide wavefronts); a workgroup holds one or more teams, all running
 * the same plan, so the stage loop ends with one workgroup barrier per
 * stage.  A stage of radix R gives each thread butterflies j, j + T, ...
 * (T threads per team), whose (group, offset) = (j / Ns, j % Ns) indices
 * advance by a constant step: no integer division inside the loops.
 * Butterflies are register codelets of compile-time size: 2, 3, 4, 8 and the
 * odd primes up to 47, the primes as a symmetric-pair DFT whose cos/sin
 * coefficients are read from the transform's own twiddle table
 * (tw[t n/R] = e^{2 pi i t/R}).  Any larger prime factor runs a direct stage
 * (inputs twiddled in place, then two outputs per thread and step from
 * symmetric pairs).
 *
 * Direction DIR = +1 computes y_k = sum_m x_m e^{+2 pi i m k / n} (the
 * unnormalised inverse, FFTW's c2r sign), DIR = -1 the forward transform.
 * The twiddle table always holds e^{+2 pi i m / n}; the forward direction
 * uses its conjugates.
 */
#pragma once
#include "amb_internal.h"

/* stages and butterflies also compile for the host, where a test runs a
   team's threads one after another */
#define FFTC_HD __host__ __device__ __forceinline__

namespace fftc {

/* plan of length n: radix-8 stages, a radix-4 stage, a radix-2 stage,
   radix-3 stages, then the other prime factors in ascending order */
inline void make_plan(int n, FftPlan *pl)
{
    pl->n = n;
    pl->nst = 0;
    int m = n;
    while (m % 8 == 0 && pl->nst < AMB_MAX_RADIX_STAGES) { pl->rad[pl->nst++] = 8; m /= 8; }
    while (m % 4 == 0 && pl->nst < AMB_MAX_RADIX_STAGES) { pl->rad[pl->nst++] = 4; m /= 4; }
    while (m % 2 == 0 && pl->nst < AMB_MAX_RADIX_STAGES) { pl->rad[pl->nst++] = 2; m /= 2; }
    while (m % 3 == 0 && pl->nst < AMB_MAX_RADIX_STAGES) { pl->rad[pl->nst++] = 3; m /= 3; }
    for (int p = 5; m > 1 && pl->nst < AMB_MAX_RADIX_STAGES; p += 2)
        while (m % p == 0) { pl->rad[pl->nst++] = p; m /= p; }
}

template <typename T>
struct CT;
template <>
struct CT<float> { typedef float2 c; };
template <>
struct CT<double> { typedef double2 c; };

FFTC_HD float2 cx(float a, float b) { return make_float2(a, b); }
FFTC_HD double2 cx(double a, double b) { return make_double2(a, b); }
template <typename C>
FFTC_HD C cadd(C a, C b) { return cx(a.x + b.x, a.y + b.y); }
template <typename C>
FFTC_HD C csub(C a, C b) { return cx(a.x - b.x, a.y - b.y); }
template <typename C>
FFTC_HD C cmul(C a, C b) { return cx(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
/* a times w (DIR > 0) or conj(w) (DIR < 0) */
template <int DIR, typename C>
FFTC_HD C twmul(C a, C w)
{
    if (DIR > 0) return cx(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
    return cx(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}
/* a times e^{DIR i pi / 2} */
template <int DIR, typename C>
FFTC_HD C rot90(C a)
{
    return DIR > 0 ? cx(-a.y, a.x) : cx(a.y, -a.x);
}

/* R-point DFT of the inputs v (already twiddled),
 * y_q = sum_r v_r e^{DIR 2 pi i r q / R}, written to yo[q * ys].
 * tw: the transform's table, stride rstep = n / R between the R-th roots. */
template <int R, int DIR, typename C>
struct Bfly {
    /* odd prime R: symmetric pairs a_m = v_m + v_{R-m}, b_m = v_m - v_{R-m};
       y_q = v_0 + sum a_m cos(2 pi m q / R) +- i sum b_m DIR sin(...); each
       output pair is stored as soon as it is formed, so the outputs never
       occupy registers beside the pairs */
    FFTC_HD static void run(C *v, const C *tw, int rstep, C *yo, int ys)
    {
        typedef decltype(v[0].x) T;
        constexpr int h = (R - 1) / 2;
        T cs[h + 1], sn[h + 1];
#pragma unroll
        for (int m = 1; m <= h; ++m) {
            const C w = tw[m * rstep];
            cs[m] = w.x;
            sn[m] = DIR > 0 ? w.y : -w.y;
        }
        C y0 = v[0];
#pragma unroll
        for (int m = 1; m <= h; ++m) {
            const C a = cadd(v[m], v[R - m]), b = csub(v[m], v[R - m]);
            v[m] = a;
            v[R - m] = b;
            y0 = cadd(y0, a);
        }
#pragma unroll
        for (int q = 1; q <= h; ++q) {
            C A = v[0], B = cx((T)0, (T)0);
#pragma unroll
            for (int m = 1; m <= h; ++m) {
                const int t = (m * q) % R;
                const T c = t <= h ? cs[t] : cs[R - t];
                const T s = t <= h ? sn[t] : -sn[R - t];
                A.x += v[m].x * c;
                A.y += v[m].y * c;
                B.x += v[R - m].x * s;
                B.y += v[R - m].y * s;
            }
            /* i B = (-B.y, B.x) */
            yo[q * ys] = cx(A.x - B.y, A.y + B.x);
            yo[(R - q) * ys] = cx(A.x + B.y, A.y - B.x);
        }
        yo[0] = y0;
    }
};
template <int DIR, typename C>
struct Bfly<2, DIR, C> {
    FFTC_HD static void run(C *v, const C *, int, C *yo, int ys)
    {
        yo[0] = cadd(v[0], v[1]);
        yo[ys] = csub(v[0], v[1]);
    }
};
template <int DIR, typename C>
struct Bfly<3, DIR, C> {
    FFTC_HD static void run(C *v, const C *, int, C *yo, int ys)
    {
        typedef decltype(v[0].x) T;
        const T h = (T)0.86602540378443864676; /* sin(2 pi / 3) */
        const C sm = cadd(v[1], v[2]), df = csub(v[1], v[2]);
        const C m = cx(v[0].x - (T)0.5 * sm.x, v[0].y - (T)0.5 * sm.y);
        const C r = rot90<DIR>(cx(h * df.x, h * df.y));
        yo[0] = cadd(v[0], sm);
        yo[ys] = cadd(m, r);
        yo[2 * ys] = csub(m, r);
    }
};
template <int DIR, typename C>
struct Bfly<4, DIR, C> {
    FFTC_HD static void run(C *v, const C *, int, C *yo, int ys)
    {
        const C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const C s13 = cadd(v[1], v[3]), r13 = rot90<DIR>(csub(v[1], v[3]));
        yo[0] = cadd(s02, s13);
        yo[ys] = cadd(d02, r13);
        yo[2 * ys] = csub(s02, s13);
        yo[3 * ys] = csub(d02, r13);
    }
};

template <int DIR, typename C>
struct Bfly<8, DIR, C> {
    /* two 4-point DFTs of the even and odd inputs, combined with w^q,
       w = e^{DIR 2 pi i / 8}: y_q = E_q + w^q O_q, y_{q+4} = E_q - w^q O_q */
    FFTC_HD static void run(C *v, const C *, int, C *yo, int ys)
    {
        typedef decltype(v[0].x) T;
        const T c = (T)0.70710678118654752440; /* cos(pi / 4) */
        const C es02 = cadd(v[0], v[4]), ed02 = csub(v[0], v[4]);
        const C es13 = cadd(v[2], v[6]), er13 = rot90<DIR>(csub(v[2], v[6]));
        const C os02 = cadd(v[1], v[5]), od02 = csub(v[1], v[5]);
        const C os13 = cadd(v[3], v[7]), or13 = rot90<DIR>(csub(v[3], v[7]));
        const C e0 = cadd(es02, es13), e1 = cadd(ed02, er13), e2 = csub(es02, es13), e3 = csub(ed02, er13);
        const C o0 = cadd(os02, os13), o1 = cadd(od02, or13), o2 = csub(os02, os13), o3 = csub(od02, or13);
        const C r1 = rot90<DIR>(o1), r3 = rot90<DIR>(o3);
        const C w1 = cx(c * (o1.x + r1.x), c * (o1.y + r1.y));  /* w o1 */
        const C w2 = rot90<DIR>(o2);                            /* w^2 o2 */
        const C w3 = cx(c * (r3.x - o3.x), c * (r3.y - o3.y));  /* w^3 o3 */
        yo[0] = cadd(e0, o0);
        yo[ys] = cadd(e1, w1);
        yo[2 * ys] = cadd(e2, w2);
        yo[3 * ys] = cadd(e3, w3);
        yo[4 * ys] = csub(e0, o0);
        yo[5 * ys] = csub(e1, w1);
        yo[6 * ys] = csub(e2, w2);
        yo[7 * ys] = csub(e3, w3);
    }
};

/* One radix-R stage of a team's transform: x -> y (Stockham order). */
template <int R, int DIR, typename C>
FFTC_HD void stage(const C *__restrict__ x, C *__restrict__ y, const C *__restrict__ tw, int n,
                                      int Ns, int t, int T)
{
    const int nR = n / R, tstep = n / (Ns * R);
    int g = t / Ns, k = t - g * Ns;
    const int dg = T / Ns, dk = T - dg * Ns;
    for (int j = t; j < nR; j += T) {
        C v[R];
        v[0] = x[j];
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = twmul<DIR>(x[j + r * nR], tw[r * k * tstep]);
        Bfly<R, DIR, C>::run(v, tw, n / R, y + g * Ns * R + k, Ns);
        k += dk;
        g += dg;
        if (k >= Ns) { k -= Ns; ++g; }
    }
}

/* Stage of a prime R without a codelet, in two steps with a workgroup
 * barrier between them.  First the inputs r >= 1 of every butterfly are
 * twiddled in place, x[j + r nR] *= w^{r k n/(Ns R)}; then each step of a
 * thread forms outputs q and R - q of one butterfly from the symmetric
 * pairs a_m = v_m + v_{R-m}, b_m = v_m - v_{R-m}:
 *   y_q = v_0 + sum a_m cos(2 pi m q/R) + i sum b_m DIR sin(2 pi m q/R),
 *   y_{R-q} = the same with -i,
 * cos/sin read from the transform's table, and q = 0 sums all inputs. */
template <int DIR, typename C>
FFTC_HD void stage_direct_tw(C *__restrict__ x, const C *__restrict__ tw, int n, int R, int Ns, int t, int T)
{
    /* element u of the n - nR inputs r >= 1: j = u mod nR, r = 1 + u / nR,
       k = j mod Ns (Ns divides nR), advanced by constant steps */
    const int nR = n / R, tstep = n / (Ns * R), tot = n - nR;
    const int dj = T % nR, dr = T / nR, dk = dj % Ns;
    int j = t % nR, r = 1 + t / nR, k = j % Ns;
    for (int u = t; u < tot; u += T) {
        x[j + r * nR] = twmul<DIR>(x[j + r * nR], tw[r * k * tstep]);
        j += dj;
        r += dr;
        k += dk;
        if (k >= Ns) k -= Ns;
        if (j >= nR) { j -= nR; ++r; }
    }
}
template <int DIR, typename C>
FFTC_HD void stage_direct_dft(const C *__restrict__ x, C *__restrict__ y, const C *__restrict__ tw, int n, int R,
                              int Ns, int t, int T)
{
    typedef decltype(x[0].x) Tr;
    /* item u of nR (h + 1): butterfly j = u mod nR = g Ns + k, output pair q = u / nR */
    const int nR = n / R, h = (R - 1) / 2, rstep = n / R, tot = nR * (h + 1);
    const int dj = T % nR, dq = T / nR, dgj = dj / Ns, dkj = dj - dgj * Ns, gw = nR / Ns;
    int j = t % nR, q = t / nR, g = j / Ns, k = j - g * Ns;
    for (int u = t; u < tot; u += T) {
        const C *xi = x + j;
        C *yo = y + g * Ns * R + k;
        if (q == 0) {
            C sum = xi[0];
            for (int r = 1; r < R; ++r) sum = cadd(sum, xi[r * nR]);
            yo[0] = sum;
        } else {
            C A = xi[0], B = cx((Tr)0, (Tr)0);
            int tq = 0;
#pragma unroll 4
            for (int m = 1; m <= h; ++m) {
                tq += q;
                if (tq >= R) tq -= R;
                const C w = tw[tq * rstep];
                const Tr c = w.x, sn = DIR > 0 ? w.y : -w.y;
                const C vm = xi[m * nR], vr = xi[(R - m) * nR];
                A.x += (vm.x + vr.x) * c;
                A.y += (vm.y + vr.y) * c;
                B.x += (vm.x - vr.x) * sn;
                B.y += (vm.y - vr.y) * sn;
            }
            yo[q * Ns] = cx(A.x - B.y, A.y + B.x);
            yo[(R - q) * Ns] = cx(A.x + B.y, A.y - B.x);
        }
        j += dj;
        q += dq;
        k += dkj;
        g += dgj;
        if (k >= Ns) { k -= Ns; ++g; }
        if (j >= nR) { j -= nR; ++q; g -= gw; }
    }
}

/* Between the stages of a team's transform.  A team is one 64-wide
 * wavefront (T = 64 in every kernel) and its stages touch only its own
 * buffers, so the wavefront's own LDS order suffices: the fences keep the
 * compiler from moving the team's LDS accesses across the point, and a
 * wavefront's LDS operations complete in issue order.  The workgroup
 * barriers stay where teams exchange data (gathers, scatters) and around
 * the twiddle table a workgroup loads. */
__device__ __forceinline__ void team_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Full transform of a team's buffer x (scratch y of the same size); every
 * thread of the team calls this with the same plan, the team a whole
 * wavefront (T = 64; one team_sync per stage).  t: thread index in the team,
 * T: team size.  The caller's workgroup barrier precedes the call when
 * other teams wrote x or the twiddles.  Returns the buffer holding the
 * result (x or y). */
template <int DIR, int MAXP = 47, typename C>
__device__ C *transform(C *x, C *y, const C *tw, const FftPlan &pl, int t, int T)
{
    static_assert(MAXP >= 5, "radices 2, 3, 4, 5 and 8 are codelets in every kernel");
    const int n = pl.n;
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.rad[s];
        /* prime codelets above MAXP are not instantiated (register pressure
           of the kernel): such factors take the direct stage, which is for
           odd primes only; radix 8 is always a codelet */
        switch (R > MAXP && R != 8 ? 0 : R) {
        case 2: stage<2, DIR>(x, y, tw, n, Ns, t, T); break;
        case 3: stage<3, DIR>(x, y, tw, n, Ns, t, T); break;
        case 4: stage<4, DIR>(x, y, tw, n, Ns, t, T); break;
        case 8: stage<8, DIR>(x, y, tw, n, Ns, t, T); break;
        case 5: if constexpr (MAXP >= 5) stage<5, DIR>(x, y, tw, n, Ns, t, T); break;
        case 7: if constexpr (MAXP >= 7) stage<7, DIR>(x, y, tw, n, Ns, t, T); break;
        case 11: if constexpr (MAXP >= 11) stage<11, DIR>(x, y, tw, n, Ns, t, T); break;
        case 13: if constexpr (MAXP >= 13) stage<13, DIR>(x, y, tw, n, Ns, t, T); break;
        case 17: if constexpr (MAXP >= 17) stage<17, DIR>(x, y, tw, n, Ns, t, T); break;
        case 19: if constexpr (MAXP >= 19) stage<19, DIR>(x, y, tw, n, Ns, t, T); break;
        case 23: if constexpr (MAXP >= 23) stage<23, DIR>(x, y, tw, n, Ns, t, T); break;
        case 29: if constexpr (MAXP >= 29) stage<29, DIR>(x, y, tw, n, Ns, t, T); break;
        case 31: if constexpr (MAXP >= 31) stage<31, DIR>(x, y, tw, n, Ns, t, T); break;
        case 37: if constexpr (MAXP >= 37) stage<37, DIR>(x, y, tw, n, Ns, t, T); break;
        case 41: if constexpr (MAXP >= 41) stage<41, DIR>(x, y, tw, n, Ns, t, T); break;
        case 43: if constexpr (MAXP >= 43) stage<43, DIR>(x, y, tw, n, Ns, t, T); break;
        case 47: if constexpr (MAXP >= 47) stage<47, DIR>(x, y, tw, n, Ns, t, T); break;
        default:
            stage_direct_tw<DIR>(x, tw, n, R, Ns, t, T);
            team_sync();
            stage_direct_dft<DIR>(x, y, tw, n, R, Ns, t, T);
            break;
        }
        team_sync();
        C *tmp = x;
        x = y;
        y = tmp;
        Ns *= R;
    }
    return x;
}

/* ------------------------------------------------------------- in place */
/* Position of input element c in the digit-reversed order that the
 * in-place transform below takes: with the plan's radices R_0 .. R_{S-1},
 * c = d_{S-1} + R_{S-1} (d_{S-2} + R_{S-2} (... + R_1 d_0)) is stored at
 * p = d_0 + R_0 (d_1 + R_1 (... + R_{S-2} d_{S-1})). */
inline int digit_reversed(int c, const FftPlan &pl)
{
    int d[AMB_MAX_RADIX_STAGES];
    for (int s = pl.nst - 1; s >= 0; --s) {
        d[s] = c % pl.rad[s];
        c /= pl.rad[s];
    }
    int p = 0;
    for (int s = pl.nst - 1; s >= 0; --s) p = p * pl.rad[s] + d[s];
    return p;
}

/* Index map of an in-place buffer: position i of the transform is held at
 * x[ix(i)]; ix is a bijection of [0, n). */
struct IdxLinear {
    FFTC_HD int operator()(int i) const { return i; }
};
/* Bank swizzle of an 8-byte-element LDS buffer: bits 0-4 of i are XORed
 * with q = bits 5-7 and with bits 6-7 moved to bits 3-4.  It moves i only
 * inside its aligned block of 32, so it is a bijection of [0, n) when 32
 * divides n (otherwise the identity).  Against 64 banks of 4 bytes (32
 * elements a cycle) it spreads the first two radix-8 stages of a 64-thread
 * team (stride 8 and blocks of 8 at stride 64) over all banks, where the
 * identity puts 8 and 4 lanes on one bank; the later stages, whose lanes
 * read consecutive positions, stay conflict-free. */
struct IdxXor32 {
    int m;
    FFTC_HD explicit IdxXor32(int n) : m(n % 32 == 0 ? 31 : 0) {}
    FFTC_HD int operator()(int i) const
    {
        const int q = (i >> 5) & 7;
        return i ^ ((q | ((q >> 1) & 3) << 3) & m);
    }
};

/* One in-place decimation-in-time stage of radix R: sub-transforms of
 * length Ns are combined R at a time; butterfly (g, k) reads positions
 * g Ns R + k + r Ns, twiddles input r by w^{r k n / (Ns R)} and writes its
 * R outputs back to the same positions (they are in registers first). */
template <int R, int DIR, typename C, typename IX = IdxLinear>
FFTC_HD void stage_ip(C *__restrict__ x, const C *__restrict__ tw, int n, int Ns, int t, int T, IX ix = IX())
{
    const int nR = n / R, tstep = n / (Ns * R);
    int g = t / Ns, k = t - g * Ns;
    const int dg = T / Ns, dk = T - dg * Ns;
    for (int j = t; j < nR; j += T) {
        const int b = g * Ns * R + k;
        C v[R], o[R];
        v[0] = x[ix(b)];
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = twmul<DIR>(x[ix(b + r * Ns)], tw[r * k * tstep]);
        Bfly<R, DIR, C>::run(v, tw, n / R, o, 1);
#pragma unroll
        for (int r = 0; r < R; ++r) x[ix(b + r * Ns)] = o[r];
        k += dk;
        g += dg;
        if (k >= Ns) { k -= Ns; ++g; }
    }
}

/* In-place transform of a team's buffer x (input in digit-reversed order,
 * output in natural order, both through the index map ix) for plans of
 * radices 2, 3, 4, 5 and 8 only (the {2,3,5}-smooth lengths of the channel
 * axis); one team_sync per stage, as transform(). */
template <int DIR, typename C, typename IX = IdxLinear>
__device__ void transform_ip(C *x, const C *tw, const FftPlan &pl, int t, int T, IX ix = IX())
{
    const int n = pl.n;
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.rad[s];
        switch (R) {
        case 2: stage_ip<2, DIR>(x, tw, n, Ns, t, T, ix); break;
        case 3: stage_ip<3, DIR>(x, tw, n, Ns, t, T, ix); break;
        case 4: stage_ip<4, DIR>(x, tw, n, Ns, t, T, ix); break;
        case 8: stage_ip<8, DIR>(x, tw, n, Ns, t, T, ix); break;
        default: stage_ip<5, DIR>(x, tw, n, Ns, t, T, ix); break;
        }
        team_sync();
        Ns *= R;
    }
}

} // namespace fftc

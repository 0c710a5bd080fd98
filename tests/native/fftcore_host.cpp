/*
 * Host build of the LDS FFT core's stages (amber_amd/csrc/amb_fftcore.h):
 * runs a team's T threads one after another per stage, the order a
 * workgroup barrier separates on the device, so tests can compare every
 * plan against numpy without a GPU.
 */
#include "amb_fftcore.h"

#include <cmath>
#include <vector>

template <typename T, int DIR>
static int run(int n, int TT, int maxp, const double *in, double *out)
{
    typedef typename fftc::CT<T>::c C;
    FftPlan pl;
    fftc::make_plan(n, &pl);
    std::vector<C> x(n), y(n), tw(n);
    for (int m = 0; m < n; ++m) {
        const double a = 2.0 * M_PI * (double)m / (double)n;
        tw[m] = fftc::cx((T)cos(a), (T)sin(a));
        x[m] = fftc::cx((T)in[2 * m], (T)in[2 * m + 1]);
    }
    C *px = x.data(), *py = y.data();
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.rad[s];
        /* primes above maxp take the direct stage (its two steps are
           separated by the barrier every thread reaches) */
        const bool direct = (R > maxp && R != 8) || R > 47;
        if (direct) {
            for (int t = 0; t < TT; ++t) fftc::stage_direct_tw<DIR>(px, tw.data(), n, R, Ns, t, TT);
            for (int t = 0; t < TT; ++t) fftc::stage_direct_dft<DIR>(px, py, tw.data(), n, R, Ns, t, TT);
        }
        for (int t = 0; t < TT && !direct; ++t) {
            switch (R) {
#define CASE(r) case r: fftc::stage<r, DIR>(px, py, tw.data(), n, Ns, t, TT); break;
            CASE(2) CASE(3) CASE(4) CASE(5) CASE(7) CASE(8) CASE(11) CASE(13) CASE(17) CASE(19) CASE(23) CASE(29)
            CASE(31) CASE(37) CASE(41) CASE(43) CASE(47)
#undef CASE
            default: break;
            }
        }
        C *tmp = px; px = py; py = tmp;
        Ns *= R;
    }
    for (int m = 0; m < n; ++m) {
        out[2 * m] = (double)px[m].x;
        out[2 * m + 1] = (double)px[m].y;
    }
    return pl.nst;
}

/* maxp: largest prime run as a codelet (the device kernels' MAXP) */
extern "C" int fftc_host(int n, int dir, int dbl, int team, int maxp, const double *in, double *out)
{
    if (n < 1 || team < 1) return -1;
    if (dbl) return dir > 0 ? run<double, 1>(n, team, maxp, in, out) : run<double, -1>(n, team, maxp, in, out);
    return dir > 0 ? run<float, 1>(n, team, maxp, in, out) : run<float, -1>(n, team, maxp, in, out);
}

/* the in-place form: input placed at its digit-reversed positions, the
   buffer read and written through the index map ix */
template <typename T, int DIR, typename IX>
static int run_ip(int n, int TT, const double *in, double *out, IX ix)
{
    typedef typename fftc::CT<T>::c C;
    FftPlan pl;
    fftc::make_plan(n, &pl);
    for (int s = 0; s < pl.nst; ++s)
        if (pl.rad[s] > 5 && pl.rad[s] != 8) return -2;
    std::vector<C> x(n), tw(n);
    for (int m = 0; m < n; ++m) {
        const double a = 2.0 * M_PI * (double)m / (double)n;
        tw[m] = fftc::cx((T)cos(a), (T)sin(a));
        x[ix(fftc::digit_reversed(m, pl))] = fftc::cx((T)in[2 * m], (T)in[2 * m + 1]);
    }
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.rad[s];
        for (int t = 0; t < TT; ++t) {
            switch (R) {
            case 2: fftc::stage_ip<2, DIR>(x.data(), tw.data(), n, Ns, t, TT, ix); break;
            case 3: fftc::stage_ip<3, DIR>(x.data(), tw.data(), n, Ns, t, TT, ix); break;
            case 4: fftc::stage_ip<4, DIR>(x.data(), tw.data(), n, Ns, t, TT, ix); break;
            case 8: fftc::stage_ip<8, DIR>(x.data(), tw.data(), n, Ns, t, TT, ix); break;
            default: fftc::stage_ip<5, DIR>(x.data(), tw.data(), n, Ns, t, TT, ix); break;
            }
        }
        Ns *= R;
    }
    for (int m = 0; m < n; ++m) {
        out[2 * m] = (double)x[ix(m)].x;
        out[2 * m + 1] = (double)x[ix(m)].y;
    }
    return pl.nst;
}

template <typename IX>
static int ip_dispatch(int n, int dir, int dbl, int team, const double *in, double *out, IX ix)
{
    if (n < 1 || team < 1) return -1;
    if (dbl) return dir > 0 ? run_ip<double, 1>(n, team, in, out, ix) : run_ip<double, -1>(n, team, in, out, ix);
    return dir > 0 ? run_ip<float, 1>(n, team, in, out, ix) : run_ip<float, -1>(n, team, in, out, ix);
}

extern "C" int fftc_host_ip(int n, int dir, int dbl, int team, const double *in, double *out)
{
    return ip_dispatch(n, dir, dbl, team, in, out, fftc::IdxLinear());
}

/* the in-place form through the bank swizzle of the channel-spectra kernel */
extern "C" int fftc_host_ip_swz(int n, int dir, int dbl, int team, const double *in, double *out)
{
    return ip_dispatch(n, dir, dbl, team, in, out, fftc::IdxXor32(n));
}

/* the swizzle's image of [0, n) (a permutation when it is a bijection) */
extern "C" void fftc_host_swz_map(int n, int *out)
{
    const fftc::IdxXor32 ix(n);
    for (int i = 0; i < n; ++i) out[i] = ix(i);
}

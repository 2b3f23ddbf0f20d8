// sl_recurrence.h -- exact k-step advance of the barycentric recurrence
// s <- fl(s + D) used by the scanline raster (SL_TriRasterizer.cpp:857-910:
// bc += 4*dbc/dx per group of four pixels).
//
// Stepping k times costs k dependent adds.  Inside one binade of |s| (the
// interval [2^e, 2^(e+1)) where float spacing is u = 2^(e-23)) every step adds
// the same grid multiple: with s = sign * a * u (a an integer in [2^23, 2^24))
// and d = sign * D / u, round-to-nearest gives a <- a + round(d) as long as
// the exact sum a + d stays inside [2^23, 2^24) and d is not a half-integer
// (a tie, whose rounding depends on the parity of a).  So a whole run of
// steps within a binade is one multiply-add; only the steps that cross a
// binade boundary, hit a tie, or involve zero/subnormal/non-finite values
// are taken one at a time.  The result is bit-identical to stepping.
#pragma once

#include <stdint.h>
#include <string.h>
#include <math.h>

#ifndef SL_RECUR_HD
#if defined(__HIPCC__)
#define SL_RECUR_HD __host__ __device__ __forceinline__
#else
#define SL_RECUR_HD inline
#endif
#endif

SL_RECUR_HD uint32_t sl_recur_bits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

// Approximate 1/x for the run length below: any approximation is safe there
// (an overestimate is trimmed, an underestimate only splits a run).
SL_RECUR_HD float sl_recur_rcp(float x)
{
#if defined(SL_RECUR_RCP_OVERRIDE)
    return SL_RECUR_RCP_OVERRIDE(x);     // lets a host check run with a deliberately poor reciprocal
#elif defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// One run of the recurrence: advances s by up to k steps of s = fl(s + D)
// that stay inside s's binade (one multiply-add), or by exactly one step
// when a run is not possible.  Returns the number of steps taken (0 iff k <= 0).
SL_RECUR_HD int sl_recur_run(float& s, float D, int k)
{
    if (k <= 0) return 0;
    const double kLo = 8388608.0 + 1.0;      // 2^23 + 1
    const double kHi = 16777216.0 - 1.0;     // 2^24 - 1
    const uint32_t bits = sl_recur_bits(s);
    const uint32_t ex = (bits >> 23) & 0xFFu;
    if (k >= 3 && ex != 0u && ex != 0xFFu && D == D && D != 0.f && fabsf(D) < 3.0e38f)
    {
        const int sh = (int)ex - 150;                              // s's binade has spacing 2^sh
        const double sign = (bits >> 31) ? -1.0 : 1.0;
        const double a = (double)((bits & 0x7FFFFFu) | 0x800000u);   // |s| / 2^sh
        const double d = sign * ldexp((double)D, -sh);             // step in units of 2^sh (exact)
        const double fl = floor(d);
        if (d - fl != 0.5)                                         // not a tie
        {
            const double r = (d - fl > 0.5) ? fl + 1.0 : fl;       // round(d)
            if (r == 0.0)
            {
                // |d| < 1/2: every step rounds back to s while the sum stays in the binade
                if (a + d >= kLo && a + d <= kHi) return k;
            }
            else
            {
                // m = the largest count with a + (m-1) r + d inside [kLo, kHi]:
                // estimated high, then trimmed
                const double span = (r > 0.0) ? (kHi - d - a) : (a + d - kLo);
                double m = floor(span * (double)sl_recur_rcp((float)fabs(r))) + 3.0;
                if (m > (double)k) m = (double)k;
                while (m >= 1.0)
                {
                    const double last = a + (m - 1.0) * r + d;
                    if (last >= kLo && last <= kHi) break;
                    m -= 1.0;
                }
                if (m >= 1.0)
                {
                    s = (float)ldexp(sign * (a + m * r), sh);   // exact: an integer below 2^24 times 2^sh
                    return (int)m;
                }
            }
        }
    }
    s = s + D;
    return 1;
}

// s after k steps of s = fl(s + D) (no fused operations).
SL_RECUR_HD float sl_recur_advance(float s, float D, int k)
{
    while (k > 0) k -= sl_recur_run(s, D, k);
    return s;
}

// Three independent recurrences advanced together (one run of each per
// round, so their runs overlap).
SL_RECUR_HD void sl_recur_advance3(float& s0, float& s1, float& s2, float D0, float D1, float D2, int k)
{
    int k0 = k, k1 = k, k2 = k;
    while (k0 > 0 || k1 > 0 || k2 > 0)
    {
        k0 -= sl_recur_run(s0, D0, k0);
        k1 -= sl_recur_run(s1, D1, k1);
        k2 -= sl_recur_run(s2, D2, k2);
    }
}

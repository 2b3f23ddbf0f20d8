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

// s after k steps of s = fl(s + D) (no fused operations).
SL_RECUR_HD float sl_recur_advance(float s, float D, int k)
{
    const double kLo = 8388608.0 + 1.0;      // 2^23 + 1
    const double kHi = 16777216.0 - 1.0;     // 2^24 - 1
    while (k > 0)
    {
        const uint32_t bits = sl_recur_bits(s);
        const uint32_t ex = (bits >> 23) & 0xFFu;
        if (k >= 3 && ex != 0u && ex != 0xFFu && D == D && D != 0.f && fabsf(D) < 3.0e38f)
        {
            const double u = ldexp(1.0, (int)ex - 150);           // spacing of floats in s's binade
            const double sign = (bits >> 31) ? -1.0 : 1.0;
            const double a = (double)((bits & 0x7FFFFFu) | 0x800000u);   // |s| / u
            const double d = sign * ((double)D / u);              // step in units of u, toward |s| growing
            const double fl = floor(d);
            if (d - fl != 0.5)                                    // not a tie
            {
                const double r = (d - fl > 0.5) ? fl + 1.0 : fl;  // round(d)
                if (r == 0.0)
                {
                    // |d| < 1/2: every step rounds back to s while the sum stays in the binade
                    if (a + d >= kLo && a + d <= kHi) return s;
                }
                else
                {
                    // largest m with a + (m-1) r + d inside [kLo, kHi]
                    double m;
                    if (r > 0.0) m = floor((kHi - d - a) / r) + 1.0;
                    else m = floor((a + d - kLo) / (-r)) + 1.0;
                    if (m > (double)k) m = (double)k;
                    while (m >= 1.0)
                    {
                        const double last = a + (m - 1.0) * r + d;
                        if (last >= kLo && last <= kHi) break;
                        m -= 1.0;
                    }
                    if (m >= 1.0)
                    {
                        const double na = a + m * r;                    // exact integer, inside [2^23, 2^24]
                        s = (float)(sign * na * u);                       // exact: na * u is a float
                        k -= (int)m;
                        continue;
                    }
                }
            }
        }
        s = s + D;
        --k;
    }
    return s;
}

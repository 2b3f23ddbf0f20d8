// Compares sl_recur_advance(s, D, k) and sl_recur_advance3 with k explicit
// steps s = fl(s + D) on many seeded (s, D, k) cases; prints "cases
// mismatches" and exits non-zero on any mismatch.  Built with
// -ffp-contract=off, like the kernels.  With -DSKEWED_RCP the run-length
// reciprocal is deliberately off by up to 40% either way: the result must not
// change (the run length is trimmed or split, never trusted).
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <cmath>

#ifdef SKEWED_RCP
static uint64_t skew_state = 12345;
static float skewed_rcp(float x)
{
    skew_state = skew_state * 6364136223846793005ull + 1442695040888963407ull;
    const float f = 0.6f + 0.8f * (float)((skew_state >> 40) * (1.0 / 16777216.0));
    return f / x;
}
#define SL_RECUR_RCP_OVERRIDE skewed_rcp
#endif

#include "../../softlight_amd/csrc/sl_recurrence.h"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t nxt()
{
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float unif(float lo, float hi) { return lo + (hi - lo) * (float)((nxt() >> 40) * (1.0 / 16777216.0)); }
static float bitsf(uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; }

int main(int argc, char** argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 2000000;
    long bad = 0;
    for (long i = 0; i < n; ++i)
    {
        float s, D;
        int k;
        switch (i % 8)
        {
            case 0: s = unif(-1.f, 1.f); D = unif(-1e-3f, 1e-3f); k = (int)(nxt() % 600); break;
            case 1: s = unif(-2.f, 2.f); D = unif(-0.1f, 0.1f); k = (int)(nxt() % 200); break;
            case 2: s = unif(-1e-6f, 1e-6f); D = unif(-1e-7f, 1e-7f); k = (int)(nxt() % 400); break;       // zero crossings
            case 3: { // ties: D = (odd) * u/2 of s's binade
                s = unif(0.25f, 4.f);
                int e; std::frexp(s, &e);
                const float u = std::ldexp(1.f, e - 24);
                D = (float)((int)(nxt() % 7) * 2 + 1) * 0.5f * u * ((nxt() & 1) ? 1.f : -1.f);
                k = (int)(nxt() % 300); break; }
            case 4: { // s just below / above a power of two
                const int e = (int)(nxt() % 10) - 5;
                s = std::ldexp(1.f, e) * ((nxt() & 1) ? 1.f : -1.f);
                s = bitsf(((uint32_t*)&s)[0] + (uint32_t)((int)(nxt() % 9) - 4));
                D = unif(-1e-4f, 1e-4f); k = (int)(nxt() % 500); break; }
            case 5: s = bitsf((uint32_t)(nxt() & 0x807FFFFFu)); D = bitsf((uint32_t)(nxt() & 0x807FFFFFu)); k = (int)(nxt() % 50); break;  // subnormals
            case 6: s = unif(-100.f, 100.f); D = unif(-3.f, 3.f); k = (int)(nxt() % 1000); break;
            default: s = unif(-1.f, 1.f); D = unif(-1e-7f, 1e-7f); k = (int)(nxt() % 2000); break;     // |D| below half an ulp
        }
        float ref = s;
        for (int j = 0; j < k; ++j) ref = ref + D;
        float got = sl_recur_advance(s, D, k);
        if ((i & 3) == 0)
        {
            // the three-component form, with this case in the middle slot
            float s0 = s * 0.75f, s1 = s, s2 = -s;
            const float D0 = D * 3.f, D2 = D * 0.5f;
            float r0 = s0, r2 = s2;
            for (int j = 0; j < k; ++j) { r0 = r0 + D0; r2 = r2 + D2; }
            sl_recur_advance3(s0, s1, s2, D0, D, D2, k);
            uint32_t x0, y0, x2, y2;
            std::memcpy(&x0, &r0, 4); std::memcpy(&y0, &s0, 4);
            std::memcpy(&x2, &r2, 4); std::memcpy(&y2, &s2, 4);
            if ((x0 != y0 && !(r0 != r0 && s0 != s0)) || (x2 != y2 && !(r2 != r2 && s2 != s2))) got = s0 * 0.f + NAN;
            else got = s1;
        }
        uint32_t a, b;
        std::memcpy(&a, &ref, 4);
        std::memcpy(&b, &got, 4);
        if (a != b && !(ref != ref && got != got))
        {
            if (bad < 10) std::printf("mismatch s=%a D=%a k=%d ref=%a got=%a\n", s, D, k, ref, got);
            ++bad;
        }
    }
    std::printf("%ld %ld\n", n, bad);
    return bad ? 1 : 0;
}

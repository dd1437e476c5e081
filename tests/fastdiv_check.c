/* Test helper (built by tests/test_host.py with gcc): compares the
 * reciprocal-and-correction quotient the fixed-point P2G kernel uses,
 *     q = a*r;  e = fma(-q, b, a);  q' = fma(e, r, q),   r = RN(1/b),
 * with the IEEE quotient a/b, and the fixed-point words both give
 * (rint(quotient * 2^24)), over random and boundary numerators.  Returns the
 * number of numerators whose fixed-point word differs; *normal_mismatch counts
 * quotients that differ at all while the numerator is in [2^-60, 2^100] (the
 * range where Markstein's theorem applies for divisors in [2^-20, 2^20]).
 * Outside it the words still agree: tiny quotients round to 0 and huge ones
 * are non-finite or beyond the fixed-point limit both ways. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static uint64_t next(uint64_t* s) {
    *s ^= *s << 13;
    *s ^= *s >> 7;
    *s ^= *s << 17;
    return *s;
}

static float bits(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static float example_a;

static int one(float a, float b, float r, uint64_t* normal_mismatch) {
    volatile float av = a, bv = b;
    const float exact = av / bv;
    const float q = a * r;
    const float e = fmaf(-q, b, a);
    const float fast = fmaf(e, r, q);
    /* Markstein's theorem covers quotients whose residual does not underflow
     * and whose product does not overflow: well inside the normal range */
    if (fabsf(a) >= 0x1p-60f && fabsf(a) <= 0x1p100f && fast != exact) {
        ++*normal_mismatch;
        example_a = a;
    }
    /* the kernel maps every non-finite scaled value to the same flag */
    const float we = rintf(exact * 16777216.0f), wf = rintf(fast * 16777216.0f);
    const float limit = 2147483520.0f; /* both outside the limit: the same flag */
    const int out_e = !isfinite(we) || fabsf(we) > limit, out_f = !isfinite(wf) || fabsf(wf) > limit;
    const int same = (out_e && out_f) || (!out_e && !out_f && we == wf);
    if (!same) example_a = a;
    return !same;
}

float fastdiv_example(void) { return example_a; }

uint64_t fastdiv_check(float b, uint64_t seed, uint64_t n, uint64_t* normal_mismatch) {
    volatile float one_v = 1.0f;
    const float r = one_v / b;
    uint64_t bad = 0, s = seed | 1u;
    *normal_mismatch = 0;
    for (uint64_t i = 0; i < n; ++i) {
        /* random sign, exponent over the whole float range, random mantissa */
        uint32_t u = (uint32_t)next(&s);
        bad += one(bits(u & 0x7f7fffffu), b, r, normal_mismatch);
        /* numerators near b * k (quotients near integers and halves) */
        const float k = (float)(next(&s) % 4096u) * 0.5f;
        const float c = k * b;
        uint32_t cu;
        memcpy(&cu, &c, 4);
        for (int d = -2; d <= 2; ++d) bad += one(bits(cu + (uint32_t)d), b, r, normal_mismatch);
    }
    /* every mantissa of one binade of the numerator */
    for (uint32_t m = 0; m < (1u << 23); m += 7u) bad += one(bits(0x3f800000u | m), b, r, normal_mismatch);
    return bad;
}

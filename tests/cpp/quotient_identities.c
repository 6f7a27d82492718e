/* CPU check of the two rounding identities the fluid neighbor terms rely on, evaluated with the
 * same IEEE binary32 operations the device code issues (multiply, fused multiply-add, compare):
 *
 * 1. Division through the correctly rounded reciprocal y = RN(1/b), b > 0:
 *      q0 = RN(a*y); r = fma(-q0, b, a); q = -fma(-r, y, -q0)
 *    equals RN(a/b) bit for bit (signed zeros included) for a == 0 and for |a| >= 2^-96 with
 *    2^-100 <= |a/b| < 2^100.  Checked exhaustively over every dividend significand at six
 *    exponents (2^-96 .. 2^60) and both signs, for the divisors given on the command line
 *    (smoothing radii) and for pseudo-random divisors in [2^-30, 2^31).
 *
 * 2. Square root from an estimate s within one ulp of the root: with sm = s - ulp, sp = s + ulp,
 *    rm = fma(-sm, s, x), rp = fma(-sp, s, x), the root is (rp > 0 ? sp : (rm <= 0 ? sm : s)).
 *    Checked for every estimate in {RN - ulp, RN, RN + ulp} of RN(sqrt(x)) over every
 *    significand of x at exponents 2^-96 .. 2^100, and at x == 0.
 *
 * 3. Reciprocal refinement: for a normal b whose significand is not all ones and any estimate y0
 *    within two ulps of RN(1/b), two Newton steps in fma, y <- fma(fma(-b, y, 1), y, y), give
 *    RN(1/b).  Checked for every significand of b at exponents -20, 0 and 20 and every such y0.
 *
 * Prints "ok <checks>" and exits 0, or prints the first mismatches and exits 1. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float bf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t fb(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float div_rcp(float a, float b, float y)
{
    const float q0 = a * y;
    const float r = fmaf(-q0, b, a);
    return -fmaf(-r, y, -q0);
}

static float sqrt_sel(float x, float s)
{
    const float sm = bf(fb(s) - 1u), sp = bf(fb(s) + 1u);
    const float rm = fmaf(-sm, s, x), rp = fmaf(-sp, s, x);
    const float t = rm <= 0.0f ? sm : s;
    return rp > 0.0f ? sp : t;
}

static float rcp_refine(float b, float y)
{
    y = fmaf(fmaf(-b, y, 1.0f), y, y);
    return fmaf(fmaf(-b, y, 1.0f), y, y);
}

static long check_divisor(float b)
{
    const float y = 1.0f / b;
    static const int ea[6] = {-96, -60, -20, 0, 20, 60};
    long bad = 0;
#pragma omp parallel for reduction(+ : bad)
    for (uint32_t m = 0; m < (1u << 23); ++m)
        for (int t = 0; t < 6; ++t)
            for (uint32_t sg = 0; sg < 2; ++sg) {
                const float a = bf((sg << 31) | ((uint32_t)(ea[t] + 127) << 23) | m);
                const float q = a / b;
                if (!(fabsf(q) >= 0x1p-100f && fabsf(q) < 0x1p100f)) continue;
                if (fb(div_rcp(a, b, y)) != fb(q)) ++bad;
            }
    const float z[2] = {0.0f, -0.0f};
    for (int k = 0; k < 2; ++k)
        if (fb(div_rcp(z[k], b, y)) != fb(z[k] / b)) ++bad;
    return bad;
}

int main(int argc, char **argv)
{
    long bad = 0, checks = 0;
    int nrand = argc > 1 ? atoi(argv[1]) : 8;
    for (int i = 2; i < argc; ++i) {
        const float b = strtof(argv[i], NULL);
        const long e = check_divisor(b);
        if (e) printf("divisor %a: %ld mismatches\n", b, e);
        bad += e;
        checks += 12l << 23;
    }
    uint64_t st = 0x9e3779b97f4a7c15ull;
    for (int i = 0; i < nrand; ++i) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        const uint32_t m = (uint32_t)(st & 0x7fffffu);
        const int e = (int)((st >> 32) % 61u) - 30;
        const float b = bf(((uint32_t)(e + 127) << 23) | m);
        const long k = check_divisor(b);
        if (k) printf("divisor %a: %ld mismatches\n", b, k);
        bad += k;
        checks += 12l << 23;
    }
    long sbad = 0;
#pragma omp parallel for reduction(+ : sbad)
    for (int e = -96; e <= 100; ++e)
        for (uint32_t m = 0; m < (1u << 23); ++m) {
            const float x = bf(((uint32_t)(e + 127) << 23) | m);
            const float r = sqrtf(x);
            for (int d = -1; d <= 1; ++d)
                if (fb(sqrt_sel(x, bf(fb(r) + (uint32_t)d))) != fb(r)) ++sbad;
        }
    checks += 197l * 3 * (1l << 23);
    if (fb(sqrt_sel(0.0f, 0.0f)) != fb(0.0f) || fb(sqrt_sel(0.0f, bf(1u))) != fb(0.0f)) ++sbad;
    if (sbad) printf("sqrt selection: %ld mismatches\n", sbad);
    bad += sbad;
    long rbad = 0;
    static const int re[3] = {-20, 0, 20};
#pragma omp parallel for reduction(+ : rbad)
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        if (m == 0x7fffffu) continue;
        for (int t = 0; t < 3; ++t) {
            const float b = bf(((uint32_t)(re[t] + 127) << 23) | m);
            const float rn = 1.0f / b;
            for (int d = -2; d <= 2; ++d)
                if (fb(rcp_refine(b, bf(fb(rn) + (uint32_t)d))) != fb(rn)) ++rbad;
        }
    }
    checks += 15l * ((1l << 23) - 1);
    if (rbad) printf("reciprocal refinement: %ld mismatches\n", rbad);
    bad += rbad;
    if (bad) return 1;
    printf("ok %ld\n", checks);
    return 0;
}

/*
 * bg_detmath.h -- deterministic arithmetic shared by the HIP kernels and the CPU oracle.
 *
 * BrainGrid's CPU path draws its Gaussian noise from one sequential Marsaglia-polar stream
 * (RNG/Norm.cpp:59-86) and its CUDA path from a per-thread Mersenne Twister with Box-Muller
 * (RNG/MersenneTwister_d.cu:164-237); neither can be reproduced identically on two different
 * machines, and glibc's expf/acosf/sinf differ from any GPU math library in the last bit.
 * This header defines the *deterministic mode* of the engine:
 *
 *   - noise(seed, neuron, step): one N(0,1) float per neuron per step from Philox4x32-10
 *     (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the
 *     Box-Muller transform;
 *   - bg_exp / bg_log / bg_sin / bg_cos / bg_acos / bg_pow: double-precision evaluations built only
 *     from IEEE-754 +, -, *, / and sqrt (Cody-Waite argument reduction, Taylor / atanh / arcsine
 *     series with exactly rational coefficients), rounded to float by the *f wrappers.
 *
 * Every operation here is a correctly rounded IEEE operation, so a host compiler and the gfx950
 * device compiler produce identical bits provided floating-point contraction is disabled
 * (-ffp-contract=off, set by every build file of this project) and denormals are not flushed.
 * This file is plain C99 and compiles as HIP device code unchanged.
 */
#ifndef BG_DETMATH_H
#define BG_DETMATH_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define BG_HD __host__ __device__
#pragma clang fp contract(off)
#else
#define BG_HD
#endif

#ifdef __cplusplus
extern "C++" {
#endif

/* ---------------------------------------------------------------- Philox4x32-10 ------------ */
#define BG_PHILOX_M0 0xD2511F53u
#define BG_PHILOX_M1 0xCD9E8D57u
#define BG_PHILOX_W0 0x9E3779B9u
#define BG_PHILOX_W1 0xBB67AE85u

static inline BG_HD void bg_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)BG_PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)BG_PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += BG_PHILOX_W0;
        k1 += BG_PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ---------------------------------------------------------------- bit helpers -------------- */
static inline BG_HD double bg_from_bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static inline BG_HD uint64_t bg_to_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }
static inline BG_HD int bg_isnan(double x) { return x != x; }

/* x * 2^k for any int k in [-2100, 2100] (two exact power-of-two scalings). */
static inline BG_HD double bg_scale2(double x, int k)
{
    while (k > 1000) { x *= bg_from_bits((uint64_t)(1000 + 1023) << 52); k -= 1000; }
    while (k < -1000) { x *= bg_from_bits((uint64_t)(-1000 + 1023) << 52); k += 1000; }
    return x * bg_from_bits((uint64_t)(k + 1023) << 52);
}

/* ---------------------------------------------------------------- exp ---------------------- */
/* exp(x): x = k ln2 + r, |r| <= ln2/2 (Cody-Waite split of ln2), Taylor series to r^14. */
static inline BG_HD double bg_exp(double x)
{
    const double ln2_hi = 0x1.62e42fee00000p-1, ln2_lo = 0x1.a39ef35793c76p-33;
    const double inv_ln2 = 1.4426950408889634;
    if (bg_isnan(x)) return x;
    if (x > 709.78) return bg_from_bits(0x7ff0000000000000ull);
    if (x < -745.2) return 0.0;
    double t = x * inv_ln2;
    int k = (int)(t + (t >= 0.0 ? 0.5 : -0.5));
    double r = (x - (double)k * ln2_hi) - (double)k * ln2_lo;
    double p = 1.1470745597729725e-11;
    p = p * r + 1.6059043836821613e-10;
    p = p * r + 2.08767569878681e-09;
    p = p * r + 2.505210838544172e-08;
    p = p * r + 2.755731922398589e-07;
    p = p * r + 2.7557319223985893e-06;
    p = p * r + 2.48015873015873e-05;
    p = p * r + 0.0001984126984126984;
    p = p * r + 0.001388888888888889;
    p = p * r + 0.008333333333333333;
    p = p * r + 0.041666666666666664;
    p = p * r + 0.16666666666666666;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return bg_scale2(p, k);
}

/* ---------------------------------------------------------------- log ---------------------- */
/* log(x), x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1). */
static inline BG_HD double bg_log(double x)
{
    const double ln2_hi = 0x1.62e42fee00000p-1, ln2_lo = 0x1.a39ef35793c76p-33;
    if (bg_isnan(x) || x < 0.0) return bg_from_bits(0x7ff8000000000000ull);
    if (x == 0.0) return -bg_from_bits(0x7ff0000000000000ull);
    if (x == bg_from_bits(0x7ff0000000000000ull)) return x;
    int e = 0;
    if (x < 0x1p-1022) { x *= 0x1p54; e -= 54; }
    uint64_t b = bg_to_bits(x);
    e += (int)((b >> 52) & 0x7ff) - 1022;
    double m = bg_from_bits((b & 0x000fffffffffffffull) | 0x3fe0000000000000ull); /* [0.5, 1) */
    if (m < 0.70710678118654752) { m = m * 2.0; e -= 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double s2 = s * s;
    double p = 0.08;
    p = p * s2 + 0.08695652173913043;
    p = p * s2 + 0.09523809523809523;
    p = p * s2 + 0.10526315789473684;
    p = p * s2 + 0.11764705882352941;
    p = p * s2 + 0.13333333333333333;
    p = p * s2 + 0.15384615384615385;
    p = p * s2 + 0.18181818181818182;
    p = p * s2 + 0.2222222222222222;
    p = p * s2 + 0.2857142857142857;
    p = p * s2 + 0.4;
    p = p * s2 + 0.6666666666666666;
    double lm = 2.0 * s + s * s2 * p;
    return (double)e * ln2_hi + ((double)e * ln2_lo + lm);
}

/* ---------------------------------------------------------------- sin / cos ---------------- */
static inline BG_HD double bg_sin_kernel(double r)
{
    double r2 = r * r;
    double p = -7.647163731819816e-13;
    p = p * r2 + 1.6059043836821613e-10;
    p = p * r2 + -2.505210838544172e-08;
    p = p * r2 + 2.7557319223985893e-06;
    p = p * r2 + -0.0001984126984126984;
    p = p * r2 + 0.008333333333333333;
    p = p * r2 + -0.16666666666666666;
    return r + r * r2 * p;
}
static inline BG_HD double bg_cos_kernel(double r)
{
    double r2 = r * r;
    double p = 4.779477332387385e-14;
    p = p * r2 + -1.1470745597729725e-11;
    p = p * r2 + 2.08767569878681e-09;
    p = p * r2 + -2.755731922398589e-07;
    p = p * r2 + 2.48015873015873e-05;
    p = p * r2 + -0.001388888888888889;
    p = p * r2 + 0.041666666666666664;
    p = p * r2 + -0.5;
    return 1.0 + r2 * p;
}
/* Quadrant reduction x = q pi/2 + r, |r| <= pi/4, three-part Cody-Waite split of pi/2.
 * Accurate for |x| < 2^20 (all arguments this engine produces are below 2 pi). */
static inline BG_HD double bg_reduce_pio2(double x, int *q)
{
    const double p1 = 0x1.921fb54400000p+0, p2 = 0x1.0b4611a600000p-34, p3 = 0x1.3198a2e037073p-69;
    double t = x * 0.6366197723675814;
    int k = (int)(t + (t >= 0.0 ? 0.5 : -0.5));
    *q = k;
    return ((x - (double)k * p1) - (double)k * p2) - (double)k * p3;
}
static inline BG_HD double bg_sin(double x)
{
    if (bg_isnan(x)) return x;
    int q;
    double r = bg_reduce_pio2(x, &q);
    switch (q & 3) {
    case 0: return bg_sin_kernel(r);
    case 1: return bg_cos_kernel(r);
    case 2: return -bg_sin_kernel(r);
    default: return -bg_cos_kernel(r);
    }
}
static inline BG_HD double bg_cos(double x)
{
    if (bg_isnan(x)) return x;
    int q;
    double r = bg_reduce_pio2(x, &q);
    switch (q & 3) {
    case 0: return bg_cos_kernel(r);
    case 1: return -bg_sin_kernel(r);
    case 2: return -bg_cos_kernel(r);
    default: return bg_sin_kernel(r);
    }
}
/* cos(2 pi u) for u in [0, 1): exact quarter-turn reduction, then the kernels on [0, pi/4]. */
static inline BG_HD double bg_cos2pi(double u)
{
    double t = u * 8.0;                 /* octant, exact */
    int o = (int)t;
    double f = t - (double)o;           /* exact */
    const double pio4 = 0.78539816339744831;
    /* angle = (o + f) pi/4 */
    switch (o & 7) {
    case 0: return bg_cos_kernel(f * pio4);
    case 1: return bg_sin_kernel((1.0 - f) * pio4);
    case 2: return -bg_sin_kernel(f * pio4);
    case 3: return -bg_cos_kernel((1.0 - f) * pio4);
    case 4: return -bg_cos_kernel(f * pio4);
    case 5: return -bg_sin_kernel((1.0 - f) * pio4);
    case 6: return bg_sin_kernel(f * pio4);
    default: return bg_cos_kernel((1.0 - f) * pio4);
    }
}

/* ---------------------------------------------------------------- acos --------------------- */
/* arcsine series sum c_n y^(2n+1), c_n = (2n)! / (4^n (n!)^2 (2n+1)), used for |y| <= 1/2. */
static inline BG_HD double bg_asin_series(double y)
{
    double y2 = y * y;
    double p = 0.0017680811205154183;
    p = p * y2 + 0.0018622264064031275;
    p = p * y2 + 0.0019650336162772837;
    p = p * y2 + 0.0020776610325181676;
    p = p * y2 + 0.0022014739737101384;
    p = p * y2 + 0.002338091892111975;
    p = p * y2 + 0.0024894486782468836;
    p = p * y2 + 0.00265787063820729;
    p = p * y2 + 0.002846178401108942;
    p = p * y2 + 0.0030578216492580306;
    p = p * y2 + 0.003297059503473485;
    p = p * y2 + 0.0035692053938259347;
    p = p * y2 + 0.003880964558837669;
    p = p * y2 + 0.004240907093679363;
    p = p * y2 + 0.004660143486915096;
    p = p * y2 + 0.005153309682319905;
    p = p * y2 + 0.005740037670841924;
    p = p * y2 + 0.006447210311889649;
    p = p * y2 + 0.0073125258735988454;
    p = p * y2 + 0.008390335809616815;
    p = p * y2 + 0.009761609529194078;
    p = p * y2 + 0.011551800896139705;
    p = p * y2 + 0.01396484375;
    p = p * y2 + 0.017352764423076924;
    p = p * y2 + 0.022372159090909092;
    p = p * y2 + 0.030381944444444444;
    p = p * y2 + 0.044642857142857144;
    p = p * y2 + 0.075;
    p = p * y2 + 0.16666666666666666;
    return y + y * y2 * p;
}
/* acos(x) on [-1, 1]; NaN outside, like the C library. Half-angle identities keep the series
 * argument at most 1/2. */
static inline BG_HD double bg_acos(double x)
{
    const double pio2_hi = 1.5707963267948966, pio2_lo = 6.123233995736766e-17;
    if (bg_isnan(x) || x > 1.0 || x < -1.0) return bg_from_bits(0x7ff8000000000000ull);
    if (x > 0.5) {
        double y = __builtin_sqrt((1.0 - x) * 0.5);
        return 2.0 * bg_asin_series(y);
    }
    if (x < -0.5) {
        double y = __builtin_sqrt((1.0 + x) * 0.5);
        return (2.0 * pio2_hi - 2.0 * bg_asin_series(y)) + 2.0 * pio2_lo;
    }
    return (pio2_hi - bg_asin_series(x)) + pio2_lo;
}

/* pow(x, y) for the STDP weight exponents: exp(y log x) with the C library's special cases for
 * y == 0, x == 0 and negative x with integral y. */
static inline BG_HD double bg_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (bg_isnan(x) || bg_isnan(y)) return bg_from_bits(0x7ff8000000000000ull);
    if (x == 0.0) return y > 0.0 ? 0.0 : bg_from_bits(0x7ff0000000000000ull);
    if (x < 0.0) {
        double yi = (double)(int64_t)y;
        if (yi != y) return bg_from_bits(0x7ff8000000000000ull);
        double m = bg_exp(y * bg_log(-x));
        return (((int64_t)y) & 1) ? -m : m;
    }
    return bg_exp(y * bg_log(x));
}

/* exp for float results: the reduction of bg_exp with the series truncated after r^10
 * (remainder below 2.5e-13 relative on |r| <= ln2/2, far below float resolution). */
static inline BG_HD double bg_exp_for_float(double x)
{
    const double ln2_hi = 0x1.62e42fee00000p-1, ln2_lo = 0x1.a39ef35793c76p-33;
    const double inv_ln2 = 1.4426950408889634;
    if (bg_isnan(x)) return x;
    if (x > 709.78) return bg_from_bits(0x7ff0000000000000ull);
    if (x < -745.2) return 0.0;
    double t = x * inv_ln2;
    int k = (int)(t + (t >= 0.0 ? 0.5 : -0.5));
    double r = (x - (double)k * ln2_hi) - (double)k * ln2_lo;
    double p = 2.755731922398589e-07;
    p = p * r + 2.7557319223985893e-06;
    p = p * r + 2.48015873015873e-05;
    p = p * r + 0.0001984126984126984;
    p = p * r + 0.001388888888888889;
    p = p * r + 0.008333333333333333;
    p = p * r + 0.041666666666666664;
    p = p * r + 0.16666666666666666;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return bg_scale2(p, k);
}

/* exp in single precision for the per-spike synapse updates: x = k ln2 + r with a 16-bit ln2_hi
 * (k * ln2_hi exact for |k| < 256), |r| <= ln2/2, Taylor series to r^7 (truncation below 5e-9
 * relative) in Horner form with fused multiply-adds, scaled by 2^k through the exponent bits.
 * About one ulp from the correctly rounded value; every operation is a single correctly rounded
 * IEEE float operation (fmaf included: one rounding on the host's C library and on the device's
 * v_fma_f32 alike), so host and device agree bit for bit. */
static inline BG_HD float bg_scale2f(float x, int k)
{
    uint32_t b;
    float s;
    if (k < -126) {
        b = (uint32_t)(-126 + 127) << 23;
        memcpy(&s, &b, 4);
        x *= s;
        k += 126;
    } else if (k > 127) {
        b = (uint32_t)(127 + 127) << 23;
        memcpy(&s, &b, 4);
        x *= s;
        k -= 127;
    }
    b = (uint32_t)(k + 127) << 23;
    memcpy(&s, &b, 4);
    return x * s;
}
static inline BG_HD float bg_expf(float x)
{
    if (x != x) return x;
    if (x > 88.72283935546875f) {
        uint32_t b = 0x7f800000u;
        float inf;
        memcpy(&inf, &b, 4);
        return inf;
    }
    if (x < -103.972076416015625f) return 0.0f;
    const float ln2_hi = 0.693145751953125f, ln2_lo = 1.428606765330187e-06f;
    float t = x * 1.4426950216293335f;
    int k = (int)(t + (t >= 0.0f ? 0.5f : -0.5f));
    float kf = (float)k;
    float r = __builtin_fmaf(-kf, ln2_lo, x - kf * ln2_hi);
    float p = 0.00019841270113829523f;
    p = __builtin_fmaf(p, r, 0.0013888889225199819f);
    p = __builtin_fmaf(p, r, 0.008333333767950535f);
    p = __builtin_fmaf(p, r, 0.0416666679084301f);
    p = __builtin_fmaf(p, r, 0.1666666716337204f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return bg_scale2f(p, k);
}

/* float entry points of the double-precision functions: evaluate in double, round once. */
static inline BG_HD float bg_acosf(float x) { return (float)bg_acos((double)x); }
static inline BG_HD float bg_sinf(float x) { return (float)bg_sin((double)x); }
static inline BG_HD float bg_powf(float x, float y) { return (float)bg_pow((double)x, (double)y); }

/* ---------------------------------------------------------------- noise -------------------- */
/* Philox counter layout used by the engine: (index, step low word, step high word, stream). */
#define BG_STREAM_NOISE 0u
#define BG_STREAM_POISSON 1u

static inline BG_HD void bg_philox_draw(uint64_t seed, uint32_t index, uint64_t step, uint32_t stream,
                                        uint32_t out[4])
{
    uint32_t ctr[4] = {index, (uint32_t)step, (uint32_t)(step >> 32), stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    bg_philox4x32_10(ctr, key, out);
}

/* N(0,1) noise for (neuron, step): Box-Muller on u1 in (0,1], u2 in [0,1). */
static inline BG_HD float bg_noise_normal(uint64_t seed, uint32_t neuron, uint64_t step)
{
    uint32_t w[4];
    bg_philox_draw(seed, neuron, step, BG_STREAM_NOISE, w);
    double u1 = ((double)w[0] + 1.0) * 0x1p-32;
    double u2 = (double)w[1] * 0x1p-32;
    double r = __builtin_sqrt(-2.0 * bg_log(u1));
    return (float)(r * bg_cos2pi(u2));
}

/* Poisson stimulus input: the next inter-spike interval in steps, HostSInputPoisson::inputStimulus
 * (Inputs/HostSInputPoisson.cpp:75-84) with its MTRand draws replaced by a counter-based stream.
 * Draw j of (neuron, step) is word j mod 4 of Philox block (neuron, step, POISSON | (j / 4) << 1),
 * mapped to a float in (0, 1] (24 bits, never 0, so the log is finite). The reference's
 * rejection of intervals inside the refractoriness, `while (u <= exp(-isi^2 / 32))`, is bounded
 * at BG_POISSON_MAX_REJECT redraws; the loop is left after about 1.06 draws on average at 10 Hz,
 * so the bound only makes the worst case finite. lambda_ms is the mean interval in ms. */
#define BG_POISSON_MAX_REJECT 64

typedef struct bg_poisson_stream {
    uint64_t seed, step;
    uint32_t neuron, block, used;
    uint32_t w[4];
} bg_poisson_stream;

static inline BG_HD float bg_poisson_next(bg_poisson_stream *s)
{
    if (s->used == 4) {
        bg_philox_draw(s->seed, s->neuron, s->step, BG_STREAM_POISSON | (s->block << 1), s->w);
        s->block++;
        s->used = 0;
    }
    return (float)((s->w[s->used++] >> 8) + 1u) * 0x1p-24f;
}

static inline BG_HD int bg_poisson_isi_steps(uint64_t seed, uint32_t neuron, uint64_t step, float lambda_ms,
                                             float dt)
{
    bg_poisson_stream s;
    s.seed = seed; s.step = step; s.neuron = neuron; s.block = 0; s.used = 4;
    float isi = -lambda_ms * (float)bg_log((double)bg_poisson_next(&s));
    for (int a = 0; a < BG_POISSON_MAX_REJECT; a++) {
        if (!(bg_poisson_next(&s) <= bg_expf(-(isi * isi) / 32.0f))) break;
        isi = -lambda_ms * (float)bg_log((double)bg_poisson_next(&s));
    }
    return (int)((double)((isi / 1000.0f) / dt) + 0.5);
}

#ifdef __cplusplus
}
#endif

#endif /* BG_DETMATH_H */

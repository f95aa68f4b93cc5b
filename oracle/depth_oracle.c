/*
 * depth_oracle.c - CPU oracle (test infrastructure only; see depth_oracle.h).
 *
 * Plain-C restatement of crates/chiaro-fusion/src/depth.rs with its helpers
 * geometry.rs, math.rs, image.rs (Plane::sample) and align.rs (Warp). Each
 * function names the reference lines it follows. Build with
 * -O2 -ffp-contract=off -fno-fast-math (Rust never contracts a*b+c).
 */
#define _GNU_SOURCE
#include "depth_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* depth.rs:23-37 */
#define INFINITY_DEPTH 1.0e8
#define MISSING_COST 2.25f
#define SGM_SMALL_PENALTY 0.06f
#define SGM_LARGE_PENALTY 0.45f
#define SGM_DIRECTIONS 8.0f
#define MINIMUM_REGULARIZED_SCORE 0.30f
#define MAXIMUM_REGULARIZED_BASELINE_LOSS 0.04f
#define NEAR_DEPTH_PRIOR 0.10f
#define WARP_BOUNDARY_DELTA_PX 2.0f
#define WARP_BOUNDARY_CONTRAST 0.30f
#define MINIMUM_DIRECT_COMPONENT_NODES 24u

/* ------------------------------------------------------------ Rust helpers */

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint64_t d2u(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
static inline double u2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }

/* f32::total_cmp / f64::total_cmp keys (signed compare of the adjusted bits) */
static inline int32_t tkey32(float f) {
    int32_t k = (int32_t)f2u(f);
    k ^= (int32_t)(((uint32_t)(k >> 31)) >> 1);
    return k;
}
static inline int64_t tkey64(double d) {
    int64_t k = (int64_t)d2u(d);
    k ^= (int64_t)(((uint64_t)(k >> 63)) >> 1);
    return k;
}
static inline int tcmp32(float a, float b) {
    int32_t x = tkey32(a), y = tkey32(b);
    return (x > y) - (x < y);
}
static inline int tcmp64(double a, double b) {
    int64_t x = tkey64(a), y = tkey64(b);
    return (x > y) - (x < y);
}
/* f32::max / f32::min (IEEE maxNum / minNum) */
static inline float rmaxf(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}
static inline float rminf(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a < b ? a : b;
}
static inline double rmax(double a, double b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}
/* f32::clamp / f64::clamp (NaN passes through) */
static inline float rclampf(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
static inline double rclamp(double v, double lo, double hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
/* `as usize` for floats saturates: NaN and negatives give 0. */
static inline size_t f32_usize(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551616.0f) return SIZE_MAX;
    return (size_t)v;
}
static inline size_t f64_usize(double v) {
    if (!(v > 0.0)) return 0;
    if (v >= 18446744073709551616.0) return SIZE_MAX;
    return (size_t)v;
}
static inline uint16_t f64_u16(double v) {
    if (!(v > 0.0)) return 0;
    if (v >= 65535.0) return 65535;
    return (uint16_t)v;
}
static inline size_t div_ceil(size_t a, size_t b) { return (a + b - 1) / b; }
static inline double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---------------------------------------------------- transcendental funcs */

/*
 * expf as glibc (>= 2.28, sysdeps/ieee754/flt-32/e_expf.c) computes it: the
 * ARM optimized-routines method. exp(x) = 2^(k/32) * 2^(r/32) with
 * k = round(x * 32/ln2), r the remainder, 2^(i/32) from a 32-entry table
 * whose entries are the correctly rounded doubles of 2^(i/32) less i<<47,
 * and a cubic for 2^(r/32). The x86_64 FMA build fuses the multiply-adds;
 * the fused form below reproduces it bit for bit.
 */
static const uint64_t EXP2F_TAB[32] = {
    0x3ff0000000000000ULL, 0x3fefd9b0d3158574ULL, 0x3fefb5586cf9890fULL, 0x3fef9301d0125b51ULL,
    0x3fef72b83c7d517bULL, 0x3fef54873168b9aaULL, 0x3fef387a6e756238ULL, 0x3fef1e9df51fdee1ULL,
    0x3fef06fe0a31b715ULL, 0x3feef1a7373aa9cbULL, 0x3feedea64c123422ULL, 0x3feece086061892dULL,
    0x3feebfdad5362a27ULL, 0x3feeb42b569d4f82ULL, 0x3feeab07dd485429ULL, 0x3feea47eb03a5585ULL,
    0x3feea09e667f3bcdULL, 0x3fee9f75e8ec5f74ULL, 0x3feea11473eb0187ULL, 0x3feea589994cce13ULL,
    0x3feeace5422aa0dbULL, 0x3feeb737b0cdc5e5ULL, 0x3feec49182a3f090ULL, 0x3feed503b23e255dULL,
    0x3feee89f995ad3adULL, 0x3feeff76f2fb5e47ULL, 0x3fef199bdd85529cULL, 0x3fef3720dcef9069ULL,
    0x3fef5818dcfba487ULL, 0x3fef7c97337b9b5fULL, 0x3fefa4afa2a490daULL, 0x3fefd0765b6e4540ULL};

float or_expf(float x) {
    const double inv_ln2_n = 0x1.71547652b82fep+0 * 32.0;
    const double shift = 0x1.8p+52;
    const double c0 = 0x1.c6af84b912394p-5 / 32.0 / 32.0 / 32.0;
    const double c1 = 0x1.ebfce50fac4f3p-3 / 32.0 / 32.0;
    const double c2 = 0x1.62e42ff0c52d6p-1 / 32.0;
    uint32_t abstop = (f2u(x) >> 20) & 0x7ff;
    if (abstop >= (f2u(88.0f) >> 20)) {
        if (f2u(x) == f2u(-INFINITY)) return 0.0f;
        if (abstop >= (f2u(INFINITY) >> 20)) return x + x;
        if (x > 0x1.62e42ep6f) return INFINITY;
        if (x < -0x1.9fe368p6f) return 0.0f;
    }
    double xd = (double)x;
    double kd = fma(inv_ln2_n, xd, shift);
    uint64_t ki = d2u(kd);
    kd -= shift;
    double r = fma(inv_ln2_n, xd, -kd);
    uint64_t t = EXP2F_TAB[ki % 32];
    t += ki << (52 - 5);
    double s = u2d(t);
    double z = fma(c0, r, c1);
    double r2 = r * r;
    double y = fma(c2, r, 1.0);
    y = fma(z, r2, y);
    y = y * s;
    return (float)y;
}

/*
 * f64 exp as glibc (sysdeps/ieee754/dbl-64/e_exp.c) computes it for the
 * arguments the path uses (|x| < 512): ARM optimized-routines exp with a
 * 128-entry table 2^(k/128) = H[k] (1 + T[k]) (H the correctly rounded
 * double, T the rounded relative remainder; both derived from the
 * definition) and a degree-5 polynomial, FMA-fused as in the x86_64 FMA
 * build.
 */
/* EXP_TAB[2k] = bits of T[k], EXP_TAB[2k+1] = bits of H[k] - (k << 45), with
 * H[k] the correctly rounded double of 2^(k/128) and T[k] the double nearest
 * (2^(k/128) - H[k]) / H[k]; generated with 80-digit decimal arithmetic
 * (tests/test_transcendentals.py recomputes and compares it). */
static const uint64_t EXP_TAB[256] = {
    0x0000000000000000ULL, 0x3ff0000000000000ULL,
    0x3c9b3b4f1a88bf6eULL, 0x3feff63da9fb3335ULL,
    0xbc7160139cd8dc5dULL, 0x3fefec9a3e778061ULL,
    0xbc905e7a108766d1ULL, 0x3fefe315e86e7f85ULL,
    0x3c8cd2523567f613ULL, 0x3fefd9b0d3158574ULL,
    0xbc8bce8023f98efaULL, 0x3fefd06b29ddf6deULL,
    0x3c60f74e61e6c861ULL, 0x3fefc74518759bc8ULL,
    0x3c90a3e45b33d399ULL, 0x3fefbe3ecac6f383ULL,
    0x3c979aa65d837b6dULL, 0x3fefb5586cf9890fULL,
    0x3c8eb51a92fdeffcULL, 0x3fefac922b7247f7ULL,
    0x3c3ebe3d702f9cd1ULL, 0x3fefa3ec32d3d1a2ULL,
    0xbc6a033489906e0bULL, 0x3fef9b66affed31bULL,
    0xbc9556522a2fbd0eULL, 0x3fef9301d0125b51ULL,
    0xbc5080ef8c4eea55ULL, 0x3fef8abdc06c31ccULL,
    0xbc91c923b9d5f416ULL, 0x3fef829aaea92de0ULL,
    0x3c80d3e3e95c55afULL, 0x3fef7a98c8a58e51ULL,
    0xbc801b15eaa59348ULL, 0x3fef72b83c7d517bULL,
    0xbc8f1ff055de323dULL, 0x3fef6af9388c8deaULL,
    0x3c8b898c3f1353bfULL, 0x3fef635beb6fcb75ULL,
    0xbc96d99c7611eb26ULL, 0x3fef5be084045cd4ULL,
    0x3c9aecf73e3a2f60ULL, 0x3fef54873168b9aaULL,
    0xbc8fe782cb86389dULL, 0x3fef4d5022fcd91dULL,
    0x3c8a6f4144a6c38dULL, 0x3fef463b88628cd6ULL,
    0x3c807a05b0e4047dULL, 0x3fef3f49917ddc96ULL,
    0x3c968efde3a8a894ULL, 0x3fef387a6e756238ULL,
    0x3c875e18f274487dULL, 0x3fef31ce4fb2a63fULL,
    0x3c80472b981fe7f2ULL, 0x3fef2b4565e27cddULL,
    0xbc96b87b3f71085eULL, 0x3fef24dfe1f56381ULL,
    0x3c82f7e16d09ab31ULL, 0x3fef1e9df51fdee1ULL,
    0xbc3d219b1a6fbffaULL, 0x3fef187fd0dad990ULL,
    0x3c8b3782720c0ab4ULL, 0x3fef1285a6e4030bULL,
    0x3c6e149289cecb8fULL, 0x3fef0cafa93e2f56ULL,
    0x3c834d754db0abb6ULL, 0x3fef06fe0a31b715ULL,
    0x3c864201e2ac744cULL, 0x3fef0170fc4cd831ULL,
    0x3c8fdd395dd3f84aULL, 0x3feefc08b26416ffULL,
    0xbc86a3803b8e5b04ULL, 0x3feef6c55f929ff1ULL,
    0xbc924aedcc4b5068ULL, 0x3feef1a7373aa9cbULL,
    0xbc9907f81b512d8eULL, 0x3feeecae6d05d866ULL,
    0xbc71d1e83e9436d2ULL, 0x3feee7db34e59ff7ULL,
    0xbc991919b3ce1b15ULL, 0x3feee32dc313a8e5ULL,
    0x3c859f48a72a4c6dULL, 0x3feedea64c123422ULL,
    0xbc9312607a28698aULL, 0x3feeda4504ac801cULL,
    0xbc58a78f4817895bULL, 0x3feed60a21f72e2aULL,
    0xbc7c2c9b67499a1bULL, 0x3feed1f5d950a897ULL,
    0x3c4363ed60c2ac11ULL, 0x3feece086061892dULL,
    0x3c9666093b0664efULL, 0x3feeca41ed1d0057ULL,
    0x3c6ecce1daa10379ULL, 0x3feec6a2b5c13cd0ULL,
    0x3c93ff8e3f0f1230ULL, 0x3feec32af0d7d3deULL,
    0x3c7690cebb7aafb0ULL, 0x3feebfdad5362a27ULL,
    0x3c931dbdeb54e077ULL, 0x3feebcb299fddd0dULL,
    0xbc8f94340071a38eULL, 0x3feeb9b2769d2ca7ULL,
    0xbc87deccdc93a349ULL, 0x3feeb6daa2cf6642ULL,
    0xbc78dec6bd0f385fULL, 0x3feeb42b569d4f82ULL,
    0xbc861246ec7b5cf6ULL, 0x3feeb1a4ca5d920fULL,
    0x3c93350518fdd78eULL, 0x3feeaf4736b527daULL,
    0x3c7b98b72f8a9b05ULL, 0x3feead12d497c7fdULL,
    0x3c9063e1e21c5409ULL, 0x3feeab07dd485429ULL,
    0x3c34c7855019c6eaULL, 0x3feea9268a5946b7ULL,
    0x3c9432e62b64c035ULL, 0x3feea76f15ad2148ULL,
    0xbc8ce44a6199769fULL, 0x3feea5e1b976dc09ULL,
    0xbc8c33c53bef4da8ULL, 0x3feea47eb03a5585ULL,
    0xbc845378892be9aeULL, 0x3feea34634ccc320ULL,
    0xbc93cedd78565858ULL, 0x3feea23882552225ULL,
    0x3c5710aa807e1964ULL, 0x3feea155d44ca973ULL,
    0xbc93b3efbf5e2228ULL, 0x3feea09e667f3bcdULL,
    0xbc6a12ad8734b982ULL, 0x3feea012750bdabfULL,
    0xbc6367efb86da9eeULL, 0x3fee9fb23c651a2fULL,
    0xbc80dc3d54e08851ULL, 0x3fee9f7df9519484ULL,
    0xbc781f647e5a3ecfULL, 0x3fee9f75e8ec5f74ULL,
    0xbc86ee4ac08b7db0ULL, 0x3fee9f9a48a58174ULL,
    0xbc8619321e55e68aULL, 0x3fee9feb564267c9ULL,
    0x3c909ccb5e09d4d3ULL, 0x3feea0694fde5d3fULL,
    0xbc7b32dcb94da51dULL, 0x3feea11473eb0187ULL,
    0x3c94ecfd5467c06bULL, 0x3feea1ed0130c132ULL,
    0x3c65ebe1abd66c55ULL, 0x3feea2f336cf4e62ULL,
    0xbc88a1c52fb3cf42ULL, 0x3feea427543e1a12ULL,
    0xbc9369b6f13b3734ULL, 0x3feea589994cce13ULL,
    0xbc805e843a19ff1eULL, 0x3feea71a4623c7adULL,
    0xbc94d450d872576eULL, 0x3feea8d99b4492edULL,
    0x3c90ad675b0e8a00ULL, 0x3feeaac7d98a6699ULL,
    0x3c8db72fc1f0eab4ULL, 0x3feeace5422aa0dbULL,
    0xbc65b6609cc5e7ffULL, 0x3feeaf3216b5448cULL,
    0x3c7bf68359f35f44ULL, 0x3feeb1ae99157736ULL,
    0xbc93091fa71e3d83ULL, 0x3feeb45b0b91ffc6ULL,
    0xbc5da9b88b6c1e29ULL, 0x3feeb737b0cdc5e5ULL,
    0xbc6c23f97c90b959ULL, 0x3feeba44cbc8520fULL,
    0xbc92434322f4f9aaULL, 0x3feebd829fde4e50ULL,
    0xbc85ca6cd7668e4bULL, 0x3feec0f170ca07baULL,
    0x3c71affc2b91ce27ULL, 0x3feec49182a3f090ULL,
    0x3c6dd235e10a73bbULL, 0x3feec86319e32323ULL,
    0xbc87c50422622263ULL, 0x3feecc667b5de565ULL,
    0x3c8b1c86e3e231d5ULL, 0x3feed09bec4a2d33ULL,
    0xbc91bbd1d3bcbb15ULL, 0x3feed503b23e255dULL,
    0x3c90cc319cee31d2ULL, 0x3feed99e1330b358ULL,
    0x3c8469846e735ab3ULL, 0x3feede6b5579fdbfULL,
    0xbc82dfcd978e9db4ULL, 0x3feee36bbfd3f37aULL,
    0x3c8c1a7792cb3387ULL, 0x3feee89f995ad3adULL,
    0xbc907b8f4ad1d9faULL, 0x3feeee07298db666ULL,
    0xbc55c3d956dcaebaULL, 0x3feef3a2b84f15fbULL,
    0xbc90a40e3da6f640ULL, 0x3feef9728de5593aULL,
    0xbc68d6f438ad9334ULL, 0x3feeff76f2fb5e47ULL,
    0xbc91eee26b588a35ULL, 0x3fef05b030a1064aULL,
    0x3c74ffd70a5fddcdULL, 0x3fef0c1e904bc1d2ULL,
    0xbc91bdfbfa9298acULL, 0x3fef12c25bd71e09ULL,
    0x3c736eae30af0cb3ULL, 0x3fef199bdd85529cULL,
    0x3c8ee3325c9ffd94ULL, 0x3fef20ab5fffd07aULL,
    0x3c84e08fd10959acULL, 0x3fef27f12e57d14bULL,
    0x3c63cdaf384e1a67ULL, 0x3fef2f6d9406e7b5ULL,
    0x3c676b2c6c921968ULL, 0x3fef3720dcef9069ULL,
    0xbc808a1883ccb5d2ULL, 0x3fef3f0b555dc3faULL,
    0xbc8fad5d3ffffa6fULL, 0x3fef472d4a07897cULL,
    0xbc900dae3875a949ULL, 0x3fef4f87080d89f2ULL,
    0x3c74a385a63d07a7ULL, 0x3fef5818dcfba487ULL,
    0xbc82919e2040220fULL, 0x3fef60e316c98398ULL,
    0x3c8e5a50d5c192acULL, 0x3fef69e603db3285ULL,
    0x3c843a59ac016b4bULL, 0x3fef7321f301b460ULL,
    0xbc82d52107b43e1fULL, 0x3fef7c97337b9b5fULL,
    0xbc892ab93b470dc9ULL, 0x3fef864614f5a129ULL,
    0x3c74b604603a88d3ULL, 0x3fef902ee78b3ff6ULL,
    0x3c83c5ec519d7271ULL, 0x3fef9a51fbc74c83ULL,
    0xbc8ff7128fd391f0ULL, 0x3fefa4afa2a490daULL,
    0xbc8dae98e223747dULL, 0x3fefaf482d8e67f1ULL,
    0x3c8ec3bc41aa2008ULL, 0x3fefba1bee615a27ULL,
    0x3c842b94c3a9eb32ULL, 0x3fefc52b376bba97ULL,
    0x3c8a64a931d185eeULL, 0x3fefd0765b6e4540ULL,
    0xbc8e37bae43be3edULL, 0x3fefdbfdad9cbe14ULL,
    0x3c77893b4d91cd9dULL, 0x3fefe7c1819e90d8ULL,
    0x3c5305c14160cc89ULL, 0x3feff3c22b8f71f1ULL,
};

double or_exp(double x) {
    const double inv_ln2_n = 0x1.71547652b82fep0 * 128.0;
    const double neg_ln2_hi_n = -0x1.62e42fefa0000p-8;
    const double neg_ln2_lo_n = -0x1.cf79abc9e3b3ap-47;
    const double shift = 0x1.8p52;
    const double c2 = 0x1.ffffffffffdbdp-2, c3 = 0x1.555555555543cp-3;
    const double c4 = 0x1.55555cf172b91p-5, c5 = 0x1.1111167a4d017p-7;
    uint32_t abstop = (uint32_t)(d2u(x) >> 52) & 0x7ff;
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {
        if (abstop - 0x3c9u >= 0x80000000u) return 1.0 + x;
        /* |x| >= 512 is outside the path's domain; defer to the libm. */
        return exp(x);
    }
    double kd = fma(inv_ln2_n, x, shift);
    uint64_t ki = d2u(kd);
    kd -= shift;
    double r = fma(kd, neg_ln2_lo_n, fma(kd, neg_ln2_hi_n, x));
    uint64_t idx = 2 * (ki % 128);
    uint64_t top = ki << (52 - 7);
    double tail = u2d(EXP_TAB[idx]);
    uint64_t sbits = EXP_TAB[idx + 1] + top;
    double r2 = r * r;
    double tmp = fma(r2 * r2, fma(r, c5, c4), fma(r2, fma(r, c3, c2), tail + r));
    double scale = u2d(sbits);
    return fma(scale, tmp, scale);
}

float or_libm_expf(float x) { return expf(x); }
double or_libm_exp(double x) { return exp(x); }

static int g_use_libm = 0;
void or_set_libm_mode(int use_libm) { g_use_libm = use_libm; }
static inline float r_expf(float x) { return g_use_libm ? expf(x) : or_expf(x); }
static inline double r_exp(double x) { return g_use_libm ? exp(x) : or_exp(x); }

typedef struct {
    uint64_t lo, hi, stride, start;
    uint64_t mismatches;
} expf_job;

static void *expf_worker(void *arg) {
    expf_job *job = (expf_job *)arg;
    uint64_t bad = 0;
    for (uint64_t u = job->lo + job->start; u < job->hi; u += job->stride) {
        float x = u2f((uint32_t)u);
        float a = or_expf(x), b = expf(x);
        if (f2u(a) != f2u(b) && !(a != a && b != b)) bad++;
    }
    job->mismatches = bad;
    return NULL;
}

typedef struct {
    uint64_t lo, begin, end;
    const float *got;
    uint64_t mismatches;
} expf_compare_job;

static void *expf_compare_worker(void *arg) {
    expf_compare_job *job = (expf_compare_job *)arg;
    uint64_t bad = 0;
    for (uint64_t i = job->begin; i < job->end; i++) {
        float a = or_expf(u2f((uint32_t)(job->lo + i))), b = job->got[i];
        if (f2u(a) != f2u(b) && !(a != a && b != b)) bad++;
    }
    job->mismatches = bad;
    return NULL;
}

uint64_t or_expf_compare(uint32_t lo_bits, uint64_t n, const float *got, uint32_t threads) {
    if (threads == 0) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t tid[64];
    expf_compare_job jobs[64];
    uint64_t per = (n + threads - 1) / threads;
    for (uint32_t i = 0; i < threads; i++) {
        jobs[i].lo = lo_bits;
        jobs[i].begin = per * i < n ? per * i : n;
        jobs[i].end = per * (i + 1) < n ? per * (i + 1) : n;
        jobs[i].got = got;
        jobs[i].mismatches = 0;
        pthread_create(&tid[i], NULL, expf_compare_worker, &jobs[i]);
    }
    uint64_t total = 0;
    for (uint32_t i = 0; i < threads; i++) {
        pthread_join(tid[i], NULL);
        total += jobs[i].mismatches;
    }
    return total;
}

uint64_t or_expf_mismatches(uint32_t lo_bits, uint32_t hi_bits, uint32_t stride,
                            uint32_t threads) {
    if (threads == 0) threads = 1;
    if (threads > 64) threads = 64;
    if (stride == 0) stride = 1;
    pthread_t tid[64];
    expf_job jobs[64];
    for (uint32_t i = 0; i < threads; i++) {
        jobs[i].lo = lo_bits;
        jobs[i].hi = hi_bits;
        jobs[i].stride = (uint64_t)stride * threads;
        jobs[i].start = (uint64_t)stride * i;
        jobs[i].mismatches = 0;
        pthread_create(&tid[i], NULL, expf_worker, &jobs[i]);
    }
    uint64_t total = 0;
    for (uint32_t i = 0; i < threads; i++) {
        pthread_join(tid[i], NULL);
        total += jobs[i].mismatches;
    }
    return total;
}

/* ------------------------------------------------------------- math.rs */

static inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; /* math.rs:11-13 */
}
static inline void mul_vec3(const double *m, const double *v, double *out) {
    double r0 = dot3(m + 0, v), r1 = dot3(m + 3, v), r2 = dot3(m + 6, v);
    out[0] = r0; out[1] = r1; out[2] = r2; /* math.rs:48-50 */
}
static inline void normalize3(const double *a, double *out) {
    double n = sqrt(dot3(a, a)); /* math.rs:39-46 */
    if (n == 0.0) {
        out[0] = a[0]; out[1] = a[1]; out[2] = a[2];
        return;
    }
    double s = 1.0 / n;
    out[0] = a[0] * s; out[1] = a[1] * s; out[2] = a[2] * s;
}
static inline void transpose3(const double *m, double *out) {
    double t[9] = {m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8]};
    memcpy(out, t, sizeof t);
}

double or_determinant3(const double *m) { /* math.rs:71-75 */
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

int or_inverse3(const double *m, double *out) { /* math.rs:77-97 */
    double det = or_determinant3(m);
    if (fabs(det) < 1e-300) return 0;
    double inv = 1.0 / det;
#define C(r0, c0, r1, c1) (m[(r0)*3 + (c0)] * m[(r1)*3 + (c1)] - m[(r0)*3 + (c1)] * m[(r1)*3 + (c0)])
    double r[9] = {
        C(1, 1, 2, 2) * inv, -C(0, 1, 2, 2) * inv, C(0, 1, 1, 2) * inv,
        -C(1, 0, 2, 2) * inv, C(0, 0, 2, 2) * inv, -C(0, 0, 1, 2) * inv,
        C(1, 0, 2, 1) * inv, -C(0, 0, 2, 1) * inv, C(0, 0, 1, 1) * inv,
    };
#undef C
    memcpy(out, r, sizeof r);
    return 1;
}

/* ---------------------------------------------------------- geometry.rs */

int or_camera_canonical(const double k[9], const double rotation_wc[9],
                        const double translation_wc[3], uint32_t width, uint32_t height,
                        const chd_distortion *distortion, chd_camera *out) {
    /* ResolvedCamera::new, canonical branch (geometry.rs:64-133);
     * CanonicalPose::center_world (calibration.rs:192-197). */
    memset(out, 0, sizeof *out);
    out->width = width;
    out->height = height;
    memcpy(out->k, k, sizeof out->k);
    if (!or_inverse3(k, out->k_inverse)) return 0;
    if (distortion) {
        out->has_distortion = 1;
        out->distortion = *distortion;
    }
    out->flip_around_x = -1;
    out->pose_kind = 0;
    memcpy(out->rotation_wc, rotation_wc, sizeof out->rotation_wc);
    transpose3(rotation_wc, out->rotation_cw);
    double c[3];
    mul_vec3(out->rotation_cw, translation_wc, c);
    out->center[0] = c[0] * -1.0;
    out->center[1] = c[1] * -1.0;
    out->center[2] = c[2] * -1.0;
    memcpy(out->translation_wc, translation_wc, sizeof out->translation_wc);
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(out->orientation_correction, I, sizeof I);
    out->focal_px = k[0];
    return 1;
}

static inline double coeff(const chd_distortion *d, uint32_t i) {
    return i < d->coeff_count ? d->coeffs[i] : 0.0;
}

static void distort_normalized(const chd_distortion *d, const double xy[2], double out[2]) {
    /* geometry.rs:315-326 */
    double k1 = coeff(d, 0), k2 = coeff(d, 1), p1 = coeff(d, 2), p2 = coeff(d, 3), k3 = coeff(d, 4);
    double x = xy[0], y = xy[1];
    double r2 = x * x + y * y;
    double radial = 1.0 + k1 * r2 + k2 * r2 * r2 + k3 * r2 * r2 * r2;
    double ox = x * radial + 2.0 * p1 * x * y + p2 * (r2 + 2.0 * x * x);
    double oy = y * radial + p1 * (r2 + 2.0 * y * y) + 2.0 * p2 * x * y;
    out[0] = ox;
    out[1] = oy;
}

void or_distort(const chd_distortion *d, const double in[2], double out[2]) {
    /* geometry.rs:298-313, 328-333 */
    double n[2] = {(in[0] - d->center[0]) / d->normalization[0],
                   (in[1] - d->center[1]) / d->normalization[1]};
    double dn[2];
    distort_normalized(d, n, dn);
    out[0] = dn[0] * d->normalization[0] + d->center[0];
    out[1] = dn[1] * d->normalization[1] + d->center[1];
}

void or_undistort(const chd_distortion *d, const double in[2], double out[2]) {
    /* geometry.rs:335-356 */
    double observed[2] = {(in[0] - d->center[0]) / d->normalization[0],
                          (in[1] - d->center[1]) / d->normalization[1]};
    double est[2] = {observed[0], observed[1]};
    for (int i = 0; i < 20; i++) {
        double dist[2];
        distort_normalized(d, est, dist);
        double next[2] = {est[0] + (observed[0] - dist[0]), est[1] + (observed[1] - dist[1])};
        double delta = rmax(fabs(next[0] - est[0]), fabs(next[1] - est[1]));
        est[0] = next[0];
        est[1] = next[1];
        if (delta < 1e-12) break;
    }
    out[0] = est[0] * d->normalization[0] + d->center[0];
    out[1] = est[1] * d->normalization[1] + d->center[1];
}

static inline void reflect_axis(const chd_camera *c, double *xy) {
    /* geometry.rs:143-150 */
    if (c->flip_around_x == 1) xy[1] = -xy[1];
    else if (c->flip_around_x == 0) xy[0] = -xy[0];
}

void or_pixel_to_ray(const chd_camera *c, const double pixel[2], double origin[3],
                     double direction[3]) {
    /* geometry.rs:152-181 */
    double ideal[2] = {pixel[0], pixel[1]};
    if (c->has_distortion) or_undistort(&c->distortion, pixel, ideal);
    double h[3] = {ideal[0], ideal[1], 1.0}, d[3], t[3];
    mul_vec3(c->k_inverse, h, t);
    normalize3(t, d);
    if (c->flip_around_x >= 0) {
        double xy[2] = {d[0], d[1]};
        reflect_axis(c, xy);
        double f[3] = {xy[0], xy[1], d[2]};
        normalize3(f, d);
    }
    double world[3];
    if (c->pose_kind == 0) {
        memcpy(origin, c->center, 3 * sizeof(double));
        mul_vec3(c->rotation_cw, d, world);
    } else {
        memcpy(origin, c->virtual_center, 3 * sizeof(double));
        double a[3];
        mul_vec3(c->real_cw, d, a);
        mul_vec3(c->reflect, a, world);
    }
    double n1[3], n2[3];
    normalize3(world, n1);
    mul_vec3(c->orientation_correction, n1, n2);
    normalize3(n2, direction);
}

static int is_identity(const double *m) {
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 9; i++)
        if (!(m[i] == I[i])) return 0;
    return 1;
}

int or_project(const chd_camera *c, const double point_in[3], int bounded, double out[2]) {
    /* geometry.rs:196-243 */
    double point[3] = {point_in[0], point_in[1], point_in[2]};
    if (!is_identity(c->orientation_correction)) {
        const double *center = c->pose_kind == 0 ? c->center : c->virtual_center;
        double tr[9], s[3] = {point[0] - center[0], point[1] - center[1], point[2] - center[2]}, r[3];
        transpose3(c->orientation_correction, tr);
        mul_vec3(tr, s, r);
        point[0] = center[0] + r[0];
        point[1] = center[1] + r[1];
        point[2] = center[2] + r[2];
    }
    double cam[3];
    if (c->pose_kind == 0) {
        double r[3];
        mul_vec3(c->rotation_wc, point, r);
        cam[0] = r[0] + c->translation_wc[0];
        cam[1] = r[1] + c->translation_wc[1];
        cam[2] = r[2] + c->translation_wc[2];
    } else {
        double s[3] = {point[0] - c->virtual_center[0], point[1] - c->virtual_center[1],
                       point[2] - c->virtual_center[2]};
        double a[3], tr[9];
        mul_vec3(c->reflect, s, a);
        transpose3(c->real_cw, tr);
        mul_vec3(tr, a, cam);
    }
    if (cam[2] <= 0.0) return 0;
    double n[2] = {cam[0] / cam[2], cam[1] / cam[2]};
    reflect_axis(c, n);
    const double *k = c->k;
    double ideal[2] = {k[0] * n[0] + k[1] * n[1] + k[2], k[4] * n[1] + k[5]};
    double mx = (double)c->width * 0.5, my = (double)c->height * 0.5;
    if (bounded && (ideal[0] < -mx || ideal[1] < -my || ideal[0] > (double)c->width + mx ||
                    ideal[1] > (double)c->height + my))
        return 0;
    if (c->has_distortion) {
        or_distort(&c->distortion, ideal, out);
    } else {
        out[0] = ideal[0];
        out[1] = ideal[1];
    }
    return 1;
}

int or_map_from(const chd_camera *self, const chd_camera *reference, const double pixel[2],
                double depth, double out[2]) {
    /* geometry.rs:249-252 */
    double o[3], d[3];
    or_pixel_to_ray(reference, pixel, o, d);
    double p[3] = {o[0] + d[0] * depth, o[1] + d[1] * depth, o[2] + d[2] * depth};
    return or_project(self, p, 1, out);
}

/* -------------------------------------------------- image.rs / align.rs */

int or_plane_sample(const float *data, uint32_t width, uint32_t height, float x, float y,
                    float *out) {
    /* image.rs:451-466 */
    if (!(x >= 0.0f && y >= 0.0f && x <= (float)(width - 1) && y <= (float)(height - 1)))
        return 0;
    size_t x0 = f32_usize(floorf(x)), y0 = f32_usize(floorf(y));
    size_t x1 = x0 + 1 < width - 1 ? x0 + 1 : width - 1;
    size_t y1 = y0 + 1 < height - 1 ? y0 + 1 : height - 1;
    float tx = x - (float)x0, ty = y - (float)y0;
    float top = data[y0 * width + x0] * (1.0f - tx) + data[y0 * width + x1] * tx;
    float bottom = data[y1 * width + x0] * (1.0f - tx) + data[y1 * width + x1] * tx;
    *out = top * (1.0f - ty) + bottom * ty;
    return 1;
}

int or_warp_map(const chd_warp *w, float x, float y, float out[2]) {
    /* align.rs:84-111 */
    float fx = x / (float)w->step, fy = y / (float)w->step;
    if (fx < 0.0f || fy < 0.0f) return 0;
    size_t c0 = f32_usize(floorf(fx)), r0 = f32_usize(floorf(fy));
    if (c0 > w->columns - 1) c0 = w->columns - 1;
    if (r0 > w->rows - 1) r0 = w->rows - 1;
    size_t c1 = c0 + 1 < w->columns - 1 ? c0 + 1 : w->columns - 1;
    size_t r1 = r0 + 1 < w->rows - 1 ? r0 + 1 : w->rows - 1;
    float tx = fx - (float)c0, ty = fy - (float)r0;
    const float *a = w->points + 2 * (r0 * w->columns + c0);
    const float *b = w->points + 2 * (r0 * w->columns + c1);
    const float *c = w->points + 2 * (r1 * w->columns + c0);
    const float *d = w->points + 2 * (r1 * w->columns + c1);
    float o[2];
    for (int k = 0; k < 2; k++) {
        float top = a[k] * (1.0f - tx) + b[k] * tx;
        float bottom = c[k] * (1.0f - tx) + d[k] * tx;
        o[k] = top * (1.0f - ty) + bottom * ty;
    }
    if (o[0] != o[0] || o[1] != o[1]) return 0;
    out[0] = o[0];
    out[1] = o[1];
    return 1;
}

int or_global_warp(const chd_camera *reference, const chd_camera *target, uint32_t width,
                   uint32_t height, uint32_t grid_step, chd_warp *out) {
    /* Warp::from_fn (align.rs:50-82) of the calibration initial grid
     * (align.rs:317-324, step 8, depth FAR_DEPTH = 1e8) and the final
     * resampling with an identity correction (align.rs:456-459). */
    chd_warp initial;
    initial.step = 8;
    initial.columns = (uint32_t)div_ceil(width, 8) + 1;
    initial.rows = (uint32_t)div_ceil(height, 8) + 1;
    size_t n0 = (size_t)initial.columns * initial.rows;
    initial.points = (float *)malloc(n0 * 2 * sizeof(float));
    initial.confidence = (float *)malloc(n0 * sizeof(float));
    for (uint32_t row = 0; row < initial.rows; row++)
        for (uint32_t column = 0; column < initial.columns; column++) {
            size_t i = (size_t)row * initial.columns + column;
            double p[2] = {(double)((size_t)column * 8), (double)((size_t)row * 8)}, q[2];
            if (or_map_from(target, reference, p, 1.0e8, q) && isfinite(q[0]) && isfinite(q[1])) {
                initial.points[2 * i] = (float)q[0];
                initial.points[2 * i + 1] = (float)q[1];
                initial.confidence[i] = 1.0f;
            } else {
                initial.points[2 * i] = NAN;
                initial.points[2 * i + 1] = NAN;
                initial.confidence[i] = 0.0f;
            }
        }
    out->step = grid_step;
    out->columns = (uint32_t)div_ceil(width, grid_step) + 1;
    out->rows = (uint32_t)div_ceil(height, grid_step) + 1;
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (uint32_t row = 0; row < out->rows; row++)
        for (uint32_t column = 0; column < out->columns; column++) {
            size_t i = (size_t)row * out->columns + column;
            double p[2] = {(double)((size_t)column * grid_step), (double)((size_t)row * grid_step)};
            /* apply_homography(IDENTITY, p) (math.rs:131-138) */
            double h[3] = {p[0], p[1], 1.0}, v[3];
            mul_vec3(I, h, v);
            int ok = !(fabs(v[2]) < 1e-12);
            float q[2];
            if (ok) ok = or_warp_map(&initial, (float)(v[0] / v[2]), (float)(v[1] / v[2]), q);
            if (ok) {
                /* initial() widens to f64; from_fn narrows back to f32 */
                out->points[2 * i] = (float)(double)q[0];
                out->points[2 * i + 1] = (float)(double)q[1];
                out->confidence[i] = 1.0f;
            } else {
                out->points[2 * i] = NAN;
                out->points[2 * i + 1] = NAN;
                out->confidence[i] = 0.0f;
            }
        }
    free(initial.points);
    free(initial.confidence);
    return 1;
}

/* -------------------------------------------------------------- depth.rs */

void or_default_options(chd_options *o) { /* depth.rs:69-91 */
    o->enabled = 1;
    o->grid_step = 8;
    o->near_depth = 500.0;
    o->far_depth = 10000000.0;
    o->planes = 96;
    o->patch_radius = 4;
    o->minimum_support = 2;
    o->best_view_count = 3;
    o->minimum_score = 0.45f;
    o->minimum_margin = 0.01f;
    o->minimum_improvement = 0.01f;
    o->minimum_neighbour_support = 2;
    o->completion_iterations = 0;
}

int or_options_valid(const chd_options *o) { /* depth.rs:1879-1891 */
    return o->enabled && o->grid_step > 0 && isfinite(o->near_depth) &&
           isfinite(o->far_depth) && o->near_depth > 0.0 && o->far_depth > o->near_depth &&
           o->planes >= 3 && o->patch_radius > 0 && o->minimum_support > 0 &&
           o->best_view_count >= o->minimum_support && o->minimum_neighbour_support > 0;
}

void or_inverse_depth_samples(double near_depth, double far_depth, uint32_t count,
                              double *out) { /* depth.rs:1834-1843 */
    double near_inverse = 1.0 / near_depth, far_inverse = 1.0 / far_depth;
    for (uint32_t i = 0; i < count; i++) {
        double w = (double)i / (double)(count - 1);
        out[i] = 1.0 / (near_inverse * (1.0 - w) + far_inverse * w);
    }
}

double or_sublabel_depth(const double *labels, uint32_t label_count, const float *costs,
                         uint32_t best) { /* depth.rs:1469-1488 */
    double centre_depth = labels[best];
    if (best <= 1 || best + 1 >= label_count) return centre_depth;
    float left = costs[best - 1], centre = costs[best], right = costs[best + 1];
    float curvature = left - 2.0f * centre + right;
    if (!isfinite(curvature) || curvature <= 1.0e-6f) return centre_depth;
    double offset = (double)rclampf(0.5f * (left - right) / curvature, -0.75f, 0.75f);
    double lower_inverse = 1.0 / labels[best - 1];
    double upper_inverse = 1.0 / labels[best + 1];
    double inverse_step = (upper_inverse - lower_inverse) * 0.5;
    return 1.0 / (1.0 / centre_depth + offset * inverse_step);
}

int or_aggregate(const float *scores, uint32_t count, const chd_options *o, float *out) {
    /* depth.rs:1761-1770: sort descending by total_cmp, mean of the best
     * best_view_count, summed from -0.0 in sorted order. */
    if (count < o->minimum_support) return 0;
    float v[CHD_MAX_VIEWS];
    for (uint32_t i = 0; i < count; i++) {
        float s = scores[i];
        uint32_t j = i;
        while (j > 0 && tcmp32(v[j - 1], s) < 0) {
            v[j] = v[j - 1];
            j--;
        }
        v[j] = s;
    }
    uint32_t n = o->best_view_count < count ? o->best_view_count : count;
    if (n < o->minimum_support) return 0;
    float sum = -0.0f;
    for (uint32_t i = 0; i < n; i++) sum += v[i];
    *out = sum / (float)n;
    return 1;
}

int or_warped_patch_zncc(const float *reference, uint32_t rw, uint32_t rh,
                         const float *target, uint32_t tw, uint32_t th,
                         const chd_warp *global, double cx, double cy, double dx_,
                         double dy_, uint32_t radius, float *out) {
    /* depth.rs:1772-1822 */
    float centre_reference;
    if (!or_plane_sample(reference, rw, rh, (float)((cx - 0.5) * 0.5), (float)((cy - 0.5) * 0.5),
                         &centre_reference))
        return 0;
    float sigma = rmaxf((float)radius * 0.75f, 1.0f);
    float count = 0.0f, sum_reference = 0.0f, sum_target = 0.0f;
    float sum_reference_sq = 0.0f, sum_target_sq = 0.0f, sum_product = 0.0f;
    int r = (int)radius;
    for (int dy = -r; dy <= r; dy++) {
        for (int dx = -r; dx <= r; dx++) {
            double px = cx + (double)dx * 2.0, py = cy + (double)dy * 2.0;
            float reference_value;
            if (!or_plane_sample(reference, rw, rh, (float)((px - 0.5) * 0.5),
                                 (float)((py - 0.5) * 0.5), &reference_value))
                return 0;
            float global_q[2];
            if (!or_warp_map(global, (float)px, (float)py, global_q)) return 0;
            double qx = (double)global_q[0] + dx_, qy = (double)global_q[1] + dy_;
            float target_value;
            if (!or_plane_sample(target, tw, th, (float)((qx - 0.5) * 0.5),
                                 (float)((qy - 0.5) * 0.5), &target_value))
                return 0;
            float distance_sq = (float)(dx * dx + dy * dy);
            float spatial = r_expf(-distance_sq / (2.0f * sigma * sigma));
            float range = r_expf(-1.2f * fabsf(reference_value - centre_reference));
            float weight = spatial * range;
            count += weight;
            sum_reference += weight * reference_value;
            sum_target += weight * target_value;
            sum_reference_sq += weight * reference_value * reference_value;
            sum_target_sq += weight * target_value * target_value;
            sum_product += weight * reference_value * target_value;
        }
    }
    float covariance = sum_product - sum_reference * sum_target / count;
    float reference_energy = sum_reference_sq - sum_reference * sum_reference / count;
    float target_energy = sum_target_sq - sum_target * sum_target / count;
    float denominator = sqrtf(rmaxf(reference_energy, 0.0f) * rmaxf(target_energy, 0.0f));
    if (!(denominator > 1.0e-6f)) return 0;
    *out = rclampf(covariance / denominator, -1.0f, 1.0f);
    return 1;
}

/* Per-call view of the inputs. */
typedef struct {
    const chd_view *views;
    uint32_t count;
    uint32_t ref;
    const chd_options *o;
} ctx_t;

static inline int view_active(const ctx_t *c, uint32_t i) { /* depth.rs:1574-1579 */
    return i != c->ref && c->views[i].accepted && c->views[i].camera != NULL;
}

static int map_at_depth(const chd_camera *rc, const chd_camera *tc, const chd_warp *global,
                        const double pixel[2], double depth, double out[2]) {
    /* depth.rs:1738-1759 */
    float gq[2];
    if (!or_warp_map(global, (float)pixel[0], (float)pixel[1], gq)) return 0;
    double g[2] = {(double)gq[0], (double)gq[1]}, corrected[2], q[2];
    if (!or_map_from(rc, tc, g, INFINITY_DEPTH, corrected)) return 0;
    if (!or_map_from(tc, rc, corrected, depth, q)) return 0;
    if (!(isfinite(q[0]) && isfinite(q[1]))) return 0;
    out[0] = q[0];
    out[1] = q[1];
    return 1;
}

static int score_one_view(const chd_view *reference, const chd_view *target,
                          const chd_warp *global, const double centre[2], int has_depth,
                          double depth, uint32_t radius, float *score) {
    /* depth.rs:1593-1627 */
    if (!reference->camera || !target->camera) return 0;
    float gc[2];
    if (!or_warp_map(global, (float)centre[0], (float)centre[1], gc)) return 0;
    double mapped[2];
    if (has_depth) {
        if (!map_at_depth(reference->camera, target->camera, global, centre, depth, mapped))
            return 0;
    } else {
        mapped[0] = (double)gc[0];
        mapped[1] = (double)gc[1];
    }
    double delta[2] = {mapped[0] - (double)gc[0], mapped[1] - (double)gc[1]};
    if (fabs(delta[0]) > 128.0 || fabs(delta[1]) > 128.0) return 0;
    return or_warped_patch_zncc(reference->luminance, reference->plane_width,
                                reference->plane_height, target->luminance,
                                target->plane_width, target->plane_height, global, centre[0],
                                centre[1], delta[0], delta[1], radius, score);
}

static uint32_t score_views(const ctx_t *c, const double centre[2], int has_depth,
                            double depth, uint32_t radius, float *scores) {
    /* depth.rs:1566-1591 */
    uint32_t n = 0;
    for (uint32_t i = 0; i < c->count; i++) {
        if (!view_active(c, i)) continue;
        float s;
        if (score_one_view(&c->views[c->ref], &c->views[i], &c->views[i].global_warp, centre,
                           has_depth, depth, radius, &s))
            scores[n++] = s;
    }
    return n;
}

static int aggregate_views(const ctx_t *c, const double centre[2], int has_depth, double depth,
                           const chd_options *o, float *out) {
    float s[CHD_MAX_VIEWS];
    uint32_t n = score_views(c, centre, has_depth, depth, o->patch_radius, s);
    return or_aggregate(s, n, o, out);
}

static float reference_guidance(const chd_view *reference, const double pixel[2]) {
    /* depth.rs:1824-1832 */
    float v;
    if (!or_plane_sample(reference->luminance, reference->plane_width, reference->plane_height,
                         (float)((pixel[0] - 0.5) * 0.5), (float)((pixel[1] - 0.5) * 0.5), &v))
        return 0.0f;
    return v;
}

/* build_cost_volume (depth.rs:1187-1277) over a row band */
typedef struct {
    const ctx_t *c;
    const double *labels; /* label 0 unused */
    uint32_t label_count, columns, step;
    uint32_t row_begin, row_end, row_base;
    float *scores, *costs, *guidance;
    uint8_t *tested;
} cost_job;

static void *cost_worker(void *arg) {
    cost_job *j = (cost_job *)arg;
    uint32_t L = j->label_count;
    for (uint32_t row = j->row_begin; row < j->row_end; row++) {
        for (uint32_t column = 0; column < j->columns; column++) {
            size_t node = (size_t)(row - j->row_base) * j->columns + column;
            double p[2] = {(double)((size_t)column * j->step), (double)((size_t)row * j->step)};
            j->guidance[node] = reference_guidance(&j->c->views[j->c->ref], p);
            int node_tested = 0;
            for (uint32_t label = 0; label < L; label++) {
                float score;
                if (!aggregate_views(j->c, p, label != 0, j->labels[label], j->c->o, &score))
                    score = NAN;
                node_tested |= isfinite(score);
                j->scores[node * L + label] = score;
                float cost;
                if (isfinite(score)) {
                    float near_prior = label == 0 ? 0.0f
                                                  : NEAR_DEPTH_PRIOR * (float)(label - 1) /
                                                        (float)(L - 2);
                    cost = rclampf(1.0f - score, 0.0f, 2.0f) + near_prior;
                } else {
                    cost = MISSING_COST;
                }
                j->costs[node * L + label] = cost;
            }
            j->tested[node] = (uint8_t)node_tested;
        }
    }
    return NULL;
}

static uint32_t worker_count(uint32_t threads, uint32_t rows) {
    if (threads == 0) {
        long n = sysconf(_SC_NPROCESSORS_ONLN);
        threads = n > 0 ? (uint32_t)n : 1;
    }
    if (threads > rows) threads = rows;
    if (threads < 1) threads = 1;
    return threads;
}

static double *label_table(const chd_options *o, uint32_t *count) {
    /* depth.rs:1196-1201: label 0 = None (global), 1.. = far -> near */
    uint32_t L = o->planes + 1;
    double *labels = (double *)malloc(L * sizeof(double));
    double *finite = (double *)malloc(o->planes * sizeof(double));
    or_inverse_depth_samples(o->near_depth, o->far_depth, o->planes, finite);
    labels[0] = NAN;
    for (uint32_t i = 0; i < o->planes; i++) labels[1 + i] = finite[o->planes - 1 - i];
    free(finite);
    *count = L;
    return labels;
}

static void run_bands(void *(*fn)(void *), void *jobs, size_t job_size, uint32_t n) {
    pthread_t *tid = (pthread_t *)malloc(n * sizeof(pthread_t));
    for (uint32_t i = 0; i < n; i++) pthread_create(&tid[i], NULL, fn, (char *)jobs + i * job_size);
    for (uint32_t i = 0; i < n; i++) pthread_join(tid[i], NULL);
    free(tid);
}

int or_cost_volume_rows(const chd_view *views, uint32_t count, uint32_t reference_index,
                        const chd_options *o, uint32_t row_begin, uint32_t row_end,
                        uint32_t threads, float *scores, float *costs, float *guidance,
                        uint8_t *tested) {
    ctx_t c = {views, count, reference_index, o};
    uint32_t step = o->grid_step > 4 ? o->grid_step : 4;
    uint32_t columns = (uint32_t)div_ceil(views[reference_index].width, step) + 1;
    uint32_t L;
    double *labels = label_table(o, &L);
    uint32_t rows = row_end - row_begin;
    uint32_t workers = worker_count(threads, rows);
    uint32_t per = (uint32_t)div_ceil(rows, workers);
    cost_job jobs[256];
    uint32_t n = 0;
    for (uint32_t first = row_begin; first < row_end && n < 256; first += per) {
        cost_job *j = &jobs[n++];
        j->c = &c;
        j->labels = labels;
        j->label_count = L;
        j->columns = columns;
        j->step = step;
        j->row_begin = first;
        j->row_end = first + per < row_end ? first + per : row_end;
        j->row_base = row_begin;
        j->scores = scores;
        j->costs = costs;
        j->guidance = guidance;
        j->tested = tested;
    }
    run_bands(cost_worker, jobs, sizeof(cost_job), n);
    free(labels);
    return 0;
}

static void accumulate_path(const float *data, const float *guidance, float *output,
                            uint32_t labels, const uint32_t *path, uint32_t length,
                            float *previous, float *current) {
    /* depth.rs:1340-1383 */
    for (uint32_t pi = 0; pi < length; pi++) {
        uint32_t node = path[pi];
        size_t offset = (size_t)node * labels;
        if (pi == 0) {
            memcpy(previous, data + offset, labels * sizeof(float));
            for (uint32_t l = 0; l < labels; l++) output[offset + l] += previous[l];
            continue;
        }
        uint32_t previous_node = path[pi - 1];
        float edge = fabsf(guidance[node] - guidance[previous_node]);
        float edge_scale = 1.0f / (1.0f + 4.0f * edge);
        float p1 = SGM_SMALL_PENALTY * rmaxf(edge_scale, 0.35f);
        float p2 = SGM_LARGE_PENALTY * rmaxf(edge_scale, 0.12f);
        float minimum_previous = INFINITY;
        for (uint32_t l = 0; l < labels; l++) minimum_previous = rminf(minimum_previous, previous[l]);
        for (uint32_t l = 0; l < labels; l++) {
            float same = previous[l];
            float lower = l > 0 ? previous[l - 1] + p1 : INFINITY;
            float higher = l + 1 < labels ? previous[l + 1] + p1 : INFINITY;
            float jump = minimum_previous + p2;
            float m = rminf(rminf(rminf(same, lower), higher), jump);
            current[l] = data[offset + l] + m - minimum_previous;
            output[offset + l] += current[l];
        }
        float *t = previous;
        previous = current;
        current = t;
    }
}

uint32_t or_diagonal_paths(uint32_t columns, uint32_t rows, int mirrored, uint32_t *offsets,
                           uint32_t *nodes) { /* depth.rs:1313-1338 */
    uint32_t paths = 0, n = 0;
    offsets[0] = 0;
    for (uint32_t start_column = 0; start_column < columns; start_column++) {
        int64_t column = start_column;
        uint32_t row = 0;
        while (column >= 0 && column < (int64_t)columns && row < rows) {
            nodes[n++] = row * columns + (uint32_t)column;
            column += mirrored ? -1 : 1;
            row++;
        }
        offsets[++paths] = n;
    }
    for (uint32_t start_row = 1; start_row < rows; start_row++) {
        int64_t column = mirrored ? (int64_t)columns - 1 : 0;
        uint32_t row = start_row;
        while (column >= 0 && column < (int64_t)columns && row < rows) {
            nodes[n++] = row * columns + (uint32_t)column;
            column += mirrored ? -1 : 1;
            row++;
        }
        offsets[++paths] = n;
    }
    return paths;
}

void or_semi_global_costs(const float *data, const float *guidance, uint32_t columns,
                          uint32_t rows, uint32_t labels, float *output) {
    /* depth.rs:1279-1311 */
    size_t nodes = (size_t)columns * rows;
    memset(output, 0, nodes * labels * sizeof(float));
    float *previous = (float *)malloc(labels * sizeof(float));
    float *current = (float *)malloc(labels * sizeof(float));
    uint32_t *path = (uint32_t *)malloc((columns > rows ? columns : rows) * sizeof(uint32_t));
    uint32_t *reverse = (uint32_t *)malloc((columns > rows ? columns : rows) * sizeof(uint32_t));
    for (uint32_t row = 0; row < rows; row++) {
        for (uint32_t column = 0; column < columns; column++) {
            path[column] = row * columns + column;
            reverse[columns - 1 - column] = row * columns + column;
        }
        accumulate_path(data, guidance, output, labels, path, columns, previous, current);
        accumulate_path(data, guidance, output, labels, reverse, columns, previous, current);
    }
    for (uint32_t column = 0; column < columns; column++) {
        for (uint32_t row = 0; row < rows; row++) {
            path[row] = row * columns + column;
            reverse[rows - 1 - row] = row * columns + column;
        }
        accumulate_path(data, guidance, output, labels, path, rows, previous, current);
        accumulate_path(data, guidance, output, labels, reverse, rows, previous, current);
    }
    uint32_t *offsets = (uint32_t *)malloc((columns + rows + 1) * sizeof(uint32_t));
    uint32_t *dnodes = (uint32_t *)malloc(nodes * sizeof(uint32_t));
    for (int mirrored = 0; mirrored < 2; mirrored++) {
        uint32_t count = or_diagonal_paths(columns, rows, mirrored, offsets, dnodes);
        for (uint32_t p = 0; p < count; p++) {
            uint32_t len = offsets[p + 1] - offsets[p];
            const uint32_t *pp = dnodes + offsets[p];
            accumulate_path(data, guidance, output, labels, pp, len, previous, current);
            for (uint32_t i = 0; i < len; i++) reverse[i] = pp[len - 1 - i];
            accumulate_path(data, guidance, output, labels, reverse, len, previous, current);
        }
    }
    free(offsets);
    free(dnodes);
    free(previous);
    free(current);
    free(path);
    free(reverse);
}

void or_select_depths(const double *label_depths, uint32_t labels, const float *scores,
                      const uint8_t *tested, const float *regularised, uint32_t nodes,
                      const chd_options *o, chd_node_depth *field, uint8_t *fillable,
                      uint16_t *best_label_out) {
    /* depth.rs:1385-1467 */
    for (uint32_t node = 0; node < nodes; node++) {
        memset(&field[node], 0, sizeof field[node]);
        fillable[node] = 0;
        if (best_label_out) best_label_out[node] = 0;
        if (!tested[node]) continue;
        size_t offset = (size_t)node * labels;
        if (labels < 2) continue;
        uint32_t best_label = 1;
        float best_cost = regularised[offset + 1];
        for (uint32_t l = 2; l < labels; l++)
            if (tcmp32(regularised[offset + l], best_cost) < 0) {
                best_cost = regularised[offset + l];
                best_label = l;
            }
        if (best_label_out) best_label_out[node] = (uint16_t)best_label;
        double depth = or_sublabel_depth(label_depths, labels, regularised + offset, best_label);
        float score = scores[offset + best_label];
        if (!isfinite(score) || score < MINIMUM_REGULARIZED_SCORE) {
            fillable[node] = 1;
            continue;
        }
        int have = 0;
        float competing = 0.0f;
        for (uint32_t l = 1; l < labels; l++) {
            uint32_t diff = l > best_label ? l - best_label : best_label - l;
            if (diff <= 2) continue;
            if (!have || tcmp32(regularised[offset + l], competing) < 0) {
                competing = regularised[offset + l];
                have = 1;
            }
        }
        if (!have) competing = best_cost;
        float margin = rmaxf((competing - best_cost) / SGM_DIRECTIONS, 0.0f);
        float baseline = scores[offset];
        float improvement = isfinite(baseline) ? score - baseline : 0.0f;
        if (isfinite(baseline) && improvement < -MAXIMUM_REGULARIZED_BASELINE_LOSS) continue;
        int measured = score >= o->minimum_score &&
                       improvement >= o->minimum_improvement * 0.5f &&
                       (margin >= o->minimum_margin || improvement >= o->minimum_improvement);
        float score_confidence =
            rclampf((score - o->minimum_score) / rmaxf(1.0f - o->minimum_score, 1.0e-3f), 0.0f, 1.0f);
        float margin_confidence = rclampf(margin / rmaxf(o->minimum_margin, 1.0e-3f), 0.0f, 1.0f);
        float improvement_confidence =
            rclampf(improvement / rmaxf(o->minimum_improvement, 1.0e-3f), 0.0f, 1.0f);
        field[node].valid = 1;
        field[node].depth = depth;
        field[node].confidence =
            measured ? rclampf(0.30f + 0.45f * score_confidence + 0.25f * margin_confidence, 0.0f, 1.0f)
                     : rclampf(0.10f + 0.25f * score_confidence + 0.15f * margin_confidence +
                                   0.10f * improvement_confidence,
                               0.10f, 0.55f);
        field[node].improvement = improvement;
        field[node].regularized = (uint8_t)!measured;
    }
}

void or_complete_depth_field(chd_node_depth *field, const float *guidance,
                             const uint8_t *fillable, uint32_t columns, uint32_t rows,
                             const chd_options *o) {
    /* depth.rs:1490-1564 */
    size_t n = (size_t)columns * rows;
    chd_node_depth *previous = (chd_node_depth *)malloc(n * sizeof(chd_node_depth));
    for (uint32_t it = 0; it < o->completion_iterations; it++) {
        memcpy(previous, field, n * sizeof(chd_node_depth));
        int changed = 0;
        for (uint32_t row = 0; row < rows; row++)
            for (uint32_t column = 0; column < columns; column++) {
                size_t index = (size_t)row * columns + column;
                if (previous[index].valid || !fillable[index]) continue;
                /* candidates are (1.0 / neighbour.depth: f64, weight: f32),
                 * kept sorted by total_cmp of the inverse depth (stable) */
                double inv[8];
                float w[8];
                uint32_t m = 0;
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        if (dx == 0 && dy == 0) continue;
                        int64_t x = (int64_t)column + dx, y = (int64_t)row + dy;
                        if (x < 0 || y < 0 || x >= columns || y >= rows) continue;
                        size_t ni = (size_t)y * columns + (size_t)x;
                        if (!previous[ni].valid) continue;
                        float edge = fabsf(guidance[index] - guidance[ni]);
                        if (edge > 0.35f) continue;
                        float spatial = (dx != 0 && dy != 0) ? 0.707f : 1.0f;
                        float weight = previous[ni].confidence * spatial * r_expf(-3.0f * edge);
                        if (weight >= 0.05f) {
                            double v = 1.0 / previous[ni].depth;
                            /* stable insertion by total_cmp */
                            uint32_t j = m;
                            while (j > 0 && tcmp64(inv[j - 1], v) > 0) {
                                inv[j] = inv[j - 1];
                                w[j] = w[j - 1];
                                j--;
                            }
                            inv[j] = v;
                            w[j] = weight;
                            m++;
                        }
                    }
                if (m < o->minimum_neighbour_support) continue;
                float total = -0.0f;
                for (uint32_t i = 0; i < m; i++) total += w[i];
                float accumulated = 0.0f;
                double inverse_depth = inv[m / 2];
                for (uint32_t i = 0; i < m; i++) {
                    accumulated += w[i];
                    if (accumulated >= total * 0.5f) {
                        inverse_depth = inv[i];
                        break;
                    }
                }
                float maxw = 0.0f;
                for (uint32_t i = 0; i < m; i++) maxw = rmaxf(maxw, w[i]);
                field[index].valid = 1;
                field[index].depth = 1.0 / inverse_depth;
                field[index].confidence = rclampf(maxw * 0.92f, 0.15f, 0.75f);
                field[index].improvement = 0.0f;
                field[index].regularized = 1;
                changed = 1;
            }
        if (!changed) break;
    }
    free(previous);
}

static chd_node_depth nearest_coarse_depth(const chd_node_depth *coarse, uint32_t columns,
                                           uint32_t rows, uint32_t step,
                                           const double pixel[2]) { /* depth.rs:1062-1076 */
    size_t column = f64_usize(rclamp(round(pixel[0] / (double)step), 0.0, (double)(columns - 1)));
    size_t row = f64_usize(rclamp(round(pixel[1] / (double)step), 0.0, (double)(rows - 1)));
    return coarse[row * columns + column];
}

uint32_t or_direct_depth_candidates(int seeded, double seed_depth, double inverse_step,
                                    const chd_options *o, double *out) {
    /* depth.rs:1078-1107 */
    double far_inverse = 1.0 / o->far_depth, near_inverse = 1.0 / o->near_depth;
    uint32_t n = 0;
    if (seeded) {
        for (int offset = -6; offset <= 6; offset++) {
            double inverse = rclamp(1.0 / seed_depth + (double)offset * inverse_step * 0.5,
                                    far_inverse, near_inverse);
            double depth = 1.0 / inverse;
            if (n == 0 || fabs(1.0 / out[n - 1] - 1.0 / depth) > inverse_step * 0.1)
                out[n++] = depth;
        }
    } else {
        double d[32];
        or_inverse_depth_samples(o->near_depth, o->far_depth, 32, d);
        for (uint32_t i = 0; i < 32; i++) out[i] = d[31 - i];
        n = 32;
    }
    return n;
}

int or_select_direct_depth(const double *depths, const float *scores,
                           const uint8_t *has_score, uint32_t count, int has_baseline,
                           float baseline, int seeded, const chd_options *o,
                           chd_node_depth *out) {
    /* depth.rs:1109-1165 */
    double far_inverse = 1.0 / o->far_depth;
    double inverse_range = 1.0 / o->near_depth - far_inverse;
    uint32_t idx[64];
    float obj[64];
    uint32_t n = 0;
    for (uint32_t i = 0; i < count; i++) {
        if (!has_score[i]) continue;
        double near_fraction = rclamp((1.0 / depths[i] - far_inverse) / inverse_range, 0.0, 1.0);
        float objective = scores[i] - NEAR_DEPTH_PRIOR * (float)near_fraction;
        /* stable sort, descending by total_cmp */
        uint32_t j = n;
        while (j > 0 && tcmp32(obj[j - 1], objective) < 0) {
            obj[j] = obj[j - 1];
            idx[j] = idx[j - 1];
            j--;
        }
        obj[j] = objective;
        idx[j] = i;
        n++;
    }
    memset(out, 0, sizeof *out);
    if (n == 0) return 0;
    uint32_t best_index = idx[0];
    double depth = depths[best_index];
    float score = scores[best_index];
    float best_objective = obj[0];
    float competing = best_objective;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t d = idx[k] > best_index ? idx[k] - best_index : best_index - idx[k];
        if (d > 1) {
            competing = obj[k];
            break;
        }
    }
    float margin = rmaxf(best_objective - competing, 0.0f);
    float improvement = has_baseline ? score - baseline : 0.0f;
    float score_floor = rmaxf(o->minimum_score - 0.03f, MINIMUM_REGULARIZED_SCORE);
    float required_margin = seeded ? o->minimum_margin * 0.6f : o->minimum_margin;
    int improves_global = !has_baseline || improvement >= o->minimum_improvement * 0.5f;
    int supported = score >= score_floor && improves_global &&
                    (margin >= required_margin || improvement >= o->minimum_improvement);
    if (!supported) return 0;
    float score_confidence =
        rclampf((score - score_floor) / rmaxf(1.0f - score_floor, 1.0e-3f), 0.0f, 1.0f);
    float margin_confidence = rclampf(margin / rmaxf(required_margin, 1.0e-3f), 0.0f, 1.0f);
    out->valid = 1;
    out->depth = depth;
    out->confidence = rclampf(0.30f + 0.45f * score_confidence + 0.25f * margin_confidence, 0.0f, 1.0f);
    out->improvement = improvement;
    out->regularized = 0;
    return 1;
}

int or_best_depth_candidate(const double *depths, const float *scores,
                            const uint8_t *has_score, uint32_t count,
                            const chd_options *o) { /* depth.rs:1167-1185 (last max) */
    double far_inverse = 1.0 / o->far_depth;
    double inverse_range = 1.0 / o->near_depth - far_inverse;
    int best = -1;
    float best_value = 0.0f;
    for (uint32_t i = 0; i < count; i++) {
        if (!has_score[i]) continue;
        double near_fraction = rclamp((1.0 / depths[i] - far_inverse) / inverse_range, 0.0, 1.0);
        float v = scores[i] - NEAR_DEPTH_PRIOR * (float)near_fraction;
        if (best < 0 || tcmp32(v, best_value) >= 0) {
            best = (int)i;
            best_value = v;
        }
    }
    return best;
}

void or_reject_isolated_direct_depths(chd_node_depth *field, const float *guidance,
                                      uint32_t columns, uint32_t rows,
                                      double inverse_tolerance) { /* depth.rs:881-924 */
    size_t n = (size_t)columns * rows;
    chd_node_depth *source = (chd_node_depth *)malloc(n * sizeof(chd_node_depth));
    memcpy(source, field, n * sizeof(chd_node_depth));
    for (uint32_t row = 0; row < rows; row++)
        for (uint32_t column = 0; column < columns; column++) {
            size_t index = (size_t)row * columns + column;
            if (!source[index].valid) continue;
            double inverse = 1.0 / source[index].depth;
            uint32_t agreeing = 0;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    if (dx == 0 && dy == 0) continue;
                    int64_t x = (int64_t)column + dx, y = (int64_t)row + dy;
                    if (x < 0 || y < 0 || x >= columns || y >= rows) continue;
                    size_t ni = (size_t)y * columns + (size_t)x;
                    if (!source[ni].valid) continue;
                    float image_edge = fabsf(guidance[index] - guidance[ni]);
                    if (image_edge <= 0.6f && fabs(1.0 / source[ni].depth - inverse) <= inverse_tolerance)
                        agreeing++;
                }
            uint32_t minimum = source[index].confidence >= 0.75f ? 1 : 2;
            if (agreeing < minimum) memset(&field[index], 0, sizeof field[index]);
        }
    free(source);
}

void or_reject_small_direct_components(chd_node_depth *field, const float *guidance,
                                       uint32_t columns, uint32_t rows,
                                       double inverse_tolerance, uint32_t minimum_nodes) {
    /* depth.rs:933-989 */
    if (minimum_nodes <= 1) return;
    size_t n = (size_t)columns * rows;
    chd_node_depth *source = (chd_node_depth *)malloc(n * sizeof(chd_node_depth));
    memcpy(source, field, n * sizeof(chd_node_depth));
    uint8_t *visited = (uint8_t *)calloc(n, 1);
    uint32_t *pending = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *component = (uint32_t *)malloc(n * sizeof(uint32_t));
    for (size_t seed = 0; seed < n; seed++) {
        if (visited[seed] || !source[seed].valid) continue;
        visited[seed] = 1;
        size_t np = 0, nc = 0;
        pending[np++] = (uint32_t)seed;
        while (np > 0) {
            uint32_t index = pending[--np];
            component[nc++] = index;
            uint32_t column = index % columns, row = index / columns;
            double inverse = 1.0 / source[index].depth;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    if (dx == 0 && dy == 0) continue;
                    int64_t x = (int64_t)column + dx, y = (int64_t)row + dy;
                    if (x < 0 || y < 0 || x >= columns || y >= rows) continue;
                    size_t ni = (size_t)y * columns + (size_t)x;
                    if (visited[ni] || !source[ni].valid) continue;
                    float image_edge = fabsf(guidance[index] - guidance[ni]);
                    double inverse_difference = fabs(1.0 / source[ni].depth - inverse);
                    if (image_edge <= 0.6f && inverse_difference <= inverse_tolerance) {
                        visited[ni] = 1;
                        pending[np++] = (uint32_t)ni;
                    }
                }
        }
        if (nc < minimum_nodes)
            for (size_t k = 0; k < nc; k++) memset(&field[component[k]], 0, sizeof(chd_node_depth));
    }
    free(source);
    free(visited);
    free(pending);
    free(component);
}

void or_fit_local_depth_planes(chd_node_depth *field, const float *guidance,
                               uint32_t columns, uint32_t rows,
                               double inverse_tolerance) { /* depth.rs:995-1060 */
    size_t n = (size_t)columns * rows;
    chd_node_depth *source = (chd_node_depth *)malloc(n * sizeof(chd_node_depth));
    memcpy(source, field, n * sizeof(chd_node_depth));
    for (uint32_t row = 0; row < rows; row++)
        for (uint32_t column = 0; column < columns; column++) {
            size_t index = (size_t)row * columns + column;
            if (!source[index].valid) continue;
            chd_node_depth node = source[index];
            double centre_inverse = 1.0 / node.depth;
            double normal[9] = {0};
            double rhs[3] = {0};
            uint32_t support = 0;
            for (int dy = -2; dy <= 2; dy++)
                for (int dx = -2; dx <= 2; dx++) {
                    int64_t x = (int64_t)column + dx, y = (int64_t)row + dy;
                    if (x < 0 || y < 0 || x >= columns || y >= rows) continue;
                    size_t ni = (size_t)y * columns + (size_t)x;
                    if (!source[ni].valid) continue;
                    double inverse = 1.0 / source[ni].depth;
                    float image_edge = fabsf(guidance[index] - guidance[ni]);
                    if (image_edge > 0.6f || fabs(inverse - centre_inverse) > inverse_tolerance)
                        continue;
                    double spatial = r_exp((double)(-(dx * dx + dy * dy)) / 6.0);
                    double weight = (double)source[ni].confidence * spatial *
                                    r_exp(-2.5 * (double)image_edge);
                    double basis[3] = {1.0, (double)dx, (double)dy};
                    for (int r = 0; r < 3; r++) {
                        rhs[r] += weight * basis[r] * inverse;
                        for (int c = 0; c < 3; c++) normal[r * 3 + c] += weight * basis[r] * basis[c];
                    }
                    support++;
                }
            if (support < 6) continue;
            double inv[9];
            if (!or_inverse3(normal, inv)) continue;
            double sol[3];
            mul_vec3(inv, rhs, sol);
            double inverse = sol[0];
            if (!(isfinite(inverse) && inverse > 0.0)) continue;
            double maximum_adjustment = inverse_tolerance * 0.25;
            inverse = rclamp(inverse, centre_inverse - maximum_adjustment,
                             centre_inverse + maximum_adjustment);
            node.depth = 1.0 / inverse;
            field[index] = node;
        }
    free(source);
}

/* NodeWarp helpers (depth.rs:674-712) */
static int warp_mapping(const chd_node_warp *d, float global[2], float point[2]) {
    if (d->code == CHD_WARP_GLOBAL) {
        global[0] = point[0] = d->global[0];
        global[1] = point[1] = d->global[1];
        return 1;
    }
    if (d->code == CHD_WARP_REFINED) {
        global[0] = d->global[0];
        global[1] = d->global[1];
        point[0] = d->point[0];
        point[1] = d->point[1];
        return 1;
    }
    return 0;
}

int or_warp_decisions_agree(uint32_t kind, const chd_node_warp *c, const chd_node_warp *n) {
    if (kind == 0 && c->code == CHD_WARP_REFINED && n->code == CHD_WARP_REFINED) {
        float cd[2] = {c->point[0] - c->global[0], c->point[1] - c->global[1]};
        float nd[2] = {n->point[0] - n->global[0], n->point[1] - n->global[1]};
        float dx = cd[0] - nd[0], dy = cd[1] - nd[1];
        return dx * dx + dy * dy <= 9.0f;
    }
    if (kind == 1 && c->code == CHD_WARP_OCCLUDED && n->code == CHD_WARP_OCCLUDED) return 1;
    return 0;
}

void or_enforce_warp_consensus(chd_node_warp *decisions, uint32_t columns,
                               uint32_t rows) { /* depth.rs:582-618 */
    size_t n = (size_t)columns * rows;
    chd_node_warp *source = (chd_node_warp *)malloc(n * sizeof(chd_node_warp));
    memcpy(source, decisions, n * sizeof(chd_node_warp));
    for (uint32_t row = 0; row < rows; row++)
        for (uint32_t column = 0; column < columns; column++) {
            size_t index = (size_t)row * columns + column;
            uint32_t kind, minimum;
            if (source[index].code == CHD_WARP_REFINED) {
                kind = 0;
                minimum = source[index].measured ? 2 : 4;
            } else if (source[index].code == CHD_WARP_OCCLUDED) {
                kind = 1;
                minimum = source[index].measured ? 3 : 5;
            } else {
                continue;
            }
            uint32_t neighbours = 0;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    if (dx == 0 && dy == 0) continue;
                    int64_t x = (int64_t)column + dx, y = (int64_t)row + dy;
                    if (x < 0 || y < 0 || x >= columns || y >= rows) continue;
                    neighbours += (uint32_t)or_warp_decisions_agree(
                        kind, &source[index], &source[(size_t)y * columns + (size_t)x]);
                }
            if (neighbours < minimum) {
                chd_node_warp g;
                memset(&g, 0, sizeof g);
                g.code = CHD_WARP_GLOBAL;
                g.global[0] = source[index].global[0];
                g.global[1] = source[index].global[1];
                decisions[index] = g;
            }
        }
    free(source);
}

void or_suppress_warp_boundaries(chd_node_warp *decisions, const float *guidance,
                                 uint32_t columns, uint32_t rows) { /* depth.rs:620-672 */
    size_t n = (size_t)columns * rows;
    chd_node_warp *source = (chd_node_warp *)malloc(n * sizeof(chd_node_warp));
    memcpy(source, decisions, n * sizeof(chd_node_warp));
    uint8_t *boundary = (uint8_t *)calloc(n, 1);
    for (uint32_t row = 0; row < rows; row++)
        for (uint32_t column = 0; column < columns; column++) {
            size_t index = (size_t)row * columns + column;
            float g[2], p[2];
            if (!warp_mapping(&source[index], g, p)) continue;
            float delta[2] = {p[0] - g[0], p[1] - g[1]};
            uint32_t nx[2] = {column + 1, column}, ny[2] = {row, row + 1};
            for (int k = 0; k < 2; k++) {
                if (nx[k] >= columns || ny[k] >= rows) continue;
                size_t ni = (size_t)ny[k] * columns + nx[k];
                float ng[2], np[2];
                if (!warp_mapping(&source[ni], ng, np)) continue;
                float nd[2] = {np[0] - ng[0], np[1] - ng[1]};
                float diff[2] = {delta[0] - nd[0], delta[1] - nd[1]};
                int mapping_edge = diff[0] * diff[0] + diff[1] * diff[1] >
                                   WARP_BOUNDARY_DELTA_PX * WARP_BOUNDARY_DELTA_PX;
                int image_edge = fabsf(guidance[index] - guidance[ni]) > WARP_BOUNDARY_CONTRAST;
                if (mapping_edge && image_edge) {
                    boundary[index] = 1;
                    boundary[ni] = 1;
                }
            }
        }
    for (size_t index = 0; index < n; index++) {
        if (!boundary[index]) continue;
        float g[2], p[2];
        if (!warp_mapping(&source[index], g, p)) continue;
        chd_node_warp b;
        memset(&b, 0, sizeof b);
        b.code = CHD_WARP_BOUNDARY;
        b.global[0] = g[0];
        b.global[1] = g[1];
        decisions[index] = b;
    }
    free(source);
    free(boundary);
}

static void scale_color(const uint16_t in[3], double scale, uint16_t out[3]) {
    for (int c = 0; c < 3; c++) out[c] = f64_u16(round((double)in[c] * scale)); /* depth.rs:1857-1859 */
}

void or_diagnostic_samples(const chd_depth_map *map, uint16_t *inverse_depth,
                           uint16_t *provenance) { /* depth.rs:200-222 */
    double near_inverse = 1.0 / map->near_depth, far_inverse = 1.0 / map->far_depth;
    double range = rmax(near_inverse - far_inverse, 2.2250738585072014e-308);
    size_t n = (size_t)map->columns * map->rows;
    for (size_t i = 0; i < n; i++) {
        double depth = map->depth[i];
        uint16_t encoded = 0;
        if (!(depth != depth)) {
            double normalized = rclamp((1.0 / depth - far_inverse) / range, 0.0, 1.0);
            encoded = (uint16_t)(1 + f64_u16(round(normalized * 65534.0)));
        }
        inverse_depth[i] = encoded;
        double confidence = (double)(0.35f + 0.65f * rclampf(map->confidence[i], 0.0f, 1.0f));
        static const uint16_t green[3] = {0, 65535, 0}, amber[3] = {65535, 32768, 0};
        uint16_t *c = provenance + 3 * i;
        switch (map->provenance[i]) {
        case CHD_PROVENANCE_UNSUPPORTED: c[0] = c[1] = c[2] = 0; break;
        case CHD_PROVENANCE_GLOBAL: c[0] = 0; c[1] = 0; c[2] = 32768; break;
        case CHD_PROVENANCE_MEASURED: scale_color(green, confidence, c); break;
        default: scale_color(amber, confidence, c); break;
        }
    }
}

static void depth_color(double normalized, uint16_t out[3]) { /* depth.rs:1861-1877 */
    static const double stops[5][3] = {
        {0.0, 0.1, 0.7}, {0.0, 0.9, 1.0}, {0.0, 1.0, 0.2}, {1.0, 0.9, 0.0}, {1.0, 0.0, 0.0}};
    double position = rclamp(normalized, 0.0, 1.0) * 4.0;
    size_t lower = f64_usize(floor(position));
    if (lower > 4) lower = 4;
    size_t upper = lower + 1 < 4 ? lower + 1 : 4;
    double fraction = position - (double)lower;
    for (int c = 0; c < 3; c++)
        out[c] = f64_u16(round((stops[lower][c] * (1.0 - fraction) + stops[upper][c] * fraction) * 65535.0));
}

void or_luminance_half_u16(const uint16_t *samples, uint32_t width, uint32_t height,
                           float black_q6, float white_q6, int reserve_headroom, float *plane) {
    if (reserve_headroom) { /* image.rs:69-75; u16 `+ 1` wraps in a release build */
        black_q6 *= 0.5f;
        white_q6 *= 0.5f;
    }
    uint32_t w = width / 2, h = height / 2; /* image.rs:401-426 */
    float range = rmaxf(white_q6 - black_q6, 1.0f);
    for (uint32_t j = 0; j < h; j++) {
        const uint16_t *row0 = samples + (size_t)(2 * j) * width;
        const uint16_t *row1 = samples + (size_t)(2 * j + 1) * width;
        for (uint32_t i = 0; i < w; i++) {
            uint16_t v[4] = {row0[2 * i], row0[2 * i + 1], row1[2 * i], row1[2 * i + 1]};
            if (reserve_headroom)
                for (int k = 0; k < 4; k++) v[k] = (uint16_t)((uint16_t)(v[k] + 1) / 2);
            float sum = (float)v[0] + (float)v[1] + (float)v[2] + (float)v[3];
            float linear = rmaxf((sum * 0.25f - black_q6) / range, 0.0f) * 1.0f;
            plane[(size_t)j * w + i] = logf(1.0f + 1000.0f * linear);
        }
    }
}

void or_visualization_samples(const chd_depth_map *map, uint16_t *rgb) { /* depth.rs:224-245 */
    double log_range = log(map->far_depth / map->near_depth);
    size_t n = (size_t)map->columns * map->rows;
    for (size_t i = 0; i < n; i++) {
        uint16_t *c = rgb + 3 * i;
        double depth = map->depth[i];
        if (!(depth != depth)) {
            double normalized = rclamp((log(map->far_depth) - log(depth)) / log_range, 0.0, 1.0);
            double confidence = 0.45 + 0.55 * (double)rclampf(map->confidence[i], 0.0f, 1.0f);
            uint16_t col[3];
            depth_color(normalized, col);
            scale_color(col, confidence, c);
        } else if (map->provenance[i] == CHD_PROVENANCE_GLOBAL) {
            c[0] = 0; c[1] = 0; c[2] = 12000;
        } else {
            c[0] = c[1] = c[2] = 0;
        }
    }
}

/* ------------------------------------------------ measure_direct_depths */

typedef struct {
    const ctx_t *c;
    const chd_node_depth *coarse;
    uint32_t coarse_columns, coarse_rows, coarse_step;
    uint32_t columns, step, row_begin, row_end;
    double inverse_step;
    const chd_options *direct, *wide;
    chd_node_depth *nodes;
    float *guidance;
    uint8_t *tested;
} direct_job;

static void *direct_worker(void *arg) {
    /* the per-node body of measure_direct_depths (depth.rs:752-849) */
    direct_job *j = (direct_job *)arg;
    for (uint32_t row = j->row_begin; row < j->row_end; row++)
        for (uint32_t column = 0; column < j->columns; column++) {
            size_t node = (size_t)row * j->columns + column;
            double p[2] = {(double)((size_t)column * j->step), (double)((size_t)row * j->step)};
            j->guidance[node] = reference_guidance(&j->c->views[j->c->ref], p);
            chd_node_depth seed = nearest_coarse_depth(j->coarse, j->coarse_columns,
                                                       j->coarse_rows, j->coarse_step, p);
            double candidates[32];
            uint32_t nc = or_direct_depth_candidates(seed.valid, seed.depth, j->inverse_step,
                                                     j->direct, candidates);
            float baseline;
            int has_baseline = aggregate_views(j->c, p, 0, 0.0, j->direct, &baseline);
            float scores[32];
            uint8_t has[32];
            int any = 0;
            for (uint32_t k = 0; k < nc; k++) {
                has[k] = (uint8_t)aggregate_views(j->c, p, 1, candidates[k], j->direct, &scores[k]);
                any |= has[k];
            }
            int node_tested = has_baseline || any;
            chd_node_depth selected;
            int ok = or_select_direct_depth(candidates, scores, has, nc, has_baseline, baseline,
                                            seed.valid, j->direct, &selected);
            if (!ok) {
                int best = or_best_depth_candidate(candidates, scores, has, nc, j->direct);
                if (best >= 0 && has[best] && scores[best] >= MINIMUM_REGULARIZED_SCORE) {
                    uint32_t first = best >= 3 ? (uint32_t)best - 3 : 0;
                    uint32_t last = (uint32_t)best + 3 < nc - 1 ? (uint32_t)best + 3 : nc - 1;
                    float wb;
                    int has_wb = aggregate_views(j->c, p, 0, 0.0, j->wide, &wb);
                    float ws[8];
                    uint8_t wh[8];
                    for (uint32_t k = first; k <= last; k++)
                        wh[k - first] = (uint8_t)aggregate_views(j->c, p, 1, candidates[k], j->wide,
                                                                 &ws[k - first]);
                    ok = or_select_direct_depth(candidates + first, ws, wh, last - first + 1, has_wb,
                                                wb, seed.valid, j->wide, &selected);
                }
            }
            if (!ok) memset(&selected, 0, sizeof selected);
            j->nodes[node] = selected;
            j->tested[node] = (uint8_t)node_tested;
        }
    return NULL;
}

/* ------------------------------------------------- refine_one_view */

typedef struct {
    float score;
    float point[2];
} view_refinement;

static int score_one_view_at_point(const chd_view *reference, const chd_view *target,
                                   const chd_warp *global, const double centre[2],
                                   const float global_point[2], const float point[2],
                                   uint32_t radius, float *out) { /* depth.rs:1710-1736 */
    double delta[2] = {(double)(point[0] - global_point[0]), (double)(point[1] - global_point[1])};
    if (fabs(delta[0]) > 128.0 || fabs(delta[1]) > 128.0) return 0;
    return or_warped_patch_zncc(reference->luminance, reference->plane_width,
                                reference->plane_height, target->luminance,
                                target->plane_width, target->plane_height, global, centre[0],
                                centre[1], delta[0], delta[1], radius, out);
}

static int refine_one_view(const chd_view *reference, const chd_view *target,
                           const chd_warp *global, const double centre[2], double shared_depth,
                           double inverse_step, const chd_options *o,
                           view_refinement *out) { /* depth.rs:1634-1708 */
    if (!reference->camera || !target->camera) return 0;
    float global_point[2];
    if (!or_warp_map(global, (float)centre[0], (float)centre[1], global_point)) return 0;
    double min_inverse = 1.0 / o->far_depth, max_inverse = 1.0 / o->near_depth;
    double shared_inverse = 1.0 / shared_depth;
    static const double offsets[5] = {-1.0, -0.5, 0.0, 0.5, 1.0};
    int have = 0;
    float best_objective = 0.0f, best_score = 0.0f, best_point[2] = {0, 0};
    for (int i = 0; i < 5; i++) {
        double inverse = rclamp(shared_inverse + offsets[i] * inverse_step, min_inverse, max_inverse);
        double q[2];
        if (!map_at_depth(reference->camera, target->camera, global, centre, 1.0 / inverse, q))
            return 0;
        float point[2] = {(float)q[0], (float)q[1]};
        float score;
        if (!score_one_view_at_point(reference, target, global, centre, global_point, point,
                                     o->patch_radius, &score))
            return 0;
        float objective = score - (float)fabs(offsets[i]) * 0.002f;
        if (!have || objective > best_objective) {
            have = 1;
            best_objective = objective;
            best_score = score;
            best_point[0] = point[0];
            best_point[1] = point[1];
        }
    }
    float depth_score = best_score, depth_point[2] = {best_point[0], best_point[1]};
    float b_obj = depth_score, b_score = depth_score, b_point[2] = {depth_point[0], depth_point[1]};
    static const float steps[3] = {-1.5f, 0.0f, 1.5f};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            float dy = steps[a], dx = steps[b];
            if (dx == 0.0f && dy == 0.0f) continue;
            float point[2] = {depth_point[0] + dx, depth_point[1] + dy};
            float score;
            if (!score_one_view_at_point(reference, target, global, centre, global_point, point,
                                         o->patch_radius, &score))
                return 0;
            float residual_sq = dx * dx + dy * dy;
            float objective = score - residual_sq * 0.002f;
            if (objective > b_obj) {
                b_obj = objective;
                b_score = score;
                b_point[0] = point[0];
                b_point[1] = point[1];
            }
        }
    out->score = b_score;
    out->point[0] = b_point[0];
    out->point[1] = b_point[1];
    return 1;
}

/* ------------------------------------------ refine_multiview_depth */

static int cmp_f64_total(const void *a, const void *b) {
    return tcmp64(*(const double *)a, *(const double *)b);
}
static int cmp_f32_total(const void *a, const void *b) {
    return tcmp32(*(const float *)a, *(const float *)b);
}
static float fraction(uint64_t count, uint64_t total) {
    return total == 0 ? 0.0f : (float)count / (float)total;
}

int or_direct_rows(const chd_view *views, uint32_t count, uint32_t reference_index,
                   const chd_options *o, const chd_node_depth *coarse, uint32_t row_begin,
                   uint32_t row_end, uint32_t threads, chd_node_depth *field, float *guidance,
                   uint8_t *tested) {
    /* measure_direct_depths (depth.rs:715-876) over fine rows
     * [row_begin, row_end), row-banded over `threads` workers as the
     * reference bands all rows (depth.rs:729-731) */
    ctx_t c = {views, count, reference_index, o};
    const chd_view *reference = &views[reference_index];
    uint32_t coarse_step = o->grid_step > 4 ? o->grid_step : 4;
    uint32_t coarse_columns = (uint32_t)div_ceil(reference->width, coarse_step) + 1;
    uint32_t coarse_rows = (uint32_t)div_ceil(reference->height, coarse_step) + 1;
    uint32_t step = coarse_step / 2 > 4 ? coarse_step / 2 : 4;
    uint32_t columns = (uint32_t)div_ceil(reference->width, step) + 1;
    chd_options direct_options = *o, wide_options = *o;
    direct_options.patch_radius = o->patch_radius > 3 ? o->patch_radius : 3;
    wide_options.patch_radius = o->patch_radius * 2 > 8 ? o->patch_radius * 2 : 8;
    double inverse_step = (1.0 / o->near_depth - 1.0 / o->far_depth) / (double)(o->planes - 1);
    if (row_end <= row_begin) return 0;
    uint32_t workers = worker_count(threads, row_end - row_begin);
    uint32_t per = (uint32_t)div_ceil(row_end - row_begin, workers);
    direct_job jobs[256];
    uint32_t n = 0;
    for (uint32_t first = row_begin; first < row_end && n < 256; first += per) {
        direct_job *j = &jobs[n++];
        j->c = &c;
        j->coarse = coarse;
        j->coarse_columns = coarse_columns;
        j->coarse_rows = coarse_rows;
        j->coarse_step = coarse_step;
        j->columns = columns;
        j->step = step;
        j->row_begin = first;
        j->row_end = first + per < row_end ? first + per : row_end;
        j->inverse_step = inverse_step;
        j->direct = &direct_options;
        j->wide = &wide_options;
        j->nodes = field;
        j->guidance = guidance;
        j->tested = tested;
    }
    run_bands(direct_worker, jobs, sizeof(direct_job), n);
    return 0;
}

int or_refine_multiview_depth(const chd_view *views, uint32_t count,
                              uint32_t reference_index, const chd_options *o,
                              chd_depth_map *map, chd_target_result *targets,
                              or_debug *debug, uint32_t threads, int stop_after_sgm) {
    /* depth.rs:312-580 */
    if (!or_options_valid(o) || reference_index >= count || count > CHD_MAX_VIEWS ||
        views[reference_index].camera == NULL)
        return CHD_NONE;
    const chd_view *reference = &views[reference_index];
    uint32_t coarse_step = o->grid_step > 4 ? o->grid_step : 4;
    uint32_t coarse_columns = (uint32_t)div_ceil(reference->width, coarse_step) + 1;
    uint32_t coarse_rows = (uint32_t)div_ceil(reference->height, coarse_step) + 1;
    size_t coarse_nodes = (size_t)coarse_columns * coarse_rows;
    uint32_t L;
    double *labels = label_table(o, &L);
    double t0 = now_seconds();
    float *scores = (float *)malloc(coarse_nodes * L * sizeof(float));
    float *costs = (float *)malloc(coarse_nodes * L * sizeof(float));
    float *guidance = (float *)malloc(coarse_nodes * sizeof(float));
    uint8_t *tested = (uint8_t *)malloc(coarse_nodes);
    or_cost_volume_rows(views, count, reference_index, o, 0, coarse_rows, threads, scores, costs,
                        guidance, tested);
    double t1 = now_seconds();
    float *regularised = (float *)malloc(coarse_nodes * L * sizeof(float));
    or_semi_global_costs(costs, guidance, coarse_columns, coarse_rows, L, regularised);
    chd_node_depth *coarse = (chd_node_depth *)malloc(coarse_nodes * sizeof(chd_node_depth));
    uint8_t *fillable = (uint8_t *)malloc(coarse_nodes);
    uint16_t *best = (uint16_t *)malloc(coarse_nodes * sizeof(uint16_t));
    or_select_depths(labels, L, scores, tested, regularised, (uint32_t)coarse_nodes, o, coarse,
                     fillable, best);
    double inverse_step = (1.0 / o->near_depth - 1.0 / o->far_depth) / (double)(o->planes - 1);
    if (o->completion_iterations > 0)
        or_complete_depth_field(coarse, guidance, fillable, coarse_columns, coarse_rows, o);
    double t2 = now_seconds();
    if (debug) {
        debug->seconds[0] = t1 - t0;
        debug->seconds[1] = t2 - t1;
        size_t vol = coarse_nodes * L * sizeof(float);
        if (debug->scores) memcpy(debug->scores, scores, vol);
        if (debug->costs) memcpy(debug->costs, costs, vol);
        if (debug->guidance) memcpy(debug->guidance, guidance, coarse_nodes * sizeof(float));
        if (debug->tested) memcpy(debug->tested, tested, coarse_nodes);
        if (debug->aggregated) memcpy(debug->aggregated, regularised, vol);
        if (debug->coarse) memcpy(debug->coarse, coarse, coarse_nodes * sizeof(chd_node_depth));
        if (debug->fillable) memcpy(debug->fillable, fillable, coarse_nodes);
        if (debug->best_label) memcpy(debug->best_label, best, coarse_nodes * sizeof(uint16_t));
    }
    free(regularised);
    free(scores);
    free(costs);
    free(tested);
    free(fillable);
    free(best);
    free(guidance);
    free(labels);
    if (stop_after_sgm) {
        free(coarse);
        return CHD_OK;
    }

    /* measure_direct_depths (depth.rs:715-876) */
    uint32_t step = coarse_step / 2 > 4 ? coarse_step / 2 : 4;
    uint32_t columns = (uint32_t)div_ceil(reference->width, step) + 1;
    uint32_t rows = (uint32_t)div_ceil(reference->height, step) + 1;
    size_t fine = (size_t)columns * rows;
    chd_node_depth *field = (chd_node_depth *)malloc(fine * sizeof(chd_node_depth));
    float *fguidance = (float *)malloc(fine * sizeof(float));
    uint8_t *ftested = (uint8_t *)malloc(fine);
    or_direct_rows(views, count, reference_index, o, coarse, 0, rows, threads, field, fguidance,
                   ftested);
    free(coarse);
    double t3 = now_seconds();
    if (debug && debug->direct[0]) memcpy(debug->direct[0], field, fine * sizeof(chd_node_depth));
    or_reject_isolated_direct_depths(field, fguidance, columns, rows, inverse_step * 2.5);
    if (debug && debug->direct[1]) memcpy(debug->direct[1], field, fine * sizeof(chd_node_depth));
    or_reject_small_direct_components(field, fguidance, columns, rows, inverse_step * 2.5,
                                      MINIMUM_DIRECT_COMPONENT_NODES);
    if (debug && debug->direct[2]) memcpy(debug->direct[2], field, fine * sizeof(chd_node_depth));
    or_fit_local_depth_planes(field, fguidance, columns, rows, inverse_step * 4.0);
    if (debug && debug->direct[3]) memcpy(debug->direct[3], field, fine * sizeof(chd_node_depth));
    if (debug && debug->fine_guidance) memcpy(debug->fine_guidance, fguidance, fine * sizeof(float));
    if (debug && debug->fine_tested) memcpy(debug->fine_tested, ftested, fine);
    double t4 = now_seconds();

    uint64_t tested_nodes = 0, measured_nodes = 0, regularized_nodes = 0;
    double *selected_depths = (double *)malloc((fine + 1) * sizeof(double));
    float *improvements = (float *)malloc((fine + 1) * sizeof(float));
    size_t nsel = 0, nimp = 0;
    for (size_t i = 0; i < fine; i++) {
        tested_nodes += ftested[i] ? 1 : 0;
        if (!field[i].valid) continue;
        if (field[i].regularized) regularized_nodes++;
        else {
            measured_nodes++;
            improvements[nimp++] = field[i].improvement;
        }
        selected_depths[nsel++] = field[i].depth;
    }
    uint64_t fallback_nodes = tested_nodes > measured_nodes + regularized_nodes
                                  ? tested_nodes - (measured_nodes + regularized_nodes)
                                  : 0;
    qsort(selected_depths, nsel, sizeof(double), cmp_f64_total);
    qsort(improvements, nimp, sizeof(float), cmp_f32_total);

    chd_node_warp *decisions = (chd_node_warp *)malloc(fine * sizeof(chd_node_warp));
    for (uint32_t t = 0; t < count; t++) {
        if (targets) targets[t].replaced = 0;
        if (t == reference_index || !views[t].accepted || views[t].camera == NULL) continue;
        const chd_warp *global = &views[t].global_warp;
        for (uint32_t row = 0; row < rows; row++)
            for (uint32_t column = 0; column < columns; column++) {
                size_t index = (size_t)row * columns + column;
                chd_node_warp *d = &decisions[index];
                memset(d, 0, sizeof *d);
                double p[2] = {(double)((size_t)column * step), (double)((size_t)row * step)};
                float gq[2];
                if (!or_warp_map(global, (float)p[0], (float)p[1], gq)) {
                    d->code = CHD_WARP_UNDEFINED;
                    continue;
                }
                d->global[0] = gq[0];
                d->global[1] = gq[1];
                if (!field[index].valid) {
                    d->code = CHD_WARP_GLOBAL;
                    continue;
                }
                view_refinement sel;
                int has_sel = refine_one_view(reference, &views[t], global, p, field[index].depth,
                                              inverse_step, o, &sel);
                float base;
                int has_base = score_one_view(reference, &views[t], global, p, 0, 0.0,
                                              o->patch_radius, &base);
                int supported = has_sel && sel.score >= o->minimum_score &&
                                (!has_base || sel.score + 0.01f >= base);
                int contradicted = has_sel && has_base && sel.score + 0.12f < base;
                if (contradicted) {
                    d->code = CHD_WARP_OCCLUDED;
                    d->measured = !field[index].regularized;
                } else if (supported) {
                    d->code = CHD_WARP_REFINED;
                    d->point[0] = sel.point[0];
                    d->point[1] = sel.point[1];
                    d->confidence = field[index].regularized ? field[index].confidence * 0.6f
                                                             : 0.5f + 0.5f * field[index].confidence;
                    d->measured = !field[index].regularized;
                } else {
                    d->code = CHD_WARP_GLOBAL;
                }
            }
        or_enforce_warp_consensus(decisions, columns, rows);
        or_suppress_warp_boundaries(decisions, fguidance, columns, rows);
        if (!targets) continue;
        chd_target_result *out = &targets[t];
        uint64_t refined = 0, occluded = 0, boundary = 0;
        for (size_t i = 0; i < fine; i++) {
            const chd_node_warp *d = &decisions[i];
            float px = NAN, py = NAN, conf = 0.0f;
            switch (d->code) {
            case CHD_WARP_GLOBAL: px = d->global[0]; py = d->global[1]; conf = 1.0f; break;
            case CHD_WARP_REFINED: px = d->point[0]; py = d->point[1]; conf = d->confidence; refined++; break;
            case CHD_WARP_OCCLUDED: px = d->global[0]; py = d->global[1]; occluded++; break;
            case CHD_WARP_BOUNDARY: px = d->global[0]; py = d->global[1]; boundary++; break;
            default: break;
            }
            if (out->points) {
                out->points[2 * i] = px;
                out->points[2 * i + 1] = py;
            }
            if (out->confidence) out->confidence[i] = conf;
            if (out->decisions) out->decisions[i] = (uint8_t)d->code;
        }
        out->replaced = 1;
        chd_depth_report *r = &out->report;
        memset(r, 0, sizeof *r);
        r->tested_nodes = tested_nodes;
        r->measured_nodes = measured_nodes;
        r->regularized_nodes = regularized_nodes;
        r->fallback_nodes = fallback_nodes;
        r->refined_nodes = refined;
        r->occluded_nodes = occluded;
        r->boundary_nodes = boundary;
        r->reconstructed_fraction = fraction(measured_nodes + regularized_nodes, tested_nodes);
        r->refined_fraction = fraction(refined, tested_nodes);
        r->occluded_fraction = fraction(occluded, tested_nodes);
        if (nsel > 0) {
            r->has_median_depth = 1;
            r->median_depth = selected_depths[nsel / 2];
        }
        if (nimp > 0) {
            r->has_median_score_improvement = 1;
            r->median_score_improvement = improvements[nimp / 2];
        }
    }
    free(decisions);
    free(selected_depths);
    free(improvements);
    double t5 = now_seconds();
    if (debug) {
        debug->seconds[2] = t3 - t2;
        debug->seconds[3] = t4 - t3;
        debug->seconds[4] = t5 - t4;
    }
    if (map) {
        map->columns = columns;
        map->rows = rows;
        map->step = step;
        map->near_depth = o->near_depth;
        map->far_depth = o->far_depth;
        for (size_t i = 0; i < fine; i++) {
            if (field[i].valid) {
                map->depth[i] = field[i].depth;
                map->confidence[i] = field[i].confidence;
                map->provenance[i] =
                    field[i].regularized ? CHD_PROVENANCE_REGULARIZED : CHD_PROVENANCE_MEASURED;
            } else if (ftested[i]) {
                map->depth[i] = NAN;
                map->confidence[i] = 1.0f;
                map->provenance[i] = CHD_PROVENANCE_GLOBAL;
            } else {
                map->depth[i] = NAN;
                map->confidence[i] = 0.0f;
                map->provenance[i] = CHD_PROVENANCE_UNSUPPORTED;
            }
        }
    }
    free(field);
    free(fguidance);
    free(ftested);
    return CHD_OK;
}

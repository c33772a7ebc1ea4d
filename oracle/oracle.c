/*
 * GPUropolis CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Serial C restatement of the reference hot path, main.cu of
 * piantado/GPUropolis (cited file:line), and of the postfix/config-mode
 * algorithm the HIP library implements.  Built with -ffp-contract=off so
 * every float expression rounds once per IEEE operation.
 *
 * Pins (tests/test_oracle_*.py):
 *   Philox4x32-10        Random123 known-answer vectors + torch's PhiloxRNGEngine.h
 *   LCG                  published BSD rand() sequence (seed 0)
 *   gpo_logf/expf/sin2pi libm within a few ulp; special cases
 *   gpo_normal_icdf      scipy.stats.norm.ppf
 *   heap fold / decode   brute force against a recursive evaluation; decode == fold
 *   length / count       closed forms (#tokens + #C-ops, #tokens - #x)
 *   likelihood           NIST certified residual sums of squares
 *   samplers             exact enumeration of proposals on tiny trees, MH on
 *                        targets with closed-form posteriors
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ literals */
#define LOGPI_F   0x1.250d06p+0f        /* logf((float)3.141592653589), main.cu:194,304 */
#define LOG2PI_F  0x1.d67f1ep+0f        /* logf((float)(2*PIf)), main.cu:300 */
#define LOG01_F   -0x1.26bb1cp+1f       /* logf(PROPOSE_GEOM_R = 0.1f), main.cu:34,308 */
#define LOG015_F  -0x1.e5a9a8p+0f       /* logf(PROPOSE_INT = 0.15f), main.cu:32,761 */
#define LOG1M01_D -0x1.af8e82826b325p-4 /* log(1.0 - 0.1f), main.cu:308 */
#define LOG1M015_D -0x1.4cd6ba6a55088p-3/* log(1.0 - 0.15f), main.cu:764 */
#define LOG2_D    0x1.62e42fefa39efp-1
#define LOG28_D   0x1.aa85ad6964a23p+1  /* log(NOPS+1), main.cu:534 */
#define LOGPI_D   0x1.250d048e7a1bdp+0
#define HALF_LOG2PI_D 0x1.d67f1c864beb4p-1

static const float POW10_F[8] = { 0x1.a36e2ep-14f, 0x1.0624dep-10f, 0x1.47ae14p-7f, 0x1.99999ap-4f,
                                  1.0f, 10.0f, 100.0f, 1000.0f };

/* ------------------------------------------------------------------ Philox4x32-10
 * Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11.  Multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key bumps 0x9E3779B9 /
 * 0xBB67AE85, ten rounds. */
void gpo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void gpo_stream_init(gpo_stream *s, uint64_t seed, uint32_t chain, uint32_t step,
                     uint32_t purpose, uint32_t slot) {
    s->key[0] = (uint32_t)seed; s->key[1] = (uint32_t)(seed >> 32);
    s->ctr[0] = chain; s->ctr[1] = step; s->ctr[2] = (purpose << 24) | (slot & 0xffffffu); s->ctr[3] = 0;
    s->idx = 4;
}

uint32_t gpo_stream_next(gpo_stream *s) {
    if (s->idx == 4) { gpo_philox4x32_10(s->ctr, s->key, s->buf); s->ctr[3]++; s->idx = 0; }
    return s->buf[s->idx++];
}

/* main.cu:213-216: rx = (rx*1103515245 + 12345) & (2^31-1), restated on uint32 */
uint32_t gpo_lcg_next(uint32_t *rx) {
    *rx = (*rx * 1103515245u + 12345u) & 0x7fffffffu;
    return *rx;
}

/* ------------------------------------------------------------------ spec math */
static float bits_f(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t f_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/* natural log: m*2^e with m in (sqrt(1/2), sqrt(2)], log m = 2 atanh((m-1)/(m+1));
 * fused multiply-adds are explicit fmaf, one rounding each, as on the GPU */
float gpo_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    uint32_t b = f_bits(x);
    int e = (int)((b >> 23) & 0xffu);
    if (e == 0) { x = x * 0x1p23f; b = f_bits(x); e = (int)((b >> 23) & 0xffu) - 23; }
    e -= 127;
    float m = bits_f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = fmaf(s2, 0x1.745d18p-4f, 0x1.c71c72p-4f);
    p = fmaf(s2, p, 0x1.24924ap-3f);
    p = fmaf(s2, p, 0x1.99999ap-3f);
    p = fmaf(s2, p, 0x1.555556p-2f);
    p = p * s2;
    float lm = fmaf(s, p, s) * 2.0f;
    float ef = (float)e;
    return fmaf(ef, 0x1.62e4p-1f, fmaf(ef, 0x1.7f7d1cp-20f, lm));
}

/* exp: 2^k * e^r, |r| <= ln2/2, Cody-Waite split of ln2 (k*LN2_HI exact) */
static float pow2i(int k) { return bits_f((uint32_t)(k + 127) << 23); }
float gpo_expf(float x) {
    if (x != x) return x;
    if (x > 0x1.62e42ep+6f) return INFINITY;
    if (x < -0x1.9fe368p+6f) return 0.0f;
    float kf = floorf(fmaf(x, 0x1.715476p+0f, 0.5f));
    float r = fmaf(-kf, 0x1.7f7d1cp-20f, fmaf(-kf, 0x1.62e4p-1f, x));
    float p = fmaf(r, 0x1.a01a02p-13f, 0x1.6c16c2p-10f);
    p = fmaf(r, p, 0x1.111112p-7f);
    p = fmaf(r, p, 0x1.555556p-5f);
    p = fmaf(r, p, 0x1.555556p-3f);
    p = fmaf(r, p, 0.5f);
    p = fmaf(r, p, 1.0f);
    p = fmaf(r, p, 1.0f);
    int k = (int)kf;
    int k1 = k / 2, k2 = k - k1;
    return (p * pow2i(k1)) * pow2i(k2);
}

/* sin(2*pi*v): quarter-turn reduction, Taylor polynomials of sin/cos(pi/2 r), |r| <= 1/2 */
float gpo_sin2pi(float v) {
    if (!isfinite(v)) return NAN;
    float v4 = v * 4.0f;
    float q = floorf(v4 + 0.5f);
    float r = v4 - q;
    int qi = ((int)fmodf(q, 4.0f) + 4) & 3;
    float r2 = r * r;
    float sn = fmaf(r2, -0x1.e3075p-19f, 0x1.507834p-13f);
    sn = fmaf(r2, sn, -0x1.32d2ccp-8f);
    sn = fmaf(r2, sn, 0x1.466bc6p-4f);
    sn = fmaf(r2, sn, -0x1.4abbcep-1f);
    sn = fmaf(r2, sn, 0x1.921fb6p+0f);
    sn = sn * r;
    float cs = fmaf(r2, 0x1.f9d38ap-22f, -0x1.a6d1f2p-16f);
    cs = fmaf(r2, cs, 0x1.e1f506p-11f);
    cs = fmaf(r2, cs, -0x1.55d3c8p-6f);
    cs = fmaf(r2, cs, 0x1.03c1fp-2f);
    cs = fmaf(r2, cs, -0x1.3bd3ccp+0f);
    cs = fmaf(r2, cs, 1.0f);
    switch (qi) { case 0: return sn; case 1: return cs; case 2: return -sn; default: return -cs; }
}

/* Standard normal quantile sqrt(2)*erfinv(2u-1), with erfinv from M. Giles,
 * "Approximating the erfinv function" (GPU Computing Gems, 2011), single-precision
 * polynomials.  Odd-symmetric in 2u-1, so u and 1-u give exactly opposite draws. */
float gpo_normal_icdf(float u) {
    float x = fmaf(2.0f, u, -1.0f);
    float w = -gpo_logf((1.0f - x) * (1.0f + x));
    float p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = fmaf(p, w, 3.43273939e-07f);
        p = fmaf(p, w, -3.5233877e-06f);
        p = fmaf(p, w, -4.39150654e-06f);
        p = fmaf(p, w, 0.00021858087f);
        p = fmaf(p, w, -0.00125372503f);
        p = fmaf(p, w, -0.00417768164f);
        p = fmaf(p, w, 0.246640727f);
        p = fmaf(p, w, 1.50140941f);
    } else {
        w = sqrtf(w) - 3.0f;
        p = -0.000200214257f;
        p = fmaf(p, w, 0.000100950558f);
        p = fmaf(p, w, 0.00134934322f);
        p = fmaf(p, w, -0.00367342844f);
        p = fmaf(p, w, 0.00573950773f);
        p = fmaf(p, w, -0.0076224613f);
        p = fmaf(p, w, 0.00943887047f);
        p = fmaf(p, w, 1.00167406f);
        p = fmaf(p, w, 2.83297682f);
    }
    return (p * x) * 1.41421356f;
}

float gpo_u01(uint32_t r) { return ((float)(r >> 9) + 0.5f) * 0x1p-23f; }

/* ------------------------------------------------------------------ reference ops */
/* dispatch_eval, main.cu:397-444 (fdividef restated as IEEE division) */
static float ref_dispatch_eval(int o, float a, float b, float C) {
    switch (o) {
    case GPO_R_ONE: return 1.0f;
    case GPO_R_CONST: return C;
    case GPO_R_INV: return 1.0f / a;
    case GPO_R_A: return a;
    case GPO_R_B: return b;
    case GPO_R_PLUS: return a + b;
    case GPO_R_CPLUS: return a + C;
    case GPO_R_MINUS: return a - b;
    case GPO_R_RMINUS: return b - a;
    case GPO_R_TIMES: return a * b;
    case GPO_R_CTIMES: return C * a;
    case GPO_R_DIV: return a / b;
    case GPO_R_RDIV: return b / a;
    case GPO_R_SQRT: return sqrtf(a);
    case GPO_R_LOG: return logf(a);
    case GPO_R_SIN: return sinf(a);
    case GPO_R_ASIN: return asinf(a);
    case GPO_R_COS: return cosf(a);
    case GPO_R_ACOS: return acosf(a);
    case GPO_R_TAN: return tanf(a);
    case GPO_R_ATAN: return atanf(a);
    case GPO_R_EXP: return expf(a);
    case GPO_R_POW: return powf(a, b);
    case GPO_R_CPOW: return powf(a, C);
    case GPO_R_RPOW: return powf(b, a);
    case GPO_R_CRPOW: return powf(C, a);
    case GPO_R_ABS: return fabsf(a);
    /* PHYSICAL_CONSTANTS, main.cu:429-441: the data_t value of each double literal */
    case GPO_R_E: return (float)1.6021766208e-19;
    case GPO_R_PI: return (float)1.14472988584915;
    case GPO_R_CLIGHT: return (float)19.5186009865452;
    case GPO_R_G: return (float)-23.4302046558627;
    case GPO_R_HBAR: return (float)-78.2347583540371;
    case GPO_R_MU0: return (float)-13.5870714043359;
    case GPO_R_EL: return (float)-43.2777536741305;
    case GPO_R_MP: return (float)-61.6554051167679;
    case GPO_R_ME: return (float)-69.1708328401338;
    case GPO_R_E0: return (float)-25.4501305691713;
    case GPO_R_KB: return (float)-52.6369038080531;
    case GPO_R_LOGOF2: return (float)log(2.0);
    case GPO_R_HALF: return (float)((double)a / 2.0);
    default: return NAN;
    }
}
static int ref_is_physical(int o) { return o == GPO_R_LOGOF2 || (o >= GPO_R_E && o <= GPO_R_KB); }

/* op sets, main.cu:51-86: the heap code of each enum value in enum order; the
 * SIMPLIFIED, POLYNOMIAL and LINEAR sets are the intended ones (their #if blocks do not
 * compile as written: duplicate enumerators, a missing comma, a second enum OPS). */
static const uint8_t OPSET_SIMPLIFIED[13] = { GPO_R_ONE, GPO_R_CONST, GPO_R_INV, GPO_R_A, GPO_R_B, GPO_R_PLUS,
    GPO_R_MINUS, GPO_R_RMINUS, GPO_R_CPLUS, GPO_R_TIMES, GPO_R_CTIMES, GPO_R_DIV, GPO_R_RDIV };
static const uint8_t OPSET_POLYNOMIAL[11] = { GPO_R_ONE, GPO_R_CONST, GPO_R_A, GPO_R_B, GPO_R_PLUS, GPO_R_MINUS,
    GPO_R_RMINUS, GPO_R_CPLUS, GPO_R_TIMES, GPO_R_CTIMES, GPO_R_CPOW };
static const uint8_t OPSET_LINEAR[9] = { GPO_R_ONE, GPO_R_CONST, GPO_R_A, GPO_R_B, GPO_R_PLUS, GPO_R_MINUS,
    GPO_R_RMINUS, GPO_R_CPLUS, GPO_R_CTIMES };
static const uint8_t OPSET_PHYSICAL[19] = { GPO_R_LOGOF2, GPO_R_HALF, GPO_R_E, GPO_R_PI, GPO_R_CLIGHT, GPO_R_G,
    GPO_R_HBAR, GPO_R_MU0, GPO_R_EL, GPO_R_MP, GPO_R_ME, GPO_R_E0, GPO_R_KB, GPO_R_A, GPO_R_B, GPO_R_PLUS,
    GPO_R_MINUS, GPO_R_RMINUS, GPO_R_TIMES };
int gpo_opset_nops(int set) {
    switch (set) {
    case 1: return 13;
    case 2: return 11;
    case 3: return 9;
    case 4: return 19;
    default: return GPO_R_NOPS;
    }
}
int gpo_opset_code(int set, int i) {
    switch (set) {
    case 1: return OPSET_SIMPLIFIED[i];
    case 2: return OPSET_POLYNOMIAL[i];
    case 3: return OPSET_LINEAR[i];
    case 4: return OPSET_PHYSICAL[i];
    default: return i;
    }
}
const char *gpo_opset_program_code(int set) {     /* main.cu:55, 61, 67, 74, 85 */
    switch (set) {
    case 1: return "1CIab+-_#*@/|";
    case 2: return "1CAB+-_#*@^";
    case 3: return "1CAB+-_#@";
    case 4: return "B2epcghm:o.~Kab+-_*";
    default: return "1CIab+-_#*@/|LE^pVPRSsCcTtA";
    }
}

/* dispatch_length, main.cu:477-526 (CONST_LENGTH_PRIOR = 1) */
static float ref_dispatch_length(int o, float a, float b) {
    switch (o) {
    case GPO_R_ONE: case GPO_R_CONST: return 1.0f;
    case GPO_R_A: return a;
    case GPO_R_B: return b;
    case GPO_R_CPLUS: case GPO_R_CTIMES: case GPO_R_CPOW: case GPO_R_CRPOW: return 1.0f + a + 1.0f;
    case GPO_R_PLUS: case GPO_R_MINUS: case GPO_R_TIMES: case GPO_R_DIV: case GPO_R_RDIV:
    case GPO_R_POW: return 1.0f + a + b;
    case GPO_R_RMINUS: case GPO_R_RPOW: return 1.0f + b + a;
    case GPO_R_INV: case GPO_R_SQRT: case GPO_R_LOG: case GPO_R_SIN: case GPO_R_ASIN:
    case GPO_R_COS: case GPO_R_ACOS: case GPO_R_TAN: case GPO_R_ATAN: case GPO_R_EXP:
    case GPO_R_ABS: case GPO_R_HALF: return 1.0f + a;
    default: return ref_is_physical(o) ? 1.0f : NAN;       /* PHYSICAL_CONSTANT_LENGTH_PRIOR */
    }
}

/* dispatch_countnodes, main.cu:553-601 */
static int ref_dispatch_count(int o, int a, int b) {
    if (ref_is_physical(o)) return 1;
    switch (o) {
    case GPO_R_ONE: case GPO_R_CONST: return 1;
    case GPO_R_A: return a;
    case GPO_R_B: return b;
    case GPO_R_CPLUS: case GPO_R_CTIMES: case GPO_R_CPOW: case GPO_R_CRPOW: return a + 1;
    case GPO_R_PLUS: case GPO_R_MINUS: case GPO_R_RMINUS: case GPO_R_TIMES: case GPO_R_DIV:
    case GPO_R_RDIV: case GPO_R_POW: case GPO_R_RPOW: return 1 + a + b;
    default: return 1 + a;
    }
}

/* dispatch_degree, main.cu:615-651 (max macro of main.cu:184-187) */
static float fmax_ref(float a, float b) { return a > b ? a : b; }
static float ref_dispatch_degree(int o, float a, float b, float c) {
    switch (o) {
    case GPO_R_ONE: case GPO_R_CONST: return 0.0f;
    case GPO_R_A: return a;
    case GPO_R_B: return b;
    case GPO_R_PLUS: case GPO_R_MINUS: case GPO_R_RMINUS: return fmax_ref(a, b);
    case GPO_R_CPLUS: return a;
    case GPO_R_TIMES: return a + b;
    case GPO_R_CTIMES: return a;
    case GPO_R_CPOW: return a * c;
    case GPO_R_POW: return (b == 0.0f) ? a * c : INFINITY;
    case GPO_R_DIV: return (b == 0.0f) ? a : INFINITY;
    case GPO_R_HALF: return a;
    default: return ref_is_physical(o) ? 0.0f : INFINITY;
    }
}

/* program_fold, main.cu:371-393: bottom-up over heap slots 31..1, children 2i, 2i+1 */
float gpo_ref_fold_eval(const uint8_t *hop, const float *hc, size_t stride, float x) {
    float buf[GPO_HEAP + 1];
    for (int i = GPO_HEAP; i >= 1; i--) {
        int l = 2 * i, r = 2 * i + 1;
        float lv = l > GPO_HEAP ? x : buf[l];
        float rv = l > GPO_HEAP ? x : buf[r];
        buf[i] = ref_dispatch_eval(hop[(i - 1) * stride], lv, rv, hc[(i - 1) * stride]);
    }
    return buf[1];
}
float gpo_ref_fold_length(const uint8_t *hop, size_t stride) {
    float buf[GPO_HEAP + 1];
    for (int i = GPO_HEAP; i >= 1; i--) {
        int l = 2 * i;
        float lv = l > GPO_HEAP ? 1.0f : buf[l];       /* X_LENGTH_PRIOR, main.cu:24,532 */
        float rv = l > GPO_HEAP ? 1.0f : buf[l + 1];
        buf[i] = ref_dispatch_length(hop[(i - 1) * stride], lv, rv);
    }
    return buf[1];
}
int gpo_ref_fold_count(const uint8_t *hop, size_t stride) {
    int buf[GPO_HEAP + 1];
    for (int i = GPO_HEAP; i >= 1; i--) {
        int l = 2 * i;
        int lv = l > GPO_HEAP ? 0 : buf[l];            /* main.cu:607 */
        int rv = l > GPO_HEAP ? 0 : buf[l + 1];
        buf[i] = ref_dispatch_count(hop[(i - 1) * stride], lv, rv);
    }
    return buf[1];
}
float gpo_ref_fold_degree(const uint8_t *hop, const float *hc, size_t stride) {
    float buf[GPO_HEAP + 1];
    for (int i = GPO_HEAP; i >= 1; i--) {
        int l = 2 * i;
        float lv = l > GPO_HEAP ? 1.0f : buf[l];       /* main.cu:653 */
        float rv = l > GPO_HEAP ? 1.0f : buf[l + 1];
        buf[i] = ref_dispatch_degree(hop[(i - 1) * stride], lv, rv, hc[(i - 1) * stride]);
    }
    return buf[1];
}

/* is_descendant, main.cu:660-677 (0-based n, k) */
int gpo_ref_is_descendant(int n, int k) {
    if (k == n) return 1;
    n = n + 1; k = k + 1;
    while (k > n) { k >>= 1; if (k == n) return 1; }
    return 0;
}

/* heap -> postfix of the reachable tree.  Value-identical to the fold because the
 * fold's root depends only on reachable slots. */
static const int8_t REF_TO_TAPE[GPO_R_NCODES] = {
    GPO_ONE, GPO_CONST, GPO_INV, -1, -1, GPO_ADD, GPO_SUB, GPO_RSUB, GPO_CADD, GPO_MUL, GPO_CMUL,
    GPO_DIV, GPO_RDIV, GPO_LOG, GPO_EXP, GPO_POW, GPO_CPOW, GPO_RPOW, GPO_CRPOW, GPO_SQRT,
    GPO_SIN, GPO_ASIN, GPO_COS, GPO_ACOS, GPO_TAN, GPO_ATAN, GPO_ABS,
    /* log 2 and the physical constants are constant tokens of their value; halving is
     * c*a with c = 0.5 (a/2.0 in double, then float, equals 0.5f*a) */
    GPO_CONST, GPO_CMUL, GPO_CONST, GPO_CONST, GPO_CONST, GPO_CONST, GPO_CONST, GPO_CONST, GPO_CONST,
    GPO_CONST, GPO_CONST, GPO_CONST, GPO_CONST };

static int decode_rec(const uint8_t *hop, const float *hc, size_t stride, int i, int pos,
                      uint8_t *ops, float *cs, uint8_t *slot) {
    /* i is 1-based; i > 31 means an x leaf */
    if (i > GPO_HEAP) { ops[pos] = GPO_X; cs[pos] = 0.0f; slot[pos] = 255; return pos + 1; }
    int o = hop[(i - 1) * stride];
    if (o == GPO_R_A) return decode_rec(hop, hc, stride, 2 * i, pos, ops, cs, slot);
    if (o == GPO_R_B) return decode_rec(hop, hc, stride, 2 * i + 1, pos, ops, cs, slot);
    int t = REF_TO_TAPE[o];
    int ar = gpo_arity(t);
    if (ar >= 1) pos = decode_rec(hop, hc, stride, 2 * i, pos, ops, cs, slot);
    if (ar == 2) pos = decode_rec(hop, hc, stride, 2 * i + 1, pos, ops, cs, slot);
    ops[pos] = (uint8_t)t;
    cs[pos] = (t == GPO_CONST || t >= GPO_CADD) ? hc[(i - 1) * stride] : 0.0f;
    if (ref_is_physical(o)) cs[pos] = ref_dispatch_eval(o, 0.0f, 0.0f, 0.0f);
    if (o == GPO_R_HALF) cs[pos] = 0.5f;
    slot[pos] = (uint8_t)(i - 1);
    return pos + 1;
}
int gpo_ref_decode(const uint8_t *hop, const float *hc, size_t stride,
                   uint8_t *ops, float *cs, uint8_t *slot) {
    return decode_rec(hop, hc, stride, 1, 0, ops, cs, slot);
}

/* lnormalpdf, lcauchypdf, lgeometricpdf: main.cu:299-309, with the mixed
 * float/double arithmetic of the reference (float logs, double combination) */
double gpo_ref_lnormalpdf(float x, float s) {
    return -(double)(x * x) / (2.0 * (double)s * (double)s) - 0.5 * ((double)LOG2PI_F + 2.0 * (double)logf(s));
}
double gpo_ref_lcauchypdf(float x, float s) {
    return -(double)LOGPI_F - (double)gpo_logf(s) - (double)gpo_logf((float)(1.0 + (double)(x * x / s)));
}
double gpo_ref_lgeometricpdf(float n, float r) {
    (void)r; /* only r = PROPOSE_GEOM_R is used (main.cu:761,770) */
    return (double)((n - 1.0f) * LOG01_F) + LOG1M01_D;
}

/* compute_likelihood, main.cu:454-468 */
double gpo_ref_likelihood(const uint8_t *hop, const float *hc, size_t stride,
                          const float *x, const float *y, const float *sd, int n) {
    double ll = 0.0;
    for (int i = 0; i < n; i++) {
        float fx = gpo_ref_fold_eval(hop, hc, stride, x[i]);
        if (!isfinite(fx)) return -INFINITY;
        double l = gpo_ref_lnormalpdf(fx - y[i], sd[i]);
        if (!isfinite(l)) return -INFINITY;
        ll += l;
    }
    return ll;
}

/* compute_prior, main.cu:529-544 (PRIOR_MULTIPLIER = CONSTANT_SCALE = 1) */
double gpo_ref_prior_ops(const uint8_t *hop, const float *hc, size_t stride, int set) {
    float length = gpo_ref_fold_length(hop, stride);
    /* log((bayes_t)NOPS+1), main.cu:536; the default set's is the literal LOG28_D */
    const double lnops = set == 0 ? LOG28_D : log((double)gpo_opset_nops(set) + 1.0);
    double prior = -(double)(1.0f * length) * lnops;
    for (int c = 0; c < GPO_HEAP; c++) prior += gpo_ref_lcauchypdf(hc[c * stride], 1.0f);
    return prior;
}
double gpo_ref_prior(const uint8_t *hop, const float *hc, size_t stride) {
    return gpo_ref_prior_ops(hop, hc, stride, 0);
}

/* ------------------------------------------------------------------ postfix tapes */
int gpo_arity(int op) {
    if (op <= GPO_ONE) return 0;
    if (op <= GPO_RPOW) return 2;
    return 1;
}
int gpo_dl_cost(int op) { return op >= GPO_CADD ? 2 : 1; }

int gpo_tape_depth(const uint8_t *ops, int len) {
    int d = 0, mx = 0;
    if (len <= 0) return -1;
    for (int t = 0; t < len; t++) {
        if (ops[t] >= GPO_NUM_OPS) return -1;
        int a = gpo_arity(ops[t]);
        if (d < a) return -1;
        d += 1 - a;
        if (d > mx) mx = d;
    }
    return d == 1 ? mx : -1;
}

int gpo_subtree_start(const uint8_t *ops, int j) {
    int need = 1, k = j;
    while (1) {
        need += gpo_arity(ops[k]) - 1;
        if (need == 0) return k;
        k--;
    }
}

static float tape_op_f32(int o, float a, float b, float c) {
    switch (o) {
    case GPO_ADD: return a + b;
    case GPO_SUB: return a - b;
    case GPO_MUL: return a * b;
    case GPO_DIV: return a / b;
    case GPO_POW: return powf(a, b);
    case GPO_RSUB: return b - a;
    case GPO_RDIV: return b / a;
    case GPO_RPOW: return powf(b, a);
    case GPO_NEG: return -a;
    case GPO_EXP: return expf(a);
    case GPO_LOG: return logf(a);
    case GPO_INV: return 1.0f / a;
    case GPO_SQRT: return sqrtf(a);
    case GPO_SIN: return sinf(a);
    case GPO_ASIN: return asinf(a);
    case GPO_COS: return cosf(a);
    case GPO_ACOS: return acosf(a);
    case GPO_TAN: return tanf(a);
    case GPO_ATAN: return atanf(a);
    case GPO_ABS: return fabsf(a);
    case GPO_CADD: return a + c;
    case GPO_CMUL: return c * a;
    case GPO_CPOW: return powf(a, c);
    case GPO_CRPOW: return powf(c, a);
    default: return NAN;
    }
}
static double tape_op_f64(int o, double a, double b, double c) {
    switch (o) {
    case GPO_ADD: return a + b;
    case GPO_SUB: return a - b;
    case GPO_MUL: return a * b;
    case GPO_DIV: return a / b;
    case GPO_POW: return pow(a, b);
    case GPO_RSUB: return b - a;
    case GPO_RDIV: return b / a;
    case GPO_RPOW: return pow(b, a);
    case GPO_NEG: return -a;
    case GPO_EXP: return exp(a);
    case GPO_LOG: return log(a);
    case GPO_INV: return 1.0 / a;
    case GPO_SQRT: return sqrt(a);
    case GPO_SIN: return sin(a);
    case GPO_ASIN: return asin(a);
    case GPO_COS: return cos(a);
    case GPO_ACOS: return acos(a);
    case GPO_TAN: return tan(a);
    case GPO_ATAN: return atan(a);
    case GPO_ABS: return fabs(a);
    case GPO_CADD: return a + c;
    case GPO_CMUL: return c * a;
    case GPO_CPOW: return pow(a, c);
    case GPO_CRPOW: return pow(c, a);
    default: return NAN;
    }
}

float gpo_eval_tape_f32(const uint8_t *ops, const float *cs, int len, float x) {
    float st[256];
    int sp = 0;
    for (int t = 0; t < len; t++) {
        int o = ops[t];
        if (o == GPO_X) st[sp++] = x;
        else if (o == GPO_CONST) st[sp++] = cs[t];
        else if (o == GPO_ONE) st[sp++] = 1.0f;
        else if (gpo_arity(o) == 2) { float b = st[--sp]; float a = st[sp - 1]; st[sp - 1] = tape_op_f32(o, a, b, cs[t]); }
        else st[sp - 1] = tape_op_f32(o, st[sp - 1], 0.0f, cs[t]);
    }
    return st[0];
}
double gpo_eval_tape_f64(const uint8_t *ops, const float *cs, int len, double x) {
    double st[256];
    int sp = 0;
    for (int t = 0; t < len; t++) {
        int o = ops[t];
        if (o == GPO_X) st[sp++] = x;
        else if (o == GPO_CONST) st[sp++] = cs[t];
        else if (o == GPO_ONE) st[sp++] = 1.0;
        else if (gpo_arity(o) == 2) { double b = st[--sp]; double a = st[sp - 1]; st[sp - 1] = tape_op_f64(o, a, b, cs[t]); }
        else st[sp - 1] = tape_op_f64(o, st[sp - 1], 0.0, cs[t]);
    }
    return st[0];
}

/* Gaussian log-likelihood of a tape: sum_i -(f_i - y_i)^2/(2 sd_i^2) - log(sd_i) - log(2 pi)/2.
 * Non-finite node output or term gives -inf, as compute_likelihood (main.cu:459-465). */
double gpo_tape_loglik(const uint8_t *ops, const float *cs, int len,
                       const float *x, const float *y, const float *sd, int n, int mode) {
    double ll = 0.0;
    if (mode == 2) {
        /* the sampler's definition: lnormalpdf of main.cu:299-301 on FP32 node values,
         * FP32 residual and square, FP64 division and sum */
        for (int i = 0; i < n; i++) {
            float f = gpo_eval_tape_f32(ops, cs, len, x[i]);
            if (!isfinite(f)) return -INFINITY;
            double l = gpo_ref_lnormalpdf(f - y[i], sd[i]);
            if (!isfinite(l)) return -INFINITY;
            ll += l;
        }
        return ll;
    }
    for (int i = 0; i < n; i++) {
        double f = mode == 0 ? (double)gpo_eval_tape_f32(ops, cs, len, x[i])
                             : gpo_eval_tape_f64(ops, cs, len, (double)x[i]);
        if (!isfinite(f)) return -INFINITY;
        double r = f - (double)y[i];
        double s = (double)sd[i];
        double term = -(r * r) / (2.0 * s * s) - log(s) - HALF_LOG2PI_D;
        if (!isfinite(term)) return -INFINITY;
        ll += term;
    }
    return ll;
}

/* Description-length prior plus Cauchy(0,1) on constants.  cauchy_all_slots_count > 0
 * adds that many Cauchy terms for slots given in cs[0..count) instead of per-token. */
static double cauchy_term(float c) {
    return -(double)LOGPI_F - (double)gpo_logf((float)(1.0 + (double)(c * c)));
}
double gpo_tape_prior(const uint8_t *ops, const float *cs, int len, double log_nsym,
                      int cauchy_all_slots_count) {
    int dl = 0;
    double p = 0.0;
    for (int t = 0; t < len; t++) {
        dl += gpo_dl_cost(ops[t]);
        if (cauchy_all_slots_count <= 0 && (ops[t] == GPO_CONST || ops[t] >= GPO_CADD)) p += cauchy_term(cs[t]);
    }
    for (int c = 0; c < cauchy_all_slots_count; c++) p += cauchy_term(cs[c]);
    return -(double)dl * log_nsym + p;
}

/* Log-likelihood plus the Cauchy(0,1) terms of the constant tokens, everything in FP64
 * with FP64 constants: the function whose central differences check the adjoints.
 * Parity pinned by the closed-form tests of tests/test_oracle_sampler.py (linear and
 * rational programs with analytic derivatives). */
double gpo_tape_logpost_f64(const uint8_t *ops, const double *cs, int len,
                            const float *x, const float *y, const float *sd, int n) {
    double lp = 0.0;
    for (int i = 0; i < n; i++) {
        double st[256];
        int sp = 0;
        for (int t = 0; t < len; t++) {
            int o = ops[t];
            if (o == GPO_X) st[sp++] = (double)x[i];
            else if (o == GPO_CONST) st[sp++] = cs[t];
            else if (o == GPO_ONE) st[sp++] = 1.0;
            else if (gpo_arity(o) == 2) { double b = st[--sp]; double a = st[sp - 1]; st[sp - 1] = tape_op_f64(o, a, b, cs[t]); }
            else st[sp - 1] = tape_op_f64(o, st[sp - 1], 0.0, cs[t]);
        }
        double r = st[0] - (double)y[i];
        double s = (double)sd[i];
        lp += -(r * r) / (2.0 * s * s) - log(s) - HALF_LOG2PI_D;
    }
    for (int t = 0; t < len; t++)
        if (ops[t] == GPO_CONST || ops[t] >= GPO_CADD) lp += -LOGPI_D - log1p(cs[t] * cs[t]);
    return lp;
}

/* Reverse-mode adjoint over the tape in FP64.  grad[t] = d lp / d cs[t]. */
int gpo_tape_grad(const uint8_t *ops, const float *cs, int len,
                  const float *x, const float *y, const float *sd, int n,
                  double *grad, double *lp_out) {
    double val[256], adj[256];
    int child[256][2];
    int stk[256], sp = 0;
    for (int t = 0; t < len; t++) {
        int a = gpo_arity(ops[t]);
        child[t][0] = child[t][1] = -1;
        if (a == 2) { child[t][1] = stk[--sp]; child[t][0] = stk[--sp]; }
        else if (a == 1) child[t][0] = stk[--sp];
        stk[sp++] = t;
        grad[t] = 0.0;
    }
    double ll = 0.0;
    for (int i = 0; i < n; i++) {
        for (int t = 0; t < len; t++) {
            int o = ops[t];
            double av = child[t][0] >= 0 ? val[child[t][0]] : 0.0;
            double bv = child[t][1] >= 0 ? val[child[t][1]] : 0.0;
            if (o == GPO_X) val[t] = x[i];
            else if (o == GPO_CONST) val[t] = cs[t];
            else if (o == GPO_ONE) val[t] = 1.0;
            else val[t] = tape_op_f64(o, av, bv, cs[t]);
            adj[t] = 0.0;
        }
        double f = val[len - 1];
        double s = sd[i];
        double r = f - (double)y[i];
        ll += -(r * r) / (2.0 * s * s) - log(s) - HALF_LOG2PI_D;
        adj[len - 1] = -r / (s * s);
        for (int t = len - 1; t >= 0; t--) {
            int o = ops[t];
            double g = adj[t], v = val[t], c = cs[t];
            int ca = child[t][0], cb = child[t][1];
            double a = ca >= 0 ? val[ca] : 0.0, b = cb >= 0 ? val[cb] : 0.0;
            switch (o) {
            case GPO_CONST: grad[t] += g; break;
            case GPO_ADD: adj[ca] += g; adj[cb] += g; break;
            case GPO_SUB: adj[ca] += g; adj[cb] -= g; break;
            case GPO_RSUB: adj[ca] -= g; adj[cb] += g; break;
            case GPO_MUL: adj[ca] += g * b; adj[cb] += g * a; break;
            case GPO_DIV: adj[ca] += g / b; adj[cb] -= g * a / (b * b); break;
            case GPO_RDIV: adj[cb] += g / a; adj[ca] -= g * b / (a * a); break;
            case GPO_POW: adj[ca] += g * b * pow(a, b - 1.0); adj[cb] += g * v * log(a); break;
            case GPO_RPOW: adj[cb] += g * a * pow(b, a - 1.0); adj[ca] += g * v * log(b); break;
            case GPO_NEG: adj[ca] -= g; break;
            case GPO_EXP: adj[ca] += g * v; break;
            case GPO_LOG: adj[ca] += g / a; break;
            case GPO_INV: adj[ca] -= g / (a * a); break;
            case GPO_SQRT: adj[ca] += g * 0.5 / v; break;
            case GPO_SIN: adj[ca] += g * cos(a); break;
            case GPO_COS: adj[ca] -= g * sin(a); break;
            case GPO_TAN: adj[ca] += g * (1.0 + v * v); break;
            case GPO_ASIN: adj[ca] += g / sqrt(1.0 - a * a); break;
            case GPO_ACOS: adj[ca] -= g / sqrt(1.0 - a * a); break;
            case GPO_ATAN: adj[ca] += g / (1.0 + a * a); break;
            case GPO_ABS: adj[ca] += a < 0 ? -g : g; break;
            case GPO_CADD: adj[ca] += g; grad[t] += g; break;
            case GPO_CMUL: adj[ca] += g * c; grad[t] += g * a; break;
            case GPO_CPOW: adj[ca] += g * c * pow(a, c - 1.0); grad[t] += g * v * log(a); break;
            case GPO_CRPOW: adj[ca] += g * v * log(c); grad[t] += g * a * pow(c, a - 1.0); break;
            default: break;
            }
        }
    }
    double prior = 0.0;
    for (int t = 0; t < len; t++) {
        if (ops[t] == GPO_CONST || ops[t] >= GPO_CADD) {
            double c = cs[t];
            grad[t] += -2.0 * c / (1.0 + c * c);
            prior += -LOGPI_D - log1p(c * c);
        }
    }
    if (lp_out) *lp_out = ll + prior;
    for (int t = 0; t < len; t++) if (!isfinite(grad[t])) return -1;
    return 0;
}

/* ------------------------------------------------------------------ config sampler */
void gpo_cfg_gen_table(uint8_t sym[64], double logp[GPO_NUM_OPS]) {
    static const uint8_t order[10] = { GPO_X, GPO_CONST, GPO_ADD, GPO_SUB, GPO_MUL, GPO_DIV, GPO_POW,
                                       GPO_EXP, GPO_LOG, GPO_NEG };
    static const int w[10] = { 16, 16, 4, 4, 4, 4, 4, 4, 4, 4 };
    int k = 0;
    for (int i = 0; i < GPO_NUM_OPS; i++) logp[i] = -INFINITY;
    for (int s = 0; s < 10; s++) {
        for (int j = 0; j < w[s]; j++) sym[k++] = order[s];
        logp[order[s]] = log((double)w[s] / 64.0);
    }
}

/* standard normal log density of a float draw, FP64 */
static double lphi(float z) { return -0.5 * (double)z * (double)z - HALF_LOG2PI_D; }

/* Draw a subtree from the generation table.  Nodes are drawn in right-child-first
 * prefix order (root, then its right subtree, then its left subtree); the reverse of
 * that sequence is the subtree's postfix.  Per node: one word for the symbol (top 6
 * bits), and for CONST one more word for its N(0,1) value.  Returns the size, or -1
 * once the size would exceed maxlen. */
static int gen_subtree(gpo_stream *st, const uint8_t sym[64], const double *logp, int maxlen,
                       uint8_t *out_ops, float *out_cs, double *logq) {
    uint8_t g_op[256]; float g_c[256];
    int need = 1, m = 0;
    double lq = 0.0;
    while (need > 0) {
        if (m >= maxlen) return -1;
        uint32_t r = gpo_stream_next(st);
        int s = sym[r >> 26];
        float c = 0.0f;
        lq += logp[s];
        if (s == GPO_CONST) { c = gpo_normal_icdf(gpo_u01(gpo_stream_next(st))); lq += lphi(c); }
        g_op[m] = (uint8_t)s; g_c[m] = c; m++;
        need += gpo_arity(s) - 1;
    }
    for (int i = 0; i < m; i++) { out_ops[i] = g_op[m - 1 - i]; out_cs[i] = g_c[m - 1 - i]; }
    *logq = lq;
    return m;
}

/* A structure move's new subtree.  Symbol k (right-child-first prefix order, as above)
 * takes word 0 of the Philox block at counter {chain, step, P_STRUCT << 24, k} and, for
 * CONST, its value from word 1 of that block; word 2 of block 0 is the node draw
 * (struct_node_word).  This layout lets the device draw every symbol at once, one lane
 * each.  The generation log probability is the sum over symbols of logp + lphi(value).
 * Returns the size, or -1 when the subtree does not close within maxlen symbols. */
static uint32_t struct_block_word(uint64_t seed, uint32_t chain, uint32_t step, uint32_t k, int word) {
    uint32_t ctr[4] = { chain, step, (uint32_t)GPO_P_STRUCT << 24, k }, key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t out[4];
    gpo_philox4x32_10(ctr, key, out);
    return out[word];
}
static uint32_t struct_node_word(uint64_t seed, uint32_t chain, uint32_t step) {
    return struct_block_word(seed, chain, step, 0, 2);
}
static int gen_subtree_struct(uint64_t seed, uint32_t chain, uint32_t step, const uint8_t sym[64],
                              const double *logp, int maxlen, uint8_t *out_ops, float *out_cs, double *logq) {
    uint8_t g_op[256]; float g_c[256];
    int need = 1, m = 0;
    double lq = 0.0;
    while (need > 0) {
        if (m >= maxlen) return -1;
        int s = sym[struct_block_word(seed, chain, step, (uint32_t)m, 0) >> 26];
        float c = 0.0f;
        double term = logp[s];
        if (s == GPO_CONST) {
            c = gpo_normal_icdf(gpo_u01(struct_block_word(seed, chain, step, (uint32_t)m, 1)));
            term += lphi(c);
        }
        lq += term;
        g_op[m] = (uint8_t)s; g_c[m] = c; m++;
        need += gpo_arity(s) - 1;
    }
    for (int i = 0; i < m; i++) { out_ops[i] = g_op[m - 1 - i]; out_cs[i] = g_c[m - 1 - i]; }
    *logq = lq;
    return m;
}

int gpo_cfg_init_chain(const gpo_cfg *cfg, uint32_t chain, uint8_t *ops, float *cs) {
    uint8_t sym[64]; double logp[GPO_NUM_OPS];
    gpo_cfg_gen_table(sym, logp);
    for (uint32_t attempt = 0; attempt < 64; attempt++) {
        gpo_stream st;
        gpo_stream_init(&st, cfg->seed, chain, attempt, GPO_P_INIT, 0);
        double lq;
        int m = gen_subtree(&st, sym, logp, cfg->cap, ops, cs, &lq);
        if (m > 0 && gpo_tape_depth(ops, m) <= cfg->dmax) return m;
    }
    ops[0] = GPO_X; cs[0] = 0.0f;
    return 1;
}

/* Batched substreams: the Philox block with counter {chain, step>>2, purpose<<24|slot,
 * row} serves four consecutive steps, word step & 3.  Constant proposals of token t
 * use slot t & 63, row t >> 6; accept draws use slot 0, row 0. */
static uint32_t step_word(uint64_t seed, uint32_t chain, uint32_t step, uint32_t purpose, uint32_t slot, uint32_t row) {
    uint32_t ctr[4] = { chain, step >> 2, (purpose << 24) | (slot & 0xffffffu), row };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t out[4];
    gpo_philox4x32_10(ctr, key, out);
    return out[step & 3];
}
static uint32_t cprop_word(uint64_t seed, uint32_t chain, uint32_t step, int t) {
    return step_word(seed, chain, step, GPO_P_CPROP, (uint32_t)(t & 63), (uint32_t)(t >> 6));
}

/* MH accept rule shared by all config moves: invalid current state always moves
 * (as main.cu:730,792); otherwise log u < a with u from the ACCEPT substream. */
static int cfg_accept(const gpo_cfg *cfg, uint32_t chain, uint32_t step, double cur, double prop, double a) {
    if (!isfinite(cur)) return 1;
    if (!isfinite(prop)) return 0;
    if (a >= 0.0) return 1;
    float u = gpo_u01(step_word(cfg->seed, chain, step, GPO_P_ACCEPT, 0, 0));
    return (double)gpo_logf(u) < a;
}

/* accept margin of a decision (gpo_chain_stats.min_margin): |a - log u| over the
 * likelihood scale, where both posteriors are finite */
static void cfg_margin(const gpo_cfg *cfg, uint32_t chain, uint32_t step, double cur, double prop, double a,
                       double ll_cur, double ll_prop, double *m) {
    if (!isfinite(cur) || !isfinite(prop)) return;
    float u = gpo_u01(step_word(cfg->seed, chain, step, GPO_P_ACCEPT, 0, 0));
    double r = fabs(a - (double)gpo_logf(u)) / (fabs(ll_cur) + fabs(ll_prop) + 1.0);
    if (r < *m) *m = r;
}

/* The proposal metric M = L L^T of the constant tokens cfg->mrows (no reference
 * counterpart).  Derived once per chain: M rounded to FP32 from FP64 sums, and L^-1 in
 * FP64 by forward substitution.  The proposal uses d = L z and the drift M g (FP32 sums
 * in increasing index order); the proposal density uses w = L^-1 r (FP64). */
typedef struct {
    int k;
    float L[256], M[256];
    double Li[256];
} metric_t;
static void metric_derive(const gpo_cfg *cfg, int k, metric_t *m) {
    m->k = k;
    memcpy(m->L, cfg->mL, sizeof(m->L));
    for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) {
            double s = 0.0;
            for (int q = 0; q <= (i < j ? i : j); q++) s = s + (double)m->L[i * 16 + q] * (double)m->L[j * 16 + q];
            m->M[i * 16 + j] = (float)s;
        }
    for (int c = 0; c < k; c++)
        for (int i = 0; i < k; i++) {
            if (i < c) { m->Li[i * 16 + c] = 0.0; continue; }
            double s = i == c ? 1.0 : 0.0;
            for (int q = c; q < i; q++) s = s - (double)m->L[i * 16 + q] * m->Li[q * 16 + c];
            m->Li[i * 16 + c] = s / (double)m->L[i * 16 + i];
        }
}
static int metric_slot(const gpo_cfg *cfg, int k, int t) {
    for (int i = 0; i < k; i++) if (cfg->mrows[i] == t) return i;
    return -1;
}
static void metric_L(const metric_t *m, const float *z, float *d) {           /* d = L z */
    for (int i = 0; i < m->k; i++) {
        float acc = 0.0f;
        for (int j = 0; j <= i; j++) acc = acc + m->L[i * 16 + j] * z[j];
        d[i] = acc;
    }
}
static void metric_M(const metric_t *m, const float *g, float *v) {           /* v = M g */
    for (int i = 0; i < m->k; i++) {
        float acc = 0.0f;
        for (int j = 0; j < m->k; j++) acc = acc + m->M[i * 16 + j] * g[j];
        v[i] = acc;
    }
}
static double metric_logq(const metric_t *m, const double *r, double two_e2) {  /* -|L^-1 r|^2 / (2 e^2) */
    double s = 0.0;
    for (int i = 0; i < m->k; i++) {
        double w = 0.0;
        for (int j = 0; j <= i; j++) w = w + m->Li[i * 16 + j] * r[j];
        s += w * w;
    }
    return -s / two_e2;
}

void gpo_cfg_run_chain(const gpo_cfg *cfg, uint32_t chain, uint64_t step0, int steps,
                       uint8_t *ops, float *cs, int *lenp,
                       const float *x, const float *y, const float *sd, int n,
                       gpo_chain_stats *stt) {
    uint8_t sym[64]; double logp[GPO_NUM_OPS];
    gpo_cfg_gen_table(sym, logp);
    const double lns = log((double)cfg->nsym);
    int len = *lenp;
    stt->min_margin = INFINITY;
    double prior = gpo_tape_prior(ops, cs, len, lns, 0);
    double ll = gpo_tape_loglik(ops, cs, len, x, y, sd, n, 2);
    double grad[256];
    int grad_ok = 0;
    uint8_t pops[256]; float pcs[256];
    for (int i = 0; i < steps; i++) {
        uint32_t step = (uint32_t)(step0 + (uint64_t)i);
        double cur = prior + ll;
        if (!cfg->fixed_structure && (step % (uint32_t)cfg->struct_every) == 0) {
            uint32_t r = struct_node_word(cfg->seed, chain, step);
            int j = isfinite(cur) ? (int)(((uint64_t)r * (uint64_t)len) >> 32) : len - 1;
            int s0 = gpo_subtree_start(ops, j);
            int oldsize = j - s0 + 1;
            int maxnew = cfg->cap - (len - oldsize);
            uint8_t nops[256]; float ncs[256]; double lq_new;
            stt->n_struct_prop++;
            int m = gen_subtree_struct(cfg->seed, chain, step, sym, logp, maxnew, nops, ncs, &lq_new);
            if (m < 0) continue;
            int plen = len - oldsize + m;
            memcpy(pops, ops, (size_t)s0); memcpy(pcs, cs, sizeof(float) * (size_t)s0);
            memcpy(pops + s0, nops, (size_t)m); memcpy(pcs + s0, ncs, sizeof(float) * (size_t)m);
            memcpy(pops + s0 + m, ops + j + 1, (size_t)(len - j - 1));
            memcpy(pcs + s0 + m, cs + j + 1, sizeof(float) * (size_t)(len - j - 1));
            if (gpo_tape_depth(pops, plen) > cfg->dmax) continue;
            double lq_old = 0.0;
            for (int t = s0; t <= j; t++) {
                double term = logp[ops[t]];
                if (ops[t] == GPO_CONST) term += lphi(cs[t]);
                lq_old += term;
            }
            double pprior = gpo_tape_prior(pops, pcs, plen, lns, 0);
            double pll = gpo_tape_loglik(pops, pcs, plen, x, y, sd, n, 2);
            double fwd = -log((double)len) + lq_new;
            double bwd = -log((double)plen) + lq_old;
            double a = (pprior + pll) - cur + bwd - fwd;
            cfg_margin(cfg, chain, step, cur, pprior + pll, a, ll, pll, &stt->min_margin);
            if (cfg_accept(cfg, chain, step, cur, pprior + pll, a)) {
                memcpy(ops, pops, (size_t)plen); memcpy(cs, pcs, sizeof(float) * (size_t)plen);
                len = plen; prior = pprior; ll = pll; grad_ok = 0;
                stt->n_struct_acc++;
            }
        } else {
            int nconst = 0;
            for (int t = 0; t < len; t++) nconst += ops[t] == GPO_CONST;
            if (nconst == 0) continue;
            stt->n_const_prop++;
            memcpy(pcs, cs, sizeof(float) * (size_t)len);
            double a_extra = 0.0, pgrad[256];
            /* metric rows: their z, current constants and (MALA) gradients.  A chain whose
             * metric rows are not all constant tokens moves every constant with the identity. */
            int K = cfg->mk;
            for (int i = 0; i < cfg->mk; i++)
                if (cfg->mrows[i] >= len || ops[cfg->mrows[i]] != GPO_CONST) K = 0;
            metric_t mt;
            metric_derive(cfg, K, &mt);
            float mz[16], md[16], mc[16];
            for (int i = 0; i < K; i++) {
                mz[i] = gpo_normal_icdf(gpo_u01(cprop_word(cfg->seed, chain, step, cfg->mrows[i])));
                mc[i] = cs[cfg->mrows[i]];
            }
            metric_L(&mt, mz, md);
            if (cfg->proposal == 1) {
                if (!grad_ok) {
                    if (gpo_tape_grad(ops, cs, len, x, y, sd, n, grad, NULL) != 0) {
                        for (int t = 0; t < len; t++) grad[t] = 0.0;
                    }
                    grad_ok = 1;
                }
                const float e = cfg->eps, h = 0.5f * e * e;
                const double two_e2 = 2.0 * (double)e * (double)e;
                const double r2 = 1.0 / two_e2;     /* the identity rows' terms: -d^2 r2 */
                double fq = 0.0;
                for (int t = 0; t < len; t++) {
                    if (ops[t] != GPO_CONST || metric_slot(cfg, K, t) >= 0) continue;
                    float z = gpo_normal_icdf(gpo_u01(cprop_word(cfg->seed, chain, step, t)));
                    float mean = cs[t] + h * (float)grad[t];
                    pcs[t] = mean + e * z;
                    double d = (double)pcs[t] - (double)mean;
                    fq += -d * d * r2;
                }
                float mg[16], mv[16], mmean[16];
                double r[16];
                if (K > 0) {
                    for (int i = 0; i < K; i++) mg[i] = (float)grad[cfg->mrows[i]];
                    metric_M(&mt, mg, mv);
                    for (int i = 0; i < K; i++) {
                        mmean[i] = mc[i] + h * mv[i];
                        pcs[cfg->mrows[i]] = mmean[i] + e * md[i];
                        r[i] = (double)pcs[cfg->mrows[i]] - (double)mmean[i];
                    }
                    fq += metric_logq(&mt, r, two_e2);
                }
                int pg_ok = gpo_tape_grad(ops, pcs, len, x, y, sd, n, pgrad, NULL) == 0;
                double bq = 0.0;
                for (int t = 0; t < len; t++) {
                    if (ops[t] != GPO_CONST || metric_slot(cfg, K, t) >= 0) continue;
                    float mean = pcs[t] + h * (float)(pg_ok ? pgrad[t] : 0.0);
                    double d = (double)cs[t] - (double)mean;
                    bq += -d * d * r2;
                }
                if (K > 0) {
                    for (int i = 0; i < K; i++) mg[i] = (float)(pg_ok ? pgrad[cfg->mrows[i]] : 0.0);
                    metric_M(&mt, mg, mv);
                    for (int i = 0; i < K; i++) {
                        float mb = pcs[cfg->mrows[i]] + h * mv[i];
                        r[i] = (double)mc[i] - (double)mb;
                    }
                    bq += metric_logq(&mt, r, two_e2);
                }
                a_extra = pg_ok ? bq - fq : -INFINITY;
            } else {
                for (int t = 0; t < len; t++) {
                    if (ops[t] != GPO_CONST || metric_slot(cfg, K, t) >= 0) continue;
                    float z = gpo_normal_icdf(gpo_u01(cprop_word(cfg->seed, chain, step, t)));
                    pcs[t] = cs[t] + cfg->sigma * z;
                }
                for (int i = 0; i < K; i++) pcs[cfg->mrows[i]] = mc[i] + cfg->sigma * md[i];
            }
            double pprior = gpo_tape_prior(ops, pcs, len, lns, 0);
            double pll = gpo_tape_loglik(ops, pcs, len, x, y, sd, n, 2);
            double a = (pprior + pll) - cur + a_extra;
            cfg_margin(cfg, chain, step, cur, pprior + pll, a, ll, pll, &stt->min_margin);
            if (cfg_accept(cfg, chain, step, cur, pprior + pll, a)) {
                memcpy(cs, pcs, sizeof(float) * (size_t)len);
                prior = pprior; ll = pll;
                if (cfg->proposal == 1) memcpy(grad, pgrad, sizeof(double) * (size_t)len);
                stt->n_const_acc++;
            }
        }
    }
    *lenp = len;
    stt->prior = prior; stt->ll = ll;
}

void gpo_cfg_run_block(const gpo_cfg *cfg, int nchains, uint32_t chain0, uint64_t step0, int steps,
                       uint8_t *ops, float *cs, int32_t *len,
                       const float *x, const float *y, const float *sd, int n,
                       double *prior, double *ll, long *counters, int nthreads) {
    gpo_cfg_run_block_m(cfg, nchains, chain0, step0, steps, ops, cs, len, x, y, sd, n, prior, ll, counters, nthreads,
                        NULL);
}

void gpo_cfg_run_block_m(const gpo_cfg *cfg, int nchains, uint32_t chain0, uint64_t step0, int steps,
                         uint8_t *ops, float *cs, int32_t *len,
                         const float *x, const float *y, const float *sd, int n,
                         double *prior, double *ll, long *counters, int nthreads, double *margin) {
    long c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) reduction(+:c0,c1,c2,c3)
#endif
    for (int c = 0; c < nchains; c++) {
        uint8_t o[256]; float v[256];
        int L = len[c];
        for (int t = 0; t < L; t++) { o[t] = ops[(size_t)t * nchains + c]; v[t] = cs[(size_t)t * nchains + c]; }
        gpo_chain_stats st; memset(&st, 0, sizeof st);
        gpo_cfg_run_chain(cfg, chain0 + (uint32_t)c, step0, steps, o, v, &L, x, y, sd, n, &st);
        len[c] = L;
        for (int t = 0; t < cfg->cap; t++) {
            ops[(size_t)t * nchains + c] = t < L ? o[t] : 0;
            cs[(size_t)t * nchains + c] = t < L ? v[t] : 0.0f;
        }
        prior[c] = st.prior; ll[c] = st.ll;
        if (margin) margin[c] = st.min_margin;
        c0 += st.n_struct_prop; c1 += st.n_struct_acc; c2 += st.n_const_prop; c3 += st.n_const_acc;
    }
    (void)nthreads;
    if (counters) { counters[0] += c0; counters[1] += c1; counters[2] += c2; counters[3] += c3; }
}

int gpo_cfg_resample(int nchains, uint32_t chain0, uint32_t block, uint64_t seed, int cap,
                     double threshold, uint8_t *ops, float *cs, int32_t *len,
                     double *prior, double *ll, int32_t *src_out) {
    double map = -INFINITY;
    for (int c = 0; c < nchains; c++) { double p = prior[c] + ll[c]; if (p == p && p > map) map = p; }
    int *good = (int *)malloc(sizeof(int) * (size_t)nchains), ng = 0, nbad = 0;
    for (int c = 0; c < nchains; c++) { double p = prior[c] + ll[c]; if (p >= map - threshold) good[ng++] = c; }
    for (int c = 0; c < nchains; c++) {
        double p = prior[c] + ll[c];
        src_out[c] = c;
        if (p >= map - threshold || ng == 0) continue;
        gpo_stream st;
        gpo_stream_init(&st, seed, chain0 + (uint32_t)c, block, GPO_P_RESAMPLE, 0);
        uint32_t r = gpo_stream_next(&st);
        int j = good[(int)(((uint64_t)r * (uint64_t)ng) >> 32)];
        src_out[c] = j;
        nbad++;
    }
    /* gather from the pre-resample state: sources are good chains, never overwritten */
    for (int c = 0; c < nchains; c++) {
        int j = src_out[c];
        if (j == c) continue;
        len[c] = len[j];
        for (int t = 0; t < cap; t++) {
            ops[(size_t)t * nchains + c] = ops[(size_t)t * nchains + j];
            cs[(size_t)t * nchains + c] = cs[(size_t)t * nchains + j];
        }
        prior[c] = prior[j]; ll[c] = ll[j];
    }
    free(good);
    return nbad;
}

/* ------------------------------------------------------------------ reference sampler */
typedef struct { int lcg; uint32_t *rx; gpo_stream st; } src_t;
static uint32_t src_rand31(src_t *s) { return s->lcg ? gpo_lcg_next(s->rx) : (gpo_stream_next(&s->st) >> 1); }
static void src_sub(src_t *s, const gpo_refcfg *cfg, uint32_t *rx, uint32_t chain, uint32_t step,
                    uint32_t purpose, uint32_t slot) {
    s->lcg = cfg->rng == 1; s->rx = rx;
    if (!s->lcg) gpo_stream_init(&s->st, cfg->seed, chain, step, purpose, slot);
}

/* random_int, random_float, random_normal, random_geometric: main.cu:218-254 */
static int dev_random_int(int n, src_t *s) {
    int divisor = 0x7fffffff / (n + 1);
    int r;
    do { r = (int)(src_rand31(s) / (uint32_t)divisor); } while (r >= n);
    return r;
}
static float dev_random_float(src_t *s) { return (float)((double)(float)dev_random_int(1000000, s) / 1000000.0); }
static float dev_random_normal(src_t *s) {
    float u = dev_random_float(s);
    float v = dev_random_float(s);
    return sqrtf(-2.0f * gpo_logf(u)) * gpo_sin2pi(v);
}
static float dev_random_geometric(src_t *s) { return ceilf(gpo_logf(dev_random_float(s)) / LOG01_F); }

/* host random_int / random_float / random_normal: main.cu:262-284 */
static int host_random_int(int limit, src_t *s) {
    int divisor = 0x7fffffff / limit;
    int r;
    do { r = (int)(src_rand31(s) / (uint32_t)divisor); } while (r > limit - 1);
    return r;
}
static float host_random_float(src_t *s) { return (float)src_rand31(s) / (float)0x7fffffff; }
static float host_random_normal(src_t *s) {
    float u = host_random_float(s);
    float v = host_random_float(s);
    return sqrtf(-2.0f * gpo_logf(u)) * gpo_sin2pi(v);
}

static int ref_is_int(float x) {                        /* main.cu:192 */
    if (!isfinite(x)) return 0;
    return fabsf(x - roundf(x)) < 1.0e-6;
}
static double ref_logplusexp(double a, double b) {      /* main.cu:190 */
    double m = a > b ? a : b;
    return (double)gpo_logf(gpo_expf((float)(a - m)) + gpo_expf((float)(b - m))) + m;
}

static int ref_valid(double x) { return isfinite(x); }

/* n sequential draws of one reference transform from one source (tests):
 * kind 0 random_int(arg), 1 random_float, 2 random_normal, 3 random_geometric(0.1),
 * 4 host random_int(arg), 5 host random_float, 6 host random_normal */
static void ref_draws_from(src_t *s, int kind, int n, int arg, double *out);
void gpo_ref_draws(int kind, int rng, uint64_t seed, uint32_t lcg_seed, int n, int arg, double *out) {
    gpo_refcfg cfg = { rng, seed, 0 };
    uint32_t rx = lcg_seed;
    src_t s;
    src_sub(&s, &cfg, &rx, 0, 0, 0, 0);
    ref_draws_from(&s, kind, n, arg, out);
}
/* the same draws from the Philox substream (chain, step, purpose, slot) that the sampler
 * uses at that step, e.g. the accept draw or a constant's proposal draws */
void gpo_ref_stream_draws(int kind, uint64_t seed, uint32_t chain, uint32_t step, uint32_t purpose,
                          uint32_t slot, int n, int arg, double *out) {
    gpo_refcfg cfg = { 0, seed, 0 };
    src_t s;
    src_sub(&s, &cfg, NULL, chain, step, purpose, slot);
    ref_draws_from(&s, kind, n, arg, out);
}
static void ref_draws_from(src_t *s, int kind, int n, int arg, double *out) {
    for (int i = 0; i < n; i++) {
        switch (kind) {
        case 0: out[i] = dev_random_int(arg, s); break;
        case 1: out[i] = dev_random_float(s); break;
        case 2: out[i] = dev_random_normal(s); break;
        case 3: out[i] = dev_random_geometric(s); break;
        case 4: out[i] = host_random_int(arg, s); break;
        case 5: out[i] = host_random_float(s); break;
        default: out[i] = host_random_normal(s); break;
        }
    }
}

/* MH_simple_kernel, main.cu:679-813, one chain (AoS heap arrays) */
/* accept margin of a reference decision: the step's random_float u drawn from a copy of
 * the draw source (the LCG state included), so the chain's own draws are untouched */
static void ref_margin(const src_t *s, double cur, double prop, double acc, double ll_cur, double ll_prop,
                       double *m) {
    if (!ref_valid(cur) || !ref_valid(prop)) return;
    src_t c = *s;
    uint32_t rx2 = s->lcg ? *s->rx : 0;
    if (s->lcg) c.rx = &rx2;
    float u = dev_random_float(&c);
    double r = fabs(acc - (u > 0.0f ? log((double)u) : -1e300)) / (fabs(ll_cur) + fabs(ll_prop) + 1.0);
    if (r < *m) *m = r;
}

static void ref_run_chain(const gpo_refcfg *cfg, uint32_t chain, uint64_t step0, int steps,
                          int32_t random_seed, uint8_t *P, float *C,
                          const float *x, const float *y, const float *sd, int n,
                          double *prior_out, double *ll_out, double *margin) {
    double cur_prior = gpo_ref_prior_ops(P, C, 1, cfg->ops);
    double cur_ll = gpo_ref_likelihood(P, C, 1, x, y, sd, n);
    double cur_post = cur_prior + cur_ll;
    uint32_t rx = (uint32_t)random_seed + chain;         /* main.cu:691 */
    uint8_t old_P[GPO_HEAP]; float old_C[GPO_HEAP];
    double forward = 0.0, backward = 0.0;                 /* main.cu:695: once per launch */
    src_t s;
    for (int mcmci = 0; mcmci < steps; mcmci++) {
        uint32_t step = (uint32_t)(step0 + (uint64_t)mcmci);
        if (mcmci % 5 == 0) {
            forward = -(double)gpo_logf((float)gpo_ref_fold_count(P, 1));
            src_sub(&s, cfg, &rx, chain, step, GPO_P_STRUCT, 0);
            int nn = !ref_valid(cur_post) ? 0 : dev_random_int(GPO_HEAP, &s);
            for (int i = nn; i < GPO_HEAP; i++) {
                old_P[i] = P[i];
                if (gpo_ref_is_descendant(nn, i))
                    P[i] = (uint8_t)gpo_opset_code(cfg->ops, dev_random_int(gpo_opset_nops(cfg->ops), &s));
            }
            backward = -(double)gpo_logf((float)gpo_ref_fold_count(P, 1));
            double pp = gpo_ref_prior_ops(P, C, 1, cfg->ops);
            double pl = gpo_ref_likelihood(P, C, 1, x, y, sd, n);
            double ppost = pp + pl;
            double acc = ppost - cur_post + backward - forward;
            src_sub(&s, cfg, &rx, chain, step, GPO_P_ACCEPT, 0);
            if (margin) ref_margin(&s, cur_post, ppost, acc, cur_ll, pl, margin);
            if ((ref_valid(ppost) && (acc > 0.0 || dev_random_float(&s) < gpo_expf((float)acc))) || !ref_valid(cur_post)) {
                cur_post = ppost; cur_ll = pl; cur_prior = pp;
            } else {
                for (int i = nn; i < GPO_HEAP; i++) P[i] = old_P[i];
            }
        } else {
            for (int c = 0; c < GPO_HEAP; c++) {
                src_sub(&s, cfg, &rx, chain, step, GPO_P_CPROP, (uint32_t)c);
                old_C[c] = C[c];
                if (dev_random_float(&s) < 1.0f) {                  /* P_PROPOSE_C_GEOMETRIC */
                    if (dev_random_float(&s) < 0.15f) {             /* PROPOSE_INT */
                        float sgn = dev_random_float(&s) < 0.5f ? -1.0f : 1.0f;
                        C[c] = sgn * dev_random_geometric(&s) - 1.0f;
                        forward += gpo_ref_lgeometricpdf(C[c] + 1.0f, 0.1f) + (C[c] == 0.0f ? 0.0 : LOG2_D) + (double)LOG015_F;
                    } else {
                        C[c] = 1.0f * dev_random_normal(&s);
                        forward += gpo_ref_lnormalpdf(C[c] / 1.0f, 1.0f) + LOG1M015_D;
                    }
                    if (ref_is_int(old_C[c])) {
                        backward += ref_logplusexp(gpo_ref_lgeometricpdf(old_C[c] + 1.0f, 0.1f) + (C[c] == 0.0f ? 0.0 : LOG2_D) + (double)LOG015_F,
                                                   gpo_ref_lnormalpdf(old_C[c] / 1.0f, 1.0f) + LOG1M015_D);
                    } else {
                        backward += gpo_ref_lnormalpdf(old_C[c] / 1.0f, 1.0f) + LOG1M015_D;
                    }
                } else {                                            /* unreachable, main.cu:779-784 */
                    int order = dev_random_int(8, &s) - 4;
                    C[c] = C[c] + 1.0f * POW10_F[order + 4] * dev_random_normal(&s);
                }
            }
            double pp = gpo_ref_prior_ops(P, C, 1, cfg->ops);
            double pl = gpo_ref_likelihood(P, C, 1, x, y, sd, n);
            double ppost = pp + pl;
            double acc = ppost - cur_post + backward - forward;
            src_sub(&s, cfg, &rx, chain, step, GPO_P_ACCEPT, 0);
            if (margin) ref_margin(&s, cur_post, ppost, acc, cur_ll, pl, margin);
            if ((ref_valid(ppost) && (acc > 0.0 || dev_random_float(&s) < gpo_expf((float)acc))) || !ref_valid(cur_post)) {
                cur_post = ppost; cur_ll = pl; cur_prior = pp;
            } else {
                for (int c = 0; c < GPO_HEAP; c++) C[c] = old_C[c];
            }
        }
    }
    *prior_out = cur_prior; *ll_out = cur_ll;
}

void gpo_ref_run_block(const gpo_refcfg *cfg, int nchains, uint32_t chain0, uint64_t step0, int steps,
                       int32_t random_seed, uint8_t *hop, float *hc,
                       const float *x, const float *y, const float *sd, int n,
                       double *prior, double *ll, int nthreads) {
    gpo_ref_run_block_m(cfg, nchains, chain0, step0, steps, random_seed, hop, hc, x, y, sd, n, prior, ll, nthreads,
                        NULL);
}

void gpo_ref_run_block_m(const gpo_refcfg *cfg, int nchains, uint32_t chain0, uint64_t step0, int steps,
                         int32_t random_seed, uint8_t *hop, float *hc,
                         const float *x, const float *y, const float *sd, int n,
                         double *prior, double *ll, int nthreads, double *margin) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int c = 0; c < nchains; c++) {
        uint8_t P[GPO_HEAP]; float C[GPO_HEAP];
        for (int i = 0; i < GPO_HEAP; i++) { P[i] = hop[(size_t)i * nchains + c]; C[i] = hc[(size_t)i * nchains + c]; }
        double m = INFINITY;
        ref_run_chain(cfg, chain0 + (uint32_t)c, step0, steps, random_seed, P, C, x, y, sd, n, &prior[c], &ll[c],
                      margin ? &m : NULL);
        if (margin) margin[c] = m;
        for (int i = 0; i < GPO_HEAP; i++) { hop[(size_t)i * nchains + c] = P[i]; hc[(size_t)i * nchains + c] = C[i]; }
    }
    (void)nthreads;
}

/* initial draw, main.cu:1133-1137, with Philox in place of rand() */
static void ref_draw_chain(src_t *s, uint8_t *hop, float *hc, size_t stride, int set) {
    for (int i = 0; i < GPO_HEAP; i++)
        hop[i * stride] = (uint8_t)gpo_opset_code(set, host_random_int(gpo_opset_nops(set) - 1, s));
    for (int i = 0; i < GPO_HEAP; i++) {
        int order = host_random_int(8, s) - 4;
        hc[i * stride] = 1.0f * POW10_F[order + 4] * host_random_normal(s);
    }
}
void gpo_ref_init(uint64_t seed, int nchains, uint32_t chain0, uint8_t *hop, float *hc, int set) {
    gpo_refcfg cfg = { 0, seed, set };
    for (int c = 0; c < nchains; c++) {
        src_t s;
        src_sub(&s, &cfg, NULL, chain0 + (uint32_t)c, 0, GPO_P_INIT, 0);
        ref_draw_chain(&s, hop + c, hc + c, (size_t)nchains, set);
    }
}

/* cmp, main.cu:987-995: NaN sorts below everything */
static int cmp_ref(const void *pa, const void *pb) {
    double a = *(const double *)pa, b = *(const double *)pb;
    int an = a != a, bn = b != b;
    if (an || bn) return an && bn ? 0 : (an ? -1 : 1);
    return a < b ? -1 : (a > b ? 1 : 0);
}
int gpo_ref_resample(uint64_t seed, uint32_t block, int nchains, uint32_t chain0,
                     uint8_t *hop, float *hc, const double *prior, const double *ll,
                     double *cutoff_out, int set) {
    double *post = (double *)malloc(sizeof(double) * (size_t)nchains);
    for (int c = 0; c < nchains; c++) post[c] = prior[c] + ll[c];
    qsort(post, (size_t)nchains, sizeof(double), cmp_ref);
    double cutoff = post[(int)(nchains * 0.5f)];         /* RESAMPLE_QUANTILE, main.cu:36,1219 */
    free(post);
    gpo_refcfg cfg = { 0, seed, set };
    int nre = 0;
    for (int c = 0; c < nchains; c++) {
        if (prior[c] + ll[c] < cutoff) {
            src_t s;
            src_sub(&s, &cfg, NULL, chain0 + (uint32_t)c, block, GPO_P_RESAMPLE, 0);
            ref_draw_chain(&s, hop + c, hc + c, (size_t)nchains, set);
            nre++;
        }
    }
    if (cutoff_out) *cutoff_out = cutoff;
    return nre;
}

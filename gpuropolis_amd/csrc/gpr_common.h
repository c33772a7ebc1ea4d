// Shared definitions for the GPUropolis MI355X kernels: op table, Philox4x32-10
// streams and the draw transforms.  Everything here is __host__ __device__ so the
// host-side C ABI can run the same init/decode code when it needs to.
//
// The draw transforms (gpr_logf, gpr_expf, gpr_sin2pi, gpr_normal) use only IEEE
// +,-,*,/,sqrt with contraction disabled, so a draw is bit-identical wherever it
// is computed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gpuropolis.h"

#define GPR_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------- op table
// Postfix op codes (GPR_OP_* in gpuropolis.h).  Leaves push, binary ops pop rhs
// then lhs, unary ops rewrite the top of the stack.
GPR_HD int gpr_arity(int op) { return op <= GPR_OP_ONE ? 0 : (op <= GPR_OP_RPOW ? 2 : 1); }
GPR_HD int gpr_dl_cost(int op) { return op >= GPR_OP_CADD ? 2 : 1; }
GPR_HD bool gpr_has_const(int op) { return op == GPR_OP_CONST || op >= GPR_OP_CADD; }

// Philox purposes: counter word 2 = purpose << 24 | slot
enum { GPR_P_STRUCT = 1, GPR_P_CPROP = 2, GPR_P_ACCEPT = 3, GPR_P_INIT = 4, GPR_P_RESAMPLE = 5 };

// ---------------------------------------------------------------- Philox4x32-10
// Salmon et al., SC'11: two 32x32->64 multiplies per round, ten rounds, Weyl key.
struct gpr_u4 { uint32_t x, y, z, w; };

// The key is the same at every call site of a kernel, so the compiler hoists its ten-
// round schedule out of every loop and keeps twenty scalar registers live (spilled to
// vector lanes in the large kernels).  On the device the key passes through an empty
// asm, so each call recomputes the schedule with ten scalar adds instead.
__host__ inline uint32_t gpr_keep(uint32_t k) { return k; }
__device__ __forceinline__ uint32_t gpr_keep(uint32_t k) {
    __asm__ volatile("" : "+s"(k));
    return k;
}

GPR_HD gpr_u4 gpr_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    k0 = gpr_keep(k0);
    k1 = gpr_keep(k1);
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}

// A substream: key = seed, counter = {chain, step, purpose<<24|slot, block}; words are
// consumed x, y, z, w of block 0, then of block 1, ...  (selects, not a buffer, so
// wave-uniform code keeps the state in scalar registers)
struct gpr_rstream {
    uint32_t k0, k1, c0, c1, c2, c3;
    gpr_u4 b;
    int idx;
    GPR_HD void init(uint64_t seed, uint32_t chain, uint32_t step, uint32_t purpose, uint32_t slot) {
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
        c0 = chain; c1 = step; c2 = (purpose << 24) | (slot & 0xffffffu); c3 = 0;
        idx = 4;
    }
    GPR_HD uint32_t next() {
        if (idx == 4) {
            b = gpr_philox(c0, c1, c2, c3, k0, k1);
            c3++;
            idx = 0;
        }
        uint32_t r = idx == 0 ? b.x : (idx == 1 ? b.y : (idx == 2 ? b.z : b.w));
        idx++;
        return r;
    }
};

// first word of a substream (single-draw purposes)
GPR_HD uint32_t gpr_first(uint64_t seed, uint32_t chain, uint32_t step, uint32_t purpose, uint32_t slot) {
    return gpr_philox(chain, step, (purpose << 24) | (slot & 0xffffffu), 0, (uint32_t)seed, (uint32_t)(seed >> 32)).x;
}

// ---------------------------------------------------------------- draw transforms
GPR_HD float gpr_bits_f(uint32_t b) { return __builtin_bit_cast(float, b); }
GPR_HD uint32_t gpr_f_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

GPR_HD float gpr_u01(uint32_t r) { return ((float)(r >> 9) + 0.5f) * 0x1p-23f; }

// natural log: m*2^e, m in (sqrt(1/2), sqrt(2)], log m = 2 atanh((m-1)/(m+1)); fused
// multiply-adds are written as fmaf so host and device round identically.  No branches:
// the main path runs for every input and special inputs select their result at the end
// (on the GPU a data-dependent branch costs more than the whole polynomial).
GPR_HD float gpr_logf(float x) {
#pragma clang fp contract(off)
    const bool sub = x > 0.0f && x < 0x1p-126f;
    const float xs = sub ? x * 0x1p23f : x;
    const uint32_t b = gpr_f_bits(xs);
    int e = (int)((b >> 23) & 0xffu) - (sub ? 23 : 0) - 127;
    float m = gpr_bits_f((b & 0x007fffffu) | 0x3f800000u);
    const bool big = m > 0x1.6a09e6p+0f;
    m = big ? m * 0.5f : m;
    e += big ? 1 : 0;
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = __builtin_fmaf(s2, 0x1.745d18p-4f, 0x1.c71c72p-4f);
    p = __builtin_fmaf(s2, p, 0x1.24924ap-3f);
    p = __builtin_fmaf(s2, p, 0x1.99999ap-3f);
    p = __builtin_fmaf(s2, p, 0x1.555556p-2f);
    p = p * s2;
    const float lm = __builtin_fmaf(s, p, s) * 2.0f;
    const float ef = (float)e;
    float r = __builtin_fmaf(ef, 0x1.62e4p-1f, __builtin_fmaf(ef, 0x1.7f7d1cp-20f, lm));
    r = x == __builtin_inff() ? x : r;
    r = x == 0.0f ? -__builtin_inff() : r;
    r = x < 0.0f ? __builtin_nanf("") : r;
    return x != x ? x : r;
}

GPR_HD float gpr_pow2i(int k) { return gpr_bits_f((uint32_t)(k + 127) << 23); }

// exp: 2^k e^r, |r| <= ln2/2, Cody-Waite split of ln2
GPR_HD float gpr_expf(float x) {
#pragma clang fp contract(off)
    if (x != x) return x;
    if (x > 0x1.62e42ep+6f) return __builtin_inff();
    if (x < -0x1.9fe368p+6f) return 0.0f;
    const float kf = __builtin_floorf(__builtin_fmaf(x, 0x1.715476p+0f, 0.5f));
    const float r = __builtin_fmaf(-kf, 0x1.7f7d1cp-20f, __builtin_fmaf(-kf, 0x1.62e4p-1f, x));
    float p = __builtin_fmaf(r, 0x1.a01a02p-13f, 0x1.6c16c2p-10f);
    p = __builtin_fmaf(r, p, 0x1.111112p-7f);
    p = __builtin_fmaf(r, p, 0x1.555556p-5f);
    p = __builtin_fmaf(r, p, 0x1.555556p-3f);
    p = __builtin_fmaf(r, p, 0.5f);
    p = __builtin_fmaf(r, p, 1.0f);
    p = __builtin_fmaf(r, p, 1.0f);
    const int k = (int)kf;
    const int k1 = k / 2, k2 = k - k1;
    return (p * gpr_pow2i(k1)) * gpr_pow2i(k2);
}

// sin(2 pi v): quarter-turn reduction, Taylor polynomials of sin/cos(pi r / 2), |r| <= 1/2
GPR_HD float gpr_sin2pi(float v) {
#pragma clang fp contract(off)
    if (!(v - v == 0.0f)) return __builtin_nanf("");
    const float v4 = v * 4.0f;
    const float q = __builtin_floorf(v4 + 0.5f);
    const float r = v4 - q;
    const int qi = ((int)__builtin_fmodf(q, 4.0f) + 4) & 3;
    const float r2 = r * r;
    float sn = __builtin_fmaf(r2, -0x1.e3075p-19f, 0x1.507834p-13f);
    sn = __builtin_fmaf(r2, sn, -0x1.32d2ccp-8f);
    sn = __builtin_fmaf(r2, sn, 0x1.466bc6p-4f);
    sn = __builtin_fmaf(r2, sn, -0x1.4abbcep-1f);
    sn = __builtin_fmaf(r2, sn, 0x1.921fb6p+0f);
    sn = sn * r;
    float cs = __builtin_fmaf(r2, 0x1.f9d38ap-22f, -0x1.a6d1f2p-16f);
    cs = __builtin_fmaf(r2, cs, 0x1.e1f506p-11f);
    cs = __builtin_fmaf(r2, cs, -0x1.55d3c8p-6f);
    cs = __builtin_fmaf(r2, cs, 0x1.03c1fp-2f);
    cs = __builtin_fmaf(r2, cs, -0x1.3bd3ccp+0f);
    cs = __builtin_fmaf(r2, cs, 1.0f);
    return qi == 0 ? sn : (qi == 1 ? cs : (qi == 2 ? -sn : -cs));
}

// standard normal quantile sqrt(2) erfinv(2u-1), erfinv by M. Giles (GPU Computing
// Gems, 2011) single-precision polynomials; odd in 2u-1.  Both branches are evaluated
// and selected (w < 5 for all but the outer 0.7 % of u), so lanes never diverge.
GPR_HD float gpr_normal(float u) {
#pragma clang fp contract(off)
    const float x = __builtin_fmaf(2.0f, u, -1.0f);
    const float w = -gpr_logf((1.0f - x) * (1.0f + x));
    const float wc = w - 2.5f;
    float p = 2.81022636e-08f;
    p = __builtin_fmaf(p, wc, 3.43273939e-07f);
    p = __builtin_fmaf(p, wc, -3.5233877e-06f);
    p = __builtin_fmaf(p, wc, -4.39150654e-06f);
    p = __builtin_fmaf(p, wc, 0.00021858087f);
    p = __builtin_fmaf(p, wc, -0.00125372503f);
    p = __builtin_fmaf(p, wc, -0.00417768164f);
    p = __builtin_fmaf(p, wc, 0.246640727f);
    p = __builtin_fmaf(p, wc, 1.50140941f);
    const float wt = __builtin_sqrtf(w) - 3.0f;
    float q = -0.000200214257f;
    q = __builtin_fmaf(q, wt, 0.000100950558f);
    q = __builtin_fmaf(q, wt, 0.00134934322f);
    q = __builtin_fmaf(q, wt, -0.00367342844f);
    q = __builtin_fmaf(q, wt, 0.00573950773f);
    q = __builtin_fmaf(q, wt, -0.0076224613f);
    q = __builtin_fmaf(q, wt, 0.00943887047f);
    q = __builtin_fmaf(q, wt, 1.00167406f);
    q = __builtin_fmaf(q, wt, 2.83297682f);
    return ((w < 5.0f ? p : q) * x) * 1.41421356f;
}

// Words of the batched substreams: a Philox block at counter step>>2 serves four
// consecutive steps, word step & 3.  Constant proposals of token t use slot t & 63 and
// block word-row t >> 6; accept draws use slot 0.
GPR_HD uint32_t gpr_u4_word(const gpr_u4 &w, int i) { return i == 0 ? w.x : (i == 1 ? w.y : (i == 2 ? w.z : w.w)); }
GPR_HD gpr_u4 gpr_step_block(uint64_t seed, uint32_t chain, uint32_t step, uint32_t purpose, uint32_t slot, uint32_t row) {
    return gpr_philox(chain, step >> 2, (purpose << 24) | (slot & 0xffffffu), row, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// ---------------------------------------------------------------- densities
#define GPR_LOGPI_F 0x1.250d06p+0f          /* logf((float)3.141592653589) */
#define GPR_HALF_LOG2PI_D 0x1.d67f1c864beb4p-1
// Cauchy(0,1) log density with the reference's float arithmetic (main.cu:303-305)
GPR_HD float gpr_lcauchy_logf(float c) {
#pragma clang fp contract(off)
    return gpr_logf((float)(1.0 + (double)(c * c)));
}
GPR_HD double gpr_lcauchy(float c) {
#pragma clang fp contract(off)
    return -(double)GPR_LOGPI_F - (double)gpr_lcauchy_logf(c);
}
GPR_HD double gpr_lphi(float z) {
#pragma clang fp contract(off)
    return -0.5 * (double)z * (double)z - GPR_HALF_LOG2PI_D;
}
// y + a*x with two roundings (no FMA contraction), as the host computes it
GPR_HD float gpr_axpy(float a, float x, float y) {
#pragma clang fp contract(off)
    return y + a * x;
}
// -dl*lnsym + s with two roundings
GPR_HD double gpr_dl_prior(int dl, double lnsym, double s) {
#pragma clang fp contract(off)
    return -(double)dl * lnsym + s;
}

// ---------------------------------------------------------------- reference op sets
// The reference chooses an op set at compile time (main.cu:51-86); here it is a run-time
// table: the heap code of each enum value in enum order and the size NOPS.  SIMPLIFIED,
// POLYNOMIAL and LINEAR are the intended sets (their #if blocks do not compile as
// written: duplicate enumerators, a missing comma, a second enum OPS).
GPR_HD int gpr_opset_nops(int set) {
    return set == GPR_OPS_SIMPLIFIED ? 13 : set == GPR_OPS_POLYNOMIAL ? 11 : set == GPR_OPS_LINEAR ? 9
         : set == GPR_OPS_PHYSICAL ? 19 : GPR_REF_NOPS;
}
GPR_HD int gpr_opset_code(int set, int i) {
    // one 40-byte row per op set, in the order of GPR_OPS_* (a load, no selects: on the
    // device i is per lane)
    const char *t =
        // default
        "\x00\x01\x02\x03\x04\x05\x06\x07\x08\x09\x0a\x0b\x0c\x0d\x0e\x0f\x10\x11\x12\x13\x14\x15\x16\x17\x18\x19\x1a\x1b\x1c\x1d\x1e\x1f\x20\x21\x22\x23\x24\x25\x26\x27"
        // SIMPLIFIED
        "\x00\x01\x02\x03\x04\x05\x06\x07\x08\x09\x0a\x0b\x0c\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00"
        // POLYNOMIAL
        "\x00\x01\x03\x04\x05\x06\x07\x08\x09\x0a\x10\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00"
        // LINEAR
        "\x00\x01\x03\x04\x05\x06\x07\x08\x0a\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00"
        // PHYSICAL
        "\x1b\x1c\x1d\x1e\x1f\x20\x21\x22\x23\x24\x25\x26\x27\x03\x04\x05\x06\x07\x09\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00\x00";
    return (unsigned char)t[set * 40 + i];
}
// PHYSICAL_CONSTANTS values (main.cu:429-441): the data_t value of each double literal;
// log 2 is log(2.0) in double, then float
GPR_HD bool gpr_ref_is_physical(int o) { return o == GPR_REF_LOGOF2 || (o >= GPR_REF_E && o <= GPR_REF_KB); }
GPR_HD float gpr_ref_physical_value(int o) {
    switch (o) {
    case GPR_REF_E: return (float)1.6021766208e-19;
    case GPR_REF_PI: return (float)1.14472988584915;
    case GPR_REF_CLIGHT: return (float)19.5186009865452;
    case GPR_REF_G: return (float)-23.4302046558627;
    case GPR_REF_HBAR: return (float)-78.2347583540371;
    case GPR_REF_MU0: return (float)-13.5870714043359;
    case GPR_REF_EL: return (float)-43.2777536741305;
    case GPR_REF_MP: return (float)-61.6554051167679;
    case GPR_REF_ME: return (float)-69.1708328401338;
    case GPR_REF_E0: return (float)-25.4501305691713;
    case GPR_REF_KB: return (float)-52.6369038080531;
    default: return (float)0.69314718055994530942;   // GPR_REF_LOGOF2: log(2.0)
    }
}

// ---------------------------------------------------------------- config generation table
// 64 buckets of the top 6 bits of a Philox word: X 16, CONST 16, ADD SUB MUL DIV POW
// EXP LOG NEG 4 each.
GPR_HD int gpr_gen_symbol(uint32_t r) {
    int k = (int)(r >> 26);
    if (k < 16) return GPR_OP_X;
    if (k < 32) return GPR_OP_CONST;
    const int tail = (k - 32) >> 2;  // 0..7
    const int order[8] = {GPR_OP_ADD, GPR_OP_SUB, GPR_OP_MUL, GPR_OP_DIV, GPR_OP_POW,
                          GPR_OP_EXP, GPR_OP_LOG, GPR_OP_NEG};
    return order[tail];
}
#define GPR_LOG_QUARTER (-0x1.62e42fefa39efp+0)
#define GPR_LOG_SIXTEENTH (-0x1.62e42fefa39efp+1)
GPR_HD double gpr_gen_logp(int op) {
    if (op == GPR_OP_X || op == GPR_OP_CONST) return GPR_LOG_QUARTER;
    switch (op) {
    case GPR_OP_ADD: case GPR_OP_SUB: case GPR_OP_MUL: case GPR_OP_DIV: case GPR_OP_POW:
    case GPR_OP_EXP: case GPR_OP_LOG: case GPR_OP_NEG: return GPR_LOG_SIXTEENTH;
    default: return -__builtin_inf();
    }
}


// Device building blocks shared by the config and reference kernels: wave-level
// primitives (DPP reductions and scans, readlane/writelane), FP32 op semantics, the
// wave-uniform tape interpreter with its VGPR value stack, the Gaussian
// log-likelihood and description-length prior of a wave-resident program.
#pragma once
#include "gpr_internal.h"

#define GPR_INF (__builtin_inf())

// ---------------------------------------------------------------- wave primitives
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }
__device__ __forceinline__ int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rl_f(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double rl_d(double v, int l) {
    uint64_t b = __builtin_bit_cast(uint64_t, v);
    int lo = __builtin_amdgcn_readlane((int)(uint32_t)b, l);
    int hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ int wl_i(int old, int v, int l) { return lane_id() == l ? v : old; }
__device__ __forceinline__ float wl_f(float old, float v, int l) { return lane_id() == l ? v : old; }
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni_d(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned long long hi = (uint32_t)uni((int)(b >> 32)), lo = (uint32_t)uni((int)(uint32_t)b);
    return __longlong_as_double((long long)((hi << 32) | lo));
}
// A dataset record with every field wave-uniform (scalar registers): a record loaded by
// the whole wave otherwise lands in vector registers, where values used at every step
// take registers from the step loop.
__device__ __forceinline__ GprDataset uni_dataset(const GprDataset &g) {
    GprDataset d;
    d.n = uni(g.n);
    d.off = uni(g.off);
    d.chain0 = uni(g.chain0);
    d.nchains = uni(g.nchains);
    d.cap = uni(g.cap);
    d.kc = uni(g.kc);
    d.lconst = uni_d(g.lconst);
    d.wconst = uni_d(g.wconst);
    return d;
}
// A copy of a wave-uniform value that the compiler cannot prove equal to the original:
// an address computed from it at a late use is computed there, instead of being kept
// live (and spilled) from an earlier use of the same address.
__device__ __forceinline__ uint32_t opaque_u32(uint32_t v) { __asm__ volatile("" : "+s"(v)); return v; }
__device__ __forceinline__ int opaque_lane_i32(int v) { __asm__ volatile("" : "+v"(v)); return v; }

// Lane gathers that stay in wave-uniform control flow: lane l receives v of lane
// src mod 64 (ds_bpermute reads the byte address' lane bits, so no index masking is
// needed).  ds_bpermute only sees the values of lanes active when it executes; the
// empty asm pins the result at this point so the compiler cannot sink the gather into a
// branch that deactivates source lanes.
__device__ __forceinline__ int shfl_all(int v, int src) {
    int r = __builtin_amdgcn_ds_bpermute(src << 2, v);
    __asm__ volatile("" : "+v"(r));
    return r;
}
__device__ __forceinline__ float shfl_all(float v, int src) {
    return __builtin_bit_cast(float, shfl_all(__builtin_bit_cast(int, v), src));
}
__device__ __forceinline__ double shfl_all_d(double v, int src) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)shfl_all((int)(uint32_t)b, src), hi = (uint32_t)shfl_all((int)(uint32_t)(b >> 32), src);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// DPP controls (gfx9 encoding): quad_perm, row_shr:n, row_mirror, row_half_mirror,
// row_bcast:15, row_bcast:31.  DPP moves are VALU operand modifiers: no LDS round trip.
#define DPP_QP_1032 0xB1
#define DPP_QP_2301 0x4E
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_MIRROR 0x140
#define DPP_ROW_HALF_MIRROR 0x141
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, false);
}

// Sum over the 64 lanes, returned uniform.  Butterflies inside each 16-lane row, then
// the four row sums combined in a fixed order.
__device__ __forceinline__ float wave_sum_f(float v) {
    v += dpp_f<DPP_QP_1032>(v);
    v += dpp_f<DPP_QP_2301>(v);
    v += dpp_f<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_f<DPP_ROW_MIRROR>(v);
    return (rl_f(v, 0) + rl_f(v, 16)) + (rl_f(v, 32) + rl_f(v, 48));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, 0xf, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double wave_sum_d(double v) {
    v += dpp_d<DPP_QP_1032>(v);
    v += dpp_d<DPP_QP_2301>(v);
    v += dpp_d<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_d<DPP_ROW_MIRROR>(v);
    return (rl_d(v, 0) + rl_d(v, 16)) + (rl_d(v, 32) + rl_d(v, 48));
}
__device__ __forceinline__ int wave_max_i(int v) {
    v = max(v, dpp_i<DPP_QP_1032>(v));
    v = max(v, dpp_i<DPP_QP_2301>(v));
    v = max(v, dpp_i<DPP_ROW_HALF_MIRROR>(v));
    v = max(v, dpp_i<DPP_ROW_MIRROR>(v));
    return max(max(rl_i(v, 0), rl_i(v, 16)), max(rl_i(v, 32), rl_i(v, 48)));
}
// Inclusive prefix sum over the 64 lanes: Kogge-Stone inside rows (row_shr with zero
// fill), then row 15 -> row 1,3 and lane 31 -> rows 2,3 broadcasts.
__device__ __forceinline__ int wave_scan_i(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(1), 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(2), 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(4), 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(8), 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_BCAST15, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_BCAST31, 0xc, 0xf, false);
    return v;
}

// ---------------------------------------------------------------- op semantics (FP32)
// Values of dispatch_eval (main.cu:397-444) with IEEE division and accurate libm.
__device__ __forceinline__ float op_bin(int op, float a, float b) {
    switch (op) {
    case GPR_OP_ADD: return a + b;
    case GPR_OP_SUB: return a - b;
    case GPR_OP_MUL: return a * b;
    case GPR_OP_DIV: return a / b;
    case GPR_OP_POW: return powf(a, b);
    case GPR_OP_RSUB: return b - a;
    case GPR_OP_RDIV: return b / a;
    default: return powf(b, a);  // RPOW
    }
}
__device__ __forceinline__ float op_un(int op, float a, float c) {
    switch (op) {
    case GPR_OP_NEG: return -a;
    case GPR_OP_EXP: return expf(a);
    case GPR_OP_LOG: return logf(a);
    case GPR_OP_INV: return 1.0f / a;
    case GPR_OP_SQRT: return sqrtf(a);
    case GPR_OP_SIN: return sinf(a);
    case GPR_OP_ASIN: return asinf(a);
    case GPR_OP_COS: return cosf(a);
    case GPR_OP_ACOS: return acosf(a);
    case GPR_OP_TAN: return tanf(a);
    case GPR_OP_ATAN: return atanf(a);
    case GPR_OP_ABS: return fabsf(a);
    case GPR_OP_CADD: return a + c;
    case GPR_OP_CMUL: return c * a;
    case GPR_OP_CPOW: return powf(a, c);
    case GPR_OP_CRPOW: return powf(c, a);
    default: return __builtin_nanf("");
    }
}

// ---------------------------------------------------------------- wave program state
// Lane l holds token l in (o0, c0) and token l+64 in (o1, c1); tokens >= len are 0.
struct WaveProg {
    int o0, o1;
    float c0, c1;
    int len;
};

// Per-lane value stack: KC vector values of GPR_MAX_DEPTH floats.  Element access
// with the uniform stack pointer lowers to s_set_gpr_idx moves; nothing goes through
// an alloca, so nothing is spilled to scratch.
typedef float gpr_f16v __attribute__((ext_vector_type(GPR_MAX_DEPTH)));
template <int KC>
struct VStack {
    gpr_f16v s0, s1, s2, s3;
    __device__ __forceinline__ float get(int k, int i) const {
        return k == 0 ? s0[i] : (k == 1 ? s1[i] : (k == 2 ? s2[i] : s3[i]));
    }
    __device__ __forceinline__ void put(int k, int i, float v) {
        if (k == 0) s0[i] = v; else if (k == 1) s1[i] = v; else if (k == 2) s2[i] = v; else s3[i] = v;
    }
};

#define GPR_KLOOP(expr)                              \
    do {                                             \
        _Pragma("unroll") for (int k = 0; k < KC; k++) T[k] = (expr); \
    } while (0)

// Interpret the tape at KC points per lane.  op is wave-uniform: every branch below is
// a scalar branch and only the selected operation executes.  Taken branches cost far
// more than VALU work here, so the two common token classes take one branch each and
// no nested switch: leaves (x, constant, one) select their value, and +, -, * compute
// a + s*b (s = +-1, exact, so one rounding like a +- b) or a*b and select.  Every other
// op goes through the full switch.
template <int KC, bool HI>
__device__ __forceinline__ void wave_eval(const WaveProg &p, const float (&xs)[KC], float (&T)[KC]) {
    VStack<KC> S;
    int sp = 0;
#pragma unroll
    for (int k = 0; k < KC; k++) T[k] = 0.0f;
    for (int t = 0; t < p.len; t++) {
        const int l = t & 63;
        const bool h = HI && t >= 64;
        const int op = h ? rl_i(p.o1, l) : rl_i(p.o0, l);
        if (op <= GPR_OP_ONE) {
            // the write index is bounded for malformed tapes; valid ones (depth <= 16) never hit it
            const int wi = min(sp, GPR_MAX_DEPTH - 1);
#pragma unroll
            for (int k = 0; k < KC; k++) S.put(k, wi, T[k]);
            sp = min(sp + 1, GPR_MAX_DEPTH);
            const float cv = h ? rl_f(p.c1, l) : rl_f(p.c0, l);
            const float sv = op == GPR_OP_CONST ? cv : 1.0f;
            const bool isx = op == GPR_OP_X;
            GPR_KLOOP(isx ? xs[k] : sv);
        } else if (op <= GPR_OP_MUL) {
            sp = max(sp - 1, 0);
            const bool mul = op == GPR_OP_MUL;
            const float sgn = op == GPR_OP_SUB ? -1.0f : 1.0f;
#pragma unroll
            for (int k = 0; k < KC; k++) {
                const float a = S.get(k, sp);
                const float pm = __builtin_fmaf(sgn, T[k], a);
                const float pr = a * T[k];
                T[k] = mul ? pr : pm;
            }
        } else if (op <= GPR_OP_RPOW) {
            sp = max(sp - 1, 0);
            switch (op) {
            case GPR_OP_DIV: GPR_KLOOP(S.get(k, sp) / T[k]); break;
            case GPR_OP_RSUB: GPR_KLOOP(T[k] - S.get(k, sp)); break;
            case GPR_OP_RDIV: GPR_KLOOP(T[k] / S.get(k, sp)); break;
            case GPR_OP_POW: GPR_KLOOP(powf(S.get(k, sp), T[k])); break;
            default: GPR_KLOOP(powf(T[k], S.get(k, sp))); break;
            }
        } else {
            const float c = op >= GPR_OP_CADD ? (h ? rl_f(p.c1, l) : rl_f(p.c0, l)) : 0.0f;
            switch (op) {
            case GPR_OP_NEG: GPR_KLOOP(-T[k]); break;
            case GPR_OP_EXP: GPR_KLOOP(expf(T[k])); break;
            case GPR_OP_LOG: GPR_KLOOP(logf(T[k])); break;
            case GPR_OP_INV: GPR_KLOOP(1.0f / T[k]); break;
            case GPR_OP_SQRT: GPR_KLOOP(sqrtf(T[k])); break;
            case GPR_OP_SIN: GPR_KLOOP(sinf(T[k])); break;
            case GPR_OP_ASIN: GPR_KLOOP(asinf(T[k])); break;
            case GPR_OP_COS: GPR_KLOOP(cosf(T[k])); break;
            case GPR_OP_ACOS: GPR_KLOOP(acosf(T[k])); break;
            case GPR_OP_TAN: GPR_KLOOP(tanf(T[k])); break;
            case GPR_OP_ATAN: GPR_KLOOP(atanf(T[k])); break;
            case GPR_OP_ABS: GPR_KLOOP(fabsf(T[k])); break;
            case GPR_OP_CADD: GPR_KLOOP(T[k] + c); break;
            case GPR_OP_CMUL: GPR_KLOOP(c * T[k]); break;
            case GPR_OP_CPOW: GPR_KLOOP(powf(T[k], c)); break;
            case GPR_OP_CRPOW: GPR_KLOOP(powf(c, T[k])); break;
            default: GPR_KLOOP(__builtin_nanf("")); break;
            }
        }
    }
}

// The same interpreter with the value stack in LDS: the top two entries stay in VGPRs
// (T, U) and the rest live in this wave's slab ls[level][lane][KC] (GPR_MAX_DEPTH + 2
// levels; levels 0 and 1 take the stores and loads of an empty U, so the pushes and
// pops need no test).  A push spills U to the slab, a binary op refills U from it; both
// are one 16-byte LDS access per lane at KC = 4.  Saves the 16*KC VGPRs of the indexed
// stack (occupancy) and the indexing-mode switches.  Same operations on the same
// operands as wave_eval, so the values are bit-identical.  The tape must be well formed
// with stack depth at most GPR_MAX_DEPTH: the sampler's programs are (init, proposals
// and gpr_write_state check it), and the stack pointer is not clamped.
template <int KC>
__device__ __forceinline__ void ls_store(float *q, const float (&v)[KC]) {
    if (KC == 4) *(float4 *)q = make_float4(v[0], v[1], v[2], v[3]);
    else if (KC == 2) *(float2 *)q = make_float2(v[0], v[KC > 1 ? 1 : 0]);
    else q[0] = v[0];
}
template <int KC>
__device__ __forceinline__ void ls_load(const float *q, float (&v)[KC]) {
    if (KC == 4) { const float4 f = *(const float4 *)q; v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w; }
    else if (KC == 2) { const float2 f = *(const float2 *)q; v[0] = f.x; v[KC > 1 ? 1 : 0] = f.y; }
    else v[0] = q[0];
}

// Token classes as wave-uniform bit masks (bit t of the low word, bit t-64 of the high
// word): the common classes dispatch on a scalar bit test, with no vector-to-scalar
// read of the op on their path.
// (x and constant leaves; +, -, *; every other op takes the generic path)
struct OpMasks {
    uint64_t leaf[2], x[2], cheap[2], mul[2], sub[2];
};
template <bool HI>
__device__ __forceinline__ OpMasks op_masks(const WaveProg &p) {
    const int l = lane_id();
    OpMasks m;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const bool v = (h == 0 || HI) && l + 64 * h < p.len;
        const int o = h ? p.o1 : p.o0;
        m.leaf[h] = ballot(v && o <= GPR_OP_CONST);
        m.x[h] = ballot(v && o == GPR_OP_X);
        m.cheap[h] = ballot(v && o >= GPR_OP_ADD && o <= GPR_OP_MUL);
        m.mul[h] = ballot(v && o == GPR_OP_MUL);
        m.sub[h] = ballot(v && o == GPR_OP_SUB);
    }
    return m;
}

template <int KC, bool HI>
__device__ __forceinline__ void wave_eval_ls(const WaveProg &p, const float (&xs)[KC], float (&T)[KC], float *ls,
                                             const OpMasks &m) {
    float U[KC];
    float *lane = ls + 2 * 64 * KC + lane_id() * KC;   // level -2 of the stack below T, U
    int sp = 0;   // entries on the stack, T and U included
#pragma unroll
    for (int k = 0; k < KC; k++) { T[k] = 0.0f; U[k] = 0.0f; }
    for (int t = 0; t < p.len; t++) {
        const int l = t & 63;
        const int h = (HI && t >= 64) ? 1 : 0;
        const uint64_t bit = 1ull << l;
        // a leaf followed by +, - or * within the same half: push and pop cancel, so
        // T = T op leaf with the stack untouched (the same operation on the same operands)
        if (m.leaf[h] & (m.cheap[h] >> 1) & bit) {
            // the leaf and the op are wave-uniform: scalar branches pick one vector
            // instruction per point (the same operation on the same operands)
            if (m.x[h] & bit) {
                if ((m.mul[h] >> 1) & bit) {
#pragma unroll
                    for (int k = 0; k < KC; k++) T[k] = T[k] * xs[k];
                } else if ((m.sub[h] >> 1) & bit) {
#pragma unroll
                    for (int k = 0; k < KC; k++) T[k] = __builtin_fmaf(-1.0f, xs[k], T[k]);
                } else {
#pragma unroll
                    for (int k = 0; k < KC; k++) T[k] = __builtin_fmaf(1.0f, xs[k], T[k]);
                }
            } else {
                const float cv = h ? rl_f(p.c1, l) : rl_f(p.c0, l);
                if ((m.mul[h] >> 1) & bit) {
#pragma unroll
                    for (int k = 0; k < KC; k++) T[k] = T[k] * cv;
                } else if ((m.sub[h] >> 1) & bit) {
#pragma unroll
                    for (int k = 0; k < KC; k++) T[k] = __builtin_fmaf(-1.0f, cv, T[k]);
                } else {
#pragma unroll
                    for (int k = 0; k < KC; k++) T[k] = __builtin_fmaf(1.0f, cv, T[k]);
                }
            }
            t++;
            continue;
        }
        if (m.leaf[h] & bit) {
            // level sp-2 receives U (an unused level while U is empty)
            ls_store<KC>(lane + (sp - 2) * 64 * KC, U);
#pragma unroll
            for (int k = 0; k < KC; k++) U[k] = T[k];
            sp++;
            const float cv = h ? rl_f(p.c1, l) : rl_f(p.c0, l);
            const bool isx = (m.x[h] & bit) != 0;
            GPR_KLOOP(isx ? xs[k] : cv);
        } else if (m.cheap[h] & bit) {
            if (m.mul[h] & bit) {
#pragma unroll
                for (int k = 0; k < KC; k++) T[k] = U[k] * T[k];
            } else if (m.sub[h] & bit) {
#pragma unroll
                for (int k = 0; k < KC; k++) T[k] = __builtin_fmaf(-1.0f, T[k], U[k]);
            } else {
#pragma unroll
                for (int k = 0; k < KC; k++) T[k] = __builtin_fmaf(1.0f, T[k], U[k]);
            }
            ls_load<KC>(lane + (sp - 3) * 64 * KC, U);   // the new second entry (level sp-3)
            sp--;
            continue;
        } else {
        const int op = h ? rl_i(p.o1, l) : rl_i(p.o0, l);
        if (op <= GPR_OP_ONE) {
            ls_store<KC>(lane + (sp - 2) * 64 * KC, U);
#pragma unroll
            for (int k = 0; k < KC; k++) U[k] = T[k];
            sp++;
            const float cv = h ? rl_f(p.c1, l) : rl_f(p.c0, l);
            const float sv = op == GPR_OP_CONST ? cv : 1.0f;
            const bool isx = op == GPR_OP_X;
            GPR_KLOOP(isx ? xs[k] : sv);
        } else if (op <= GPR_OP_RPOW) {
            if (op <= GPR_OP_MUL) {
                const bool mul = op == GPR_OP_MUL;
                const float sgn = op == GPR_OP_SUB ? -1.0f : 1.0f;
#pragma unroll
                for (int k = 0; k < KC; k++) {
                    const float pm = __builtin_fmaf(sgn, T[k], U[k]);
                    const float pr = U[k] * T[k];
                    T[k] = mul ? pr : pm;
                }
            } else {
                switch (op) {
                case GPR_OP_DIV: GPR_KLOOP(U[k] / T[k]); break;
                case GPR_OP_RSUB: GPR_KLOOP(T[k] - U[k]); break;
                case GPR_OP_RDIV: GPR_KLOOP(T[k] / U[k]); break;
                case GPR_OP_POW: GPR_KLOOP(powf(U[k], T[k])); break;
                default: GPR_KLOOP(powf(T[k], U[k])); break;
                }
            }
            ls_load<KC>(lane + (sp - 3) * 64 * KC, U);   // the new second entry (level sp-3)
            sp--;
        } else {
            const float c = op >= GPR_OP_CADD ? (h ? rl_f(p.c1, l) : rl_f(p.c0, l)) : 0.0f;
            switch (op) {
            case GPR_OP_NEG: GPR_KLOOP(-T[k]); break;
            case GPR_OP_EXP: GPR_KLOOP(expf(T[k])); break;
            case GPR_OP_LOG: GPR_KLOOP(logf(T[k])); break;
            case GPR_OP_INV: GPR_KLOOP(1.0f / T[k]); break;
            case GPR_OP_SQRT: GPR_KLOOP(sqrtf(T[k])); break;
            case GPR_OP_SIN: GPR_KLOOP(sinf(T[k])); break;
            case GPR_OP_ASIN: GPR_KLOOP(asinf(T[k])); break;
            case GPR_OP_COS: GPR_KLOOP(cosf(T[k])); break;
            case GPR_OP_ACOS: GPR_KLOOP(acosf(T[k])); break;
            case GPR_OP_TAN: GPR_KLOOP(tanf(T[k])); break;
            case GPR_OP_ATAN: GPR_KLOOP(atanf(T[k])); break;
            case GPR_OP_ABS: GPR_KLOOP(fabsf(T[k])); break;
            case GPR_OP_CADD: GPR_KLOOP(T[k] + c); break;
            case GPR_OP_CMUL: GPR_KLOOP(c * T[k]); break;
            case GPR_OP_CPOW: GPR_KLOOP(powf(T[k], c)); break;
            case GPR_OP_CRPOW: GPR_KLOOP(powf(c, T[k])); break;
            default: GPR_KLOOP(__builtin_nanf("")); break;
            }
        }
        }
    }
}

// Dataset points of one chunk: point index base + k*64 + lane.
template <int KC>
struct Chunk {
    float x[KC], y[KC], w[KC];
    __device__ __forceinline__ void load(const float *dx, const float *dy, const float *dw, int n, int base) {
        const int l = opaque_lane_i32(lane_id());   // the point's offset is computed here, not kept live
#pragma unroll
        for (int k = 0; k < KC; k++) {
            const int i = base + k * 64 + l;
            const bool v = i < n;
            x[k] = v ? dx[i] : 0.0f;
            y[k] = v ? dy[i] : 0.0f;
            w[k] = v ? dw[i] : 0.0f;
        }
    }
};

// Gaussian log-likelihood of the wave's program: lconst - sum_i (f_i - y_i)^2 w_i.
// Per-lane Kahan sums, wave tree reduction; -inf if any term is not finite.
// LS: the value stack in LDS, ls = this wave's slab of (GPR_MAX_DEPTH + 2) * 64 * KC floats,
// pm = the program's token-class masks; else the VGPR stack.
// reject_below: the caller rejects the proposal whenever the LL lies below this value.
// Every term of the sum is >= 0, so after each chunk but the last the sum so far bounds
// the LL from above; when that bound (with a relative margin of 1e-4 on the partial sum,
// far above its FP32 rounding) is below reject_below, or the partial sum is not finite,
// the evaluation stops and returns -inf: the proposal is rejected either way.  *points
// receives the number of data points evaluated.
// Work-counter values of a program: operator tokens + 4 (the per-point FP32 ops of its
// evaluation and Gaussian term) and its constant-carrying tokens.
template <bool HI>
__device__ __forceinline__ void prog_counts(const WaveProg &p, int &opq, int &nc) {
    const int l = lane_id();
    const int nleaf = __popcll(ballot(l < p.len && p.o0 <= GPR_OP_ONE)) +
                      (HI ? __popcll(ballot(l + 64 < p.len && p.o1 <= GPR_OP_ONE)) : 0);
    nc = __popcll(ballot(l < p.len && gpr_has_const(p.o0))) +
         (HI ? __popcll(ballot(l + 64 < p.len && gpr_has_const(p.o1))) : 0);
    opq = p.len - nleaf + 4;
}

// The words of a Philox substream (chain, step, purpose, slot) as gpr_rstream yields
// them, computed on the vector unit 64 at a time: lane k holds word 64 b + k (block
// 16 b + k / 4, word k % 4) of the current batch b, and next() reads them in order with
// v_readlane.  One vector Philox evaluation replaces up to 16 scalar ones, which takes
// the structure move's draws off the scalar unit.
struct WaveStream {
    uint64_t seed;
    uint32_t chain, step, c2;
    uint32_t w;     // this lane's word of the current batch
    int idx, base;  // next word within the batch, batch index
    __device__ __forceinline__ void fill() {
        const int k = lane_id();
        const gpr_u4 b = gpr_philox(chain, step, c2, (uint32_t)(16 * base + (k >> 2)), (uint32_t)seed,
                                    (uint32_t)(seed >> 32));
        w = gpr_u4_word(b, k & 3);
    }
    __device__ __forceinline__ void init(uint64_t s, uint32_t ch, uint32_t st, uint32_t purpose, uint32_t slot) {
        seed = s; chain = ch; step = st; c2 = (purpose << 24) | (slot & 0xffffffu);
        idx = 0; base = 0;
        fill();
    }
    __device__ __forceinline__ uint32_t next() {
        if (idx == 64) { base++; idx = 0; fill(); }
        return (uint32_t)rl_i((int)w, idx++);
    }
};

// Constant tokens of a program.
template <bool HI>
__device__ __forceinline__ int const_count(const WaveProg &p) {
    const int l = lane_id();
    return __popcll(ballot(l < p.len && p.o0 == GPR_OP_CONST)) +
           (HI ? __popcll(ballot(l + 64 < p.len && p.o1 == GPR_OP_CONST)) : 0);
}

// Random-walk moves of a program can be batched (four steps per Philox block) when
// BATCH holds, the chain has no proposal metric (mk = 0), the tape fits the low half (at
// most 64 tokens), it has at most 16 constant tokens and no constant-operand token.
template <bool BATCH, bool HI>
__device__ __forceinline__ bool batch_ok(const WaveProg &p, int mk) {
    if (!BATCH || mk != 0 || (HI && p.len > 64)) return false;
    const int l = lane_id();
    const uint64_t m = ballot(l < p.len && p.o0 == GPR_OP_CONST);
    return __popcll(m) <= 16 && ballot(l < p.len && gpr_has_const(p.o0)) == m;
}

template <int KC, bool HI, bool LS = false>
__device__ __forceinline__ double wave_loglik(const WaveProg &p, const GprDataset &d, const float *dx, const float *dy,
                              const float *dw, const Chunk<KC> &c0, int nchunks, float *ls = nullptr,
                              const OpMasks *pm = nullptr, double reject_below = -GPR_INF, int *points = nullptr) {
    float acc = 0.0f, comp = 0.0f;
    const int l = lane_id();
    if (points) *points = d.n;
    for (int ch = 0; ch < nchunks; ch++) {
        Chunk<KC> c;
        if (ch == 0) c = c0; else c.load(dx, dy, dw, d.n, ch * 64 * KC);
        float T[KC];
        if (LS) wave_eval_ls<KC, HI>(p, c.x, T, ls, *pm);
        else wave_eval<KC, HI>(p, c.x, T);
#pragma unroll
        for (int k = 0; k < KC; k++) {
            const float r = T[k] - c.y[k];
            const float term = (ch * 64 * KC + k * 64 + l < d.n) ? (r * r) * c.w[k] : 0.0f;
            const float yv = term - comp;
            const float tv = acc + yv;
            comp = (tv - acc) - yv;
            acc = tv;
        }
        if (ch + 1 < nchunks && reject_below > -GPR_INF) {
            const double lb = (double)wave_sum_f(acc - comp) * (1.0 - 1e-4);
            const double ub = d.lconst - (d.wconst > 0.0 ? d.wconst * lb : lb);
            if (!(ub >= reject_below)) {
                if (points) *points = (ch + 1) * 64 * KC;
                return -GPR_INF;
            }
        }
    }
    // each lane's compensated partial, combined across the wave in FP64
    const double s = wave_sum_d((double)acc - (double)comp);
    if (!__builtin_isfinite(s)) return -GPR_INF;
    return d.lconst - (d.wconst > 0.0 ? d.wconst * s : s);
}

// ---------------------------------------------------------------- forward-mode gradient
// The MALA proposals of short programs with few constants get their gradient in the same
// interpreter pass as their likelihood: every stack entry carries the value and its
// tangents with respect to the program's constant tokens (constant of rank j: tangent j),
// W = 1 + NT floats, pushed to the wave's LDS stack as one 16-byte store (NT =
// GPR_DUAL_NT = 3, and NT = 2 in the same layout), one 8-byte store (NT = 1, programs of a
// single constant), for programs of four constants (NT = GPR_DUAL_NT4) a 16-byte store of
// the value and the first three tangents and a 4-byte store of the fourth into a second
// plane, and for five or six (NT = GPR_DUAL_NT6) a 16-byte, an 8-byte and a 4-byte store
// into three planes.  Programs of up to GPR_DUAL_NTMAX constants that fit none of these take
// two passes of four tangents each (constants of rank 0-3, then 4-7: the second pass only
// when the first completes).  The values take the same
// operations on the same operands as wave_eval_ls, so the LL is
// bit-identical to it (and early rejection stops a pass at the same chunk); each lane
// adds the per-point gradient (-2 r w) f' of its points, and one wave sum per constant
// ends the pass.  Programs qualify when they use only the config grammar (x, constant,
// one, + - * / pow exp log neg), hold no constant-operand token, and hold at most
// GPR_DUAL_NT constant tokens within stack depth GPR_DUAL_DEPTH or at most GPR_DUAL_NTMAX
// within GPR_DUAL_DEPTH4 (dual_nt: one pass of GPR_DUAL_NT4 tangents for four constants,
// of GPR_DUAL_NT6 for five or six within GPR_DUAL_DEPTH6, else two passes).
#define GPR_DUAL_NT 3
#define GPR_DUAL_NT4 4
#define GPR_DUAL_NT6 6
#define GPR_DUAL_NTMAX 8
#define GPR_DUAL_DEPTH 9
#define GPR_DUAL_DEPTH4 8
#define GPR_DUAL_DEPTH6 6
#define GPR_DUAL_W (1 + GPR_DUAL_NT)
// Stack levels: a push at stack size sp stores U at level sp - 2 and a binary op at sp
// loads level sp - 3, so a program of depth at most D uses levels -1 .. D - 3 (D - 1 of
// them), stored from offset 0 (level -1): token 0, always a leaf, starts the stack in T
// without the store of level -2 (U holds no entry yet).  Levels -1 and below hold no
// stack entry (the store at sp = 1 and the load at sp = 2 move the empty U).
#define GPR_DUAL_PACKED_FLOATS ((GPR_DUAL_DEPTH - 1) * 64 * GPR_DUAL_W)
#define GPR_DUAL_SPLIT_FLOATS ((GPR_DUAL_DEPTH4 - 1) * 64 * (GPR_DUAL_NT4 + 1))
#define GPR_DUAL_SPLIT6_FLOATS ((GPR_DUAL_DEPTH6 - 1) * 64 * (GPR_DUAL_NT6 + 1))
// floats of the dual stack of one wave
#define GPR_DUAL_LDS_FLOATS \
    (GPR_DUAL_PACKED_FLOATS > GPR_DUAL_SPLIT_FLOATS                                      \
         ? (GPR_DUAL_PACKED_FLOATS > GPR_DUAL_SPLIT6_FLOATS ? GPR_DUAL_PACKED_FLOATS : GPR_DUAL_SPLIT6_FLOATS) \
         : (GPR_DUAL_SPLIT_FLOATS > GPR_DUAL_SPLIT6_FLOATS ? GPR_DUAL_SPLIT_FLOATS : GPR_DUAL_SPLIT6_FLOATS))

// One stack entry of the dual stack at level lv >= -2 (lane, laneb, lanec: this lane's
// level 0 in each plane): packed W floats per lane; for NT = 4 four floats per lane in the
// first plane and the last tangent in laneb's; for NT = 6 also two in lanec's.
template <int NT>
__device__ __forceinline__ void dual_st(float *lane, float *laneb, float *lanec, int lv, const float (&v)[1 + NT]) {
    if constexpr (NT == GPR_DUAL_NT6) {
        *(float4 *)(lane + lv * 256) = make_float4(v[0], v[1], v[2], v[3]);
        *(float2 *)(lanec + lv * 128) = make_float2(v[4], v[5]);
        laneb[lv * 64] = v[NT];
    } else if constexpr (NT == GPR_DUAL_NT4) {
        *(float4 *)(lane + lv * 256) = make_float4(v[0], v[1], v[2], v[3]);
        laneb[lv * 64] = v[NT];
    } else if constexpr (NT == 2) {
        *(float4 *)(lane + lv * 256) = make_float4(v[0], v[1], v[NT], 0.0f);   // the NT = 3 layout
    } else {
        ls_store<1 + NT>(lane + lv * 64 * (1 + NT), v);
    }
}
template <int NT>
__device__ __forceinline__ void dual_ld(const float *lane, const float *laneb, const float *lanec, int lv,
                                        float (&v)[1 + NT]) {
    if constexpr (NT == GPR_DUAL_NT6) {
        const float4 f = *(const float4 *)(lane + lv * 256);
        const float2 e = *(const float2 *)(lanec + lv * 128);
        v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
        v[4] = e.x; v[5] = e.y;
        v[NT] = laneb[lv * 64];
    } else if constexpr (NT == GPR_DUAL_NT4) {
        const float4 f = *(const float4 *)(lane + lv * 256);
        v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
        v[NT] = laneb[lv * 64];
    } else if constexpr (NT == 2) {
        const float4 f = *(const float4 *)(lane + lv * 256);
        v[0] = f.x; v[1] = f.y; v[NT] = f.z;
    } else {
        ls_load<1 + NT>(lane + lv * 64 * (1 + NT), v);
    }
}

// a^(b-1), the partial derivative of a^b in a over b, given v = a^b: v / a where a is
// finite and non-zero and v finite on every lane of the wave (a division instead of a
// second pow), else the pow itself.
__device__ __forceinline__ float pow_bm1(float a, float b, float v) {
    if (ballot(a == 0.0f || !__builtin_isfinite(a) || !__builtin_isfinite(v))) return powf(a, b - 1.0f);
    return v / a;
}

// A zero tangent contributes nothing, whatever the partial derivative (an infinite or NaN
// partial of a constant-free operand must not reach the constants, as in the adjoint):
// the vector unit's legacy multiply, whose product is +0 when either factor is 0 and the
// IEEE product otherwise, one instruction for the test, the product and the select.
__device__ __forceinline__ float tmul(float t, float partial) {
    float r;
    __asm__("v_mul_legacy_f32 %0, %1, %2" : "=v"(r) : "v"(t), "v"(partial));
    return r;
}

// tangents of the forward-mode pass of a program of k constant tokens and stack depth dmax
// (GPR_DUAL_NTMAX: two passes of GPR_DUAL_NT4)
__device__ __forceinline__ int dual_nt(int k, int dmax) {
    return k <= 2 ? (k <= 1 ? 1 : 2)
                  : (k <= GPR_DUAL_NT ? GPR_DUAL_NT
                                      : (k <= GPR_DUAL_NT4 ? GPR_DUAL_NT4
                                                           : (k <= GPR_DUAL_NT6 && dmax <= GPR_DUAL_DEPTH6 ? GPR_DUAL_NT6
                                                                                                            : GPR_DUAL_NTMAX)));
}

template <bool HI>
__device__ __forceinline__ bool dual_ok(const WaveProg &p, int dmax) {
    constexpr uint32_t GRAMMAR = (1u << GPR_OP_X) | (1u << GPR_OP_CONST) | (1u << GPR_OP_ONE) | (1u << GPR_OP_ADD) |
                                 (1u << GPR_OP_SUB) | (1u << GPR_OP_MUL) | (1u << GPR_OP_DIV) | (1u << GPR_OP_POW) |
                                 (1u << GPR_OP_EXP) | (1u << GPR_OP_LOG) | (1u << GPR_OP_NEG);
    const int l = lane_id();
    const bool v0 = l < p.len, v1 = HI && l + 64 < p.len;
    const bool bad = (v0 && !((GRAMMAR >> (p.o0 & 31)) & 1u)) || (v1 && !((GRAMMAR >> (p.o1 & 31)) & 1u));
    const int k = const_count<HI>(p);
    return ballot(bad) == 0 && ((k <= GPR_DUAL_NT && dmax <= GPR_DUAL_DEPTH) || (k <= GPR_DUAL_NTMAX && dmax <= GPR_DUAL_DEPTH4));
}

// One token (two for a leaf fused with the next +, - or *) of the forward-mode
// interpreter, tokens of half H of the tape (t < 64 for H = 0): the half is a
// compile-time constant, so the token-class masks and the tape words are read without
// per-token selects.  NT tangents (1, 2, GPR_DUAL_NT, GPR_DUAL_NT4 or GPR_DUAL_NT6) for the
// constants of rank j0 .. j0 + NT - 1.  lane, laneb: this lane's level 0 of the stack (dual_st).
template <int H, int NT>
__device__ __forceinline__ void dual_token(const WaveProg &p, const OpMasks &m, const uint64_t (&cm)[2], int nlo,
                                           int j0, float xv, float *lane, float *laneb, float *lanec, float (&T)[1 + NT],
                                           float (&U)[1 + NT], int &sp, int &t) {
    constexpr int W = 1 + NT;
    const int li = t & 63;
    constexpr int h = H;
    const uint64_t bit = 1ull << li;
    if (m.leaf[h] & bit) {
        // the leaf's value and tangent: x (0), or constant of rank j0 + rk (e_rk)
        const bool isx = (m.x[h] & bit) != 0;
        const float cv = h ? rl_f(p.c1, li) : rl_f(p.c0, li);
        const int rk = isx ? -1 : (h ? nlo : 0) + __popcll(cm[h] & (bit - 1)) - j0;
        const float b = isx ? xv : cv;
        if (m.cheap[h] & (bit << 1)) {
            // leaf followed by +, - or *: T = T op leaf, stack untouched
            const uint64_t nb = bit << 1;
            if (m.mul[h] & nb) {
                const float tv = T[0];
                T[0] = T[0] * b;
#pragma unroll
                for (int j = 1; j < W; j++) T[j] = tmul(T[j], b) + (rk == j - 1 ? tv : 0.0f);
            } else {
                const float sgn = (m.sub[h] & nb) ? -1.0f : 1.0f;
                T[0] = __builtin_fmaf(sgn, b, T[0]);
#pragma unroll
                for (int j = 1; j < W; j++) T[j] = rk == j - 1 ? T[j] + sgn : T[j];
            }
            t++;
            return;
        }
        dual_st<NT>(lane, laneb, lanec, sp - 2, U);
#pragma unroll
        for (int j = 0; j < W; j++) U[j] = T[j];
        sp++;
        T[0] = b;
#pragma unroll
        for (int j = 1; j < W; j++) T[j] = rk == j - 1 ? 1.0f : 0.0f;
        return;
    }
    if (m.cheap[h] & bit) {
        // U op T for +, -, *
        if (m.mul[h] & bit) {
            const float uv = U[0], tv = T[0];
            T[0] = U[0] * T[0];
#pragma unroll
            for (int j = 1; j < W; j++) T[j] = tmul(U[j], tv) + tmul(T[j], uv);
        } else {
            const float sgn = (m.sub[h] & bit) ? -1.0f : 1.0f;
            T[0] = __builtin_fmaf(sgn, T[0], U[0]);
#pragma unroll
            for (int j = 1; j < W; j++) T[j] = __builtin_fmaf(sgn, T[j], U[j]);
        }
        dual_ld<NT>(lane, laneb, lanec, sp - 3, U);
        sp--;
        return;
    }
    const int op = h ? rl_i(p.o1, li) : rl_i(p.o0, li);
    if (op == GPR_OP_ONE) {
        dual_st<NT>(lane, laneb, lanec, sp - 2, U);
#pragma unroll
        for (int j = 0; j < W; j++) { U[j] = T[j]; T[j] = j == 0 ? 1.0f : 0.0f; }
        sp++;
    } else if (op == GPR_OP_DIV) {
        const float a = U[0], b = T[0];
        const float q = a / b;
        const float ib = 1.0f / b, nq = -(a / (b * b));
#pragma unroll
        for (int j = 1; j < W; j++) T[j] = tmul(U[j], ib) + tmul(T[j], nq);
        T[0] = q;
        dual_ld<NT>(lane, laneb, lanec, sp - 3, U);
        sp--;
    } else if (op == GPR_OP_POW) {
        const float a = U[0], b = T[0];
        const float v = powf(a, b);
        const float pa = b * pow_bm1(a, b, v), pb = v * logf(a);
#pragma unroll
        for (int j = 1; j < W; j++) T[j] = tmul(U[j], pa) + tmul(T[j], pb);
        T[0] = v;
        dual_ld<NT>(lane, laneb, lanec, sp - 3, U);
        sp--;
    } else if (op == GPR_OP_EXP) {
        const float v = expf(T[0]);
#pragma unroll
        for (int j = 1; j < W; j++) T[j] = tmul(T[j], v);
        T[0] = v;
    } else if (op == GPR_OP_LOG) {
        const float a = T[0];
        const float ia = 1.0f / a;
        T[0] = logf(a);
#pragma unroll
        for (int j = 1; j < W; j++) T[j] = tmul(T[j], ia);
    } else {   // GPR_OP_NEG (dual_ok admits no other op)
#pragma unroll
        for (int j = 0; j < W; j++) T[j] = -T[j];
    }
}

// LL of the wave's program (as wave_loglik<1, HI, true>) and, when the pass completes,
// gs[j] = sum over points of (-2 r w) df/dc_{j0+j} for the constant token of rank j0 + j
// (token order; cm = the constant-token masks).  ls: the wave's dual stack region.
// NT: tangents carried (1, 2, GPR_DUAL_NT, GPR_DUAL_NT4 or GPR_DUAL_NT6), at least the
// program's constant count less j0 unless a second pass follows; gs[j] is 0 for j >= NT
// and for ranks past the program's constants.
template <bool HI, int NT>
__device__ __forceinline__ double wave_loglik_dual(const WaveProg &p, const GprDataset &d, const float *dx, const float *dy,
                                   const float *dw, const Chunk<1> &c0, int nchunks, float *ls, const OpMasks &m,
                                   const uint64_t (&cm)[2], double reject_below, int *points, int j0,
                                   float (&gs)[GPR_DUAL_NTMAX]) {
    static_assert(NT == 1 || NT == 2 || NT == GPR_DUAL_NT || NT == GPR_DUAL_NT4 || NT == GPR_DUAL_NT6, "tangent count");
    constexpr int W = 1 + NT;
    const int l = lane_id();
    // this lane's level 0 of the stack below T, U (dual_st); level -1 lies below it
    // NT = 4: planes of 4 and 1 floats, GPR_DUAL_DEPTH4 - 1 levels each; NT = 6: planes of
    // 4, 2 and 1 floats, GPR_DUAL_DEPTH6 - 1 levels each
    constexpr int LV = NT == GPR_DUAL_NT6 ? GPR_DUAL_DEPTH6 - 1 : GPR_DUAL_DEPTH4 - 1;
    float *lane = (NT >= GPR_DUAL_NT4 || NT == 2) ? ls + 256 + l * 4 : ls + 64 * W + l * W;
    float *lanec = ls + LV * 256 + 128 + l * 2;
    float *laneb = ls + LV * (NT == GPR_DUAL_NT6 ? 384 : 256) + 64 + l;
    float acc = 0.0f, comp = 0.0f;
    float ga[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) ga[j] = 0.0f;
    const bool folded = d.wconst > 0.0;
    const float wf = (float)d.wconst;
    const int nlo = __popcll(cm[0]);
    if (points) *points = d.n;
    for (int ch = 0; ch < nchunks; ch++) {
        Chunk<1> c;
        if (ch == 0) c = c0; else c.load(dx, dy, dw, d.n, ch * 64);
        const float xv = c.x[0];
        float T[W], U[W];
        // token 0 (x, a constant of rank 0 or one) starts the stack in T
        const bool t0x = (m.x[0] & 1ull) != 0, t0c = (cm[0] & 1ull) != 0;
        T[0] = t0x ? xv : (t0c ? rl_f(p.c0, 0) : 1.0f);
#pragma unroll
        for (int j = 1; j < W; j++) T[j] = (t0c && j0 == 0 && j == 1) ? 1.0f : 0.0f;
#pragma unroll
        for (int j = 0; j < W; j++) U[j] = 0.0f;
        int sp = 1;
        // tokens of the low half, then (two-register tapes) the high half
        const int lo_end = HI ? min(p.len, 64) : p.len;
        for (int t = 1; t < lo_end; t++) dual_token<0, NT>(p, m, cm, nlo, j0, xv, lane, laneb, lanec, T, U, sp, t);
        if (HI)
            for (int t = 64; t < p.len; t++) dual_token<1, NT>(p, m, cm, nlo, j0, xv, lane, laneb, lanec, T, U, sp, t);
        const int i = ch * 64 + l;
        const bool valid = i < d.n;
        const float r = T[0] - c.y[0];
        const float term = valid ? (r * r) * c.w[0] : 0.0f;
        const float yv = term - comp;
        const float tv = acc + yv;
        comp = (tv - acc) - yv;
        acc = tv;
        const float G = (-2.0f * r) * (folded ? wf : c.w[0]);
#pragma unroll
        for (int j = 0; j < NT; j++) ga[j] = valid ? __builtin_fmaf(G, T[j + 1], ga[j]) : ga[j];
        if (ch + 1 < nchunks && reject_below > -GPR_INF) {
            const double lb = (double)wave_sum_f(acc - comp) * (1.0 - 1e-4);
            const double ub = d.lconst - (d.wconst > 0.0 ? d.wconst * lb : lb);
            if (!(ub >= reject_below)) {
                if (points) *points = (ch + 1) * 64;
                return -GPR_INF;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < GPR_DUAL_NTMAX; j++)
        gs[j] = (j < NT && j0 + j < nlo + (HI ? __popcll(cm[1]) : 0)) ? wave_sum_f(ga[j < NT ? j : 0]) : 0.0f;
    const double s = wave_sum_d((double)acc - (double)comp);
    if (!__builtin_isfinite(s)) return -GPR_INF;
    return d.lconst - (d.wconst > 0.0 ? d.wconst * s : s);
}

// npass (1 or 2) passes of GPR_DUAL_NT4 tangents: the constants of rank 0-3, then, when the
// first pass completes, those of rank 4-7 (its LL is the first's: same values, same
// order).
template <bool HI>
__device__ __forceinline__ double dual_passes4(const WaveProg &p, const GprDataset &d, const float *dx, const float *dy,
                                               const float *dw, const Chunk<1> &c0, int nchunks, float *ls,
                                               const OpMasks &m, const uint64_t (&cm)[2], double reject_below,
                                               int *points, int npass, float (&gs)[GPR_DUAL_NTMAX]) {
    const double ll =
        wave_loglik_dual<HI, GPR_DUAL_NT4>(p, d, dx, dy, dw, c0, nchunks, ls, m, cm, reject_below, points, 0, gs);
    if (npass > 1 && __builtin_isfinite(ll)) {
        float g[GPR_DUAL_NTMAX];
        (void)wave_loglik_dual<HI, GPR_DUAL_NT4>(p, d, dx, dy, dw, c0, nchunks, ls, m, cm, -GPR_INF, nullptr,
                                                 GPR_DUAL_NT4, g);
#pragma unroll
        for (int j = 0; j < GPR_DUAL_NT4; j++) gs[GPR_DUAL_NT4 + j] = g[j];
    }
    return ll;
}

// The gradient of the log posterior in the lane layout of wave_loglik_grad: lane t of g0
// (token t < 64) or g1 holds d/dc_t (from gs by the constant's rank) plus the Cauchy
// prior's -2c/(1+c^2); ok is false when one of them is not finite.
template <bool HI>
__device__ __forceinline__ void dual_grad_lanes(const WaveProg &p, const uint64_t (&cm)[2],
                                                const float (&gs)[GPR_DUAL_NTMAX], double &g0, double &g1, bool &ok) {
    const int l = lane_id();
    const uint64_t below = (1ull << l) - 1ull;
    const int r0 = __popcll(cm[0] & below);
    const int r1 = __popcll(cm[0]) + (HI ? __popcll(cm[1] & below) : 0);
    const bool c0 = (cm[0] >> l) & 1ull, c1 = HI && ((cm[1] >> l) & 1ull);
    float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int j = 0; j < GPR_DUAL_NTMAX; j++) {
        s0 = r0 == j ? gs[j] : s0;
        s1 = r1 == j ? gs[j] : s1;
    }
    g0 = c0 ? (double)s0 - 2.0 * (double)p.c0 / (1.0 + (double)p.c0 * (double)p.c0) : 0.0;
    g1 = c1 ? (double)s1 - 2.0 * (double)p.c1 / (1.0 + (double)p.c1 * (double)p.c1) : 0.0;
    ok = ballot((c0 && !__builtin_isfinite(g0)) || (c1 && !__builtin_isfinite(g1))) == 0;
}

// ---------------------------------------------------------------- speculative constant steps
// Datasets of at most 32 points fill a fraction of a wave.  The random-walk kernel then
// evaluates the proposals of G consecutive constant steps in one interpreter pass, one
// per lane group (G = 4 groups of 16 lanes for n <= 16, 2 of 32 for n <= 32): lane l
// takes point l mod (64 / G) of the proposal of step u0 + l / (64 / G).  All of them are
// proposals from the same current state, which is the state each step sees as long as
// the steps before it reject; the kernel uses them in step order and drops the rest
// after an accept.  The dispatch of every token serves the G proposals.
//
// Proposal constants come from the draw batch: the constant of rank r of step u is lane
// ks u + r of cb.  The value of each proposal is bit-identical to its own evaluation by
// wave_loglik<1, false>: the same operations per point, a one-chunk Kahan sum (the term
// itself), and a group sum whose FP64 additions are those of wave_sum_d over the lanes
// it would occupy alone (rows of 16 lanes summed by the same DPP steps; a group of two
// rows adds them as wave_sum_d adds rows 0 and 1, and adding the zero rows of a lone
// evaluation changes nothing).  ls: the wave's value-stack region (KC = 1 layout).
// out[g] receives the LL of group g's proposal, for g < gn (groups g >= gn repeat the
// proposal of group gn - 1).
__device__ __forceinline__ void wave_loglik_groups(const WaveProg &p, const GprDataset &d, const float *dx,
                                                   const float *dy, const float *dw, float *ls, const OpMasks &m,
                                                   uint64_t cm, float cb, int ks, int u0, int G, int gn,
                                                   double (&out)[4]) {
    const int l = opaque_lane_i32(lane_id());       // what follows is computed here, not kept live
    const int lgs = G == 4 ? 4 : 5;                  // log2 of the group size
    const int gl = l >> lgs;
    const int i = l & ((1 << lgs) - 1);
    const bool valid = i < d.n;
    const float xv = valid ? dx[i] : 0.0f, yv = valid ? dy[i] : 0.0f, wv = valid ? dw[i] : 0.0f;
    const int src = ks * (u0 + min(gl, gn - 1));    // lane of cb holding rank 0 of this group's step
    float *lane = ls + 2 * 64 + l;                   // level -2 of the stack below T, U
    float T = 0.0f, U = 0.0f;
    int sp = 0;
    for (int t = 0; t < p.len; t++) {
        const uint64_t bit = 1ull << t;
        if (m.leaf[0] & bit) {
            const bool isx = (m.x[0] & bit) != 0;
            float b = xv;
            if (!isx) b = shfl_all(cb, src + __popcll(cm & (bit - 1)));
            if (m.cheap[0] & (bit << 1)) {
                // leaf followed by +, - or *: T = T op leaf, stack untouched
                const uint64_t nb = bit << 1;
                const bool mul = (m.mul[0] & nb) != 0;
                const float sgn = (m.sub[0] & nb) ? -1.0f : 1.0f;
                const float pm = __builtin_fmaf(sgn, b, T), pr = T * b;
                T = mul ? pr : pm;
                t++;
                continue;
            }
            lane[(sp - 2) * 64] = U;
            U = T;
            sp++;
            T = b;
            continue;
        }
        if (m.cheap[0] & bit) {
            const bool mul = (m.mul[0] & bit) != 0;
            const float sgn = (m.sub[0] & bit) ? -1.0f : 1.0f;
            const float pm = __builtin_fmaf(sgn, T, U), pr = U * T;
            T = mul ? pr : pm;
            U = lane[(sp - 3) * 64];
            sp--;
            continue;
        }
        const int op = rl_i(p.o0, t);
        if (op == GPR_OP_ONE) {
            lane[(sp - 2) * 64] = U;
            U = T;
            sp++;
            T = 1.0f;
        } else if (op <= GPR_OP_RPOW) {
            switch (op) {
            case GPR_OP_DIV: T = U / T; break;
            case GPR_OP_RSUB: T = T - U; break;
            case GPR_OP_RDIV: T = T / U; break;
            case GPR_OP_POW: T = powf(U, T); break;
            default: T = powf(T, U); break;
            }
            U = lane[(sp - 3) * 64];
            sp--;
        } else {
            switch (op) {
            case GPR_OP_NEG: T = -T; break;
            case GPR_OP_EXP: T = expf(T); break;
            case GPR_OP_LOG: T = logf(T); break;
            case GPR_OP_INV: T = 1.0f / T; break;
            case GPR_OP_SQRT: T = sqrtf(T); break;
            case GPR_OP_SIN: T = sinf(T); break;
            case GPR_OP_ASIN: T = asinf(T); break;
            case GPR_OP_COS: T = cosf(T); break;
            case GPR_OP_ACOS: T = acosf(T); break;
            case GPR_OP_TAN: T = tanf(T); break;
            case GPR_OP_ATAN: T = atanf(T); break;
            case GPR_OP_ABS: T = fabsf(T); break;
            default: T = __builtin_nanf(""); break;   // constant-operand ops never reach here
            }
        }
    }
    const float r = T - yv;
    const float term = valid ? (r * r) * wv : 0.0f;
    double v = (double)term;
    v += dpp_d<DPP_QP_1032>(v);
    v += dpp_d<DPP_QP_2301>(v);
    v += dpp_d<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_d<DPP_ROW_MIRROR>(v);
#pragma unroll
    for (int g = 0; g < 4; g++) {
        double s;
        if (G == 4) s = rl_d(v, 16 * g);
        else s = g < 2 ? rl_d(v, 32 * g) + rl_d(v, 32 * g + 16) : 0.0;
        out[g] = !__builtin_isfinite(s) ? -GPR_INF : d.lconst - (d.wconst > 0.0 ? d.wconst * s : s);
    }
}

// Description-length prior plus Cauchy(0,1) on every constant token, summed in token
// order in FP64 (the same order as the oracle, so the value is bit-reproducible).
template <bool HI>
__device__ __forceinline__ double wave_prior(const WaveProg &p, double lnsym) {
    const int l = lane_id();
    const bool v0 = l < p.len, v1 = HI && (l + 64) < p.len;
    int dl = p.len + __popcll(ballot(v0 && p.o0 >= GPR_OP_CADD)) + (HI ? __popcll(ballot(v1 && p.o1 >= GPR_OP_CADD)) : 0);
    uint64_t m0 = ballot(v0 && gpr_has_const(p.o0));
    const double lc0 = gpr_lcauchy(p.c0);
    double s = 0.0;
    while (m0) {
        const int b = __builtin_ctzll(m0);
        m0 &= m0 - 1;
        s += rl_d(lc0, b);
    }
    if (HI) {
        uint64_t m1 = ballot(v1 && gpr_has_const(p.o1));
        if (m1) {
            const double lc1 = gpr_lcauchy(p.c1);
            while (m1) {
                const int b = __builtin_ctzll(m1);
                m1 &= m1 - 1;
                s += rl_d(lc1, b);
            }
        }
    }
    return gpr_dl_prior(dl, lnsym, s);
}

// Subtree [s0, j] ending at token j: s0-1 is the last u < j whose stack depth after u is
// below the depth after j.  Also returns the maximum depth of the tape.
template <bool HI>
__device__ __forceinline__ void wave_depths(const WaveProg &p, int &D0, int &D1, int &dmax) {
    const int l = lane_id();
    const int d0 = l < p.len ? 1 - gpr_arity(p.o0) : 0;
    D0 = wave_scan_i(d0);
    dmax = wave_max_i(l < p.len ? D0 : 0);
    D1 = 0;
    if (HI) {
        const int tot = rl_i(D0, 63);
        const int d1 = l + 64 < p.len ? 1 - gpr_arity(p.o1) : 0;
        D1 = wave_scan_i(d1) + tot;
        dmax = max(dmax, wave_max_i(l + 64 < p.len ? D1 : 0));
    }
}

template <bool HI>
__device__ __forceinline__ int subtree_start(int j, int D0, int D1) {
    const int l = lane_id();
    const int Dj = (HI && j >= 64) ? rl_i(D1, j - 64) : rl_i(D0, j);
    uint64_t m1 = HI ? ballot(l + 64 < j && D1 < Dj) : 0;
    if (m1) return 64 + (63 - __builtin_clzll(m1)) + 1;
    uint64_t m0 = ballot(l < j && D0 < Dj);
    if (m0) return (63 - __builtin_clzll(m0)) + 1;
    return 0;
}

// ---------------------------------------------------------------- subtree generation
// A structure move draws its new subtree one symbol per lane.  Symbol k of the
// right-child-first prefix order takes word 0 of the Philox block at counter
// {chain, step, GPR_P_STRUCT << 24, k} and, when it is CONST, its value from word 1 of
// that block; word 2 of block 0 picks the node to replace.  The subtree ends at the
// first symbol after which no operand is missing: a prefix sum of (arity - 1) over the
// lanes finds it.
//
// gpr_gen_symbol per lane: the top 6 bits of a word pick X (16 of 64 buckets), CONST
// (16), then ADD SUB MUL DIV POW EXP LOG NEG (4 each), packed one code per byte.
__device__ __forceinline__ int gen_symbol_lane(uint32_t r) {
    constexpr uint64_t ORDER = (uint64_t)GPR_OP_ADD | ((uint64_t)GPR_OP_SUB << 8) | ((uint64_t)GPR_OP_MUL << 16) |
                               ((uint64_t)GPR_OP_DIV << 24) | ((uint64_t)GPR_OP_POW << 32) |
                               ((uint64_t)GPR_OP_EXP << 40) | ((uint64_t)GPR_OP_LOG << 48) |
                               ((uint64_t)GPR_OP_NEG << 56);
    const int k = (int)(r >> 26);
    const int op = (int)((ORDER >> (8 * ((k >> 2) & 7))) & 0xff);
    return k < 16 ? GPR_OP_X : (k < 32 ? GPR_OP_CONST : op);
}
// gpr_gen_logp per lane: log 1/4 for the leaves, log 1/16 for the generated operators,
// -inf for ops the table never draws.
__device__ __forceinline__ double gen_logp_lane(int op) {
    constexpr uint32_t LEAF = (1u << GPR_OP_X) | (1u << GPR_OP_CONST);
    constexpr uint32_t GEN = (1u << GPR_OP_ADD) | (1u << GPR_OP_SUB) | (1u << GPR_OP_MUL) | (1u << GPR_OP_DIV) |
                             (1u << GPR_OP_POW) | (1u << GPR_OP_EXP) | (1u << GPR_OP_LOG) | (1u << GPR_OP_NEG);
    const uint32_t b = 1u << (op & 31);
    return (b & LEAF) ? GPR_LOG_QUARTER : ((b & GEN) ? GPR_LOG_SIXTEENTH : -__builtin_inf());
}
// generation term of one token: its symbol's log probability and, for CONST, the
// standard normal log density of its value
__device__ __forceinline__ double gen_term_lane(int op, float c) {
    return gen_logp_lane(op) + (op == GPR_OP_CONST ? gpr_lphi(c) : 0.0);
}

// value of a 128-token lane array at token index src (src may differ per lane)
__device__ __forceinline__ int gather_i(int v0, int v1, int src, bool hi) {
    const int a = shfl_all(v0, src & 63);
    if (!hi) return a;
    const int b = shfl_all(v1, src & 63);
    return src >= 64 ? b : a;
}
__device__ __forceinline__ float gather_f(float v0, float v1, int src, bool hi) {
    const float a = shfl_all(v0, src & 63);
    if (!hi) return a;
    const float b = shfl_all(v1, src & 63);
    return src >= 64 ? b : a;
}

// ---------------------------------------------------------------- reverse-mode adjoint
// One point per lane (KC = 1), no value stack: the forward sweep stores the value of
// every operator token in this wave's LDS slab lv[t * 64 + lane]; leaves (x, constant,
// one) are not stored, their values come from the token, and a binary token finds its
// left operand through a per-structure plan.  The reverse sweep walks the tape backwards
// carrying the adjoint of the current token.  Reverse postfix order visits a node, then
// its right subtree (ending at t-1), then its left subtree, so a binary token carries its
// right operand's adjoint and leaves its left operand's in a mailbox: the slot of the
// first token of its own subtree, always a leaf, which the left subtree's root reads
// when the sweep arrives there (the next token after a leaf).  Pending mailboxes are
// distinct slots (nested subtrees pending at once start at distinct tokens), and a
// constant leaf's mailbox is read before the sweep reaches the leaf itself, whose slot
// then receives the leaf's per-point gradient; a token with a constant operand (c+, c*,
// ^c, c^) overwrites its value slot with its gradient once it is done.

// Per-structure plan of the two sweeps: token-class masks (bit t of word 0, bit t-64 of
// word 1); for every token t (lane t of the *0 words, lane t-64 of the *1 words) the left
// operand lc of a binary token and the first token st of the subtree ending at t.
struct GradPlan {
    uint64_t leaf[2], x[2], cnst[2], cheap[2], mul[2], sub[2];
    int lc0, lc1, st0, st1;
};

template <bool HI>
__device__ __forceinline__ GradPlan grad_plan(const WaveProg &p) {
    const int l = lane_id();
    GradPlan g;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const bool v = (h == 0 || HI) && l + 64 * h < p.len;
        const int o = h ? p.o1 : p.o0;
        g.leaf[h] = ballot(v && o <= GPR_OP_ONE);
        g.x[h] = ballot(v && o == GPR_OP_X);
        g.cnst[h] = ballot(v && o == GPR_OP_CONST);
        g.cheap[h] = ballot(v && o >= GPR_OP_ADD && o <= GPR_OP_MUL);
        g.mul[h] = ballot(v && o == GPR_OP_MUL);
        g.sub[h] = ballot(v && o == GPR_OP_SUB);
    }
    // replay the stack on token indices: lane k of idx / sdx holds the token whose value
    // is at level k and the first token of its subtree (levels are bounded for malformed
    // tapes; valid ones stay below GPR_MAX_DEPTH)
    int idx = 0, sdx = 0, sp = 0;
    g.lc0 = 0;
    g.lc1 = 0;
    g.st0 = 0;
    g.st1 = 0;
    for (int t = 0; t < p.len; t++) {
        const int li = t & 63;
        const bool h = HI && t >= 64;
        const int ar = gpr_arity(h ? rl_i(p.o1, li) : rl_i(p.o0, li));
        int lev, st = t;
        if (ar == 0) {
            lev = min(sp, 63);
            sp = min(sp + 1, 64);
        } else {
            lev = max(sp - ar, 0);
            st = rl_i(sdx, lev);
            if (ar == 2) {
                const int lc = rl_i(idx, lev);
                if (h) g.lc1 = l == li ? lc : g.lc1; else g.lc0 = l == li ? lc : g.lc0;
                sp = max(sp - 1, 1);
            }
        }
        if (h) g.st1 = l == li ? st : g.st1; else g.st0 = l == li ? st : g.st0;
        idx = l == lev ? t : idx;
        sdx = l == lev ? st : sdx;
    }
    return g;
}

// value of leaf token u (x, its constant, or one) at this lane's point
template <bool HI>
__device__ __forceinline__ float leaf_val(const WaveProg &p, const GradPlan &g, int u, float x) {
    const int li = u & 63;
    const int h = (HI && u >= 64) ? 1 : 0;
    const uint64_t bit = 1ull << li;
    const float c = h ? rl_f(p.c1, li) : rl_f(p.c0, li);
    return (g.x[h] & bit) ? x : ((g.cnst[h] & bit) ? c : 1.0f);
}
// value of token u: a leaf from the token, an operator from the slab (lvl = lv + lane)
template <bool HI>
__device__ __forceinline__ float tok_val(const WaveProg &p, const GradPlan &g, int u, float x, const float *lvl) {
    const int h = (HI && u >= 64) ? 1 : 0;
    const float s = lvl[u * 64];
    return (g.leaf[h] & (1ull << (u & 63))) ? leaf_val<HI>(p, g, u, x) : s;
}
template <bool HI>
__device__ __forceinline__ int left_of(const GradPlan &g, int t) {
    return (HI && t >= 64) ? rl_i(g.lc1, t & 63) : rl_i(g.lc0, t & 63);
}
template <bool HI>
__device__ __forceinline__ int start_of(const GradPlan &g, int t) {
    return (HI && t >= 64) ? rl_i(g.st1, t & 63) : rl_i(g.st0, t & 63);
}

// Forward sweep: the same operations on the same operands as wave_eval, so the value is
// bit-identical.  A leaf followed by +, - or * in the same half is applied at once.
template <bool HI>
__device__ __forceinline__ float grad_forward(const WaveProg &p, const GradPlan &g, float x, float *lvl) {
    float T = 0.0f;
    for (int t = 0; t < p.len; t++) {
        const int li = t & 63;
        const int h = (HI && t >= 64) ? 1 : 0;
        const uint64_t bit = 1ull << li;
        if (g.leaf[h] & bit) {
            const float b = leaf_val<HI>(p, g, t, x);
            const uint64_t nb = bit << 1;
            if (g.cheap[h] & nb) {
                const bool mul = (g.mul[h] & nb) != 0;
                const float sgn = (g.sub[h] & nb) ? -1.0f : 1.0f;
                const float pm = __builtin_fmaf(sgn, b, T), pr = T * b;
                T = mul ? pr : pm;
                t++;
                lvl[t * 64] = T;
            } else {
                T = b;
            }
            continue;
        }
        if (g.cheap[h] & bit) {
            const float a = tok_val<HI>(p, g, left_of<HI>(g, t), x, lvl);
            const bool mul = (g.mul[h] & bit) != 0;
            const float sgn = (g.sub[h] & bit) ? -1.0f : 1.0f;
            const float pm = __builtin_fmaf(sgn, T, a), pr = a * T;
            T = mul ? pr : pm;
        } else {
            const int op = h ? rl_i(p.o1, li) : rl_i(p.o0, li);
            if (op <= GPR_OP_RPOW) {
                T = op_bin(op, tok_val<HI>(p, g, left_of<HI>(g, t), x, lvl), T);
            } else {
                const float c = op >= GPR_OP_CADD ? (h ? rl_f(p.c1, li) : rl_f(p.c0, li)) : 0.0f;
                T = op_un(op, T, c);
            }
        }
        lvl[t * 64] = T;
    }
    return T;
}

// Reverse sweep from adjoint G of the program's value.
template <bool HI>
__device__ __forceinline__ void grad_reverse(const WaveProg &p, const GradPlan &g, float x, float *lvl, float G) {
    for (int t = p.len - 1; t >= 0; t--) {
        const int li = t & 63;
        const int h = (HI && t >= 64) ? 1 : 0;
        const uint64_t bit = 1ull << li;
        if (g.leaf[h] & bit) {
            if (g.cnst[h] & bit) lvl[t * 64] = G;
            // the next token roots a left operand: its adjoint waits in its mailbox
            if (t > 0) G = lvl[start_of<HI>(g, t - 1) * 64];
            continue;
        }
        float ga, gb;
        if (g.cheap[h] & bit) {
            const bool mul = (g.mul[h] & bit) != 0;
            ga = G;
            gb = (g.sub[h] & bit) ? -G : G;
            const uint64_t rb = bit >> 1;
            if (g.leaf[h] & rb) {
                // right operand is a leaf: its adjoint is final now, and the left operand's
                // subtree ends at t-2 and comes next, so its adjoint is carried
                if (mul) {
                    const float b = leaf_val<HI>(p, g, t - 1, x);
                    const float a = tok_val<HI>(p, g, max(t - 2, 0), x, lvl);
                    ga = G * b;
                    gb = G * a;
                }
                if (g.cnst[h] & rb) lvl[(t - 1) * 64] = gb;
                G = ga;
                t--;
                continue;
            }
            if (mul) {
                const float a = tok_val<HI>(p, g, left_of<HI>(g, t), x, lvl);
                const float b = tok_val<HI>(p, g, t - 1, x, lvl);
                ga = G * b;
                gb = G * a;
            }
        } else {
            const int op = h ? rl_i(p.o1, li) : rl_i(p.o0, li);
            const float v = lvl[t * 64];
            if (op <= GPR_OP_RPOW) {
                const float a = tok_val<HI>(p, g, left_of<HI>(g, t), x, lvl);
                const float b = tok_val<HI>(p, g, t - 1, x, lvl);
                switch (op) {
                case GPR_OP_DIV: ga = G / b; gb = -(G * a) / (b * b); break;
                case GPR_OP_RSUB: ga = -G; gb = G; break;
                case GPR_OP_RDIV: gb = G / a; ga = -(G * b) / (a * a); break;
                case GPR_OP_POW: ga = G * b * pow_bm1(a, b, v); gb = G * v * logf(a); break;
                default: gb = G * a * pow_bm1(b, a, v); ga = G * v * logf(b); break;  // RPOW: b^a
                }
            } else {
                const float a = tok_val<HI>(p, g, t - 1, x, lvl);
                const float c = op >= GPR_OP_CADD ? (h ? rl_f(p.c1, li) : rl_f(p.c0, li)) : 0.0f;
                float gc = 0.0f;
                switch (op) {
                case GPR_OP_NEG: ga = -G; break;
                case GPR_OP_EXP: ga = G * v; break;
                case GPR_OP_LOG: ga = G / a; break;
                case GPR_OP_INV: ga = -G / (a * a); break;
                case GPR_OP_SQRT: ga = G * 0.5f / v; break;
                case GPR_OP_SIN: ga = G * cosf(a); break;
                case GPR_OP_COS: ga = -G * sinf(a); break;
                case GPR_OP_TAN: ga = G * (1.0f + v * v); break;
                case GPR_OP_ASIN: ga = G / sqrtf(1.0f - a * a); break;
                case GPR_OP_ACOS: ga = -G / sqrtf(1.0f - a * a); break;
                case GPR_OP_ATAN: ga = G / (1.0f + a * a); break;
                case GPR_OP_ABS: ga = a < 0.0f ? -G : G; break;
                case GPR_OP_CADD: ga = G; gc = G; break;
                case GPR_OP_CMUL: ga = G * c; gc = G * a; break;
                case GPR_OP_CPOW: ga = G * c * pow_bm1(a, c, v); gc = G * v * logf(a); break;
                case GPR_OP_CRPOW: ga = G * v * logf(c); gc = G * a * pow_bm1(c, a, v); break;
                default: ga = __builtin_nanf(""); break;
                }
                if (op >= GPR_OP_CADD) lvl[t * 64] = gc;
                G = ga;
                continue;
            }
        }
        // binary: the left operand's adjoint goes to the mailbox, the right one is carried
        lvl[start_of<HI>(g, t) * 64] = ga;
        G = gb;
    }
}

// FP32 wave sums of the per-point gradients in the slots of the constant-carrying tokens
// (lanes with valid points only), placed at lane t of s0 (t < 64) or s1.  Two tokens per
// pass so their reductions overlap.
template <bool HI>
__device__ __forceinline__ void grad_chunk_sums(uint64_t m0, uint64_t m1, const float *lvl, bool valid, float &s0,
                                                float &s1) {
    const int l = lane_id();
#pragma unroll
    for (int h = 0; h < (HI ? 2 : 1); h++) {
        uint64_t m = h ? m1 : m0;
        float s = 0.0f;
        while (m) {
            const int ta = __builtin_ctzll(m);
            m &= m - 1;
            const int tb = m ? __builtin_ctzll(m) : ta;
            if (m) m &= m - 1;
            const float va = lvl[(ta + 64 * h) * 64], vb = lvl[(tb + 64 * h) * 64];
            const float sa = wave_sum_f(valid ? va : 0.0f), sb = wave_sum_f(valid ? vb : 0.0f);
            s = l == ta ? sa : s;
            s = l == tb ? sb : s;
        }
        if (h) s1 = s; else s0 = s;
    }
}

// Log-likelihood and d(log posterior)/d(constant) of the wave's program (g: its plan).
// The LL is bit-identical to wave_loglik<KC, HI> for every KC: lane l visits points
// c*64 + l in increasing c, the same per-lane order and Kahan updates.  The gradient
// holds the wave sums of the per-point adjoints (FP32) accumulated over chunks in FP64,
// plus the Cauchy prior's -2c/(1+c^2); token t sits at lane t of g0 (t < 64) or g1.
// lv: this wave's LDS slab of at least len * 64 floats.  Returns false in ok when the
// LL or a constant's gradient is not finite.
template <bool HI>
__device__ __forceinline__ double wave_loglik_grad(const WaveProg &p, const GradPlan &g, const GprDataset &d, const float *dx,
                                   const float *dy, const float *dw, float *lv, double &g0, double &g1, bool &ok) {
    const int l = lane_id();
    float *lvl = lv + l;
    const bool c0 = l < p.len && gpr_has_const(p.o0), c1 = HI && l + 64 < p.len && gpr_has_const(p.o1);
    const uint64_t m0 = ballot(c0), m1 = HI ? ballot(c1) : 0;
    g0 = 0.0;
    g1 = 0.0;
    float acc = 0.0f, comp = 0.0f;
    const bool folded = d.wconst > 0.0;
    const float wf = (float)d.wconst;
    const int nch = (d.n + 63) / 64;
    for (int ch = 0; ch < nch; ch++) {
        const int i = ch * 64 + l;
        const bool valid = i < d.n;
        const float x = valid ? dx[i] : 0.0f, y = valid ? dy[i] : 0.0f, w = valid ? dw[i] : 0.0f;
        const float f = grad_forward<HI>(p, g, x, lvl);
        const float r = f - y;
        const float term = valid ? (r * r) * w : 0.0f;
        const float yv = term - comp;
        const float tv = acc + yv;
        comp = (tv - acc) - yv;
        acc = tv;
        // d/df of -(f - y)^2 w, with w = wconst when the weights are folded into it
        const float G = valid ? (-2.0f * r) * (folded ? wf : w) : 0.0f;
        grad_reverse<HI>(p, g, x, lvl, G);
        float s0 = 0.0f, s1 = 0.0f;
        grad_chunk_sums<HI>(m0, m1, lvl, valid, s0, s1);
        g0 += (double)s0;
        if (HI) g1 += (double)s1;
    }
    const double s = wave_sum_d((double)acc - (double)comp);
    if (c0) g0 += -2.0 * (double)p.c0 / (1.0 + (double)p.c0 * (double)p.c0);
    if (c1) g1 += -2.0 * (double)p.c1 / (1.0 + (double)p.c1 * (double)p.c1);
    const bool bad = ballot((c0 && !__builtin_isfinite(g0)) || (c1 && !__builtin_isfinite(g1))) != 0;
    ok = !bad && __builtin_isfinite(s);
    if (!__builtin_isfinite(s)) return -GPR_INF;
    return d.lconst - (folded ? d.wconst * s : s);
}

// FP64 sum of the Gaussian log-density terms -f^2 r2 (r2 = 1 / (2 e^2)) over the lanes where m holds, in
// token order (lanes of the low half, then the high half), as a serial loop over tokens
// would add them.  A half without such lanes computes no terms.
template <bool HI>
__device__ __forceinline__ double wave_gauss_sum_d(double f0, bool m0, double f1, bool m1, double r2) {
    double s = 0.0;
    uint64_t b0 = ballot(m0);
    if (b0) {
        const double v0 = -f0 * f0 * r2;
        while (b0) {
            const int b = __builtin_ctzll(b0);
            b0 &= b0 - 1;
            s += rl_d(v0, b);
        }
    }
    if (HI) {
        uint64_t b1 = ballot(m1);
        if (b1) {
            const double v1 = -f1 * f1 * r2;
            while (b1) {
                const int b = __builtin_ctzll(b1);
                b1 &= b1 - 1;
                s += rl_d(v1, b);
            }
        }
    }
    return s;
}

// ---------------------------------------------------------------- proposal metric
// Lanes 0..k-1 hold metric rows; the matrices are row-major in global memory, and each
// lane reads its row with 16-byte loads issued together, before the products.  Products
// accumulate in FP32 in increasing index order, the density's forward substitution in
// FP64, both without contraction (the oracle's operation order).

// lane l's row of a 16-wide float matrix, columns below k (lanes past k read row k-1)
__device__ __forceinline__ void metric_row(const float *A, int k, float (&r)[GPR_METRIC_MAX]) {
    // (the row address is computed here, not kept live across the steps)
    const float4 *q = (const float4 *)(A + min(opaque_lane_i32(lane_id()), k - 1) * GPR_METRIC_MAX);
#pragma unroll
    for (int j = 0; j < GPR_METRIC_MAX; j += 4) {
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (j < k) v = q[j / 4];
        r[j] = v.x; r[j + 1] = v.y; r[j + 2] = v.z; r[j + 3] = v.w;
    }
}

// d_i = sum_{j <= i} L_ij v_j, v and d at lanes 0..k-1
__device__ __forceinline__ float metric_Lv(const float *L, int k, float v) {
    const int l = lane_id();
    float r[GPR_METRIC_MAX];
    metric_row(L, k, r);
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < GPR_METRIC_MAX; j++) {
        if (j < k) {
            const float vj = rl_f(v, j);
            if (l < k && j <= l) acc = gpr_axpy(r[j], vj, acc);
        }
    }
    return acc;
}
// v_i = sum_j M_ij g_j, g and v at lanes 0..k-1
__device__ __forceinline__ float metric_Mv(const float *M, int k, float g) {
    const int l = lane_id();
    float r[GPR_METRIC_MAX];
    metric_row(M, k, r);
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < GPR_METRIC_MAX; j++) {
        if (j < k) {
            const float gj = rl_f(g, j);
            if (l < k) acc = gpr_axpy(r[j], gj, acc);
        }
    }
    return acc;
}
// log density of a Gaussian proposal with covariance e^2 L L^T at offset r (lanes
// 0..k-1), up to its constant: -|w|^2 / (2 e^2) with w = L^-1 r; |w|^2 is summed in row
// order, then divided once.  The row of Li is read in two halves of 8 columns.
__device__ __forceinline__ double metric_logq(const double *Li, int k, double r, double two_e2) {
#pragma clang fp contract(off)
    const int l = lane_id();
    const double2 *q = (const double2 *)(Li + min(opaque_lane_i32(l), k - 1) * GPR_METRIC_MAX);
    double w = 0.0;
#pragma unroll
    for (int half = 0; half < 2; half++) {
        if (half * 8 >= k) continue;
        double row[8];
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
            double2 v = make_double2(0.0, 0.0);
            if (half * 8 + j < k) v = q[(half * 8 + j) / 2];
            row[j] = v.x;
            row[j + 1] = v.y;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int jj = half * 8 + j;
            if (jj < k) {
                const double rj = rl_d(r, jj);
                if (l < k && jj <= l) w = w + row[j] * rj;
            }
        }
    }
    const double w2 = w * w;
    double s = 0.0;
    for (int i = 0; i < k; i++) s += rl_d(w2, i);
    return -s / two_e2;
}

// config accept rule: invalid current always moves; log u < a otherwise.  u lies on the
// grid ((k + 1/2) 2^-23), so log u >= -16.64: any a < -17 rejects without drawing.
// The accept draw of step s is word s & 3 of the ACCEPT block s >> 2, so one Philox
// block serves four consecutive steps: a chain loop keeps the last block it computed.
struct AcceptDraws {
    gpr_u4 w;
    uint32_t blk;
    __device__ __forceinline__ AcceptDraws() : w{0, 0, 0, 0}, blk(0xffffffffu) {}
};
__device__ __forceinline__ bool cfg_accept(uint64_t seed, uint32_t chain, uint32_t step, double cur, double prop,
                                           double a, AcceptDraws &ad) {
    if (!__builtin_isfinite(cur)) return true;
    if (!__builtin_isfinite(prop)) return false;
    if (a >= 0.0) return true;
    if (a < -17.0) return false;
    if ((step >> 2) != ad.blk) {
        ad.w = gpr_step_block(seed, chain, step, GPR_P_ACCEPT, 0, 0);
        ad.blk = step >> 2;
    }
    const float u = gpr_u01(gpr_u4_word(ad.w, (int)(step & 3)));
    return (double)gpr_logf(u) < a;
}
__device__ __forceinline__ bool cfg_accept(uint64_t seed, uint32_t chain, uint32_t step, double cur, double prop,
                                           double a) {
    AcceptDraws ad;
    return cfg_accept(seed, chain, step, cur, prop, a, ad);
}


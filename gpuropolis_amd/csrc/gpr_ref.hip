// Reference-semantics kernels: the reference's own sampler (MH_simple_kernel,
// main.cu:679-813) on heap-31 programs, with its proposals, quirks and median
// resampling (main.cu:1215-1234), on the GPU.
//
// Draw sources: Philox substreams {chain, step, purpose, slot} (default) or the
// reference's 31-bit LCG seeded block_seed + chain (main.cu:213-216, 691), which
// replays the reference's per-thread draw sequence.  The draw transforms restate
// random_int / random_float / random_normal / random_geometric (main.cu:218-254) and
// the host variants (main.cu:262-284) with the shared spec math.
//
//   k_mh_ref_lane  one lane per chain, the literal bottom-up fold over 31 slots per
//                  datum (as the reference evaluates); Philox or LCG.
//   k_mh_ref_wave  one wave per chain: heap in VGPRs (lane = slot), lane-parallel
//                  heap -> postfix decode of the reachable tree, the shared wave
//                  evaluator, per-constant proposals in parallel lanes.
#include "gpr_device.h"

// gpr_logf of the node counts 0 .. 31 (-inf at 0), filled once per device by
// gpr_launch_fill_ref_logint: the wave kernel's -log count terms become scalar loads.
static __device__ float gpr_ref_logint[GPR_REF_SLOTS + 1];
__global__ void k_fill_ref_logint() {
    const int i = (int)threadIdx.x;
    if (i <= GPR_REF_SLOTS) gpr_ref_logint[i] = gpr_logf((float)i);
}
hipError_t gpr_launch_fill_ref_logint(hipStream_t s) {
    hipLaunchKernelGGL(k_fill_ref_logint, dim3(1), dim3(64), 0, s);
    return hipGetLastError();
}

#define R_LOG01_F (-0x1.26bb1cp+1f)        /* logf(PROPOSE_GEOM_R = 0.1f) */
#define R_LOG015_F (-0x1.e5a9a8p+0f)       /* logf(PROPOSE_INT = 0.15f) */
#define R_LOG1M01_D (-0x1.af8e82826b325p-4) /* log(1.0 - 0.1f) */
#define R_LOG1M015_D (-0x1.4cd6ba6a55088p-3)/* log(1.0 - 0.15f) */
#define R_LOG2_D 0x1.62e42fefa39efp-1
#define R_LOG28_D 0x1.aa85ad6964a23p+1    /* log(NOPS + 1) of the default op set */
#define R_LOG2PI_F 0x1.d67f1ep+0f

__device__ __forceinline__ float r_pow10(int order) {  // powf(10, order), order in [-4, 3]
    switch (order) {
    case -4: return 0x1.a36e2ep-14f;
    case -3: return 0x1.0624dep-10f;
    case -2: return 0x1.47ae14p-7f;
    case -1: return 0x1.99999ap-4f;
    case 0: return 1.0f;
    case 1: return 10.0f;
    case 2: return 100.0f;
    default: return 1000.0f;
    }
}

// ---------------------------------------------------------------- draw sources
struct LcgSrc {
    uint32_t rx;
    __device__ __forceinline__ uint32_t rand31() { rx = (rx * 1103515245u + 12345u) & 0x7fffffffu; return rx; }
};
struct PhiloxSrc {
    gpr_rstream st;
    __device__ __forceinline__ uint32_t rand31() { return st.next() >> 1; }
};
// the same words, drawn 64 at a time on the vector unit (wave-uniform draws)
struct WavePhiloxSrc {
    WaveStream st;
    __device__ __forceinline__ uint32_t rand31() { return st.next() >> 1; }
};

template <class S> __device__ __forceinline__ int r_random_int(int n, S &s) {
    const uint32_t div = (uint32_t)(0x7fffffff / (n + 1));
    int r;
    do { r = (int)(s.rand31() / div); } while (r >= n);
    return r;
}
template <class S> __device__ __forceinline__ float r_random_float(S &s) {
    return (float)((double)(float)r_random_int(1000000, s) / 1000000.0);
}
template <class S> __device__ __forceinline__ float r_random_normal(S &s) {
#pragma clang fp contract(off)
    const float u = r_random_float(s);
    const float v = r_random_float(s);
    return sqrtf(-2.0f * gpr_logf(u)) * gpr_sin2pi(v);
}
template <class S> __device__ __forceinline__ float r_random_geometric(S &s) {
    return ceilf(gpr_logf(r_random_float(s)) / R_LOG01_F);
}
template <class S> __device__ __forceinline__ int r_host_random_int(int limit, S &s) {
    const uint32_t div = (uint32_t)(0x7fffffff / limit);
    int r;
    do { r = (int)(s.rand31() / div); } while (r > limit - 1);
    return r;
}
template <class S> __device__ __forceinline__ float r_host_random_float(S &s) {
    return (float)s.rand31() / (float)0x7fffffff;
}
template <class S> __device__ __forceinline__ float r_host_random_normal(S &s) {
#pragma clang fp contract(off)
    const float u = r_host_random_float(s);
    const float v = r_host_random_float(s);
    return sqrtf(-2.0f * gpr_logf(u)) * gpr_sin2pi(v);
}

// ---------------------------------------------------------------- densities (main.cu:299-309)
__device__ __forceinline__ double r_lgeom(float n) {
#pragma clang fp contract(off)
    return (double)((n - 1.0f) * R_LOG01_F) + R_LOG1M01_D;
}
__device__ __forceinline__ double r_lnormal1(float x) {  // lnormalpdf(x, 1)
#pragma clang fp contract(off)
    return -(double)(x * x) / 2.0 - 0.5 * ((double)R_LOG2PI_F + 2.0 * (double)0.0f);
}
__device__ __forceinline__ bool r_is_int(float x) {      // main.cu:192
    if (!__builtin_isfinite(x)) return false;
    return (double)fabsf(x - roundf(x)) < 1.0e-6;
}
__device__ __forceinline__ double r_logplusexp(double a, double b) {  // main.cu:190
#pragma clang fp contract(off)
    const double m = a > b ? a : b;
    return (double)gpr_logf(gpr_expf((float)(a - m)) + gpr_expf((float)(b - m))) + m;
}
__device__ __forceinline__ bool r_valid(double x) { return __builtin_isfinite(x); }

// One constant proposal of the constant step (main.cu:746-785).  Returns the forward and
// backward terms this constant adds (forward/backward accumulate over constants).
template <class S>
__device__ __forceinline__ void r_const_prop(S &s, float &C, double &ft, double &bt) {
#pragma clang fp contract(off)
    const float oldC = C;
    ft = 0.0;
    bt = 0.0;
    if (r_random_float(s) < 1.0f) {                 // P_PROPOSE_C_GEOMETRIC = 1: always
        if (r_random_float(s) < 0.15f) {            // PROPOSE_INT
            const float sgn = r_random_float(s) < 0.5f ? -1.0f : 1.0f;
            C = sgn * r_random_geometric(s) - 1.0f;
            ft = r_lgeom(C + 1.0f) + (C == 0.0f ? 0.0 : R_LOG2_D) + (double)R_LOG015_F;
        } else {
            C = 1.0f * r_random_normal(s);
            ft = r_lnormal1(C / 1.0f) + R_LOG1M015_D;
        }
        if (r_is_int(oldC)) {
            bt = r_logplusexp(r_lgeom(oldC + 1.0f) + (C == 0.0f ? 0.0 : R_LOG2_D) + (double)R_LOG015_F,
                              r_lnormal1(oldC / 1.0f) + R_LOG1M015_D);
        } else {
            bt = r_lnormal1(oldC / 1.0f) + R_LOG1M015_D;
        }
    } else {                                        // unreachable, main.cu:779-784
        const int order = r_random_int(8, s) - 4;
        C = C + 1.0f * r_pow10(order) * r_random_normal(s);
    }
}

__device__ __forceinline__ bool r_is_descendant(int n, int k) {  // main.cu:660-677
    if (k == n) return true;
    n = n + 1;
    k = k + 1;
    while (k > n) {
        k >>= 1;
        if (k == n) return true;
    }
    return false;
}

// reference op code -> postfix op code (A, B: -1, pass-through)
__device__ __forceinline__ int r_to_tape(int o) {
    switch (o) {
    case GPR_REF_ONE: return GPR_OP_ONE;
    case GPR_REF_CONST: return GPR_OP_CONST;
    case GPR_REF_INV: return GPR_OP_INV;
    case GPR_REF_PLUS: return GPR_OP_ADD;
    case GPR_REF_MINUS: return GPR_OP_SUB;
    case GPR_REF_RMINUS: return GPR_OP_RSUB;
    case GPR_REF_CPLUS: return GPR_OP_CADD;
    case GPR_REF_TIMES: return GPR_OP_MUL;
    case GPR_REF_CTIMES: return GPR_OP_CMUL;
    case GPR_REF_DIV: return GPR_OP_DIV;
    case GPR_REF_RDIV: return GPR_OP_RDIV;
    case GPR_REF_LOG: return GPR_OP_LOG;
    case GPR_REF_EXP: return GPR_OP_EXP;
    case GPR_REF_POW: return GPR_OP_POW;
    case GPR_REF_CPOW: return GPR_OP_CPOW;
    case GPR_REF_RPOW: return GPR_OP_RPOW;
    case GPR_REF_CRPOW: return GPR_OP_CRPOW;
    case GPR_REF_SQRT: return GPR_OP_SQRT;
    case GPR_REF_SIN: return GPR_OP_SIN;
    case GPR_REF_ASIN: return GPR_OP_ASIN;
    case GPR_REF_COS: return GPR_OP_COS;
    case GPR_REF_ACOS: return GPR_OP_ACOS;
    case GPR_REF_TAN: return GPR_OP_TAN;
    case GPR_REF_ATAN: return GPR_OP_ATAN;
    case GPR_REF_ABS: return GPR_OP_ABS;
    case GPR_REF_HALF: return GPR_OP_CMUL;     // a/2.0 in double, then float == 0.5f * a
    default: return gpr_ref_is_physical(o) ? GPR_OP_CONST : -1;
    }
}
// the constant a token of heap op o carries: its heap constant, a physical constant's
// value, or the factor of halving
__device__ __forceinline__ float r_tape_const(int o, float hc) {
    return gpr_ref_is_physical(o) ? gpr_ref_physical_value(o) : (o == GPR_REF_HALF ? 0.5f : hc);
}

// ---------------------------------------------------------------- lane schedule
typedef float gpr_f32v __attribute__((ext_vector_type(32)));
typedef int gpr_i32v __attribute__((ext_vector_type(32)));

// dispatch_eval (main.cu:397-444), one lane
__device__ __forceinline__ float r_dispatch_eval(int o, float a, float b, float C) {
    switch (o) {
    case GPR_REF_ONE: return 1.0f;
    case GPR_REF_CONST: return C;
    case GPR_REF_INV: return 1.0f / a;
    case GPR_REF_A: return a;
    case GPR_REF_B: return b;
    case GPR_REF_PLUS: return a + b;
    case GPR_REF_CPLUS: return a + C;
    case GPR_REF_MINUS: return a - b;
    case GPR_REF_RMINUS: return b - a;
    case GPR_REF_TIMES: return a * b;
    case GPR_REF_CTIMES: return C * a;
    case GPR_REF_DIV: return a / b;
    case GPR_REF_RDIV: return b / a;
    case GPR_REF_SQRT: return sqrtf(a);
    case GPR_REF_LOG: return logf(a);
    case GPR_REF_SIN: return sinf(a);
    case GPR_REF_ASIN: return asinf(a);
    case GPR_REF_COS: return cosf(a);
    case GPR_REF_ACOS: return acosf(a);
    case GPR_REF_TAN: return tanf(a);
    case GPR_REF_ATAN: return atanf(a);
    case GPR_REF_EXP: return expf(a);
    case GPR_REF_POW: return powf(a, b);
    case GPR_REF_CPOW: return powf(a, C);
    case GPR_REF_RPOW: return powf(b, a);
    case GPR_REF_CRPOW: return powf(C, a);
    case GPR_REF_ABS: return fabsf(a);
    case GPR_REF_HALF: return (float)((double)a / 2.0);
    default: return gpr_ref_is_physical(o) ? gpr_ref_physical_value(o) : __builtin_nanf("");
    }
}

// program_fold (main.cu:371-393) of the eval, length and count dispatchers, one lane.
// The slot index is uniform across the wave, so the register arrays are indexed with
// s_set_gpr_idx.
__device__ float r_fold_eval(const gpr_i32v &P, const gpr_f32v &C, float x) {
    gpr_f32v buf;
    for (int i = GPR_REF_SLOTS; i >= 1; i--) {
        const int l = 2 * i;
        const float lv = l > GPR_REF_SLOTS ? x : buf[l & 31];
        const float rv = l > GPR_REF_SLOTS ? x : buf[(l + 1) & 31];
        buf[i] = r_dispatch_eval(P[i - 1], lv, rv, C[i - 1]);
    }
    return buf[1];
}
__device__ __forceinline__ int r_len_kind(int o) {  // 0 leaf, 1 A, 2 B, 3 unary, 4 C-op, 5 binary
    if (gpr_ref_is_physical(o)) return 0;              // PHYSICAL_CONSTANT_LENGTH_PRIOR = 1
    switch (o) {
    case GPR_REF_ONE: case GPR_REF_CONST: return 0;
    case GPR_REF_A: return 1;
    case GPR_REF_B: return 2;
    case GPR_REF_CPLUS: case GPR_REF_CTIMES: case GPR_REF_CPOW: case GPR_REF_CRPOW: return 4;
    case GPR_REF_PLUS: case GPR_REF_MINUS: case GPR_REF_RMINUS: case GPR_REF_TIMES: case GPR_REF_DIV:
    case GPR_REF_RDIV: case GPR_REF_POW: case GPR_REF_RPOW: return 5;
    default: return 3;
    }
}
// dispatch_length (main.cu:477-526) and dispatch_countnodes (main.cu:553-601) folds
__device__ void r_fold_len_count(const gpr_i32v &P, float &length, int &count) {
    gpr_f32v bl;
    gpr_i32v bc;
    for (int i = GPR_REF_SLOTS; i >= 1; i--) {
        const int l = 2 * i;
        const float la = l > GPR_REF_SLOTS ? 1.0f : bl[l & 31], lb = l > GPR_REF_SLOTS ? 1.0f : bl[(l + 1) & 31];
        const int ca = l > GPR_REF_SLOTS ? 0 : bc[l & 31], cb = l > GPR_REF_SLOTS ? 0 : bc[(l + 1) & 31];
        const int k = r_len_kind(P[i - 1]);
        float lv;
        int cv;
        switch (k) {
        case 0: lv = 1.0f; cv = 1; break;
        case 1: lv = la; cv = ca; break;
        case 2: lv = lb; cv = cb; break;
        case 3: lv = 1.0f + la; cv = 1 + ca; break;
        case 4: lv = 1.0f + la + 1.0f; cv = ca + 1; break;
        default: lv = 1.0f + la + lb; cv = 1 + ca + cb; break;
        }
        bl[i] = lv;
        bc[i] = cv;
    }
    length = bl[1];
    count = bc[1];
}

// compute_prior (main.cu:529-544) and compute_likelihood (main.cu:454-468), one lane.
// lnc[i] = 0.5*(logf(2 pi) + 2 logf(sd_i)) is precomputed on the host with the host libm.
__device__ double r_prior(const gpr_i32v &P, const gpr_f32v &C, double lnops) {
#pragma clang fp contract(off)
    float length;
    int count;
    r_fold_len_count(P, length, count);
    double prior = -(double)(1.0f * length) * lnops;
    for (int c = 0; c < GPR_REF_SLOTS; c++) prior += gpr_lcauchy(C[c]);
    return prior;
}
__device__ double r_likelihood(const gpr_i32v &P, const gpr_f32v &C, const float *x, const float *y, const float *sd,
                               const double *lnc, int n) {
#pragma clang fp contract(off)
    double ll = 0.0;
    for (int i = 0; i < n; i++) {
        const float fx = r_fold_eval(P, C, x[i]);
        if (!__builtin_isfinite(fx)) return -GPR_INF;
        const float r = fx - y[i];
        const double s = (double)sd[i];
        const double l = -(double)(r * r) / (2.0 * s * s) - lnc[i];
        if (!__builtin_isfinite(l)) return -GPR_INF;
        ll += l;
    }
    return ll;
}

template <class S>
__device__ __forceinline__ bool r_accept(S &s, double ppost, double cur, double acc) {
    return (r_valid(ppost) && (acc > 0.0 || r_random_float(s) < gpr_expf((float)acc))) || !r_valid(cur);
}

template <int RNG>
__global__ void __launch_bounds__(128) k_mh_ref_lane(GprMHArgs a) {
    const uint32_t step0 = a.clock ? a.clock[0] : a.step0;
    const int nops = gpr_opset_nops(a.opset);
    const int chain = blockIdx.x * blockDim.x + threadIdx.x;
    if (chain >= a.nchains) return;
    int di = 0;
    while (a.ds[di].chain0 + a.ds[di].nchains <= chain) di++;
    const GprDataset d = a.ds[di];
    const float *x = a.dx + d.off, *y = a.dy + d.off, *sdv = a.dsd + d.off;
    const double *lc = a.lnc + d.off;
    const int N = a.nchains;
    gpr_i32v P;
    gpr_f32v C;
    for (int i = 0; i < GPR_REF_SLOTS; i++) {
        P[i] = a.ops[(size_t)i * N + chain];
        C[i] = a.cs[(size_t)i * N + chain];
    }
    double cur_prior = r_prior(P, C, a.ref_lnops);
    double cur_ll = r_likelihood(P, C, x, y, sdv, lc, d.n);
    double cur_post = cur_prior + cur_ll;
    LcgSrc lcg{(uint32_t)a.block_seed + (uint32_t)chain};
    gpr_i32v oldP;
    gpr_f32v oldC;
    double forward = 0.0, backward = 0.0;
    unsigned long long n_sprop = 0, n_sacc = 0, n_cprop = 0, n_cacc = 0;
    for (int mcmci = 0; mcmci < a.steps; mcmci++) {
        const uint32_t step = step0 + (uint32_t)mcmci;
        if (mcmci % 5 == 0) {
            n_sprop++;
            float length;
            int count;
            r_fold_len_count(P, length, count);
            forward = -(double)gpr_logf((float)count);
            PhiloxSrc ph;
            ph.st.init(a.seed, (uint32_t)chain, step, GPR_P_STRUCT, 0);
            int nn;
            if (RNG == GPR_RNG_LCG) nn = !r_valid(cur_post) ? 0 : r_random_int(GPR_REF_SLOTS, lcg);
            else nn = !r_valid(cur_post) ? 0 : r_random_int(GPR_REF_SLOTS, ph);
            for (int i = 0; i < GPR_REF_SLOTS; i++) {
                oldP[i] = P[i];
                if (i >= nn && r_is_descendant(nn, i)) {
                    P[i] = gpr_opset_code(a.opset, RNG == GPR_RNG_LCG ? r_random_int(nops, lcg) : r_random_int(nops, ph));
                }
            }
            r_fold_len_count(P, length, count);
            backward = -(double)gpr_logf((float)count);
            const double pp = r_prior(P, C, a.ref_lnops);
            const double pl = r_likelihood(P, C, x, y, sdv, lc, d.n);
            const double ppost = pp + pl;
            const double acc = ppost - cur_post + backward - forward;
            bool ok;
            if (RNG == GPR_RNG_LCG) ok = r_accept(lcg, ppost, cur_post, acc);
            else { PhiloxSrc pa; pa.st.init(a.seed, (uint32_t)chain, step, GPR_P_ACCEPT, 0); ok = r_accept(pa, ppost, cur_post, acc); }
            if (ok) { cur_post = ppost; cur_ll = pl; cur_prior = pp; n_sacc++; }
            else P = oldP;
        } else {
            n_cprop++;
            for (int c = 0; c < GPR_REF_SLOTS; c++) {
                oldC[c] = C[c];
                float cv = C[c];
                double ft, bt;
                if (RNG == GPR_RNG_LCG) r_const_prop(lcg, cv, ft, bt);
                else { PhiloxSrc pc; pc.st.init(a.seed, (uint32_t)chain, step, GPR_P_CPROP, (uint32_t)c); r_const_prop(pc, cv, ft, bt); }
                C[c] = cv;
                forward += ft;
                backward += bt;
            }
            const double pp = r_prior(P, C, a.ref_lnops);
            const double pl = r_likelihood(P, C, x, y, sdv, lc, d.n);
            const double ppost = pp + pl;
            const double acc = ppost - cur_post + backward - forward;
            bool ok;
            if (RNG == GPR_RNG_LCG) ok = r_accept(lcg, ppost, cur_post, acc);
            else { PhiloxSrc pa; pa.st.init(a.seed, (uint32_t)chain, step, GPR_P_ACCEPT, 0); ok = r_accept(pa, ppost, cur_post, acc); }
            if (ok) { cur_post = ppost; cur_ll = pl; cur_prior = pp; n_cacc++; }
            else C = oldC;
        }
    }
    for (int i = 0; i < GPR_REF_SLOTS; i++) {
        a.ops[(size_t)i * N + chain] = (uint8_t)P[i];
        a.cs[(size_t)i * N + chain] = C[i];
    }
    a.prior[chain] = cur_prior;
    a.ll[chain] = cur_ll;
    if (a.counters) {
        atomicAdd(&a.counters[GPR_CNT_SPROP], n_sprop);
        atomicAdd(&a.counters[GPR_CNT_SACC], n_sacc);
        atomicAdd(&a.counters[GPR_CNT_CPROP], n_cprop);
        atomicAdd(&a.counters[GPR_CNT_CACC], n_cacc);
        atomicAdd(&a.counters[GPR_CNT_STEPS], (unsigned long long)a.steps);
        atomicAdd(&a.counters[GPR_CNT_EVALS], (unsigned long long)a.steps * (unsigned long long)(GPR_REF_SLOTS * d.n));
    }
}

hipError_t gpr_launch_mh_ref_lane(const GprMHArgs &a, hipStream_t s) {
    const int grid = (a.nchains + 127) / 128;
    if (a.rng == GPR_RNG_LCG) hipLaunchKernelGGL((k_mh_ref_lane<GPR_RNG_LCG>), dim3(grid), dim3(128), 0, s, a);
    else hipLaunchKernelGGL((k_mh_ref_lane<GPR_RNG_PHILOX>), dim3(grid), dim3(128), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- wave schedule
// Lane-parallel decode of the heap held one slot per lane (lane i = slot i+1, i < 31)
// into the postfix tape of the reachable tree, one token per lane.  Subtree token
// counts are folded bottom-up and start positions pushed top-down, level by level;
// every emitted token (the op last in its subtree, x leaves below slots 16..31) is
// scattered to its lane with ds_permute.
struct RefTape {
    WaveProg p;     // ops and constants of the tape, o0/c0 per lane
    int slot;       // heap slot of each token (255 for x)
    int count;      // reachable non-pass-through ops: count_nodes (main.cu:604-608)
    int dl;         // description length: tokens + C-op tokens (dispatch_length fold)
};

__device__ RefTape r_decode_wave(int hop, float hc) {
    const int l = lane_id();
    const int s = l + 1;                                // 1-based slot
    const bool live = l < GPR_REF_SLOTS;
    const int k = live ? r_len_kind(hop) : 0;           // 0 leaf 1 A 2 B 3 unary 4 C-op 5 binary
    const bool bottom = s >= 16;                        // children are x leaves
    // subtree token counts, bottom-up (5 levels)
    // children of bottom slots are x leaves: they read lane 63, which always holds 1
    const int li = bottom ? 63 : 2 * s - 1, ri = bottom ? 63 : 2 * s;
    int size = 1;
#pragma unroll
    for (int it = 0; it < 5; it++) {
        const int lc = shfl_all(size, li);
        const int rc = shfl_all(size, ri);
        switch (k) {
        case 0: size = 1; break;
        case 1: size = lc; break;
        case 2: size = rc; break;
        case 3: case 4: size = 1 + lc; break;
        default: size = 1 + lc + rc; break;
        }
    }
    // start positions and reachability, top-down (4 levels below the root)
    int start = 0, reach = s == 1 ? 1 : 0;
#pragma unroll
    for (int it = 0; it < 4; it++) {
        const int par = (s >> 1) - 1;                   // parent lane
        const int pk = shfl_all(k, par & 63), ps = shfl_all(start, par & 63), pr = shfl_all(reach, par & 63);
        const int sib = shfl_all(size, (s - 2) & 63);   // left sibling size (for right children)
        if (s >= 2 && live) {
            const bool left = (s & 1) == 0;
            int r = 0, st = 0;
            if (pk == 1) { r = left; st = ps; }
            else if (pk == 2) { r = !left; st = ps; }
            else if (pk == 3 || pk == 4) { r = left; st = ps; }
            else if (pk == 5) { r = 1; st = left ? ps : ps + sib; }
            reach = pr && r;
            start = st;
        }
    }
    const int len = rl_i(size, 0);
    // scatter: op token at start + size - 1; x leaves below bottom slots
    const int top = r_to_tape(hop);
    const bool emit_op = live && reach && k != 1 && k != 2;
    const int v_op = emit_op ? (((top + 1) << 8) | (s - 1)) : 0;
    const int t_op = emit_op ? start + size - 1 : 63;
    const bool x1 = live && reach && bottom && k != 0;                // first x child
    const bool x2 = live && reach && bottom && k == 5;                // second x child
    const int v_x = (GPR_OP_X + 1) << 8 | 255;
    int t1 = __builtin_amdgcn_ds_permute(t_op * 4, v_op);
    int t2 = __builtin_amdgcn_ds_permute((x1 ? start : 63) * 4, x1 ? v_x : 0);
    int t3 = __builtin_amdgcn_ds_permute((x2 ? start + 1 : 63) * 4, x2 ? v_x : 0);
    __asm__ volatile("" : "+v"(t1), "+v"(t2), "+v"(t3));
    const int tok = t1 | t2 | t3;
    RefTape rt;
    const bool valid = l < len;
    rt.p.len = len;
    rt.p.o0 = valid ? (tok >> 8) - 1 : 0;
    rt.slot = valid ? (tok & 255) : 255;
    rt.p.o1 = 0;
    rt.p.c1 = 0.0f;
    const float cs = shfl_all(r_tape_const(hop, hc), rt.slot & 31);
    rt.p.c0 = (valid && gpr_has_const(rt.p.o0)) ? cs : 0.0f;
    rt.count = len - __popcll(ballot(valid && rt.p.o0 == GPR_OP_X));
    // description length: tokens plus the C-op tokens (dispatch_length kind 4; halving is
    // a c* token of length 1 + a)
    const int tk = shfl_all(k, rt.slot & 31);
    rt.dl = len + __popcll(ballot(valid && tk == 4));
    return rt;
}

// refresh the tape's constants from the heap constants
__device__ __forceinline__ void r_tape_consts(RefTape &rt, int hop, float hc) {
    const float cs = shfl_all(r_tape_const(hop, hc), rt.slot & 31);
    rt.p.c0 = (lane_id() < rt.p.len && gpr_has_const(rt.p.o0)) ? cs : 0.0f;
}

// prior: -length*log(NOPS+1) + sum over all 31 constants in slot order (main.cu:529-544)
__device__ __forceinline__ double r_prior_wave(int dl, float hc, double lnops) {
#pragma clang fp contract(off)
    const double lc = gpr_lcauchy(hc);
    double prior = -(double)(1.0f * (float)dl) * lnops;
    for (int c = 0; c < GPR_REF_SLOTS; c++) prior += rl_d(lc, c);
    return prior;
}

template <int KC>
__global__ void __launch_bounds__(GPR_WG) __attribute__((amdgpu_waves_per_eu(KC == 1 ? 4 : 1)))
k_mh_ref_wave(GprMHArgs a) {
    const uint32_t step0 = a.clock ? a.clock[0] : a.step0;
    const int nops = gpr_opset_nops(a.opset);
    const int wave = uni((int)(threadIdx.x >> 6));
    const int l = lane_id();
    const int N = a.nchains;
    unsigned long long n_sprop = 0, n_sacc = 0, n_cprop = 0, n_cacc = 0, n_evals = 0, n_steps = 0, n_flops = 0;
    for (;;) {
        __shared__ int s_tile;
        if (threadIdx.x == 0) s_tile = atomicAdd(a.tile_counter, 1);
        __syncthreads();
        const int tile_idx = uni(s_tile);
        __syncthreads();
        if (tile_idx >= a.ntiles) break;
        const GprTile tile = a.tiles[tile_idx];
        if (wave >= tile.count) continue;
        const GprDataset d = a.ds[tile.dataset];
        const uint32_t chain = (uint32_t)uni(tile.chain0 + wave);
        const float *dx = a.dx + d.off, *dy = a.dy + d.off, *dw = a.dw + d.off;
        const int nchunks = (d.n + 64 * KC - 1) / (64 * KC);
        Chunk<KC> c0;
        c0.load(dx, dy, dw, d.n, 0);
        const bool live = l < GPR_REF_SLOTS;
        int hop = live ? a.ops[(size_t)l * N + chain] : 0;
        float hc = live ? a.cs[(size_t)l * N + chain] : 0.0f;
        RefTape cur = r_decode_wave(hop, hc);
        double cur_prior = r_prior_wave(cur.dl, hc, a.ref_lnops);
        double cur_ll = wave_loglik<KC, false>(cur.p, d, dx, dy, dw, c0, nchunks);
        double cur_post = cur_prior + cur_ll;
        double forward = 0.0, backward = 0.0;
        for (int mcmci = 0; mcmci < a.steps; mcmci++) {
            const uint32_t step = step0 + (uint32_t)mcmci;
            int php = hop;
            float phc = hc;
            RefTape prop = cur;
            bool structural = (mcmci % 5) == 0;
            if (structural) {
                n_sprop++;
                forward = -(double)gpr_ref_logint[cur.count];
                WavePhiloxSrc ph;
                ph.st.init(a.seed, chain, step, GPR_P_STRUCT, 0);
                const int nn = !r_valid(cur_post) ? 0 : r_random_int(GPR_REF_SLOTS, ph);
                for (int i = nn; i < GPR_REF_SLOTS; i++) {
                    if (r_is_descendant(nn, i)) php = wl_i(php, gpr_opset_code(a.opset, r_random_int(nops, ph)), i);
                }
                prop = r_decode_wave(php, phc);
                backward = -(double)gpr_ref_logint[prop.count];
            } else {
                n_cprop++;
                double ft = 0.0, bt = 0.0;
                if (live) {
                    PhiloxSrc pc;
                    pc.st.init(a.seed, chain, step, GPR_P_CPROP, (uint32_t)l);
                    r_const_prop(pc, phc, ft, bt);
                }
                for (int c = 0; c < GPR_REF_SLOTS; c++) {   // slot order, as the sequential loop
                    forward += rl_d(ft, c);
                    backward += rl_d(bt, c);
                }
                r_tape_consts(prop, php, phc);
            }
            const double pp = r_prior_wave(prop.dl, phc, a.ref_lnops);
            // The accept draw u is the first random_float of the step's ACCEPT substream,
            // the same value whenever it is taken, so it is drawn before the evaluation.  A
            // valid current state accepts only if acc > 0 or u < expf(acc), and either needs
            // acc >= log u - 1 (u > 0) or acc >= -105 (u = 0: expf(acc) underflows to 0
            // below -103.98).  That bounds the proposal's LL from below: the evaluation
            // stops once its partial sum proves the LL under the bound (wave_loglik), and
            // the proposal is rejected as it would be after the full sum.
            float u = 0.0f;
            double reject_below = -GPR_INF;
            if (r_valid(cur_post)) {
                WavePhiloxSrc pa;
                pa.st.init(a.seed, chain, step, GPR_P_ACCEPT, 0);
                u = r_random_float(pa);
                const double amin = u > 0.0f ? (double)gpr_logf(u) - 1.0 : -105.0;
                if (r_valid(pp) && r_valid(forward) && r_valid(backward))
                    reject_below = cur_post - backward + forward - pp + amin -
                                   1e-12 * (__builtin_fabs(cur_post) + __builtin_fabs(backward) +
                                            __builtin_fabs(forward) + __builtin_fabs(pp));
            }
            int pts = d.n;
            const double pl = wave_loglik<KC, false>(prop.p, d, dx, dy, dw, c0, nchunks, nullptr, nullptr,
                                                     reject_below, &pts);
            pts = min(pts, d.n);
            n_evals += (unsigned long long)prop.p.len * (unsigned long long)pts;
            {
                const int nleaf = __popcll(ballot(l < prop.p.len && prop.p.o0 <= GPR_OP_ONE));
                const int nconst = __popcll(ballot(l < prop.p.len && gpr_has_const(prop.p.o0)));
                n_flops += (unsigned long long)(prop.p.len - nleaf + 4) * (unsigned long long)pts + (unsigned long long)nconst;
            }
            const double ppost = pp + pl;
            const double acc = ppost - cur_post + backward - forward;
            if ((r_valid(ppost) && (acc > 0.0 || u < gpr_expf((float)acc))) || !r_valid(cur_post)) {
                cur_post = ppost; cur_ll = pl; cur_prior = pp;
                hop = php; hc = phc; cur = prop;
                if (structural) n_sacc++; else n_cacc++;
            }
        }
        n_steps += (unsigned long long)a.steps;
        if (live) { a.ops[(size_t)l * N + chain] = (uint8_t)hop; a.cs[(size_t)l * N + chain] = hc; }
        if (l == 0) { a.prior[chain] = cur_prior; a.ll[chain] = cur_ll; }
    }
    if (l == 0 && a.counters) {
        atomicAdd(&a.counters[GPR_CNT_SPROP], n_sprop);
        atomicAdd(&a.counters[GPR_CNT_SACC], n_sacc);
        atomicAdd(&a.counters[GPR_CNT_CPROP], n_cprop);
        atomicAdd(&a.counters[GPR_CNT_CACC], n_cacc);
        atomicAdd(&a.counters[GPR_CNT_EVALS], n_evals);
        atomicAdd(&a.counters[GPR_CNT_STEPS], n_steps);
        atomicAdd(&a.counters[GPR_CNT_FLOPS], n_flops);
    }
}

hipError_t gpr_launch_mh_ref_wave(const GprMHArgs &a, int kc, int grid, hipStream_t s) {
    const dim3 wg(64 * a.wpg);
    if (kc == 1) hipLaunchKernelGGL((k_mh_ref_wave<1>), dim3(grid), wg, 0, s, a);
    else if (kc == 2) hipLaunchKernelGGL((k_mh_ref_wave<2>), dim3(grid), wg, 0, s, a);
    else hipLaunchKernelGGL((k_mh_ref_wave<4>), dim3(grid), wg, 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- init and median resampling
// initial draw (main.cu:1133-1137): 31 ops ~ U{0..NOPS-2}, then 31 constants
// 10^U{-4..3} N(0,1), host transforms, stream (chain, step, purpose, 0)
// The op set is a template parameter of the draw kernels: each one-lane-per-chain draw
// loop sees a constant table row.
template <int SET>
__device__ void r_draw_chain(uint64_t seed, uint32_t chain, uint32_t step, uint32_t purpose, uint8_t *ops, float *cs,
                             int N) {
#pragma clang fp contract(off)
    PhiloxSrc s;
    s.st.init(seed, chain, step, purpose, 0);
    for (int i = 0; i < GPR_REF_SLOTS; i++)
        ops[(size_t)i * N + chain] = (uint8_t)gpr_opset_code(SET, r_host_random_int(gpr_opset_nops(SET) - 1, s));
    for (int i = 0; i < GPR_REF_SLOTS; i++) {
        const int order = r_host_random_int(8, s) - 4;
        cs[(size_t)i * N + chain] = 1.0f * r_pow10(order) * r_host_random_normal(s);
    }
}

template <int SET>
__global__ void k_init_ref(GprMHArgs a) {
    const int chain = blockIdx.x * blockDim.x + threadIdx.x;
    if (chain >= a.nchains) return;
    r_draw_chain<SET>(a.seed, (uint32_t)chain, 0, GPR_P_INIT, a.ops, a.cs, a.nchains);
}

#define GPR_OPSET_LAUNCH(KERNEL, SET, GRID, BLOCK, S, ARGS)                                                        \
    switch (SET) {                                                                                                \
    case GPR_OPS_SIMPLIFIED: hipLaunchKernelGGL(KERNEL<GPR_OPS_SIMPLIFIED>, GRID, BLOCK, 0, S, ARGS); break;      \
    case GPR_OPS_POLYNOMIAL: hipLaunchKernelGGL(KERNEL<GPR_OPS_POLYNOMIAL>, GRID, BLOCK, 0, S, ARGS); break;      \
    case GPR_OPS_LINEAR: hipLaunchKernelGGL(KERNEL<GPR_OPS_LINEAR>, GRID, BLOCK, 0, S, ARGS); break;              \
    case GPR_OPS_PHYSICAL: hipLaunchKernelGGL(KERNEL<GPR_OPS_PHYSICAL>, GRID, BLOCK, 0, S, ARGS); break;          \
    default: hipLaunchKernelGGL(KERNEL<GPR_OPS_DEFAULT>, GRID, BLOCK, 0, S, ARGS); break;                         \
    }

hipError_t gpr_launch_init_ref(const GprMHArgs &a, hipStream_t s) {
    GPR_OPSET_LAUNCH(k_init_ref, a.opset, dim3((a.nchains + 255) / 256), dim3(256), s, a);
    return hipGetLastError();
}

// Order-preserving key of a double with NaN below everything (cmp, main.cu:987-995).
__device__ __forceinline__ uint64_t r_key(double v) {
    if (v != v) return 0;
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double r_unkey(uint64_t k) {
    if (k == 0) return __builtin_nan("");
    const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __builtin_bit_cast(double, b);
}

// One workgroup per dataset: radix select (8 passes of 8 bits) of the element of rank
// floor(nchains * quantile) in the NaN-first order; cutoff[d] is its value.
#define R_SEL_WG 1024
__global__ void __launch_bounds__(R_SEL_WG) k_rs_quantile_select(GprResampleArgs a) {
    const GprDataset d = a.ds[blockIdx.x];
    __shared__ unsigned int hist[256];
    __shared__ uint64_t s_prefix;
    __shared__ int s_rank;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_prefix = 0;
        // the rank of the cutoff, cutoff = sorted[(int)(N * q)] (main.cu:1219); gpr_create
        // keeps q in [0, 1), and the clamp keeps the select inside the keys regardless
        s_rank = min(max((int)((float)d.nchains * a.quantile), 0), d.nchains - 1);
    }
    __syncthreads();
    for (int pass = 7; pass >= 0; pass--) {
        for (int b = tid; b < 256; b += R_SEL_WG) hist[b] = 0;
        __syncthreads();
        const uint64_t prefix = s_prefix;
        const uint64_t hmask = pass == 7 ? 0 : (~0ull << ((pass + 1) * 8));
        for (int i = tid; i < d.nchains; i += R_SEL_WG) {
            const uint64_t k = r_key(a.prior[d.chain0 + i] + a.ll[d.chain0 + i]);
            if ((k & hmask) == prefix) atomicAdd(&hist[(k >> (pass * 8)) & 255], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            int r = s_rank;
            int b = 0;
            for (; b < 256; b++) {
                if (r < (int)hist[b]) break;
                r -= (int)hist[b];
            }
            s_rank = r;
            s_prefix = prefix | ((uint64_t)b << (pass * 8));
        }
        __syncthreads();
    }
    if (tid == 0) a.cutoff[blockIdx.x] = r_unkey(s_prefix);
}

template <int SET>
__global__ void k_rs_quantile_redraw(GprResampleArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nchains) return;
    int di = 0;
    while (di + 1 < a.ndatasets && a.ds[di + 1].chain0 <= c) di++;
    if (!(a.prior[c] + a.ll[c] < a.cutoff[di])) return;   // NaN posteriors never resampled
    r_draw_chain<SET>(a.seed, (uint32_t)c, a.clock ? a.clock[1] : a.block, GPR_P_RESAMPLE, a.ops, a.cs, a.nchains);
    atomicAdd(a.nresampled, 1);
}

hipError_t gpr_launch_resample_quantile(const GprResampleArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_rs_quantile_select, dim3(a.ndatasets), dim3(R_SEL_WG), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    GPR_OPSET_LAUNCH(k_rs_quantile_redraw, a.opset, dim3((a.nchains + 255) / 256), dim3(256), s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- test hooks
// One wave per heap: run the wave decode and write the tape (ops, consts) and its
// length, count and description length, for comparison with the host decode.
__global__ void k_test_ref_decode(const uint8_t *hop, const float *hc, int nheaps, uint8_t *ops, float *cs,
                                  int32_t *meta) {
    const int h = uni((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    if (h >= nheaps) return;
    const int l = lane_id();
    const int o = l < GPR_REF_SLOTS ? hop[(size_t)h * GPR_REF_SLOTS + l] : 0;
    const float c = l < GPR_REF_SLOTS ? hc[(size_t)h * GPR_REF_SLOTS + l] : 0.0f;
    const RefTape rt = r_decode_wave(o, c);
    ops[(size_t)h * 64 + l] = (uint8_t)rt.p.o0;
    cs[(size_t)h * 64 + l] = rt.p.c0;
    if (l == 0) {
        meta[h * 3 + 0] = rt.p.len;
        meta[h * 3 + 1] = rt.count;
        meta[h * 3 + 2] = rt.dl;
    }
}

hipError_t gpr_launch_test_ref_decode(const uint8_t *hop, const float *hc, int nheaps, uint8_t *ops, float *cs,
                                      int32_t *meta, hipStream_t s) {
    hipLaunchKernelGGL(k_test_ref_decode, dim3((nheaps + 3) / 4), dim3(256), 0, s, hop, hc, nheaps, ops, cs, meta);
    return hipGetLastError();
}

// The draw transforms on the device, one word per thread: u01, normal quantile, logf,
// expf, sin2pi of u, and the reference's random_float / random_normal /
// random_geometric / host random_normal of a Philox substream (seed, chain = i).
__global__ void k_test_draws(const uint32_t *words, int n, uint64_t seed, float *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t w = words[i];
    const float u = gpr_u01(w);
    float *o = out + (size_t)i * 10;
    o[0] = u;
    o[1] = gpr_normal(u);
    o[2] = gpr_logf(__builtin_bit_cast(float, w & 0x7fffffffu));
    o[3] = gpr_expf((float)((int32_t)w) * 0x1p-24f);
    o[4] = gpr_sin2pi(u);
    PhiloxSrc s;
    s.st.init(seed, (uint32_t)i, 0, 0, 0);
    o[5] = r_random_float(s);
    o[6] = r_random_normal(s);
    o[7] = r_random_geometric(s);
    o[8] = r_host_random_normal(s);
    o[9] = (float)r_random_int(27, s);
}

hipError_t gpr_launch_test_draws(const uint32_t *words, int n, uint64_t seed, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_test_draws, dim3((n + 255) / 256), dim3(256), 0, s, words, n, seed, out);
    return hipGetLastError();
}

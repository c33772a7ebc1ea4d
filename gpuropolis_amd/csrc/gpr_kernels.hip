// HIP kernels of the GPUropolis MI355X sampler (gfx950, wave64).
//
// Hot path: the fused MH step (MH_simple_kernel, main.cu:679-813, of the reference)
// with its evaluator and Gaussian log-likelihood (main.cu:371-468), prior
// (main.cu:529-544) and accept test, plus the resampling between blocks
// (main.cu:1215-1234), here on the device.
//
// Two evaluator schedules:
//   wave  one wavefront per chain.  The program is wave-uniform: lane t holds token
//         t (and t+64) of the tape, the interpreter reads each op with v_readlane into
//         a scalar register and branches on it without divergence, and the 64 lanes
//         run over data points, KC points per lane per chunk, with a per-lane value
//         stack in VGPRs indexed by the uniform stack pointer.  Per-point terms are
//         Kahan-summed per lane and tree-reduced across the wave.
//   lane  one lane per chain, as the reference.  Every lane interprets its own tape;
//         the proposal tape and the value stack live in LDS ([slot][lane],
//         conflict-free).  When the workgroup's chains share one dataset its points are
//         staged through LDS 64 at a time and read as broadcasts; the evaluation stops
//         early, per lane, once a proposal's partial sum proves its rejection.
#include "gpr_device.h"

// ---------------------------------------------------------------- config MH, wave schedule
// MALA: constant moves are Langevin proposals c' = c + h g(c) + eps z (h = eps^2/2, g the
// gradient of the log posterior) with the Hastings ratio of the two Gaussian proposal
// densities.  Under MALA every likelihood is evaluated at one point per lane (KC = 1),
// with early rejection.  Programs of at most three constants and stack depth 9, or eight
// and depth 8, get the proposal's LL and gradient from one forward-mode
// interpreter pass (value and tangents per stack entry); the others are pre-checked by
// the plain interpreter, and the survivors get the stack-free adjoint sweeps, whose token values sit in the wave's LDS
// region for programs of at most GPR_MALA_LDS_ROWS tokens and in the wave's slab in
// global memory for longer ones (LDS rows for every token of a 128-token cap, 32 KB per
// wave, would leave room for five waves per CU).  The current structure's sweep plan
// stays in registers.
extern __shared__ float gpr_lds_slab[];

// gpr_logf of the integers 0 .. GPR_MAX_CAP (program lengths), filled once per device by
// gpr_launch_fill_logint: the structure move's -log len terms become scalar loads.
static __device__ float gpr_logint[GPR_MAX_CAP + 1];
__global__ void k_fill_logint() {
    const int i = (int)threadIdx.x;
    if (i <= GPR_MAX_CAP) gpr_logint[i] = gpr_logf((float)i);
}

// The plain one-point-per-lane kernel is latency bound: six waves per SIMD (80 VGPRs) run
// it faster than the four that fit unspilled, although the compiler then spills per-tile
// values to scratch at every tile start and reloads some of them in the step loop.
// The MALA kernel: unconstrained it takes about 160 VGPRs (three waves per SIMD), 15 %
// slower on the full form than at four waves (128 VGPRs, per-tile values spilled at tile
// start only).  Five waves (96 VGPRs) ran 2 % faster still, but spilled inside the step
// loop: 18 scratch stores per MH step, 2.4 TB written per 1000-step launch of the full
// form against 1.1 GB at four waves.  The metric build runs at four waves as well:
// unconstrained it takes 138 VGPRs and three waves, 20 % slower.
// MET: some dataset of the launch has a proposal metric.  The metric path holds many
// wave-uniform values (matrix rows, pointers), so launches without a metric use a
// build without it, whose scalar registers then fit with far fewer spills.
#ifndef GPR_MALA_WAVES_PER_EU
#define GPR_MALA_WAVES_PER_EU 4
#endif
#ifndef GPR_MALA_MET_WAVES_PER_EU
#define GPR_MALA_MET_WAVES_PER_EU 4
#endif
// SPEC: the launch's datasets have at most 64 points; those of at most 32 evaluate
// speculative constant steps (wave_loglik_groups).  A separate build, so the kernels of
// larger datasets keep their registers.
template <int KC, bool HI, bool MALA, bool MET, bool SPEC = false>
__global__ void __launch_bounds__(GPR_WG)
__attribute__((amdgpu_waves_per_eu(MALA ? (MET ? GPR_MALA_MET_WAVES_PER_EU : GPR_MALA_WAVES_PER_EU) : (KC == 1 ? 6 : 1))))
k_mh_config_wave(GprMHArgs a) {
    const int wave = uni((int)(threadIdx.x >> 6));
    const int l = lane_id();
    const int N = a.nchains;
    // dynamic LDS: the value stacks of the wpg waves.  Under MALA each wave's LDS region
    // (MROWS rows of 64 floats) also holds the token values of the adjoint sweeps of
    // programs of at most GPR_MALA_LDS_ROWS tokens (the interpreter's stack and the sweeps
    // are never in use at once); longer programs sweep through the wave's slab in global
    // memory, lvg.
    constexpr int SROWS = GPR_MAX_DEPTH + 2;
    constexpr int MROWS = SROWS > GPR_MALA_LDS_ROWS ? SROWS : GPR_MALA_LDS_ROWS;
    // (the forward-mode interpreter's stack of value + tangents shares it too)
    constexpr int MFLOATS = MROWS * 64 > GPR_DUAL_LDS_FLOATS ? MROWS * 64 : GPR_DUAL_LDS_FLOATS;
    float *ls = gpr_lds_slab + (size_t)wave * (MALA ? MFLOATS : SROWS * 64 * KC);
    // after the wpg value stacks: each wave's priors of the (up to 16) steps of its draw
    // batch, then the LLs of its (up to 4) speculatively evaluated steps
    double *bpri = MALA ? nullptr
                        : (double *)(gpr_lds_slab + (size_t)a.wpg * ((GPR_MAX_DEPTH + 2) * 64 * KC)) + wave * 20;
    double *sll = MALA ? nullptr : bpri + 16;
    float *lvg = (MALA && a.mala_gslab) ? a.mala_gslab + ((size_t)blockIdx.x * a.wpg + wave) * (size_t)a.dmax_cap * 64
                                        : nullptr;
    const uint32_t step0 = a.clock ? (uint32_t)uni((int)a.clock[0]) : a.step0;
    // launch counters: lane k adds counter k into its wave's LDS row once per chain (a
    // register would live across the chain loop and spill to memory); the rows are folded
    // into the global counters at the end
    __shared__ unsigned long long s_cnt[GPR_WG / 64][GPR_NCNT];
    if (l < GPR_NCNT) s_cnt[wave][l] = 0;
    for (;;) {
        __shared__ int s_tile;
        if (threadIdx.x == 0) s_tile = atomicAdd(a.tile_counter, 1);
        __syncthreads();
        const int tile_idx = uni(s_tile);
        __syncthreads();
        if (tile_idx >= a.ntiles) break;
        const GprTile tile = a.tiles[tile_idx];
        if (wave >= tile.count) continue;
        const int tds = uni(tile.dataset);
        const GprDataset d = uni_dataset(a.ds[tds]);
        const uint32_t chain = (uint32_t)uni(tile.chain0 + wave);
        const int cap = d.cap;
        const float *dx = a.dx + d.off, *dy = a.dy + d.off, *dw = a.dw + d.off;
        const int nchunks = (d.n + 64 * KC - 1) / (64 * KC);
        // speculative constant steps of small datasets: G proposals per interpreter pass
        // (wave_loglik_groups); the window of the last pass starts at step spec_lo and
        // holds spec_n cached LLs in sll, valid while the state does not change
        const int G = (SPEC && !MALA && KC == 1) ? (d.n <= 16 ? 4 : (d.n <= 32 ? 2 : 1)) : 1;
        uint32_t spec_lo = 0;
        int spec_n = 0;

        WaveProg cur;
        cur.len = uni(a.len[chain]);
        const int len0 = cur.len;
        bool changed = false;
        cur.o0 = l < cur.len ? a.ops[(size_t)l * N + chain] : 0;
        cur.c0 = l < cur.len ? a.cs[(size_t)l * N + chain] : 0.0f;
        cur.o1 = (HI && l + 64 < cur.len) ? a.ops[(size_t)(l + 64) * N + chain] : 0;
        cur.c1 = (HI && l + 64 < cur.len) ? a.cs[(size_t)(l + 64) * N + chain] : 0.0f;

        Chunk<KC> c0;
        c0.load(dx, dy, dw, d.n, 0);

        // proposal metric of this dataset; a chain whose metric rows are not all constant
        // tokens moves every constant with the identity
        const GprMetric *mt = (MET && a.metric) ? a.metric + tds : nullptr;
        int mk = mt ? uni(mt->k) : 0;
        const int mrow = (l < mk) ? mt->rows[l] : 0;
        if (mk) {
            const int orow = gather_i(cur.o0, cur.o1, mrow, HI);
            if (ballot(l < mk && (mrow >= cur.len || orow != GPR_OP_CONST))) mk = 0;
        }
        const int slot0 = (mk && l < cap) ? mt->slot[l] : -1;
        const int slot1 = (HI && mk && l + 64 < cap) ? mt->slot[l + 64] : -1;
        const float *mL = mk ? mt->L : nullptr, *mM = mk ? mt->M : nullptr;
        const double *mLi = mk ? mt->Li : nullptr;

        // iteration -1 evaluates the current state (main.cu:683-684); the prior and
        // likelihood have a single call site so the kernel body stays small
        double prior = 0.0, ll = 0.0;
        // per-lane constant-proposal block, reused by four consecutive steps
        gpr_u4 wlo = {0, 0, 0, 0}, whi = {0, 0, 0, 0};
        uint32_t wblk = 0xffffffffu, whblk = 0xffffffffu;
        // MALA: the Langevin noise batch (key: step >> lg, and whether it spans 16 steps)
        float zbm = 0.0f;
        uint32_t zmblk = 0xffffffffu;
        // Batched random-walk moves (BATCH: no MALA; per program: tape within the low half,
        // no metric, at most 16 constant tokens and no constant-operand token).  The draw of
        // step s for a constant is word s & 3 of the Philox block s >> 2 of its token, so a
        // batch covers 4 steps (16 lanes each, up to 16 constants) or, for programs of at
        // most 4 constants, 16 steps (4 lanes each).  At the first constant step of a batch,
        // lane j = ks u + r takes the normal of step u's word for the r-th constant token
        // and, from the current constants, the proposed constant and its Cauchy term; the
        // batch also sums each step's prior into bpri.  Each step then reads its proposal
        // lanes and its prior, the same values as the per-step path.  Any accepted move drops the batch (the next constant step
        // recomputes it).
        constexpr bool BATCH = !MALA;
        float cb = 0.0f;
        uint32_t zblk = 0xffffffffu;
        // MALA: gradient of the current state (lane-distributed like the constants) and its
        // metric drift M g at the metric rows
        float cg0 = 0.0f, cg1 = 0.0f, pg0 = 0.0f, pg1 = 0.0f, cmg = 0.0f, pmg = 0.0f;
        bool grad_ok = false;
        AcceptDraws adraw;
        // token classes of the current program (MALA: with the sweeps' plan for programs
        // without the forward-mode pass; only proposals that survive early rejection need
        // their gradient)
        GradPlan gplan{};
        OpMasks cmask = op_masks<HI>(cur);
        // per-program values that change only with the structure: (operator tokens + 4)
        // and constants of the current program for the work counters, and whether its
        // random-walk moves can be batched (see BATCH below)
        // MALA: whether the current program's gradient comes from the forward-mode
        // interpreter (dual_ok), with its constant-token masks
        // (cur_nt: tangents of its pass, 1, 2, GPR_DUAL_NT, GPR_DUAL_NT4 or GPR_DUAL_NT6, or
        // GPR_DUAL_NTMAX for two passes of GPR_DUAL_NT4)
        bool cur_dual = false;
        int cur_nt = GPR_DUAL_NT;
        uint64_t ccm[2] = {0, 0};
        if (MALA) {
            int D0, D1, dm;
            wave_depths<HI>(cur, D0, D1, dm);
            cur_dual = !MET && mk == 0 && dual_ok<HI>(cur, dm);
            ccm[0] = ballot(l < cur.len && cur.o0 == GPR_OP_CONST);
            ccm[1] = HI ? ballot(l + 64 < cur.len && cur.o1 == GPR_OP_CONST) : 0;
            cur_nt = dual_nt(__popcll(ccm[0]) + __popcll(ccm[1]), dm);
            // the sweeps' plan serves only programs without the forward-mode pass
            if (!cur_dual) gplan = grad_plan<HI>(cur);
        }
        int cur_opq, cur_nc;
        prog_counts<HI>(cur, cur_opq, cur_nc);
        bool cur_bok = batch_ok<BATCH, HI>(cur, mk);
        bool cur_b16 = cur_bok && const_count<HI>(cur) <= 4;
        // step % struct_every, carried as a counter (no integer division per step)
        // per-tile counts: structure proposed/accepted, constant proposed/accepted, and over
        // evaluated programs the sums of tokens x points, of (operator tokens + 4) x points
        // and of constants
        uint32_t c_sp = 0, c_sa = 0, c_cp = 0, c_ca = 0, c_nc = 0;
        unsigned long long c_ev = 0, c_fl = 0;
        // evaluations of the current program add only their points and their number; the
        // sums take them with the program's values when its structure changes and at the
        // end of the tile (a product per structure, not per step)
        unsigned long long cur_pts = 0;
        uint32_t cur_nev = 0;
        const uint32_t se = (uint32_t)a.struct_every;
        uint32_t srem = step0 % se;
        for (int i = -1; i < a.steps; i++) {
            const uint32_t step = step0 + (uint32_t)i;
            const bool struct_slot = i >= 0 && srem == 0;
            if (i >= 0) srem = srem + 1 == se ? 0 : srem + 1;
            const double post = prior + ll;
            WaveProg prop = cur;
            double extra = 0.0;
            bool have_pll = false;   // MALA constant step: the LL (pll_m, over mala_pts points) is known
            double pll_m = 0.0;
            int mala_pts = 0;
            bool have_prior = false; // MALA constant step: the prior (pre_prior) is known
            double pre_prior = 0.0;
            bool bprior = false;   // the prior comes from the batch: bpri[bu]
            int bu = 0;
            const bool structural = struct_slot && !a.fixed_structure;
            if (i < 0) {
            } else if (structural) {
                c_sp++;
                const uint32_t sc2 = (uint32_t)GPR_P_STRUCT << 24;
                const uint32_t klo = (uint32_t)a.seed, khi = (uint32_t)(a.seed >> 32);
                gpr_u4 gw = gpr_philox(chain, step, sc2, (uint32_t)l, klo, khi);
                const uint32_t r = (uint32_t)rl_i((int)gw.z, 0);
                const int j = __builtin_isfinite(post) ? (int)(((uint64_t)r * (uint64_t)cur.len) >> 32) : cur.len - 1;
                int D0, D1, dm;
                wave_depths<HI>(cur, D0, D1, dm);
                const int s0 = subtree_start<HI>(j, D0, D1);
                const int oldsize = j - s0 + 1;
                const int maxnew = cap - (cur.len - oldsize);
                // this lane's generation terms: its tokens of the old subtree minus its
                // symbols of the new one, summed over the wave below
                double tq = 0.0;
                if (l >= s0 && l <= j) tq = gen_term_lane(cur.o0, cur.c0);
                if (HI && l + 64 >= s0 && l + 64 <= j) tq += gen_term_lane(cur.o1, cur.c1);
                // generate symbols 64h + l in right-child-first prefix order; postfix is the reverse
                int g0 = 0, g1 = 0;
                float gc0 = 0.0f, gc1 = 0.0f;
                int need = 1, m = 0;
                bool closed = false;
#pragma unroll
                for (int h = 0; h < (HI ? 2 : 1); h++) {
                    if (h == 1) {
                        if (maxnew <= 64) break;
                        gw = gpr_philox(chain, step, sc2, 64u + (uint32_t)l, klo, khi);
                    }
                    const int sy = gen_symbol_lane(gw.x);
                    const int nd = need + wave_scan_i(gpr_arity(sy) - 1);   // operands missing after it
                    const uint64_t z = ballot(nd == 0);
                    const int cnt = z ? __builtin_ctzll(z) + 1 : 64;
                    const bool in = l < cnt;
                    const float cv = sy == GPR_OP_CONST ? gpr_normal(gpr_u01(gw.y)) : 0.0f;
                    if (in) tq -= gen_term_lane(sy, cv);
                    if (h == 0) { g0 = in ? sy : 0; gc0 = in ? cv : 0.0f; }
                    else { g1 = in ? sy : 0; gc1 = in ? cv : 0.0f; }
                    m += cnt;
                    if (z) { closed = true; break; }
                    need = rl_i(nd, 63);
                }
                // the sequential draw stops with a failure before symbol index maxnew
                if (!closed || m > maxnew) continue;
                const int plen = cur.len - oldsize + m;
                // splice: [0,s0) from cur, [s0, s0+m) reversed generation, rest from cur
#pragma unroll
                for (int h = 0; h < (HI ? 2 : 1); h++) {
                    const int t = l + 64 * h;
                    int src;
                    bool from_gen;
                    if (t < s0) { src = t; from_gen = false; }
                    else if (t < s0 + m) { src = m - 1 - (t - s0); from_gen = true; }
                    else { src = t - m + oldsize; from_gen = false; }
                    if (src > 127) src = 127;
                    const int oc = gather_i(cur.o0, cur.o1, src, HI);
                    const float cc = gather_f(cur.c0, cur.c1, src, HI);
                    const int og = gather_i(g0, g1, src, HI);
                    const float cg = gather_f(gc0, gc1, src, HI);
                    int ov = from_gen ? og : oc;
                    float cvv = from_gen ? cg : cc;
                    if (t >= plen) { ov = 0; cvv = 0.0f; }
                    if (h == 0) { prop.o0 = ov; prop.c0 = cvv; } else { prop.o1 = ov; prop.c1 = cvv; }
                }
                prop.len = plen;
                int P0, P1, pdm;
                wave_depths<HI>(prop, P0, P1, pdm);
                if (pdm > a.dmax) continue;
                // (-log plen + lq_old) - (-log len + lq_new)
                extra = ((double)gpr_logint[cur.len] - (double)gpr_logint[plen]) + wave_sum_d(tq);
            } else {
                const bool k0 = l < cur.len && cur.o0 == GPR_OP_CONST;
                const bool k1 = HI && l + 64 < cur.len && cur.o1 == GPR_OP_CONST;
                const uint64_t mk0 = ballot(k0), mk1 = HI ? ballot(k1) : 0;
                if ((mk0 | mk1) == 0) continue;
                c_cp++;
                const bool batch = BATCH && cur_bok;
                if (batch) {
                    // KS lanes per step: 4 for programs of at most 4 constants (a batch of 16
                    // steps, four Philox blocks per constant), else 16 (a batch of 4 steps)
                    const int lg = cur_b16 ? 4 : 2, ks = cur_b16 ? 4 : 16;
                    const int u = (int)(step & ((1u << lg) - 1));
                    if ((step >> lg) != zblk) {
                        zblk = step >> lg;
                        // lane j = ks * uj + r: step base + uj, constant of rank r, whose
                        // token is tr; the lane draws that step's word of that token's block
                        const int r = l & (ks - 1);
                        int tr = 0;
                        uint64_t m = mk0;
                        for (int q = 0; m; q++) {
                            const int b = __builtin_ctzll(m);
                            m &= m - 1;
                            tr = r == q ? b : tr;
                        }
                        const uint32_t sj = (zblk << lg) + (uint32_t)(l >> (6 - lg));
                        const gpr_u4 wb = gpr_step_block(a.seed, chain, sj, GPR_P_CPROP, (uint32_t)tr, 0);
                        const float zb = gpr_normal(gpr_u01(gpr_u4_word(wb, (int)(sj & 3))));
                        cb = gpr_axpy(a.sigma, zb, shfl_all(cur.c0, tr));
                        // the prior of each step's proposal, summed in token order as
                        // wave_prior: lanes ks u .. ks u + K - 1 hold its constants' terms
                        const double t = -(double)GPR_LOGPI_F - (double)gpr_lcauchy_logf(cb);
                        const int K = __popcll(mk0);
                        double sacc = 0.0;
                        for (int q = 0; q < K; q++) sacc += shfl_all_d(t, (l & ~(ks - 1)) + q);
                        if (r == 0) bpri[l >> (6 - lg)] = gpr_dl_prior(cur.len, a.lnsym, sacc);
                    }
                    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mk0 >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)mk0, 0u));
                    const float cv = shfl_all(cb, ks * u + rank);
                    if (k0) prop.c0 = cv;
                    bprior = true;
                    bu = u;
                    if (G > 1) {
                        if (spec_n > 0 && step >= spec_lo && step < spec_lo + (uint32_t)spec_n) {
                            pll_m = rl_d(sll[step - spec_lo], 0);   // evaluated by the window's pass
                            have_pll = true;
                            mala_pts = d.n;
                        } else {
                            // the constant steps from this one to the next structure step, the
                            // end of the block and the end of the draw batch
                            int win = G;
                            if (!a.fixed_structure) win = min(win, 1 + (int)((se - srem) % se));
                            win = min(win, a.steps - i);
                            win = min(win, (1 << lg) - u);
                            if (win > 1) {
                                double o[4];
                                wave_loglik_groups(cur, d, dx, dy, dw, ls, cmask, mk0, cb, ks, u, G, win, o);
                                if (l == 0) {
#pragma unroll
                                    for (int g = 0; g < 4; g++) sll[g] = o[g];
                                }
                                spec_lo = step;
                                spec_n = win;
                                pll_m = o[0];
                                have_pll = true;
                                mala_pts = d.n;
                            }
                        }
                    }
                } else {
                // identity-metric constants; metric rows move together below
                const bool i0 = k0 && slot0 < 0, i1 = k1 && slot1 < 0;
                float z0, z1 = 0.0f;
                const int nk0 = __popcll(mk0), nk = nk0 + (HI ? __popcll(mk1) : 0);
                if (MALA && !MET && nk <= 16) {
                    // Langevin noise from a draw batch in the random-walk batch's layout
                    // (lane ks u + r: step base + u, constant of rank r in token order), one
                    // normal per lane for 4 steps (16 for at most 4 constants).  The noise
                    // depends on the steps and the constant tokens only, so the batch holds
                    // across accepted constant moves, until the structure changes (zmblk is
                    // reset there).
                    const bool b16 = nk <= 4;
                    const int lg = b16 ? 4 : 2, ks = b16 ? 4 : 16;
                    const uint32_t key = ((step >> lg) << 1) | (b16 ? 1u : 0u);
                    if (key != zmblk) {
                        zmblk = key;
                        const int r = l & (ks - 1);
                        int tr = 0, q = 0;
                        for (uint64_t m = mk0; m; q++) {
                            const int b = __builtin_ctzll(m);
                            m &= m - 1;
                            tr = r == q ? b : tr;
                        }
                        if (HI) {
                            for (uint64_t m = mk1; m; q++) {
                                const int b = __builtin_ctzll(m);
                                m &= m - 1;
                                tr = r == q ? 64 + b : tr;
                            }
                        }
                        const uint32_t sj = ((step >> lg) << lg) + (uint32_t)(l >> (6 - lg));
                        const gpr_u4 wb = gpr_step_block(a.seed, chain, sj, GPR_P_CPROP, (uint32_t)(tr & 63),
                                                         (uint32_t)(tr >> 6));
                        zbm = gpr_normal(gpr_u01(gpr_u4_word(wb, (int)(sj & 3))));
                    }
                    const int base = ks * (int)(step & ((1u << lg) - 1));
                    const int rank0 = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mk0 >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mk0, 0u));
                    z0 = shfl_all(zbm, base + rank0);
                    if (HI && mk1) {
                        const int rank1 = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mk1 >> 32),
                                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)mk1, 0u));
                        z1 = shfl_all(zbm, base + nk0 + rank1);
                    }
                } else {
                    // per-step draws: each lane's block of the four steps (kept across them
                    // only with a proposal metric; the batch above serves the common case)
                    if (BATCH || (MALA && !MET) || (step >> 2) != wblk) {
                        wblk = step >> 2;
                        wlo = gpr_step_block(a.seed, chain, step, GPR_P_CPROP, (uint32_t)l, 0);
                    }
                    // the high half's block only for programs with constants there
                    const bool hi_live = HI && mk1 != 0;
                    if (hi_live && (BATCH || (MALA && !MET) || (step >> 2) != whblk)) {
                        whblk = step >> 2;
                        whi = gpr_step_block(a.seed, chain, step, GPR_P_CPROP, (uint32_t)l, 1);
                    }
                    const int wi = (int)(step & 3);
                    z0 = gpr_normal(gpr_u01(gpr_u4_word(wlo, wi)));
                    if (hi_live) z1 = gpr_normal(gpr_u01(gpr_u4_word(whi, wi)));
                }
                float zr = 0.0f, cr = 0.0f, dr = 0.0f;
                if (mk) {
                    zr = gather_f(z0, z1, mrow, HI);
                    cr = gather_f(cur.c0, cur.c1, mrow, HI);
                    dr = metric_Lv(mL, mk, zr);
                }
                if (!MALA) {
                    if (i0) prop.c0 = gpr_axpy(a.sigma, z0, cur.c0);
                    if (HI && i1) prop.c1 = gpr_axpy(a.sigma, z1, cur.c1);
                    if (mk) {
                        const float cn = gpr_axpy(a.sigma, dr, cr);
                        const float v0 = shfl_all(cn, max(slot0, 0));
                        if (slot0 >= 0) prop.c0 = v0;
                        if (HI) {
                            const float v1 = shfl_all(cn, max(slot1, 0));
                            if (slot1 >= 0) prop.c1 = v1;
                        }
                    }
                } else {
                    if (!grad_ok) {
                        double h0, h1;
                        bool gok;
                        if (cur_dual) {
                            float gs[GPR_DUAL_NTMAX];
                            int np_;
                            const Chunk<1> &k0c = *(const Chunk<1> *)&c0;
                            const double cl =
                                cur_nt == 1 ? wave_loglik_dual<HI, 1>(cur, d, dx, dy, dw, k0c, nchunks, ls, cmask, ccm,
                                                                      -GPR_INF, &np_, 0, gs)
                                : cur_nt == 2 ? wave_loglik_dual<HI, 2>(cur, d, dx, dy, dw, k0c, nchunks, ls, cmask, ccm,
                                                                        -GPR_INF, &np_, 0, gs)
                                : cur_nt == GPR_DUAL_NT
                                    ? wave_loglik_dual<HI, GPR_DUAL_NT>(cur, d, dx, dy, dw, k0c, nchunks, ls, cmask, ccm,
                                                                        -GPR_INF, &np_, 0, gs)
                                : cur_nt == GPR_DUAL_NT6
                                    ? wave_loglik_dual<HI, GPR_DUAL_NT6>(cur, d, dx, dy, dw, k0c, nchunks, ls, cmask,
                                                                         ccm, -GPR_INF, &np_, 0, gs)
                                    : dual_passes4<HI>(cur, d, dx, dy, dw, k0c, nchunks, ls, cmask, ccm, -GPR_INF,
                                                       &np_, cur_nt == GPR_DUAL_NTMAX ? 2 : 1, gs);
                            dual_grad_lanes<HI>(cur, ccm, gs, h0, h1, gok);
                            gok = gok && __builtin_isfinite(cl);
                        } else {
                            (void)wave_loglik_grad<HI>(cur, gplan, d, dx, dy, dw,
                                                       cur.len <= GPR_MALA_LDS_ROWS ? ls : lvg, h0, h1, gok);
                        }
                        c_fl += (unsigned long long)(cur_opq - 2 + cur_nc) * (unsigned long long)d.n;
                        cg0 = gok ? (float)h0 : 0.0f;
                        cg1 = gok ? (float)h1 : 0.0f;
                        if (mk) cmg = metric_Mv(mM, mk, gather_f(cg0, cg1, mrow, HI));
                        grad_ok = true;
                    }
                    const float e = a.eps, h = 0.5f * e * e;
                    const double inv2e2 = 2.0 * (double)e * (double)e, r2 = 1.0 / inv2e2;
                    const float m0 = gpr_axpy(h, cg0, cur.c0), m1 = gpr_axpy(h, cg1, cur.c1);
                    if (i0) prop.c0 = gpr_axpy(e, z0, m0);
                    if (HI && i1) prop.c1 = gpr_axpy(e, z1, m1);
                    const double f0 = (double)prop.c0 - (double)m0, f1 = (double)prop.c1 - (double)m1;
                    double fq = wave_gauss_sum_d<HI>(f0, i0, f1, i1, r2);
                    float cn = 0.0f;
                    if (mk) {
                        // c' = c + h M g + e L z over the metric rows
                        const float mean = gpr_axpy(h, cmg, cr);
                        cn = gpr_axpy(e, dr, mean);
                        fq += metric_logq(mLi, mk, l < mk ? (double)cn - (double)mean : 0.0, inv2e2);
                        const float v0 = shfl_all(cn, max(slot0, 0));
                        if (slot0 >= 0) prop.c0 = v0;
                        if (HI) {
                            const float v1 = shfl_all(cn, max(slot1, 0));
                            if (slot1 >= 0) prop.c1 = v1;
                        }
                    }
                    // The reverse proposal's log density bq is <= 0, so a = pprior + pll + bq -
                    // fq - post is at most pprior + pll - fq - post: a proposal whose LL lies
                    // below post - pprior + fq - 18 gives a < -18 whatever its gradient and is
                    // rejected without a draw (cfg_accept).  The pass that first evaluates the
                    // LL checks that bound after each 64-point chunk and stops such proposals
                    // early (not those from an invalid current state, which always move).
                    // Programs that qualify get the LL and the gradient from one forward-mode
                    // pass of the interpreter.  The others are pre-checked by the plain
                    // interpreter when the data spans several chunks (one chunk cannot stop
                    // early), and only the survivors get the adjoint sweeps, whose LL equals
                    // the interpreter's bit for bit (same operations, same per-lane order).
                    // (A chunk of both sweeps costs more than an interpreter chunk: on the full
                    // form, sweeps that check the bound themselves instead ran 9 % slower, and
                    // a pre-check of the first chunk only, before the sweeps or the
                    // forward-mode pass, 7 % slower.)
                    pre_prior = wave_prior<HI>(prop, a.lnsym);
                    have_prior = true;
                    double rb = -GPR_INF;
                    if (__builtin_isfinite(post) && __builtin_isfinite(pre_prior) && __builtin_isfinite(fq))
                        rb = post - pre_prior + fq - 18.0 -
                             1e-12 * (__builtin_fabs(post) + __builtin_fabs(pre_prior) + __builtin_fabs(fq));
                    float dgs[GPR_DUAL_NTMAX];
                    double h0 = 0.0, h1 = 0.0;
                    bool pok = false;
                    const bool precheck = !MET && !cur_dual && nchunks > 1;
                    if (cur_dual && cur_nt == 1)
                        pll_m = wave_loglik_dual<HI, 1>(prop, d, dx, dy, dw, *(const Chunk<1> *)&c0, nchunks, ls,
                                                        cmask, ccm, rb, &mala_pts, 0, dgs);
                    else if (cur_dual && cur_nt == 2)
                        pll_m = wave_loglik_dual<HI, 2>(prop, d, dx, dy, dw, *(const Chunk<1> *)&c0, nchunks, ls,
                                                        cmask, ccm, rb, &mala_pts, 0, dgs);
                    else if (cur_dual && cur_nt == GPR_DUAL_NT)
                        pll_m = wave_loglik_dual<HI, GPR_DUAL_NT>(prop, d, dx, dy, dw, *(const Chunk<1> *)&c0, nchunks,
                                                                  ls, cmask, ccm, rb, &mala_pts, 0, dgs);
                    else if (cur_dual && cur_nt == GPR_DUAL_NT6)
                        pll_m = wave_loglik_dual<HI, GPR_DUAL_NT6>(prop, d, dx, dy, dw, *(const Chunk<1> *)&c0,
                                                                   nchunks, ls, cmask, ccm, rb, &mala_pts, 0, dgs);
                    else if (cur_dual)
                        pll_m = dual_passes4<HI>(prop, d, dx, dy, dw, *(const Chunk<1> *)&c0, nchunks, ls, cmask, ccm,
                                                 rb, &mala_pts, cur_nt == GPR_DUAL_NTMAX ? 2 : 1, dgs);
                    else if (precheck && HI && prop.len <= 64)
                        pll_m = wave_loglik<KC, false, true>(prop, d, dx, dy, dw, c0, nchunks, ls, &cmask, rb,
                                                             &mala_pts);
                    else if (precheck)
                        pll_m = wave_loglik<KC, HI, true>(prop, d, dx, dy, dw, c0, nchunks, ls, &cmask, rb,
                                                          &mala_pts);
                    bool rejected = (cur_dual || precheck) && !__builtin_isfinite(pll_m) && __builtin_isfinite(post);
                    if (!cur_dual && !rejected) {
                        const double gl = wave_loglik_grad<HI>(prop, gplan, d, dx, dy, dw,
                                                               prop.len <= GPR_MALA_LDS_ROWS ? ls : lvg, h0, h1, pok);
                        if (!precheck) {
                            pll_m = gl;
                            mala_pts = d.n;
                            rejected = !__builtin_isfinite(pll_m) && __builtin_isfinite(post);
                        }
                    }
                    have_pll = true;
                    if (rejected) {
                        extra = -GPR_INF;      // rejected: invalid, or below the bound
                    } else {
                        if (cur_dual) {
                            dual_grad_lanes<HI>(prop, ccm, dgs, h0, h1, pok);
                            pok = pok && __builtin_isfinite(pll_m);
                        }
                        c_fl += (unsigned long long)(cur_opq - 2 + cur_nc) * (unsigned long long)mala_pts;
                        pg0 = pok ? (float)h0 : 0.0f;
                        pg1 = pok ? (float)h1 : 0.0f;
                        const float b0 = gpr_axpy(h, pg0, prop.c0), b1 = gpr_axpy(h, pg1, prop.c1);
                        const double r0 = (double)cur.c0 - (double)b0, r1 = (double)cur.c1 - (double)b1;
                        double bq = wave_gauss_sum_d<HI>(r0, i0, r1, i1, r2);
                        if (mk) {
                            pmg = metric_Mv(mM, mk, gather_f(pg0, pg1, mrow, HI));
                            const float mb = gpr_axpy(h, pmg, cn);
                            bq += metric_logq(mLi, mk, l < mk ? (double)cr - (double)mb : 0.0, inv2e2);
                        }
                        extra = pok ? bq - fq : -GPR_INF;
                    }
                }
                }
            }
            const double pprior = bprior ? bpri[bu] : (have_prior ? pre_prior : wave_prior<HI>(prop, a.lnsym));
            // token classes: the current program's, or the structural proposal's
            OpMasks pmask{};
            double pll;
            int pts = d.n;
            if (have_pll) {
                pll = pll_m;
                pts = min(mala_pts, d.n);
            } else {
                pmask = structural ? op_masks<HI>(prop) : cmask;
                // a = pprior + pll - post + extra < -17 rejects without a draw (cfg_accept);
                // an LL below this bound gives a < -18 after any FP64 rounding of a
                double reject_below = -GPR_INF;
                if (i >= 0 && __builtin_isfinite(post) && __builtin_isfinite(pprior) && __builtin_isfinite(extra))
                    reject_below = post - pprior - extra - 18.0 -
                                   1e-12 * (__builtin_fabs(post) + __builtin_fabs(pprior) + __builtin_fabs(extra));
                // tapes within 64 tokens evaluate with the one-register interpreter (the high
                // words of the tape and of its masks are zero)
                if (HI && prop.len <= 64)
                    pll = wave_loglik<KC, false, true>(prop, d, dx, dy, dw, c0, nchunks, ls, &pmask, reject_below, &pts);
                else
                    pll = wave_loglik<KC, HI, true>(prop, d, dx, dy, dw, c0, nchunks, ls, &pmask, reject_below, &pts);
                pts = min(pts, d.n);
            }
            int p_opq = cur_opq, p_nc = cur_nc;
            if (structural) {
                prog_counts<HI>(prop, p_opq, p_nc);
                c_ev += (unsigned long long)prop.len * (unsigned long long)pts;
                c_fl += (unsigned long long)p_opq * (unsigned long long)pts;
                c_nc += (uint32_t)p_nc;
            } else {
                cur_pts += (unsigned long long)pts;
                cur_nev++;
            }
            if (i < 0) { prior = pprior; ll = pll; continue; }
            const double ppost = pprior + pll;
            if (cfg_accept(a.seed, chain, step, post, ppost, ppost - post + extra, adraw)) {
                if (structural) {
                    c_ev += (unsigned long long)cur.len * cur_pts;
                    c_fl += (unsigned long long)cur_opq * cur_pts;
                    c_nc += (uint32_t)cur_nc * cur_nev;
                    cur_pts = 0;
                    cur_nev = 0;
                }
                cur = prop;
                prior = pprior;
                ll = pll;
                changed = true;
                spec_n = 0;
                if (structural) {
                    c_sa++;
                    grad_ok = false;
                    zblk = 0xffffffffu;
                    zmblk = 0xffffffffu;
                    cmask = pmask;
                    cur_opq = p_opq;
                    cur_nc = p_nc;
                    cur_bok = batch_ok<BATCH, HI>(cur, mk);
                    if (MALA) {
                        int D0, D1, dm;
                        wave_depths<HI>(cur, D0, D1, dm);
                        cur_dual = !MET && mk == 0 && dual_ok<HI>(cur, dm);
                        ccm[0] = ballot(l < cur.len && cur.o0 == GPR_OP_CONST);
                        ccm[1] = HI ? ballot(l + 64 < cur.len && cur.o1 == GPR_OP_CONST) : 0;
                        cur_nt = dual_nt(__popcll(ccm[0]) + __popcll(ccm[1]), dm);
                        if (!cur_dual) gplan = grad_plan<HI>(cur);
                    }
                    cur_b16 = cur_bok && const_count<HI>(cur) <= 4;
                } else {
                    c_ca++;
                    zblk = 0xffffffffu;
                    cg0 = pg0;
                    cg1 = pg1;
                    cmg = pmg;
                }
            }
        }
        c_ev += (unsigned long long)cur.len * cur_pts;
        c_fl += (unsigned long long)cur_opq * cur_pts;
        c_nc += (uint32_t)cur_nc * cur_nev;
        {
            unsigned long long v = 0;
            v = l == GPR_CNT_SPROP ? c_sp : v;
            v = l == GPR_CNT_SACC ? c_sa : v;
            v = l == GPR_CNT_CPROP ? c_cp : v;
            v = l == GPR_CNT_CACC ? c_ca : v;
            v = l == GPR_CNT_EVALS ? c_ev : v;
            v = l == GPR_CNT_STEPS ? (unsigned long long)a.steps : v;
            v = l == GPR_CNT_FLOPS ? c_fl + c_nc : v;
            if (l < GPR_NCNT) s_cnt[wave][l] += v;
        }
        // Write back only a changed program, and only its rows below the longer of the old
        // and new lengths: the state holds zeros past a program's length (init, write-back),
        // so rows past both lengths are already right.  One-byte column writes of a single
        // chain cost a memory transaction each, so skipping them saves most of the traffic.
        if (changed) {
            const int wlen = max(len0, cur.len);
            const uint32_t cw = opaque_u32(chain);   // the load's addresses are not kept live
            if (l < wlen) { a.ops[(size_t)l * N + cw] = (uint8_t)cur.o0; a.cs[(size_t)l * N + cw] = cur.c0; }
            if (HI && l + 64 < wlen) {
                a.ops[(size_t)(l + 64) * N + cw] = (uint8_t)cur.o1;
                a.cs[(size_t)(l + 64) * N + cw] = cur.c1;
            }
        }
        if (l == 0) { a.len[chain] = cur.len; a.prior[chain] = prior; a.ll[chain] = ll; }
    }
    if (l < GPR_NCNT && a.counters) {
        const unsigned long long t = s_cnt[wave][l];
        if (t) atomicAdd(&a.counters[l], t);
    }
}

template <int KC, bool HI, bool MALA, bool MET, bool SPEC = false>
static hipError_t launch_mh_wave(const GprMHArgs &a, int grid, hipStream_t s) {
    const size_t mfloats = std::max((size_t)std::max(GPR_MALA_LDS_ROWS, GPR_MAX_DEPTH + 2) * 64,
                                    (size_t)GPR_DUAL_LDS_FLOATS);
    const size_t sh = MALA ? sizeof(float) * mfloats * a.wpg
                           : (sizeof(float) * (GPR_MAX_DEPTH + 2) * 64 * KC + sizeof(double) * 20) * (size_t)a.wpg;
    if (sh > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_mh_config_wave<KC, HI, MALA, MET, SPEC>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_mh_config_wave<KC, HI, MALA, MET, SPEC>), dim3(grid), dim3(64 * a.wpg), sh, s, a);
    return hipGetLastError();
}

template <bool MET>
static hipError_t launch_mh_wave_met(const GprMHArgs &a, int kc, bool hi, bool mala, bool small, int grid,
                                     hipStream_t s) {
    // MALA evaluates one point per lane whatever the shape's kc
    if (mala) return hi ? launch_mh_wave<1, true, true, MET>(a, grid, s) : launch_mh_wave<1, false, true, MET>(a, grid, s);
    if (kc == 1 && small)
        return hi ? launch_mh_wave<1, true, false, MET, true>(a, grid, s)
                  : launch_mh_wave<1, false, false, MET, true>(a, grid, s);
#define GPR_L(K, H) return launch_mh_wave<K, H, false, MET>(a, grid, s)
    if (kc == 1) { if (hi) GPR_L(1, true); else GPR_L(1, false); }
    else if (kc == 2) { if (hi) GPR_L(2, true); else GPR_L(2, false); }
    else { if (hi) GPR_L(4, true); else GPR_L(4, false); }
#undef GPR_L
}

hipError_t gpr_launch_mh_config_wave(const GprMHArgs &a, int kc, bool hi, bool mala, bool small, int grid,
                                     hipStream_t s) {
    return a.metric ? launch_mh_wave_met<true>(a, kc, hi, mala, small, grid, s)
                    : launch_mh_wave_met<false>(a, kc, hi, mala, small, grid, s);
}

// ---------------------------------------------------------------- config MH, lane schedule
// One lane per chain: the per-chain sequential logic of the wave kernel (and the
// oracle) with every lane running its own program.  The current program lives in the
// global SoA state (coalesced [token][chain] accesses), the proposal tape and the
// value stack in LDS as [slot][lane]; dispatch is per lane (divergent).
#define GPR_CL_WG 64
struct LaneTape {
    uint8_t *o;     // LDS ops of this lane: o[t * GPR_CL_WG]
    float *c;       // LDS constants
    __device__ __forceinline__ uint8_t &op(int t) { return o[t * GPR_CL_WG]; }
    __device__ __forceinline__ float &cv(int t) { return c[t * GPR_CL_WG]; }
};

__device__ __forceinline__ float lane_op(int o, float a, float b, float c) {
    switch (o) {
    case GPR_OP_ADD: return a + b;
    case GPR_OP_SUB: return a - b;
    case GPR_OP_MUL: return a * b;
    case GPR_OP_DIV: return a / b;
    case GPR_OP_POW: return powf(a, b);
    case GPR_OP_RSUB: return b - a;
    case GPR_OP_RDIV: return b / a;
    case GPR_OP_RPOW: return powf(b, a);
    default: return op_un(o, a, c);
    }
}

// Sampler log-likelihood of one lane's tape (reference lnormalpdf numerics, as the
// wave kernel: FP32 residual and square, compensated sum, FP64 weight), over chunks of
// 64 points.  stage: LDS [3][64] for x, y, w, used when every lane of the (one-wave)
// workgroup evaluates the same dataset (the wave stages each chunk once and the lanes
// read it as broadcasts); null: each lane reads its points from global memory.
// reject_below: the caller rejects the proposal whenever the LL lies below it.  After
// each chunk but the last the partial sum bounds the LL from above (every term is >= 0;
// a relative margin of 1e-4 covers the FP32 sum's rounding, as in the wave kernel), and a
// lane whose bound is below reject_below stops and returns -inf: the proposal is
// rejected either way.  *points receives the number of points evaluated.  With stage set
// every lane of the wave must make the call (active false: the lane has no proposal to
// evaluate this step).
__device__ double lane_loglik(LaneTape &tp, int len, float *stk, const GprDataset &d, const float *dx,
                              const float *dy, const float *dw, float *stage, double reject_below, int *points,
                              bool active = true) {
    float acc = 0.0f, comp = 0.0f;
    bool live = active;   // an inactive lane only helps stage the chunks
    *points = active ? d.n : 0;
    for (int base = 0; base < d.n; base += 64) {
        const int m = min(64, d.n - base);
        if (stage) {
            // the whole wave stages the chunk (a wave-uniform loop: every lane has this dataset)
            const int lane = (int)threadIdx.x;
            stage[lane] = lane < m ? dx[base + lane] : 0.0f;
            stage[64 + lane] = lane < m ? dy[base + lane] : 0.0f;
            stage[128 + lane] = lane < m ? dw[base + lane] : 0.0f;
            __syncthreads();
        }
        if (live) {
            for (int k = 0; k < m; k++) {
                const int i = base + k;
                const float x = stage ? stage[k] : dx[i];
                float T = 0.0f;
                int sp = 0;
                for (int t = 0; t < len; t++) {
                    const int o = tp.op(t);
                    if (o <= GPR_OP_ONE) {
                        stk[min(sp, GPR_MAX_DEPTH - 1) * GPR_CL_WG] = T;
                        sp = min(sp + 1, GPR_MAX_DEPTH);
                        T = o == GPR_OP_X ? x : (o == GPR_OP_CONST ? tp.cv(t) : 1.0f);
                    } else if (o <= GPR_OP_RPOW) {
                        sp = max(sp - 1, 0);
                        T = lane_op(o, stk[min(sp, GPR_MAX_DEPTH - 1) * GPR_CL_WG], T, 0.0f);
                    } else {
                        T = lane_op(o, T, 0.0f, tp.cv(t));
                    }
                }
                const float r = T - (stage ? stage[64 + k] : dy[i]);
                const float term = (r * r) * (stage ? stage[128 + k] : dw[i]);
                const float yv = term - comp;
                const float tv = acc + yv;
                comp = (tv - acc) - yv;
                acc = tv;
            }
            if (base + 64 < d.n && reject_below > -GPR_INF) {
                const double lb = (double)(acc - comp) * (1.0 - 1e-4);
                const double ub = d.lconst - (d.wconst > 0.0 ? d.wconst * lb : lb);
                if (!(ub >= reject_below)) {
                    live = false;
                    *points = base + 64;
                }
            }
        }
        if (stage) {
            __syncthreads();            // the chunk is read before the next one is staged
            if (!__any(live)) break;     // wave-uniform
        } else if (!live) {
            break;
        }
    }
    if (!live) return -GPR_INF;
    const double s = (double)acc - (double)comp;
    if (!__builtin_isfinite(s)) return -GPR_INF;
    return d.lconst - (d.wconst > 0.0 ? d.wconst * s : s);
}

__device__ double lane_prior(LaneTape &tp, int len, double lnsym) {
    int dl = 0;
    double p = 0.0;
    for (int t = 0; t < len; t++) {
        const int o = tp.op(t);
        dl += gpr_dl_cost(o);
        if (gpr_has_const(o)) p += gpr_lcauchy(tp.cv(t));
    }
    return gpr_dl_prior(dl, lnsym, p);
}

__global__ void __launch_bounds__(GPR_CL_WG) k_mh_config_lane(GprMHArgs a) {
    const uint32_t step0 = a.clock ? a.clock[0] : a.step0;
    extern __shared__ unsigned char smem[];
    const int tid = threadIdx.x;
    const int chain = blockIdx.x * GPR_CL_WG + tid;
    // per-dataset cap bounds the LDS tape; the launch sizes it for the largest cap
    const int capmax = a.dmax_cap;
    float *s_stack = (float *)smem;                                  // [16][WG]
    float *s_c = s_stack + GPR_MAX_DEPTH * GPR_CL_WG;                // [cap][WG]
    float *s_stage = s_c + capmax * GPR_CL_WG;                       // [3][64]: x, y, w of a chunk
    uint8_t *s_o = (uint8_t *)(s_stage + 3 * 64);                    // [cap][WG]
    const bool full = (int)(blockIdx.x + 1) * GPR_CL_WG <= a.nchains;
    if (chain >= a.nchains) return;
    int di = 0;
    while (a.ds[di].chain0 + a.ds[di].nchains <= chain) di++;
    // chains of one dataset in every lane: its points are staged through LDS
    float *stage = (full && __all(di == __shfl(di, 0))) ? s_stage : nullptr;
    const GprDataset d = a.ds[di];
    const float *dx = a.dx + d.off, *dy = a.dy + d.off, *dw = a.dw + d.off;
    const int N = a.nchains, cap = d.cap;
    LaneTape tp{s_o + tid, s_c + tid};
    float *stk = s_stack + tid;
    uint8_t *gops = a.ops + chain;
    float *gcs = a.cs + chain;
    int len = a.len[chain];
    // the LDS tape holds the current program at the start of every step; a rejected
    // proposal restores only what it changed
    for (int t = 0; t < len; t++) { tp.op(t) = gops[(size_t)t * N]; tp.cv(t) = gcs[(size_t)t * N]; }
    double prior = lane_prior(tp, len, a.lnsym);
    int pts0;
    double ll = lane_loglik(tp, len, stk, d, dx, dy, dw, stage, -GPR_INF, &pts0);
    unsigned long long n_sprop = 0, n_sacc = 0, n_cprop = 0, n_cacc = 0, n_evals = 0, n_flops = 0;
    for (int i = 0; i < a.steps; i++) {
        const uint32_t step = step0 + (uint32_t)i;
        const double post = prior + ll;
        double extra = 0.0;
        int plen = len;
        // lowest token the proposal rewrote (a structure move), or -1: constants only
        int dirty = -1;
        bool skip = false;   // no proposal this step (every lane still joins the evaluation)
        int s0 = 0;
        const bool structural = !a.fixed_structure && (step % (uint32_t)a.struct_every) == 0;
        if (structural) {
            n_sprop++;
            // symbol k from word 0 of the Philox block at counter k, its constant from word 1,
            // the node from word 2 of block 0 (the wave kernel's layout, one lane per symbol)
            const uint32_t sc2 = (uint32_t)GPR_P_STRUCT << 24;
            const uint32_t klo = (uint32_t)a.seed, khi = (uint32_t)(a.seed >> 32);
            const uint32_t r = gpr_philox((uint32_t)chain, step, sc2, 0u, klo, khi).z;
            const int j = __builtin_isfinite(post) ? (int)(((uint64_t)r * (uint64_t)len) >> 32) : len - 1;
            int need = 1;
            s0 = j;
            while (true) {
                need += gpr_arity(tp.op(s0)) - 1;
                if (need == 0) break;
                s0--;
            }
            const int oldsize = j - s0 + 1;
            const int maxnew = cap - (len - oldsize);
            // the old subtree's generation terms, before the tape changes
            double lq_old = 0.0;
            for (int t = s0; t <= j; t++) {
                const int o = tp.op(t);
                double term = gpr_gen_logp(o);
                if (o == GPR_OP_CONST) term += gpr_lphi(tp.cv(t));
                lq_old += term;
            }
            // generate in right-child-first prefix order into [cap-m, cap): that run is the postfix
            int m = 0;
            need = 1;
            double lq_new = 0.0;
            bool fail = false;
            while (need > 0) {
                if (m >= maxnew) { fail = true; break; }
                const gpr_u4 w = gpr_philox((uint32_t)chain, step, sc2, (uint32_t)m, klo, khi);
                const int sy = gpr_gen_symbol(w.x);
                float cvv = 0.0f;
                double term = gpr_gen_logp(sy);
                if (sy == GPR_OP_CONST) { cvv = gpr_normal(gpr_u01(w.y)); term += gpr_lphi(cvv); }
                lq_new += term;
                tp.op(cap - 1 - m) = (uint8_t)sy;
                tp.cv(cap - 1 - m) = cvv;
                m++;
                need += gpr_arity(sy) - 1;
            }
            dirty = min(s0, cap - m);
            if (fail) {
                skip = true;
            } else {
                plen = len - oldsize + m;
                for (int t = 0; t < m; t++) { tp.op(s0 + t) = tp.op(cap - m + t); tp.cv(s0 + t) = tp.cv(cap - m + t); }
                // the tail after the old subtree, from the current program in global memory
                // (its LDS copy may be overwritten by the shift)
                for (int t = j + 1; t < len; t++) {
                    tp.op(t - oldsize + m) = gops[(size_t)t * N];
                    tp.cv(t - oldsize + m) = gcs[(size_t)t * N];
                }
                int depth = 0, dmx = 0;
                for (int t = 0; t < plen; t++) { depth += 1 - gpr_arity(tp.op(t)); dmx = max(dmx, depth); }
                if (dmx > a.dmax) skip = true;
                extra = (-(double)gpr_logf((float)plen) + lq_old) - (-(double)gpr_logf((float)len) + lq_new);
            }
        } else {
            int nconst = 0;
            for (int t = 0; t < len; t++) nconst += tp.op(t) == GPR_OP_CONST;
            if (nconst == 0) {
                skip = true;
            } else {
                n_cprop++;
                for (int t = 0; t < len; t++) {
                    if (tp.op(t) != GPR_OP_CONST) continue;
                    const gpr_u4 w = gpr_step_block(a.seed, (uint32_t)chain, step, GPR_P_CPROP, (uint32_t)(t & 63),
                                                    (uint32_t)(t >> 6));
                    tp.cv(t) = gpr_axpy(a.sigma, gpr_normal(gpr_u01(gpr_u4_word(w, (int)(step & 3)))), tp.cv(t));
                }
            }
        }
        double pp = 0.0;
        double reject_below = -GPR_INF;
        if (!skip) {
            pp = lane_prior(tp, plen, a.lnsym);
            // a = pp + pl - post + extra < -17 rejects without a draw (cfg_accept); an LL
            // below this bound gives a < -18 after any FP64 rounding of a
            if (__builtin_isfinite(post) && __builtin_isfinite(pp) && __builtin_isfinite(extra))
                reject_below = post - pp - extra - 18.0 -
                               1e-12 * (__builtin_fabs(post) + __builtin_fabs(pp) + __builtin_fabs(extra));
        }
        int pts;
        const double pl = lane_loglik(tp, plen, stk, d, dx, dy, dw, stage, reject_below, &pts, !skip);
        bool acc = false;
        if (!skip) {
            n_evals += (unsigned long long)plen * (unsigned long long)pts;
            int nleaf = 0, nconst = 0;
            for (int t = 0; t < plen; t++) { nleaf += tp.op(t) <= GPR_OP_ONE; nconst += gpr_has_const(tp.op(t)); }
            n_flops += (unsigned long long)(plen - nleaf + 4) * (unsigned long long)pts + (unsigned long long)nconst;
            acc = cfg_accept(a.seed, (uint32_t)chain, step, post, pp + pl, (pp + pl) - post + extra);
        }
        if (acc) {
            // write back what changed: the rewritten rows, or the constants
            if (dirty >= 0) {
                for (int t = s0; t < plen; t++) { gops[(size_t)t * N] = tp.op(t); gcs[(size_t)t * N] = tp.cv(t); }
                for (int t = max(plen, s0); t < len; t++) { gops[(size_t)t * N] = 0; gcs[(size_t)t * N] = 0.0f; }
            } else {
                for (int t = 0; t < plen; t++)
                    if (tp.op(t) == GPR_OP_CONST) gcs[(size_t)t * N] = tp.cv(t);
            }
            len = plen;
            prior = pp;
            ll = pl;
            if (structural) n_sacc++; else n_cacc++;
        } else if (dirty >= 0) {
            // a structure proposal rewrote [dirty, cap): restore the current program's rows there
            for (int t = dirty; t < len; t++) { tp.op(t) = gops[(size_t)t * N]; tp.cv(t) = gcs[(size_t)t * N]; }
        } else if (!skip) {
            for (int t = 0; t < len; t++)
                if (tp.op(t) == GPR_OP_CONST) tp.cv(t) = gcs[(size_t)t * N];
        }
    }
    a.len[chain] = len;
    a.prior[chain] = prior;
    a.ll[chain] = ll;
    if (a.counters) {
        atomicAdd(&a.counters[GPR_CNT_SPROP], n_sprop);
        atomicAdd(&a.counters[GPR_CNT_SACC], n_sacc);
        atomicAdd(&a.counters[GPR_CNT_CPROP], n_cprop);
        atomicAdd(&a.counters[GPR_CNT_CACC], n_cacc);
        atomicAdd(&a.counters[GPR_CNT_EVALS], n_evals);
        atomicAdd(&a.counters[GPR_CNT_STEPS], (unsigned long long)a.steps);
        atomicAdd(&a.counters[GPR_CNT_FLOPS], n_flops);
    }
}

hipError_t gpr_launch_mh_config_lane(const GprMHArgs &a, int cap_max, hipStream_t s) {
    const int grid = (a.nchains + GPR_CL_WG - 1) / GPR_CL_WG;
    const size_t sh = sizeof(float) * ((GPR_MAX_DEPTH + (size_t)cap_max) * GPR_CL_WG + 3 * 64) +
                      (size_t)cap_max * GPR_CL_WG;
    if (sh > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_mh_config_lane, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        if (e != hipSuccess) return e;
    }
    GprMHArgs b = a;
    b.dmax_cap = cap_max;
    hipLaunchKernelGGL(k_mh_config_lane, dim3(grid), dim3(GPR_CL_WG), sh, s, b);
    return hipGetLastError();
}

// ---------------------------------------------------------------- standalone evaluator
template <int KC, bool HI>
__global__ void __launch_bounds__(GPR_WG) k_eval_wave(GprEvalArgs a) {
    const int wave = uni((int)(threadIdx.x >> 6));
    const int l = lane_id();
    const int chain = uni((int)blockIdx.x * GPR_WAVES_PER_WG + wave);
    if (chain >= a.nchains) return;
    const int N = a.nchains;
    const GprDataset d = a.dd ? *a.dd : a.d;
    WaveProg p;
    // a length outside [0, cap] is clamped: the kernel never reads or writes past the
    // caller's cap rows or its LDS slab (a malformed tape yields a meaningless value)
    p.len = min(max(uni(a.len[chain]), 0), a.cap);
    p.o0 = l < p.len ? a.ops[(size_t)l * N + chain] : 0;
    p.c0 = l < p.len ? a.cs[(size_t)l * N + chain] : 0.0f;
    p.o1 = (HI && l + 64 < p.len) ? a.ops[(size_t)(l + 64) * N + chain] : 0;
    p.c1 = (HI && l + 64 < p.len) ? a.cs[(size_t)(l + 64) * N + chain] : 0.0f;
    const int nchunks = (d.n + 64 * KC - 1) / (64 * KC);
    Chunk<KC> c0;
    c0.load(a.dx + d.off, a.dy + d.off, a.dw + d.off, d.n, 0);
    const double ll = wave_loglik<KC, HI>(p, d, a.dx + d.off, a.dy + d.off, a.dw + d.off, c0, nchunks);
    if (l == 0) a.ll[chain] = ll;
}

hipError_t gpr_launch_eval_wave(const GprEvalArgs &a, hipStream_t s) {
    const int grid = (a.nchains + GPR_WAVES_PER_WG - 1) / GPR_WAVES_PER_WG;
    const bool hi = a.cap > 64;
    const int kc = a.d.kc;
#define GPR_L(K, H) hipLaunchKernelGGL((k_eval_wave<K, H>), dim3(grid), dim3(GPR_WG), 0, s, a)
    if (kc == 1) { if (hi) GPR_L(1, true); else GPR_L(1, false); }
    else if (kc == 2) { if (hi) GPR_L(2, true); else GPR_L(2, false); }
    else { if (hi) GPR_L(4, true); else GPR_L(4, false); }
#undef GPR_L
    return hipGetLastError();
}

// Gradient of the log posterior of each tape (one wave per tape, LDS slab of cap * 64
// floats per wave for the token values and adjoint mailboxes).  grad rows t < cap: d/dc_t, 0 where token t
// has no constant or t >= len.  ll receives the log posterior.
template <bool HI>
__global__ void __launch_bounds__(GPR_WG) k_grad_wave(GprEvalArgs a) {
    const int wave = uni((int)(threadIdx.x >> 6));
    const int l = lane_id();
    const int chain = uni((int)blockIdx.x * GPR_WAVES_PER_WG + wave);
    if (chain >= a.nchains) return;
    const int N = a.nchains;
    float *lv = gpr_lds_slab + (size_t)wave * a.cap * 64;
    WaveProg p;
    // a length outside [0, cap] is clamped: the kernel never reads or writes past the
    // caller's cap rows or its LDS slab (a malformed tape yields a meaningless value)
    p.len = min(max(uni(a.len[chain]), 0), a.cap);
    p.o0 = l < p.len ? a.ops[(size_t)l * N + chain] : 0;
    p.c0 = l < p.len ? a.cs[(size_t)l * N + chain] : 0.0f;
    p.o1 = (HI && l + 64 < p.len) ? a.ops[(size_t)(l + 64) * N + chain] : 0;
    p.c1 = (HI && l + 64 < p.len) ? a.cs[(size_t)(l + 64) * N + chain] : 0.0f;
    double g0, g1;
    bool ok;
    const float *dx = a.dx + a.d.off, *dy = a.dy + a.d.off, *dw = a.dw + a.d.off;
    const GradPlan g = grad_plan<HI>(p);
    const double ll = wave_loglik_grad<HI>(p, g, a.d, dx, dy, dw, lv, g0, g1, ok);
    const double pr = wave_prior<HI>(p, a.lnsym);
    if (l < a.cap) a.grad[(size_t)l * N + chain] = (l < p.len && gpr_has_const(p.o0)) ? (float)g0 : 0.0f;
    if (HI && l + 64 < a.cap)
        a.grad[(size_t)(l + 64) * N + chain] = (l + 64 < p.len && gpr_has_const(p.o1)) ? (float)g1 : 0.0f;
    if (l == 0) a.ll[chain] = pr + ll;
}

hipError_t gpr_launch_grad_wave(const GprEvalArgs &a, hipStream_t s) {
    const int grid = (a.nchains + GPR_WAVES_PER_WG - 1) / GPR_WAVES_PER_WG;
    const size_t sh = sizeof(float) * 64 * (size_t)a.cap * GPR_WAVES_PER_WG;
    const bool hi = a.cap > 64;
    const void *f = hi ? (const void *)k_grad_wave<true> : (const void *)k_grad_wave<false>;
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
    if (hi) hipLaunchKernelGGL(k_grad_wave<true>, dim3(grid), dim3(GPR_WG), sh, s, a);
    else hipLaunchKernelGGL(k_grad_wave<false>, dim3(grid), dim3(GPR_WG), sh, s, a);
    return hipGetLastError();
}

// Lane schedule: one lane per chain; tape staged in LDS [token][lane], value stack in
// LDS [depth][lane], data points broadcast from LDS in tiles.
#define GPR_LANE_WG 128
#define GPR_LANE_DTILE 256
__global__ void __launch_bounds__(GPR_LANE_WG) k_eval_lane(GprEvalArgs a) {
    extern __shared__ unsigned char smem[];
    float *s_stack = (float *)smem;                                   // [GPR_MAX_DEPTH][WG]
    float *s_c = s_stack + GPR_MAX_DEPTH * GPR_LANE_WG;                // [cap][WG]
    float *s_x = s_c + a.cap * GPR_LANE_WG;                            // [DTILE]
    float *s_y = s_x + GPR_LANE_DTILE;
    float *s_w = s_y + GPR_LANE_DTILE;
    uint8_t *s_o = (uint8_t *)(s_w + GPR_LANE_DTILE);                  // [cap][WG]
    const int tid = threadIdx.x;
    const int chain = blockIdx.x * GPR_LANE_WG + tid;
    const bool live = chain < a.nchains;
    const int N = a.nchains;
    const GprDataset d = a.dd ? *a.dd : a.d;
    const int len = live ? min(max(a.len[chain], 0), a.cap) : 0;
    for (int t = 0; t < len; t++) {
        s_o[t * GPR_LANE_WG + tid] = a.ops[(size_t)t * N + chain];
        s_c[t * GPR_LANE_WG + tid] = a.cs[(size_t)t * N + chain];
    }
    float acc = 0.0f, comp = 0.0f;
    for (int base = 0; base < d.n; base += GPR_LANE_DTILE) {
        const int cnt = min(GPR_LANE_DTILE, d.n - base);
        __syncthreads();
        for (int i = tid; i < cnt; i += GPR_LANE_WG) {
            s_x[i] = a.dx[d.off + base + i];
            s_y[i] = a.dy[d.off + base + i];
            s_w[i] = a.dw[d.off + base + i];
        }
        __syncthreads();
        for (int i = 0; i < cnt; i++) {
            const float x = s_x[i];
            float T = 0.0f;
            int sp = 0;
            for (int t = 0; t < len; t++) {
                const int op = s_o[t * GPR_LANE_WG + tid];
                if (op <= GPR_OP_ONE) {
                    s_stack[min(sp, GPR_MAX_DEPTH - 1) * GPR_LANE_WG + tid] = T;
                    sp = min(sp + 1, GPR_MAX_DEPTH);
                    T = op == GPR_OP_X ? x : (op == GPR_OP_CONST ? s_c[t * GPR_LANE_WG + tid] : 1.0f);
                } else if (op <= GPR_OP_RPOW) {
                    sp = max(sp - 1, 0);
                    T = op_bin(op, s_stack[min(sp, GPR_MAX_DEPTH - 1) * GPR_LANE_WG + tid], T);
                } else {
                    T = op_un(op, T, s_c[t * GPR_LANE_WG + tid]);
                }
            }
            const float r = T - s_y[i];
            const float term = (r * r) * s_w[i];
            const float yv = term - comp;
            const float tv = acc + yv;
            comp = (tv - acc) - yv;
            acc = tv;
        }
    }
    if (live) {
        const double sum = (double)acc - (double)comp;
        double ll = d.lconst - (d.wconst > 0.0 ? d.wconst * sum : sum);
        if (!__builtin_isfinite(sum)) ll = -GPR_INF;
        a.ll[chain] = ll;
    }
}

hipError_t gpr_launch_eval_lane(const GprEvalArgs &a, hipStream_t s) {
    const int grid = (a.nchains + GPR_LANE_WG - 1) / GPR_LANE_WG;
    const size_t sh = sizeof(float) * (GPR_MAX_DEPTH * GPR_LANE_WG + (size_t)a.cap * GPR_LANE_WG + 3 * GPR_LANE_DTILE) +
                      (size_t)a.cap * GPR_LANE_WG;
    if (sh > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_eval_lane, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_eval_lane, dim3(grid), dim3(GPR_LANE_WG), sh, s, a);
    return hipGetLastError();
}

// Dataset header and weights from the sd column, as gpr_set_data forms them on the host:
// lconst = -sum_i 0.5 (logf(2 pi) + 2 logf(sd_i)) in FP64, and either w_i = 1 with
// wconst = 1 / (2 sd^2) when every sd is equal, or w_i = 1 / (2 sd_i^2) rounded to float
// and wconst = 0.  logf is the float rounding of the FP64 log (both correctly rounded);
// the FP64 sum runs per thread, then as a fixed tree.  A sd that is not finite and
// positive makes lconst NaN, so every likelihood of the call is NaN.
#define GPR_PREP_WG 256
__global__ void __launch_bounds__(GPR_PREP_WG) k_dataset_prep(const float *sd, int n, int nchains, int cap, int kc,
                                                              GprDataset *hdr, float *w) {
    __shared__ double s_part[GPR_PREP_WG];
    __shared__ int s_diff, s_bad;
    const int tid = threadIdx.x;
    if (tid == 0) { s_diff = 0; s_bad = 0; }
    __syncthreads();
    const float s0 = sd[0];
    double lc = 0.0;
    bool diff = false, bad = false;
    for (int i = tid; i < n; i += GPR_PREP_WG) {
        const float s = sd[i];
        bad = bad || !(s > 0.0f) || !__builtin_isfinite(s);
        diff = diff || s != s0;
        const float ls = (float)log((double)s);
        lc -= 0.5 * ((double)0x1.d67f1ep+0f + 2.0 * (double)ls);
    }
    s_part[tid] = lc;
    if (diff) atomicOr(&s_diff, 1);
    if (bad) atomicOr(&s_bad, 1);
    __syncthreads();
    for (int o = GPR_PREP_WG / 2; o > 0; o >>= 1) {
        if (tid < o) s_part[tid] += s_part[tid + o];
        __syncthreads();
    }
    const bool same = s_diff == 0;
    for (int i = tid; i < n; i += GPR_PREP_WG) {
        const double s = sd[i];
        w[i] = same ? 1.0f : (float)(1.0 / (2.0 * s * s));
    }
    if (tid == 0) {
        GprDataset g{};
        g.n = n;
        g.off = 0;
        g.chain0 = 0;
        g.nchains = nchains;
        g.cap = cap;
        g.kc = kc;
        g.lconst = s_bad ? __builtin_nan("") : s_part[0];
        g.wconst = same ? 1.0 / (2.0 * (double)s0 * (double)s0) : 0.0;
        *hdr = g;
    }
}

hipError_t gpr_launch_dataset_prep(const float *sd, int n, int nchains, int cap, int kc, GprDataset *hdr, float *w,
                                   hipStream_t s) {
    hipLaunchKernelGGL(k_dataset_prep, dim3(1), dim3(GPR_PREP_WG), 0, s, sd, n, nchains, cap, kc, hdr, w);
    return hipGetLastError();
}

// ---------------------------------------------------------------- config init
// One thread per chain: draw from the generation distribution until the program fits
// the cap and depth bound (attempt a uses stream (chain, a, INIT, 0)).
__global__ void k_init_config(GprMHArgs a) {
    const int chain = blockIdx.x * blockDim.x + threadIdx.x;
    if (chain >= a.nchains) return;
    // dataset of this chain (few datasets: linear scan of the tile table is avoided by
    // a search over the dataset table)
    int lo = 0;
    while (a.ds[lo].chain0 + a.ds[lo].nchains <= chain) lo++;
    const int cap = a.ds[lo].cap;
    const int N = a.nchains;
    for (uint32_t attempt = 0; attempt < 64; attempt++) {
        gpr_rstream st;
        st.init(a.seed, (uint32_t)chain, attempt, GPR_P_INIT, 0);
        int need = 1, m = 0, depth = 0, dmax = 0;
        bool fail = false;
        // generation order is right-child-first prefix; write reversed = postfix
        // into rows [cap-1 .. cap-m] first, then shift down.
        while (need > 0) {
            if (m >= cap) { fail = true; break; }
            const int s = gpr_gen_symbol(st.next());
            float cv = 0.0f;
            if (s == GPR_OP_CONST) cv = gpr_normal(gpr_u01(st.next()));
            a.ops[(size_t)(cap - 1 - m) * N + chain] = (uint8_t)s;
            a.cs[(size_t)(cap - 1 - m) * N + chain] = cv;
            m++;
            need += gpr_arity(s) - 1;
        }
        if (fail) continue;
        // shift rows [cap-m, cap) to [0, m) and compute the depth
        for (int t = 0; t < m; t++) {
            const uint8_t o = a.ops[(size_t)(cap - m + t) * N + chain];
            const float cv = a.cs[(size_t)(cap - m + t) * N + chain];
            a.ops[(size_t)t * N + chain] = o;
            a.cs[(size_t)t * N + chain] = cv;
            depth += 1 - gpr_arity(o);
            dmax = max(dmax, depth);
        }
        for (int t = m; t < cap; t++) { a.ops[(size_t)t * N + chain] = 0; a.cs[(size_t)t * N + chain] = 0.0f; }
        if (dmax > a.dmax) continue;
        a.len[chain] = m;
        return;
    }
    a.ops[chain] = GPR_OP_X;
    a.cs[chain] = 0.0f;
    for (int t = 1; t < cap; t++) { a.ops[(size_t)t * N + chain] = 0; a.cs[(size_t)t * N + chain] = 0.0f; }
    a.len[chain] = 1;
}

hipError_t gpr_launch_init_config(const GprMHArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_init_config, dim3((a.nchains + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- resampling
// One workgroup per dataset.  MAP rule: max posterior (NaN ignored), good = post >=
// MAP - nats, deterministic compaction of good chains in chain order.
#define GPR_RS_WG 1024
__device__ __forceinline__ double post_of(const GprResampleArgs &a, int c) { return a.prior[c] + a.ll[c]; }

__global__ void __launch_bounds__(GPR_RS_WG) k_rs_map_scan(GprResampleArgs a) {
    const GprDataset d = a.ds[blockIdx.x];
    __shared__ double s_max[GPR_RS_WG / 64];
    __shared__ int s_arg[GPR_RS_WG / 64];
    __shared__ int s_cnt[GPR_RS_WG / 64];
    __shared__ int s_base;
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    double mx = -GPR_INF;
    int arg = 0x7fffffff;
    for (int i = tid; i < d.nchains; i += GPR_RS_WG) {
        const double p = post_of(a, d.chain0 + i);
        if (p == p && (p > mx || (p == mx && i < arg))) { mx = p; arg = i; }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double m2 = __shfl_xor(mx, o);
        const int a2 = __shfl_xor(arg, o);
        if (m2 > mx || (m2 == mx && a2 < arg)) { mx = m2; arg = a2; }
    }
    if (l == 0) { s_max[w] = mx; s_arg[w] = arg; }
    __syncthreads();
    if (tid == 0) {
        double m = s_max[0];
        int ag = s_arg[0];
        for (int k = 1; k < GPR_RS_WG / 64; k++)
            if (s_max[k] > m || (s_max[k] == m && s_arg[k] < ag)) { m = s_max[k]; ag = s_arg[k]; }
        a.map_post[blockIdx.x] = m;
        a.map_chain[blockIdx.x] = ag == 0x7fffffff ? -1 : d.chain0 + ag;
        s_max[0] = m;
        s_base = 0;
    }
    __syncthreads();
    const double thr = s_max[0] - a.nats;
    for (int base = 0; base < d.nchains; base += GPR_RS_WG) {
        const int i = base + tid;
        const bool good = i < d.nchains && post_of(a, d.chain0 + i) >= thr;
        const uint64_t b = ballot(good);
        const int before = __popcll(b & ((1ull << l) - 1ull));
        if (l == 0) s_cnt[w] = __popcll(b);
        __syncthreads();
        int off = s_base;
        for (int k = 0; k < w; k++) off += s_cnt[k];
        if (good) a.good[d.chain0 + off + before] = d.chain0 + i;
        __syncthreads();
        if (tid == 0) { int t = 0; for (int k = 0; k < GPR_RS_WG / 64; k++) t += s_cnt[k]; s_base += t; }
        __syncthreads();
    }
    if (tid == 0) a.ngood[blockIdx.x] = s_base;
}

__global__ void k_rs_map_gather(GprResampleArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nchains) return;
    int di = 0;
    while (di + 1 < a.ndatasets && a.ds[di + 1].chain0 <= c) di++;
    const GprDataset d = a.ds[di];
    const double p = post_of(a, c);
    const int ng = a.ngood[di];
    if (ng == 0 || p >= a.map_post[di] - a.nats) return;
    const uint32_t r = gpr_first(a.seed, (uint32_t)c, a.clock ? a.clock[1] : a.block, GPR_P_RESAMPLE, 0);
    const int j = a.good[d.chain0 + (int)(((uint64_t)r * (uint64_t)ng) >> 32)];
    const int N = a.nchains;
    for (int t = 0; t < a.cap_rows; t++) {
        a.ops[(size_t)t * N + c] = a.ops[(size_t)t * N + j];
        a.cs[(size_t)t * N + c] = a.cs[(size_t)t * N + j];
    }
    if (a.len) a.len[c] = a.len[j];
    a.prior[c] = a.prior[j];
    a.ll[c] = a.ll[j];
    atomicAdd(a.nresampled, 1);
}

hipError_t gpr_launch_resample(const GprResampleArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_rs_map_scan, dim3(a.ndatasets), dim3(GPR_RS_WG), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rs_map_gather, dim3((a.nchains + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- small utilities
__global__ void k_fill_i32(int *p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
hipError_t gpr_launch_fill_logint(hipStream_t s) {
    hipLaunchKernelGGL(k_fill_logint, dim3(1), dim3(GPR_MAX_CAP + 1), 0, s);
    return hipGetLastError();
}

hipError_t gpr_launch_fill_i32(int *p, int n, int v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// one MCMC block done: the next block starts `steps` later and has the next index
__global__ void k_clock_advance(uint32_t *clock, uint32_t steps) {
    if (threadIdx.x == 0) {
        clock[0] += steps;
        clock[1] += 1u;
    }
}
hipError_t gpr_launch_clock_advance(uint32_t *clock, uint32_t steps, hipStream_t s) {
    hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, s, clock, steps);
    return hipGetLastError();
}

/*
 * gpuropolis.h -- C ABI of libgpuropolis.so, the MI355X (gfx950) sampler.
 *
 * The reference (piantado/GPUropolis, one CUDA file) has no library boundary: its
 * host loop in main() calls the device kernel and copies state with cudaMemcpy.
 * This header replaces those calls one for one; each entry cites what it replaces.
 *
 * Conventions: every entry returns 0 on success or a negative GPR_E_* code and sets
 * a per-thread message readable with gpr_last_error().  Nothing here exits the
 * process.  Chain state is structure-of-arrays: element t of chain c is at
 * base[t * nchains + c].  Pointers named d_* are device pointers; all others are host.
 * Calls on one context are ordered on that context's HIP stream and return after
 * the work finishes unless the name ends in _async.  One context per host thread.
 */
#ifndef GPUROPOLIS_H
#define GPUROPOLIS_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- op codes */
/* Postfix tape ops.  Leaves push; binary ops pop rhs then lhs; unary ops rewrite the
 * top.  Values follow dispatch_eval (main.cu:397-444); CADD..CRPOW use the token's
 * constant like CPLUS/CTIMES/CPOW/CRPOW there. */
enum {
    GPR_OP_X = 0, GPR_OP_CONST, GPR_OP_ONE,
    GPR_OP_ADD, GPR_OP_SUB, GPR_OP_MUL, GPR_OP_DIV, GPR_OP_POW, GPR_OP_RSUB, GPR_OP_RDIV, GPR_OP_RPOW,
    GPR_OP_NEG, GPR_OP_EXP, GPR_OP_LOG, GPR_OP_INV, GPR_OP_SQRT, GPR_OP_SIN, GPR_OP_ASIN, GPR_OP_COS,
    GPR_OP_ACOS, GPR_OP_TAN, GPR_OP_ATAN, GPR_OP_ABS,
    GPR_OP_CADD, GPR_OP_CMUL, GPR_OP_CPOW, GPR_OP_CRPOW,
    GPR_NUM_OPS
};

/* Heap op codes of the reference's default op set, enum OPS at main.cu:53, with the
 * 27-character PROGRAM_CODE of main.cu:55.  Codes 27..39 are the further ops of the
 * PHYSICAL_CONSTANTS op set (main.cu:78-86): log 2, halving and eleven physical
 * constants (their logs, main.cu:429-441).  Heap programs hold these codes for every op
 * set; an op set only chooses which codes are drawn (gpr_ref_op_set). */
enum {
    GPR_REF_ONE = 0, GPR_REF_CONST, GPR_REF_INV, GPR_REF_A, GPR_REF_B, GPR_REF_PLUS, GPR_REF_MINUS,
    GPR_REF_RMINUS, GPR_REF_CPLUS, GPR_REF_TIMES, GPR_REF_CTIMES, GPR_REF_DIV, GPR_REF_RDIV,
    GPR_REF_LOG, GPR_REF_EXP, GPR_REF_POW, GPR_REF_CPOW, GPR_REF_RPOW, GPR_REF_CRPOW, GPR_REF_SQRT,
    GPR_REF_SIN, GPR_REF_ASIN, GPR_REF_COS, GPR_REF_ACOS, GPR_REF_TAN, GPR_REF_ATAN, GPR_REF_ABS,
    GPR_REF_NOPS,                   /* 27: size of the default op set */
    GPR_REF_LOGOF2 = GPR_REF_NOPS, GPR_REF_HALF, GPR_REF_E, GPR_REF_PI, GPR_REF_CLIGHT, GPR_REF_G,
    GPR_REF_HBAR, GPR_REF_MU0, GPR_REF_EL, GPR_REF_MP, GPR_REF_ME, GPR_REF_E0, GPR_REF_KB,
    GPR_REF_NCODES
};
#define GPR_REF_SLOTS 31            /* PROGRAM_LENGTH = NCONSTANTS, main.cu:26-27 */
#define GPR_REF_PROGRAM_CODE "1CIab+-_#*@/|LE^pVPRSsCcTtA"

/* Op sets of the reference, chosen at compile time there (main.cu:51-86) and at run
 * time here (gpr_config.ops, reference semantics): which heap codes the sampler draws,
 * NOPS for the prior, and the PROGRAM_CODE characters. */
enum {
    GPR_OPS_DEFAULT = 0,            /* main.cu:51-56, 27 ops */
    GPR_OPS_SIMPLIFIED = 1,         /* SIMPLIFIED_OPS, main.cu:58-62, 13 ops */
    GPR_OPS_POLYNOMIAL = 2,         /* POLYNOMIAL_OPS, main.cu:64-68, 11 ops */
    GPR_OPS_LINEAR = 3,             /* LINEAR_OPS, main.cu:71-75, 9 ops */
    GPR_OPS_PHYSICAL = 4,           /* PHYSICAL_CONSTANTS, main.cu:78-86, 19 ops */
    GPR_NUM_OPSETS
};

/* ---------------------------------------------------------------- configuration */
enum { GPR_SEM_CONFIG = 0, GPR_SEM_REFERENCE = 1 };
enum { GPR_SCHED_AUTO = 0, GPR_SCHED_LANE = 1, GPR_SCHED_WAVE = 2 };
enum { GPR_PROP_RW = 0, GPR_PROP_MALA = 1 };
enum { GPR_RNG_PHILOX = 0, GPR_RNG_LCG = 1 };
enum { GPR_RS_NONE = 0, GPR_RS_MAP = 1, GPR_RS_QUANTILE = 2 };

enum {
    GPR_OK = 0, GPR_E_ARG = -1, GPR_E_HIP = -2, GPR_E_STATE = -3, GPR_E_NOMEM = -4, GPR_E_UNSUPPORTED = -5
};

#define GPR_MAX_CAP 128             /* postfix tokens per program */
#define GPR_MAX_DEPTH 16            /* evaluation stack depth of any program */

typedef struct {
    int device;                 /* --gpu, main.cu:1037-1048 */
    int semantics;              /* GPR_SEM_*: postfix programs or the reference's heap-31 programs */
    int cap;                    /* max tokens per program (config semantics), <= GPR_MAX_CAP */
    int dmax;                   /* max evaluation stack depth (config semantics), <= GPR_MAX_DEPTH */
    int nsym;                   /* description-length alphabet size: prior = -len*log(nsym) */
    int proposal;               /* GPR_PROP_*: constant proposals */
    float sigma;                /* random-walk sd on constants */
    float eps;                  /* MALA step size */
    int struct_every;           /* structure move when step % struct_every == 0 (main.cu:699) */
    int fixed_structure;        /* 1: constant moves only */
    int schedule;               /* GPR_SCHED_*: evaluator schedule of the MH kernel */
    int rng;                    /* GPR_RNG_*; LCG only with reference semantics + lane schedule */
    int resample;               /* GPR_RS_* rule applied by gpr_resample */
    double resample_nats;       /* MAP rule: resample chains below MAP - resample_nats */
    float resample_quantile;    /* quantile rule: RESAMPLE_QUANTILE, main.cu:36 */
    uint64_t seed;              /* Philox key */
    int ops;                    /* GPR_OPS_*: op set of reference semantics (main.cu:51-86) */
} gpr_config;

typedef struct {
    long long struct_proposed, struct_accepted;
    long long const_proposed, const_accepted;
    long long steps;            /* chain-steps executed */
    long long node_evals;       /* executed token x datapoint evaluations */
    long long alg_flops;        /* algorithmic FLOPs: n*(non-leaf tokens + 4) + constants per evaluation,
                                   plus n*(non-leaf tokens + 2 + constants) per MALA gradient */
    double last_block_ms;       /* device time of the last MH block (HIP events) */
    int schedule_used;          /* GPR_SCHED_LANE or GPR_SCHED_WAVE */
    int map_chain;              /* chain of the MAP after the last resample (dataset 0) */
    double map_posterior;
} gpr_stats;

typedef struct gpr_ctx gpr_ctx;

/* Fill cfg with the reference's defaults (main.cu:21-47, 96-109). */
int gpr_config_default(gpr_config *cfg);

/* Create a context on cfg->device.  Replaces cudaSetDevice/cudaDeviceReset
 * (main.cu:1037-1051) and the DEVARRAY allocations (main.cu:171-175, 1143-1147). */
int gpr_create(const gpr_config *cfg, gpr_ctx **out);
int gpr_destroy(gpr_ctx *ctx);
const char *gpr_last_error(void);
const char *gpr_version(void);

/* Datasets and chains.  Dataset d has n[d] points (x, y, sd concatenated over d) and
 * chains[d] chains; chains of dataset d are numbered consecutively.  caps (may be
 * NULL) gives a per-dataset token cap <= cfg->cap.  Replaces the D upload of
 * main.cu:1143 and fixes N (main.cu:111-129 --N).  Allocates all device state. */
int gpr_set_data(gpr_ctx *ctx, int ndatasets, const int32_t *n, const float *x, const float *y,
                 const float *sd, const int32_t *chains, const int32_t *caps);

/* Draw every chain from the initial distribution on the device: the generation
 * distribution (config) or the reference's uniform ops and 10^k N(0,1) constants
 * (main.cu:1133-1137). */
int gpr_init_chains(gpr_ctx *ctx);

/* Host copies of the whole state.  Config semantics: ops/cs are [cap][N], len [N].
 * Reference semantics: ops/cs are the heap arrays P/C [31][N] and len is ignored.
 * Replaces the cudaMemcpy of P, C, prior, likelihood (main.cu:1163-1166, 1233-1234). */
int gpr_write_state(gpr_ctx *ctx, const uint8_t *ops, const float *cs, const int32_t *len);
int gpr_read_state(gpr_ctx *ctx, uint8_t *ops, float *cs, int32_t *len, double *prior, double *ll);

/* The same copies for reference semantics in the reference's own types: P holds the heap
 * op codes as int (typedef int op, main.cu:45; the pinned host arrays of main.cu:1123-1130)
 * and C the constants, both [31][N], so a caller keeping main.cu's arrays passes them
 * unchanged (the cudaMemcpy calls of main.cu:1143-1144, 1163-1164 and 1233-1234).  Codes
 * outside [0, GPR_REF_NCODES) are rejected.  prior and ll may be NULL. */
int gpr_write_heap_i32(gpr_ctx *ctx, const int32_t *P, const float *C);
int gpr_read_heap_i32(gpr_ctx *ctx, int32_t *P, float *C, double *prior, double *ll);

/* The Philox clock: `step` is the MH step index the next block starts at, `block` the
 * block index the next resample draws with.  gpr_set_data resets both to 0; every block
 * advances step by its length and block by one.  No counterpart in the reference, which
 * has no checkpoint (run-all.sh:43 runs one process to the end): a checkpoint saves the
 * clock with the chains, and gpr_write_state followed by gpr_set_clock continues the
 * same random streams, so the resumed chains are bit-identical to an uninterrupted run. */
int gpr_get_clock(const gpr_ctx *ctx, uint32_t *step, uint32_t *block);
int gpr_set_clock(gpr_ctx *ctx, uint32_t step, uint32_t block);

/* One MCMC block of `steps` MH steps on every chain: the launch of MH_simple_kernel
 * (main.cu:1160, kernel main.cu:679-813).  block_seed plays the role of the rand()
 * seed argument (used by GPR_RNG_LCG); Philox streams use cfg->seed and the global
 * step index.  Writes per-chain prior and log-likelihood (FP64). */
int gpr_run_block(gpr_ctx *ctx, int steps, int32_t block_seed);
int gpr_run_block_async(gpr_ctx *ctx, int steps, int32_t block_seed);

/* Resampling between blocks (main.cu:1215-1234), on the device.  GPR_RS_MAP: per
 * dataset, chains below MAP - resample_nats copy a uniformly drawn chain within it.
 * GPR_RS_QUANTILE: chains below the quantile of the posterior are re-drawn from the
 * initial distribution (the reference rule). */
int gpr_resample(gpr_ctx *ctx);

/* `nblocks` x (gpr_run_block + gpr_resample) as one hipGraph: captured on the first call
 * with these (nblocks, steps), replayed on every call.  The step and block index live
 * on the device inside the graph, so each replay continues the chains exactly as the
 * same eager calls would (bit-identical state).  Not for GPR_RNG_LCG (its block seeds
 * come from the host's rand()).  last_block_ms = graph time / nblocks.  The run loop of
 * main.cu:1156-1239 without its per-block host round trip. */
int gpr_run_blocks_graph(gpr_ctx *ctx, int nblocks, int steps);

/* Proposal metric of one dataset's constant moves (no reference counterpart).  With
 * fixed_structure, token rows[i] (i < k <= 16) of every chain of the dataset hold
 * constants; chol is the k x k lower-triangular Cholesky factor L (row-major, upper
 * part ignored) of the proposal covariance M = L L^T.  Random walk: c' = c + sigma L z on
 * those rows; MALA: c' = c + (eps^2/2) M g + eps L z with the matching Gaussian
 * densities in the Hastings ratio.  Other constant tokens keep the identity; a chain
 * whose rows are not all constant tokens uses the identity for every constant.  k = 0
 * removes the metric. */
int gpr_set_metric(gpr_ctx *ctx, int dataset, int k, const int32_t *rows, const float *chol);

/* ---------------------------------------------------------------- autotuner
 * No reference counterpart (the reference launches one fixed grid, main.cu:1022-1023).
 * Datasets fall into size classes (n <= 64, n <= 128, n > 128; cap <= 64 or > 64),
 * each run by its own persistent launch of the wave kernels with a launch shape:
 * points per lane per chunk (the data tile, 64*kc points), waves = chains per
 * workgroup, and workgroups per CU.  gpr_autotune times every candidate shape of each
 * class present on the context's own chains (steps MH steps each, from the same
 * state, which is restored afterwards) and keeps the fastest, where a candidate must be
 * 1 % faster than an earlier one (fewer points per lane, smaller workgroups, more
 * workgroups per CU) to replace it; with GPR_SCHED_AUTO it
 * also picks the lane or wave schedule.  gpr_set_tuning applies a saved table. */
typedef struct {
    int n_class;        /* 1: n <= 64, 2: n <= 128, 4: n > 128 */
    int hi;             /* 1: cap > 64 */
    int kc;             /* points per lane per chunk: 1, 2 or 4 */
    int waves_per_wg;   /* chains per workgroup: 1, 2 or 4 */
    int wg_per_cu;      /* persistent workgroups per CU: 1..64 */
    double ms;          /* device time of the class's launch when tuned */
} gpr_tune;
int gpr_autotune(gpr_ctx *ctx, int steps, gpr_tune *out, int max_out, int *nout);
int gpr_set_tuning(gpr_ctx *ctx, const gpr_tune *t, int n, int schedule);
int gpr_get_tuning(const gpr_ctx *ctx, gpr_tune *out, int max_out, int *nout, int *schedule);

int gpr_get_stats(const gpr_ctx *ctx, gpr_stats *out);
int gpr_reset_stats(gpr_ctx *ctx);
int gpr_sync(gpr_ctx *ctx);
/* hipStream_t of the context (as void*) */
void *gpr_stream(gpr_ctx *ctx);

/* ---------------------------------------------------------------- standalone kernels */
/* Gaussian log-likelihood of nchains postfix tapes against dataset `dataset` of ctx:
 * sum_i -(f(x_i)-y_i)^2/(2 sd_i^2) - log sd_i - log(2 pi)/2, -inf if any f(x_i) or term
 * is not finite (compute_likelihood, main.cu:454-468).  Node values FP32, compensated
 * FP32 sums.  d_ops/d_cs are [cap][nchains] device arrays, d_len [nchains].
 * Preconditions (not checked: the tapes are device memory): each tape is well-formed
 * postfix of GPR_OP_* codes, len in [1, cap], evaluation stack depth <= GPR_MAX_DEPTH.
 * A length outside [0, cap] is clamped, so no access leaves the caller's arrays; a tape
 * that breaks the other conditions yields an unspecified value. */
int gpr_eval_loglik(gpr_ctx *ctx, int dataset, const uint8_t *d_ops, const float *d_cs,
                    const int32_t *d_len, int nchains, int cap, int schedule, double *d_ll);

/* d log posterior / d cs[t] by a reverse-mode adjoint over each tape (no reference
 * counterpart: the reference has no gradients).  d_grad is [cap][nchains] FP32 (0 on
 * tokens without a constant); d_lp receives the log posterior (FP64).  Preconditions
 * as gpr_eval_loglik's (the torch wrapper gpuropolis_amd/autograd.py checks them on the
 * host before the call). */
int gpr_grad_logpost(gpr_ctx *ctx, int dataset, const uint8_t *d_ops, const float *d_cs,
                     const int32_t *d_len, int nchains, int cap, float *d_grad, double *d_lp);

/* gpr_eval_loglik without a context and without allocating anything.  No counterpart in
 * the reference (which allocates everything up front, main.cu:1123-1147); it is for
 * callers that cannot allocate device memory: code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc.  The caller passes the dataset (d_x, d_y, d_sd:
 * n device floats; sd finite and > 0, else every result is NaN), the tapes as for
 * gpr_eval_loglik, and the output.  The dataset's derived values (the normaliser and
 * the per-point weights that gpr_set_data computes on the host) are formed on the device
 * in a static device array of each GPU, sized for n <= GPR_NOALLOC_MAX_POINTS (about
 * 256 KiB, held until the process ends); a larger n is rejected with GPR_E_ARG.  Every
 * call on one GPU uses that GPU's array: calls that overlap on one GPU (different
 * streams or host threads) race, and must be ordered by the caller.  The work is
 * enqueued on `stream` (a hipStream_t; NULL is the null stream) and the call returns
 * without waiting.  No hipMalloc, hipFree or other allocating call is made. */
#define GPR_NOALLOC_MAX_POINTS 65536
int gpr_eval_loglik_noalloc_async(int device, const float *d_x, const float *d_y, const float *d_sd, int n,
                                  const uint8_t *d_ops, const float *d_cs, const int32_t *d_len, int nchains,
                                  int cap, int schedule, double *d_ll, void *stream);

/* Heap-31 -> postfix decode of the reachable tree (program_fold, main.cu:371-393):
 * ops/cs receive up to 63 tokens; returns the token count or a negative error. */
int gpr_ref_decode(const uint8_t *hop, const float *hc, size_t stride, uint8_t *ops, float *cs);

/* Op set `set` (GPR_OPS_*): its NOPS heap codes in the op set's enum order (main.cu:53,
 * 59, 65, 72, 82), and in program_code (may be NULL) its PROGRAM_CODE string, one
 * character per enum value.  Returns NOPS or a negative error. */
int gpr_ref_op_set(int set, uint8_t *codes, int max_codes, const char **program_code);

/* ---------------------------------------------------------------- test hooks */
/* No reference counterpart: device-side pieces exposed so tests can compare them with
 * the host bit for bit.  gpr_test_ref_decode runs the wave decode on heaps [nheaps][31]
 * and writes tapes [nheaps][64] and meta [nheaps][3] = {length, count_nodes, description
 * length}.  gpr_test_draws writes 10 floats per input word: u01, normal quantile,
 * logf(bits & 0x7fffffff), expf(int(word) * 2^-24), sin2pi(u01), and the reference's
 * random_float, random_normal, random_geometric, host random_normal and random_int(27)
 * drawn in sequence from the substream (seed, chain = index, 0, 0, 0). */
int gpr_test_ref_decode(gpr_ctx *ctx, const uint8_t *d_hop, const float *d_hc, int nheaps, uint8_t *d_ops,
                        float *d_cs, int32_t *d_meta);
int gpr_test_draws(gpr_ctx *ctx, const uint32_t *d_words, int n, uint64_t seed, float *d_out);

#ifdef __cplusplus
}
#endif
#endif

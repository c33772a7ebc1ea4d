/*
 * GPUropolis CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A serial C restatement of the reference sampler's hot path (main.cu of
 * piantado/GPUropolis) plus the postfix/config-mode algorithm that the HIP
 * library implements.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or as the
 * reported CPU baseline.  The product (libgpuropolis.so, the gpuropolis CLI)
 * never links it.
 *
 * Layouts follow the HIP library's structure-of-arrays convention: element t
 * of chain c lives at base[t * stride + c]; per-chain (AoS) calls pass
 * stride = 1.
 */
#ifndef GPO_ORACLE_H
#define GPO_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- unified postfix op table (same codes as include/gpuropolis.h) */
enum {
    GPO_X = 0, GPO_CONST, GPO_ONE,
    GPO_ADD, GPO_SUB, GPO_MUL, GPO_DIV, GPO_POW, GPO_RSUB, GPO_RDIV, GPO_RPOW,
    GPO_NEG, GPO_EXP, GPO_LOG, GPO_INV, GPO_SQRT, GPO_SIN, GPO_ASIN, GPO_COS,
    GPO_ACOS, GPO_TAN, GPO_ATAN, GPO_ABS,
    GPO_CADD, GPO_CMUL, GPO_CPOW, GPO_CRPOW,
    GPO_NUM_OPS
};

/* reference heap op codes, main.cu:53 (enum OPS, default op set) */
enum {
    GPO_R_ONE = 0, GPO_R_CONST, GPO_R_INV, GPO_R_A, GPO_R_B, GPO_R_PLUS, GPO_R_MINUS,
    GPO_R_RMINUS, GPO_R_CPLUS, GPO_R_TIMES, GPO_R_CTIMES, GPO_R_DIV, GPO_R_RDIV,
    GPO_R_LOG, GPO_R_EXP, GPO_R_POW, GPO_R_CPOW, GPO_R_RPOW, GPO_R_CRPOW, GPO_R_SQRT,
    GPO_R_SIN, GPO_R_ASIN, GPO_R_COS, GPO_R_ACOS, GPO_R_TAN, GPO_R_ATAN, GPO_R_ABS,
    GPO_R_NOPS,
    /* the further ops of the PHYSICAL_CONSTANTS op set, main.cu:82 */
    GPO_R_LOGOF2 = GPO_R_NOPS, GPO_R_HALF, GPO_R_E, GPO_R_PI, GPO_R_CLIGHT, GPO_R_G, GPO_R_HBAR,
    GPO_R_MU0, GPO_R_EL, GPO_R_MP, GPO_R_ME, GPO_R_E0, GPO_R_KB,
    GPO_R_NCODES
};
#define GPO_HEAP 31

/* Philox purposes (counter word 2, high byte) */
enum { GPO_P_STRUCT = 1, GPO_P_CPROP = 2, GPO_P_ACCEPT = 3, GPO_P_INIT = 4, GPO_P_RESAMPLE = 5 };

/* ---------------- random number generation */
void gpo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int idx;
} gpo_stream;
void gpo_stream_init(gpo_stream *s, uint64_t seed, uint32_t chain, uint32_t step,
                     uint32_t purpose, uint32_t slot);
uint32_t gpo_stream_next(gpo_stream *s);
uint32_t gpo_lcg_next(uint32_t *rx);          /* main.cu:213-216 */

/* ---------------- spec math (the draw transforms shared with the GPU) */
float gpo_logf(float x);
float gpo_expf(float x);
float gpo_sin2pi(float v);
float gpo_normal_icdf(float u);
float gpo_u01(uint32_t r);                    /* ((r>>9)+0.5)*2^-23 in (0,1) */

/* ---------------- reference heap programs (main.cu:371-654, 660-677) */
float gpo_ref_fold_eval(const uint8_t *hop, const float *hc, size_t stride, float x);
float gpo_ref_fold_length(const uint8_t *hop, size_t stride);
int   gpo_ref_fold_count(const uint8_t *hop, size_t stride);
float gpo_ref_fold_degree(const uint8_t *hop, const float *hc, size_t stride);
int   gpo_ref_is_descendant(int n, int k);
int   gpo_ref_decode(const uint8_t *hop, const float *hc, size_t stride,
                     uint8_t *ops, float *cs, uint8_t *slot);
double gpo_ref_lnormalpdf(float x, float s);
double gpo_ref_lcauchypdf(float x, float s);
double gpo_ref_lgeometricpdf(float n, float r);
double gpo_ref_likelihood(const uint8_t *hop, const float *hc, size_t stride,
                          const float *x, const float *y, const float *sd, int n);
double gpo_ref_prior(const uint8_t *hop, const float *hc, size_t stride);

/* ---------------- postfix tapes */
int   gpo_arity(int op);
int   gpo_dl_cost(int op);
int   gpo_tape_depth(const uint8_t *ops, int len);     /* max stack depth, -1 if malformed */
int   gpo_subtree_start(const uint8_t *ops, int j);
float gpo_eval_tape_f32(const uint8_t *ops, const float *cs, int len, float x);
double gpo_eval_tape_f64(const uint8_t *ops, const float *cs, int len, double x);
/* mode 0: FP32 node values, FP64 terms and sum (the north-star oracle)
 * mode 1: FP64 node values and sum
 * mode 2: the samplers' definition, lnormalpdf of main.cu:299-301 on FP32 node values
 *         (FP32 residual and square, FP64 division and sum) */
double gpo_tape_loglik(const uint8_t *ops, const float *cs, int len,
                       const float *x, const float *y, const float *sd, int n, int mode);
double gpo_tape_prior(const uint8_t *ops, const float *cs, int len, double log_nsym,
                      int cauchy_all_slots_count);
/* log-likelihood + Cauchy terms of the constant tokens, FP64 nodes and FP64 constants */
double gpo_tape_logpost_f64(const uint8_t *ops, const double *cs, int len,
                            const float *x, const float *y, const float *sd, int n);
/* d log posterior / d cs[t] for each token t (0 for non-constant tokens), FP64 */
int gpo_tape_grad(const uint8_t *ops, const float *cs, int len,
                  const float *x, const float *y, const float *sd, int n,
                  double *grad, double *lp_out);

/* ---------------- config-mode sampler */
typedef struct {
    int cap;            /* max tokens */
    int dmax;           /* max stack depth */
    int nsym;           /* description-length alphabet size */
    int proposal;       /* 0 = Gaussian random walk, 1 = MALA */
    float sigma;        /* RW sd */
    float eps;          /* MALA step */
    int struct_every;   /* structure move when step % struct_every == 0 */
    int fixed_structure;/* 1: never propose structure */
    uint64_t seed;
    /* proposal metric of the constant moves: tokens mrows[i] (i < mk <= 16) move with the
     * lower-triangular factor mL (row-major 16 x 16); other constants use the identity */
    int mk;
    int mrows[16];
    float mL[256];
} gpo_cfg;

/* generation table: 64 buckets of the 6 high bits of a Philox word */
void gpo_cfg_gen_table(uint8_t sym[64], double logp[GPO_NUM_OPS]);
int gpo_cfg_init_chain(const gpo_cfg *cfg, uint32_t chain, uint8_t *ops, float *cs);

/* min_margin: the smallest |a - log u| / (|LL| + |LL'| + 1) over the chain's accept
 * decisions that compared a with its uniform (both posteriors finite), where a is the
 * log acceptance ratio, u the step's accept uniform and LL, LL' the current and proposed
 * log-likelihoods: how close the chain came to a decision that a last-bit difference of
 * a likelihood could flip.  +inf when no decision compared. */
typedef struct {
    double prior, ll;
    long n_struct_prop, n_struct_acc, n_const_prop, n_const_acc;
    double min_margin;
} gpo_chain_stats;

/* run `steps` MH steps of one chain starting at global step `step0` */
void gpo_cfg_run_chain(const gpo_cfg *cfg, uint32_t chain, uint64_t step0, int steps,
                       uint8_t *ops, float *cs, int *len,
                       const float *x, const float *y, const float *sd, int n,
                       gpo_chain_stats *st);

/* SoA block over chains [0, nchains) of one dataset, OpenMP over chains */
void gpo_cfg_run_block(const gpo_cfg *cfg, int nchains, uint32_t chain0, uint64_t step0, int steps,
                       uint8_t *ops, float *cs, int32_t *len,
                       const float *x, const float *y, const float *sd, int n,
                       double *prior, double *ll, long *counters, int nthreads);
/* the same, also writing each chain's min_margin (margin may be NULL) */
void gpo_cfg_run_block_m(const gpo_cfg *cfg, int nchains, uint32_t chain0, uint64_t step0, int steps,
                         uint8_t *ops, float *cs, int32_t *len,
                         const float *x, const float *y, const float *sd, int n,
                         double *prior, double *ll, long *counters, int nthreads, double *margin);

/* MAP-relative resampling of one dataset's chains (SoA, in place) */
int gpo_cfg_resample(int nchains, uint32_t chain0, uint32_t block, uint64_t seed, int cap,
                     double threshold, uint8_t *ops, float *cs, int32_t *len,
                     double *prior, double *ll, int32_t *src_out);

/* ---------------- reference-mode sampler (main.cu:679-813)
 * The reference needs nvcc and cuda_runtime.h, so no golden trajectories of it exist
 * here.  Pinned instead: its draw transforms (known answers, moments), its folds and
 * densities (closed forms), its init distribution and its median resampling rule; its
 * moves through closed forms (the forced-accept root structure proposal is uniform over
 * the op set, the constant proposal is the 0.15 integer / 0.85 normal mixture, single
 * leaf-slot changes follow q * P(u < expf(a)) with the -log count terms, a pure-x
 * program never moves); and whole chains (stale forward/backward sums included) against
 * a second, independent transcription of main.cu:683-806 in the tests. */
typedef struct {
    int rng;            /* 0 = Philox substreams, 1 = reference LCG */
    uint64_t seed;
    int ops;            /* op set: 0 default, 1 simplified, 2 polynomial, 3 linear, 4 physical */
} gpo_refcfg;

/* op sets of main.cu:51-86: size NOPS, the heap code of enum value i, PROGRAM_CODE */
int gpo_opset_nops(int set);
int gpo_opset_code(int set, int i);
const char *gpo_opset_program_code(int set);
/* compute_prior with the op set's NOPS */
double gpo_ref_prior_ops(const uint8_t *hop, const float *hc, size_t stride, int set);

void gpo_ref_run_block(const gpo_refcfg *cfg, int nchains, uint32_t chain0, uint64_t step0, int steps,
                       int32_t random_seed, uint8_t *hop, float *hc,
                       const float *x, const float *y, const float *sd, int n,
                       double *prior, double *ll, int nthreads);
/* the same, also writing each chain's accept margin as gpo_chain_stats.min_margin, with
 * a = the reference's acceptance log-ratio and u its random_float (margin may be NULL) */
void gpo_ref_run_block_m(const gpo_refcfg *cfg, int nchains, uint32_t chain0, uint64_t step0, int steps,
                         int32_t random_seed, uint8_t *hop, float *hc,
                         const float *x, const float *y, const float *sd, int n,
                         double *prior, double *ll, int nthreads, double *margin);
void gpo_ref_init(uint64_t seed, int nchains, uint32_t chain0, uint8_t *hop, float *hc, int set);
/* n draws of a reference transform (random_int/float/normal/geometric, host variants) */
void gpo_ref_draws(int kind, int rng, uint64_t seed, uint32_t lcg_seed, int n, int arg, double *out);
/* the same transforms from the Philox substream (chain, step, purpose, slot) */
void gpo_ref_stream_draws(int kind, uint64_t seed, uint32_t chain, uint32_t step, uint32_t purpose,
                          uint32_t slot, int n, int arg, double *out);
/* median rule of main.cu:1215-1231; returns the number of re-drawn chains */
int gpo_ref_resample(uint64_t seed, uint32_t block, int nchains, uint32_t chain0,
                     uint8_t *hop, float *hc, const double *prior, const double *ll,
                     double *cutoff_out, int set);

#ifdef __cplusplus
}
#endif
#endif

// Internal structures shared by the kernels (gpr_kernels.hip) and the C ABI
// (gpr_api.hip).  Not installed.
#pragma once
#include "gpr_common.h"

// One dataset as the kernels see it.  x/y/w are offsets into the concatenated device
// arrays; w = 1/(2 sd^2); lconst = sum_i (-log sd_i - log(2 pi)/2) in FP64.
struct GprDataset {
    int n;          // points
    int off;        // first point in the concatenated arrays
    int chain0;     // first chain
    int nchains;    // chains of this dataset
    int cap;        // token cap of this dataset's programs
    int kc;         // points per lane per chunk chosen for the wave schedule
    double lconst;  // -sum_i 0.5*(logf(2 pi) + 2 logf(sd_i)), host libm, FP64 sum
    double wconst;  // 1/(2 sd^2) in FP64 when sd is constant (then w[] holds 1), else 0
};

// Proposal metric of one dataset's constant moves: tokens rows[i] (i < k <= 16) move
// with the lower-triangular factor L (row-major 16 x 16); slot[t] is the metric row of
// token t or -1.  k = 0: the identity metric.  M and Li are derived from L on the host.
// The matrices come first and the struct is 16-byte aligned, so a row is read with
// 16-byte loads.
#define GPR_METRIC_MAX 16
struct alignas(16) GprMetric {
    double Li[GPR_METRIC_MAX * GPR_METRIC_MAX];  // L^-1 (FP64): proposal density
    float L[GPR_METRIC_MAX * GPR_METRIC_MAX];    // lower factor: proposal noise L z
    float M[GPR_METRIC_MAX * GPR_METRIC_MAX];    // L L^T (FP64 sums rounded): Langevin drift M g
    int rows[GPR_METRIC_MAX];
    int k;
    int8_t slot[GPR_MAX_CAP];
};

// A unit of work for the wave kernels: up to waves-per-workgroup chains of one dataset.
struct GprTile {
    int dataset;
    int chain0;
    int count;
    int pad;
};

#define GPR_WAVES_PER_WG 4
// MALA: token-value rows each wave of the MH kernel keeps in LDS (8 KB, the size the
// forward-mode interpreter's stack takes anyway); the adjoint sweeps of longer programs
// use the wave's slab in global memory (GprMHArgs::mala_gslab)
#define GPR_MALA_LDS_ROWS 32
#define GPR_WG 256

// Counters accumulated by the MH kernels (device, u64).
// FLOPS: algorithmic FLOPs n*(N_op + 4) + N_const per evaluated program (N_op = non-leaf
// tokens; the +4 is residual, square, weight, accumulate), plus n*(N_op + 2 + N_const) per
// MALA gradient (one adjoint per operator token, the residual's derivative, and each
// constant's per-point gradient summed)
enum { GPR_CNT_SPROP = 0, GPR_CNT_SACC, GPR_CNT_CPROP, GPR_CNT_CACC, GPR_CNT_EVALS, GPR_CNT_STEPS, GPR_CNT_FLOPS, GPR_NCNT };

struct GprMHArgs {
    uint8_t *ops;           // [cap][N]
    float *cs;              // [cap][N]
    int32_t *len;           // [N]
    double *prior;          // [N]
    double *ll;             // [N]
    int nchains;            // N, the stride
    const float *dx, *dy, *dw;
    const float *dsd;       // sd per point (reference lane kernel)
    const double *lnc;      // 0.5*(logf(2 pi) + 2 logf(sd)) per point, host libm (main.cu:300)
    const GprDataset *ds;
    const GprMetric *metric;  // [ndatasets] or null (identity metric everywhere)
    const GprTile *tiles;
    int ntiles;
    int *tile_counter;      // persistent work queue (reset by the host before launch)
    int steps;
    uint32_t step0;
    uint64_t seed;
    int dmax;
    double lnsym;
    float sigma;
    float eps;
    int struct_every;
    int fixed_structure;
    unsigned long long *counters;
    int32_t block_seed;
    // device clock {step, block}: when set, step0 is read from clock[0] at launch (graph
    // replays advance it on the device); null: step0 as given
    const uint32_t *clock;
    int rng;
    int dmax_cap;           // largest per-dataset cap (LDS sizing: lane kernel tapes, MALA slabs)
    int wpg;                // waves (= chains of a tile) per workgroup of the wave kernels: 1, 2 or 4
    // MALA token slabs of tapes over GPR_MALA_LDS_ROWS tokens: dmax_cap rows per resident
    // wave (workgroup * wpg + wave) in global memory; null when no cap exceeds those rows
    float *mala_gslab;
    // reference semantics: op set (GPR_OPS_*) and log(NOPS + 1) of its prior
    int opset;
    double ref_lnops;
};

struct GprEvalArgs {
    const uint8_t *ops;
    const float *cs;
    const int32_t *len;
    int nchains;
    int cap;
    const float *dx, *dy, *dw;
    GprDataset d;
    // when set, the kernels read the dataset header from this device copy instead of d
    // (the no-allocation entry computes its normaliser on the device)
    const GprDataset *dd;
    double *ll;
    float *grad;            // gradient output (grad kernel)
    double lnsym;
};

struct GprResampleArgs {
    uint8_t *ops;
    float *cs;
    int32_t *len;
    double *prior;
    double *ll;
    int nchains;
    int cap_rows;           // rows of ops/cs to copy
    const GprDataset *ds;
    int ndatasets;
    int32_t *good;          // [N] scratch: good chains per dataset, packed at chain0
    int32_t *ngood;         // [ndatasets]
    int32_t *map_chain;     // [ndatasets]
    double *map_post;       // [ndatasets]
    double *cutoff;         // [ndatasets] quantile rule
    uint64_t seed;
    uint32_t block;
    const uint32_t *clock;  // when set, block is read from clock[1]
    double nats;
    float quantile;
    int semantics;
    int32_t *nresampled;    // [1]
    int opset;              // reference semantics: op set of the re-drawn programs
};

// launchers (gpr_kernels.hip)
// small: every dataset of the launch has at most 64 points (the speculative build)
hipError_t gpr_launch_mh_config_wave(const GprMHArgs &a, int kc, bool hi, bool mala, bool small, int grid,
                                     hipStream_t s);
hipError_t gpr_launch_mh_config_lane(const GprMHArgs &a, int cap_max, hipStream_t s);
hipError_t gpr_launch_mh_ref_lane(const GprMHArgs &a, hipStream_t s);
hipError_t gpr_launch_mh_ref_wave(const GprMHArgs &a, int kc, int grid, hipStream_t s);
hipError_t gpr_launch_eval_wave(const GprEvalArgs &a, hipStream_t s);
// header (n, kc, lconst, wconst, ...) and per-point weights of one dataset from its sd
// column, on the device: the values gpr_set_data computes on the host
hipError_t gpr_launch_dataset_prep(const float *sd, int n, int nchains, int cap, int kc, GprDataset *hdr, float *w,
                                   hipStream_t s);
hipError_t gpr_launch_eval_lane(const GprEvalArgs &a, hipStream_t s);
hipError_t gpr_launch_grad_wave(const GprEvalArgs &a, hipStream_t s);
hipError_t gpr_launch_init_config(const GprMHArgs &a, hipStream_t s);
hipError_t gpr_launch_init_ref(const GprMHArgs &a, hipStream_t s);
hipError_t gpr_launch_resample(const GprResampleArgs &a, hipStream_t s);
hipError_t gpr_launch_resample_quantile(const GprResampleArgs &a, hipStream_t s);
hipError_t gpr_launch_fill_i32(int *p, int n, int v, hipStream_t s);
hipError_t gpr_launch_fill_logint(hipStream_t s);
hipError_t gpr_launch_fill_ref_logint(hipStream_t s);
hipError_t gpr_launch_clock_advance(uint32_t *clock, uint32_t steps, hipStream_t s);
hipError_t gpr_launch_test_ref_decode(const uint8_t *hop, const float *hc, int nheaps, uint8_t *ops, float *cs,
                                      int32_t *meta, hipStream_t s);
hipError_t gpr_launch_test_draws(const uint32_t *words, int n, uint64_t seed, float *out, hipStream_t s);

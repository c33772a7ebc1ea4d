// C ABI of libgpuropolis.so (declared in include/gpuropolis.h).
//
// A context owns one HIP stream on one device and every device buffer of a run:
// the datasets, the chain state in structure-of-arrays layout, the work-tile tables
// and the counters.  Host copies happen only in gpr_set_data, gpr_write_state and
// gpr_read_state; a block of MH steps plus resampling stays on the device.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "gpr_internal.h"

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess) return fail(GPR_E_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

// Size classes of datasets: n <= 64, n <= 128, n > 128, each with cap <= 64 or > 64.
// Each class has a launch shape (tuned or default) and its own persistent launch.
static int cls_of(int n) { return n <= 64 ? 0 : (n <= 128 ? 1 : 2); }
static const int kClassN[3] = {1, 2, 4};
struct ClassTune {
    int kc = 0, wpg = 1, wgpc = 16;  // points per lane per chunk, waves per workgroup, workgroups per CU
    double ms = 0.0;                 // measured by the autotuner
};
struct KernelBucket {
    int cls;
    bool hi;
    int kc, wpg, wgpc;
    int tile0, ntiles;
    size_t gslab_off;                // this launch's MALA token slabs in d_gslab (floats)
};
// Size-class launches of one block run concurrently on these streams (at most 3 classes
// x 2 register forms), forked from and joined back to the context stream.
#define GPR_MAX_BUCKETS 6

struct gpr_ctx {
    gpr_config cfg;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipStream_t side[GPR_MAX_BUCKETS] = {};
    hipEvent_t ev_fork = nullptr, ev_join[GPR_MAX_BUCKETS] = {};
    int nd = 0, N = 0, rows = 0;
    std::vector<GprDataset> hds;
    GprDataset *d_ds = nullptr;
    float *d_x = nullptr, *d_y = nullptr, *d_w = nullptr, *d_sd = nullptr;
    double *d_lnc = nullptr;
    uint8_t *d_ops = nullptr;
    float *d_cs = nullptr;
    int32_t *d_len = nullptr;
    double *d_prior = nullptr, *d_ll = nullptr;
    std::vector<GprTile> htiles;
    std::vector<KernelBucket> buckets;
    ClassTune ct[3][2];
    int tuned_sched = -1;            // schedule picked by the autotuner for GPR_SCHED_AUTO
    GprTile *d_tiles = nullptr;
    std::vector<GprMetric> hmetric;  // per dataset; k = 0: identity
    GprMetric *d_metric = nullptr;
    bool any_metric = false;
    int *d_tile_counter = nullptr;  // one per bucket
    unsigned long long *d_counters = nullptr;
    int32_t *d_good = nullptr, *d_ngood = nullptr, *d_map_chain = nullptr, *d_nres = nullptr;
    float *d_gslab = nullptr;          // MALA token slabs of programs beyond the LDS rows
    size_t gslab_floats = 0;
    double *d_map_post = nullptr, *d_cutoff = nullptr;
    uint32_t step = 0, block = 0;
    // device copy of {step, block} read by graph replays, and its host staging
    uint32_t *d_clock = nullptr;
    uint32_t h_clock[2] = {0, 0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_nblocks = 0, g_steps = 0;
    bool gexec_stale = false;        // launch shapes changed since the capture
    bool have_data = false, have_state = false;
    double last_ms = 0.0;
    int sched_used = GPR_SCHED_WAVE;
    int cus = 256;
};

extern "C" const char *gpr_last_error(void) { return g_err.c_str(); }
extern "C" const char *gpr_version(void) { return "gpuropolis-mi355x 0.1 (gfx950)"; }

extern "C" int gpr_config_default(gpr_config *c) {
    if (!c) return fail(GPR_E_ARG, "null config");
    memset(c, 0, sizeof(*c));
    c->device = 0;
    c->semantics = GPR_SEM_CONFIG;
    c->cap = 32;
    c->dmax = GPR_MAX_DEPTH;
    c->nsym = 10;
    c->proposal = GPR_PROP_RW;
    c->sigma = 0.1f;
    c->eps = 0.05f;
    c->struct_every = 5;
    c->fixed_structure = 0;
    c->schedule = GPR_SCHED_AUTO;
    c->rng = GPR_RNG_PHILOX;
    c->resample = GPR_RS_MAP;
    c->resample_nats = 100.0;
    c->resample_quantile = 0.5f;
    c->seed = 0;
    return GPR_OK;
}

static void drop_graph(gpr_ctx *c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->g_nblocks = c->g_steps = 0;
}

static void free_state(gpr_ctx *c) {
    drop_graph(c);
    void *ps[] = {c->d_ds, c->d_x, c->d_y, c->d_w, c->d_sd, c->d_lnc, c->d_ops, c->d_cs, c->d_len, c->d_prior, c->d_ll,
                  c->d_tiles, c->d_tile_counter, c->d_counters, c->d_good, c->d_ngood, c->d_map_chain,
                  c->d_nres, c->d_map_post, c->d_cutoff, c->d_clock, c->d_metric, c->d_gslab};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    c->d_ds = nullptr; c->d_x = c->d_y = c->d_w = c->d_sd = nullptr; c->d_lnc = nullptr; c->d_ops = nullptr; c->d_cs = nullptr;
    c->d_len = nullptr; c->d_prior = c->d_ll = nullptr; c->d_tiles = nullptr;
    c->d_tile_counter = nullptr; c->d_counters = nullptr; c->d_good = c->d_ngood = c->d_map_chain = nullptr;
    c->d_nres = nullptr; c->d_map_post = c->d_cutoff = nullptr; c->d_clock = nullptr; c->d_metric = nullptr;
    c->d_gslab = nullptr;
    c->gslab_floats = 0;
    c->hmetric.clear();
    c->any_metric = false;
    c->have_data = c->have_state = false;
}

extern "C" int gpr_create(const gpr_config *cfg, gpr_ctx **out) {
    if (!cfg || !out) return fail(GPR_E_ARG, "null argument");
    *out = nullptr;
    if (cfg->semantics != GPR_SEM_CONFIG && cfg->semantics != GPR_SEM_REFERENCE)
        return fail(GPR_E_ARG, "semantics must be GPR_SEM_CONFIG or GPR_SEM_REFERENCE");
    if (cfg->semantics == GPR_SEM_CONFIG && (cfg->cap < 1 || cfg->cap > GPR_MAX_CAP))
        return fail(GPR_E_ARG, "cap must be in [1, 128]");
    if (cfg->dmax < 1 || cfg->dmax > GPR_MAX_DEPTH) return fail(GPR_E_ARG, "dmax must be in [1, 16]");
    if (cfg->struct_every < 1) return fail(GPR_E_ARG, "struct_every must be >= 1");
    if (cfg->nsym < 2) return fail(GPR_E_ARG, "nsym must be >= 2");
    if (!(cfg->resample_quantile >= 0.0f && cfg->resample_quantile < 1.0f))
        return fail(GPR_E_ARG, "resample_quantile must be in [0, 1)");
    if (!(cfg->resample_nats >= 0.0)) return fail(GPR_E_ARG, "resample_nats must be >= 0");
    if (cfg->rng == GPR_RNG_LCG && cfg->semantics != GPR_SEM_REFERENCE)
        return fail(GPR_E_ARG, "the LCG stream exists only with reference semantics");
    if (cfg->ops < 0 || cfg->ops >= GPR_NUM_OPSETS) return fail(GPR_E_ARG, "ops must be a GPR_OPS_* op set");
    if (cfg->ops != GPR_OPS_DEFAULT && cfg->semantics != GPR_SEM_REFERENCE)
        return fail(GPR_E_ARG, "op sets choose the reference semantics' heap ops: reference semantics only");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (cfg->device < 0 || cfg->device >= ndev)
        return fail(GPR_E_ARG, "invalid GPU device " + std::to_string(cfg->device) + " (" + std::to_string(ndev) +
                                   " present)");
    HIP_TRY(hipSetDevice(cfg->device));
    gpr_ctx *c = new gpr_ctx();
    c->cfg = *cfg;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) == hipSuccess) c->cus = prop.multiProcessorCount;
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess &&
              hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int b = 0; ok && b < GPR_MAX_BUCKETS; b++)
        ok = hipStreamCreateWithFlags(&c->side[b], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_join[b], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        (void)gpr_destroy(c);
        return fail(GPR_E_HIP, "stream/event creation failed");
    }
    // the kernels' table of integer logs (the same values on every call)
    if (gpr_launch_fill_logint(c->stream) != hipSuccess || gpr_launch_fill_ref_logint(c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        (void)gpr_destroy(c);
        return fail(GPR_E_HIP, "log table fill failed");
    }
    *out = c;
    return GPR_OK;
}

extern "C" int gpr_destroy(gpr_ctx *c) {
    if (!c) return GPR_OK;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_state(c);
    for (int b = 0; b < GPR_MAX_BUCKETS; b++) {
        if (c->side[b]) (void)hipStreamSynchronize(c->side[b]);
        if (c->side[b]) (void)hipStreamDestroy(c->side[b]);
        if (c->ev_join[b]) (void)hipEventDestroy(c->ev_join[b]);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GPR_OK;
}

static int kc_for(int n) { return n <= 64 ? 1 : (n <= 128 ? 2 : 4); }

template <class T>
static int dev_alloc(T **p, size_t count) {
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) return fail(GPR_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return GPR_OK;
}
#define ALLOC(p, n) do { int _r = dev_alloc(&(p), (n)); if (_r) return _r; } while (0)

// Under MALA every evaluation runs at one point per lane whatever the class's kc, so the
// size classes share one kernel: they are merged into the n > 128 class, one persistent
// launch (per register form) over every dataset, largest first.
static bool mala_merged(const gpr_ctx *c) {
    return c->cfg.semantics == GPR_SEM_CONFIG && c->cfg.proposal == GPR_PROP_MALA;
}
static int bucket_class(const gpr_ctx *c, int n) { return mala_merged(c) ? 2 : cls_of(n); }

// Tiles of wpg chains of one dataset, grouped by size class (one persistent launch per
// class); larger datasets first within a class (longest-processing-time order).
static void build_tiles(gpr_ctx *c) {
    c->htiles.clear();
    c->buckets.clear();
    for (int k = 2; k >= 0; k--)
        for (int hi = 1; hi >= 0; hi--) {
            const ClassTune &t = c->ct[k][hi];
            KernelBucket b{k, hi != 0, t.kc, t.wpg, t.wgpc, (int)c->htiles.size(), 0, 0};
            std::vector<int> order;
            for (int d = 0; d < c->nd; d++)
                if (bucket_class(c, c->hds[d].n) == k && ((c->hds[d].cap > 64) == (hi != 0))) order.push_back(d);
            std::stable_sort(order.begin(), order.end(), [&](int a, int bb) { return c->hds[a].n > c->hds[bb].n; });
            for (int d : order)
                for (int s0 = 0; s0 < c->hds[d].nchains; s0 += t.wpg)
                    c->htiles.push_back(GprTile{d, c->hds[d].chain0 + s0, std::min(t.wpg, c->hds[d].nchains - s0), 0});
            b.ntiles = (int)c->htiles.size() - b.tile0;
            if (b.ntiles) c->buckets.push_back(b);
        }
}

// MALA token slabs of programs longer than GPR_MALA_LDS_ROWS tokens live in global
// memory, one per resident wave (workgroups x waves per workgroup) of each launch; the
// launches of a block run concurrently, so each has its own region.  Sized whenever the
// tiles change, before any launch, so never inside a graph capture.
static int ensure_gslab(gpr_ctx *c) {
    if (c->cfg.semantics != GPR_SEM_CONFIG || c->cfg.proposal != GPR_PROP_MALA || c->rows <= GPR_MALA_LDS_ROWS)
        return GPR_OK;
    size_t need = 0;
    for (KernelBucket &kb : c->buckets) {
        if (kb.gslab_off != need) c->gexec_stale = true;
        kb.gslab_off = need;
        need += (size_t)std::max(1, std::min(kb.ntiles, c->cus * kb.wgpc)) * (size_t)kb.wpg * (size_t)c->rows * 64;
    }
    if (need <= c->gslab_floats) return GPR_OK;
    if (c->d_gslab) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipFree(c->d_gslab));
        c->d_gslab = nullptr;
        c->gslab_floats = 0;
    }
    ALLOC(c->d_gslab, need);
    c->gslab_floats = need;
    c->gexec_stale = true;
    return GPR_OK;
}

static int upload_tiles(gpr_ctx *c) {
    HIP_TRY(hipMemcpyAsync(c->d_tiles, c->htiles.data(), sizeof(GprTile) * c->htiles.size(), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ensure_gslab(c);
}

extern "C" int gpr_set_data(gpr_ctx *c, int nd, const int32_t *n, const float *x, const float *y, const float *sd,
                            const int32_t *chains, const int32_t *caps) {
    if (!c || nd < 1 || !n || !x || !y || !sd || !chains) return fail(GPR_E_ARG, "null or empty dataset arguments");
    HIP_TRY(hipSetDevice(c->cfg.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    free_state(c);
    const bool ref = c->cfg.semantics == GPR_SEM_REFERENCE;
    long long total_pts = 0, total_chains = 0;
    for (int d = 0; d < nd; d++) {
        if (n[d] < 1) return fail(GPR_E_ARG, "dataset " + std::to_string(d) + " has no points");
        if (chains[d] < 1) return fail(GPR_E_ARG, "dataset " + std::to_string(d) + " has no chains");
        if (caps && !ref && (caps[d] < 1 || caps[d] > c->cfg.cap))
            return fail(GPR_E_ARG, "per-dataset cap out of [1, cfg.cap]");
        total_pts += n[d];
        total_chains += chains[d];
    }
    if (total_chains > (1ll << 30)) return fail(GPR_E_ARG, "too many chains");
    std::vector<float> w((size_t)total_pts);
    std::vector<double> lnc((size_t)total_pts);
    c->hds.assign(nd, GprDataset{});
    long long off = 0, ch = 0;
    for (int d = 0; d < nd; d++) {
        GprDataset &g = c->hds[d];
        g.n = n[d];
        g.off = (int)off;
        g.chain0 = (int)ch;
        g.nchains = chains[d];
        g.cap = ref ? 63 : (caps ? caps[d] : c->cfg.cap);
        g.kc = kc_for(g.n);
        double lc = 0.0;
        bool same_sd = true;
        for (int i = 0; i < n[d]; i++) {
            const double s = sd[off + i];
            if (!(s > 0.0) || !std::isfinite(s) || !std::isfinite(x[off + i]) || !std::isfinite(y[off + i]))
                return fail(GPR_E_ARG, "dataset " + std::to_string(d) + " point " + std::to_string(i) +
                                           ": need finite x, y and sd > 0 (main.cu:332)");
            same_sd = same_sd && sd[off + i] == sd[off];
            // lnormalpdf's normalizer with float logs (main.cu:300)
            lnc[off + i] = 0.5 * ((double)0x1.d67f1ep+0f + 2.0 * (double)logf(sd[off + i]));
            lc -= lnc[off + i];
        }
        const double s0 = sd[off];
        g.wconst = same_sd ? 1.0 / (2.0 * s0 * s0) : 0.0;
        for (int i = 0; i < n[d]; i++) {
            const double s = sd[off + i];
            w[off + i] = same_sd ? 1.0f : (float)(1.0 / (2.0 * s * s));
        }
        g.lconst = lc;
        off += n[d];
        ch += chains[d];
    }
    c->nd = nd;
    c->N = (int)total_chains;
    c->rows = ref ? GPR_REF_SLOTS : c->cfg.cap;
    // default shapes: one chain per workgroup, 32 workgroups per CU; the postfix sampler at
    // one point per lane per chunk (it stops rejected proposals after any 64-point chunk and
    // its one-point kernel runs six waves per SIMD), the reference sampler with the chunk
    // widened to the class's points
    for (int k = 0; k < 3; k++)
        for (int h = 0; h < 2; h++) c->ct[k][h] = ClassTune{ref ? kClassN[k] : 1, 1, 32, 0.0};
    c->tuned_sched = -1;
    build_tiles(c);
    const size_t N = (size_t)c->N;
    ALLOC(c->d_ds, (size_t)nd);
    ALLOC(c->d_x, (size_t)total_pts);
    ALLOC(c->d_y, (size_t)total_pts);
    ALLOC(c->d_w, (size_t)total_pts);
    ALLOC(c->d_sd, (size_t)total_pts);
    ALLOC(c->d_lnc, (size_t)total_pts);
    ALLOC(c->d_ops, (size_t)c->rows * N);
    ALLOC(c->d_cs, (size_t)c->rows * N);
    ALLOC(c->d_len, N);
    ALLOC(c->d_prior, N);
    ALLOC(c->d_ll, N);
    ALLOC(c->d_tiles, N);                  // enough for one chain per tile
    ALLOC(c->d_tile_counter, GPR_MAX_BUCKETS);
    ALLOC(c->d_counters, (size_t)GPR_NCNT);
    ALLOC(c->d_good, N);
    ALLOC(c->d_ngood, (size_t)nd);
    ALLOC(c->d_map_chain, (size_t)nd);
    ALLOC(c->d_map_post, (size_t)nd);
    ALLOC(c->d_clock, 2);
    ALLOC(c->d_metric, (size_t)nd);
    c->hmetric.assign(nd, GprMetric{});
    for (GprMetric &m : c->hmetric) memset(m.slot, -1, sizeof(m.slot));
    ALLOC(c->d_cutoff, (size_t)nd);
    ALLOC(c->d_nres, 1);
    // everything on the context stream (non-blocking, so not ordered with the null
    // stream's asynchronous memsets), finished before returning
    const hipStream_t st = c->stream;
    HIP_TRY(hipMemcpyAsync(c->d_ds, c->hds.data(), sizeof(GprDataset) * nd, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_x, x, sizeof(float) * total_pts, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_y, y, sizeof(float) * total_pts, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_w, w.data(), sizeof(float) * total_pts, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_sd, sd, sizeof(float) * total_pts, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_lnc, lnc.data(), sizeof(double) * total_pts, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_tiles, c->htiles.data(), sizeof(GprTile) * c->htiles.size(), hipMemcpyHostToDevice,
                           st));
    c->gexec_stale = true;
    HIP_TRY(hipMemsetAsync(c->d_counters, 0, sizeof(unsigned long long) * GPR_NCNT, st));
    HIP_TRY(hipMemsetAsync(c->d_ops, 0, (size_t)c->rows * N, st));
    HIP_TRY(hipMemsetAsync(c->d_cs, 0, sizeof(float) * c->rows * N, st));
    HIP_TRY(hipMemsetAsync(c->d_len, 0, sizeof(int32_t) * N, st));
    HIP_TRY(hipMemsetAsync(c->d_prior, 0, sizeof(double) * N, st));
    HIP_TRY(hipMemsetAsync(c->d_ll, 0, sizeof(double) * N, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (int r = ensure_gslab(c)) return r;
    c->step = 0;
    c->block = 0;
    c->have_data = true;
    return GPR_OK;
}

static GprMHArgs mh_args(gpr_ctx *c, int steps, int32_t block_seed) {
    GprMHArgs a{};
    a.ops = c->d_ops; a.cs = c->d_cs; a.len = c->d_len; a.prior = c->d_prior; a.ll = c->d_ll;
    a.nchains = c->N;
    a.dx = c->d_x; a.dy = c->d_y; a.dw = c->d_w;
    a.dsd = c->d_sd;
    a.lnc = c->d_lnc;
    a.ds = c->d_ds;
    a.metric = c->any_metric ? c->d_metric : nullptr;
    a.tiles = c->d_tiles;
    a.ntiles = (int)c->htiles.size();
    a.tile_counter = c->d_tile_counter;
    a.steps = steps;
    a.step0 = c->step;
    a.seed = c->cfg.seed;
    a.dmax = c->cfg.dmax;
    a.lnsym = log((double)c->cfg.nsym);
    a.opset = c->cfg.ops;
    a.ref_lnops = log((double)gpr_opset_nops(c->cfg.ops) + 1.0);   // log((bayes_t)NOPS+1), main.cu:536
    a.sigma = c->cfg.sigma;
    a.eps = c->cfg.eps;
    a.struct_every = c->cfg.struct_every;
    a.fixed_structure = c->cfg.fixed_structure;
    a.counters = c->d_counters;
    a.block_seed = block_seed;
    a.rng = c->cfg.rng;
    return a;
}

extern "C" int gpr_init_chains(gpr_ctx *c) {
    if (!c || !c->have_data) return fail(GPR_E_STATE, "gpr_set_data must come first");
    HIP_TRY(hipSetDevice(c->cfg.device));
    GprMHArgs a = mh_args(c, 0, 0);
    if (c->cfg.semantics == GPR_SEM_CONFIG) {
        HIP_TRY(gpr_launch_init_config(a, c->stream));
    } else {
        HIP_TRY(gpr_launch_init_ref(a, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->have_state = true;
    return GPR_OK;
}

extern "C" int gpr_write_state(gpr_ctx *c, const uint8_t *ops, const float *cs, const int32_t *len) {
    if (!c || !c->have_data) return fail(GPR_E_STATE, "gpr_set_data must come first");
    if (!ops || !cs) return fail(GPR_E_ARG, "null state arrays");
    const bool ref = c->cfg.semantics == GPR_SEM_REFERENCE;
    const size_t N = (size_t)c->N;
    if (!ref) {
        if (!len) return fail(GPR_E_ARG, "null len");
        for (size_t i = 0; i < N; i++) {
            int ds = 0;
            while (ds + 1 < c->nd && c->hds[ds + 1].chain0 <= (int)i) ds++;
            const int L = len[i];
            if (L < 1 || L > c->hds[ds].cap) return fail(GPR_E_ARG, "chain " + std::to_string(i) + ": len out of [1, cap]");
            int depth = 0, dm = 0;
            for (int t = 0; t < L; t++) {
                const int o = ops[(size_t)t * N + i];
                if (o >= GPR_NUM_OPS) return fail(GPR_E_ARG, "chain " + std::to_string(i) + ": bad op code");
                const int ar = gpr_arity(o);
                if (depth < ar) return fail(GPR_E_ARG, "chain " + std::to_string(i) + ": malformed postfix");
                depth += 1 - ar;
                dm = std::max(dm, depth);
            }
            if (depth != 1) return fail(GPR_E_ARG, "chain " + std::to_string(i) + ": malformed postfix");
            if (dm > c->cfg.dmax) return fail(GPR_E_ARG, "chain " + std::to_string(i) + ": stack depth above dmax");
        }
    } else {
        for (size_t i = 0; i < N * GPR_REF_SLOTS; i++)
            if (ops[i] >= GPR_REF_NCODES) return fail(GPR_E_ARG, "bad heap op code");
    }
    HIP_TRY(hipSetDevice(c->cfg.device));
    HIP_TRY(hipMemcpyAsync(c->d_ops, ops, (size_t)c->rows * N, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_cs, cs, sizeof(float) * c->rows * N, hipMemcpyHostToDevice, c->stream));
    if (!ref) HIP_TRY(hipMemcpyAsync(c->d_len, len, sizeof(int32_t) * N, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->have_state = true;
    return GPR_OK;
}

extern "C" int gpr_read_state(gpr_ctx *c, uint8_t *ops, float *cs, int32_t *len, double *prior, double *ll) {
    if (!c || !c->have_data) return fail(GPR_E_STATE, "no state");
    HIP_TRY(hipSetDevice(c->cfg.device));
    const size_t N = (size_t)c->N;
    if (ops) HIP_TRY(hipMemcpyAsync(ops, c->d_ops, (size_t)c->rows * N, hipMemcpyDeviceToHost, c->stream));
    if (cs) HIP_TRY(hipMemcpyAsync(cs, c->d_cs, sizeof(float) * c->rows * N, hipMemcpyDeviceToHost, c->stream));
    if (len) HIP_TRY(hipMemcpyAsync(len, c->d_len, sizeof(int32_t) * N, hipMemcpyDeviceToHost, c->stream));
    if (prior) HIP_TRY(hipMemcpyAsync(prior, c->d_prior, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    if (ll) HIP_TRY(hipMemcpyAsync(ll, c->d_ll, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GPR_OK;
}

// Heap state in main.cu's types (int op codes), converted on the host.
extern "C" int gpr_write_heap_i32(gpr_ctx *c, const int32_t *P, const float *C) {
    if (!c || !c->have_data) return fail(GPR_E_STATE, "gpr_set_data must come first");
    if (c->cfg.semantics != GPR_SEM_REFERENCE) return fail(GPR_E_ARG, "heap arrays need reference semantics");
    if (!P || !C) return fail(GPR_E_ARG, "null heap arrays");
    const size_t n = (size_t)GPR_REF_SLOTS * (size_t)c->N;
    std::vector<uint8_t> ops(n);
    for (size_t i = 0; i < n; i++) {
        if (P[i] < 0 || P[i] >= GPR_REF_NCODES)
            return fail(GPR_E_ARG, "heap op code " + std::to_string(P[i]) + " at element " + std::to_string(i) +
                                       " is outside [0, " + std::to_string(GPR_REF_NCODES) + ")");
        ops[i] = (uint8_t)P[i];
    }
    return gpr_write_state(c, ops.data(), C, nullptr);
}

extern "C" int gpr_read_heap_i32(gpr_ctx *c, int32_t *P, float *C, double *prior, double *ll) {
    if (!c || !c->have_data) return fail(GPR_E_STATE, "no state");
    if (c->cfg.semantics != GPR_SEM_REFERENCE) return fail(GPR_E_ARG, "heap arrays need reference semantics");
    const size_t n = (size_t)GPR_REF_SLOTS * (size_t)c->N;
    std::vector<uint8_t> ops(P ? n : 0);
    if (int r = gpr_read_state(c, P ? ops.data() : nullptr, C, nullptr, prior, ll)) return r;
    for (size_t i = 0; P && i < n; i++) P[i] = ops[i];
    return GPR_OK;
}

// The host copies of the clock are authoritative between calls: graph replays upload
// them before the launch and every block advances them on return.
extern "C" int gpr_get_clock(const gpr_ctx *c, uint32_t *step, uint32_t *block) {
    if (!c || !step || !block) return fail(GPR_E_ARG, "null argument");
    *step = c->step;
    *block = c->block;
    return GPR_OK;
}

extern "C" int gpr_set_clock(gpr_ctx *c, uint32_t step, uint32_t block) {
    if (!c) return fail(GPR_E_ARG, "null context");
    if (!c->have_data) return fail(GPR_E_STATE, "gpr_set_data must come first (it resets the clock)");
    c->step = step;
    c->block = block;
    return GPR_OK;
}

// The persistent wave-schedule launch of one size class (its tile counter must be 0).
static hipError_t launch_bucket(gpr_ctx *c, const GprMHArgs &a, int b, hipStream_t s) {
    const KernelBucket &kb = c->buckets[b];
    GprMHArgs ab = a;
    ab.tiles = c->d_tiles + kb.tile0;
    ab.ntiles = kb.ntiles;
    ab.tile_counter = c->d_tile_counter + b;
    ab.dmax_cap = c->rows;
    ab.wpg = kb.wpg;
    ab.mala_gslab = c->d_gslab ? c->d_gslab + kb.gslab_off : nullptr;
    const int grid = std::max(1, std::min(kb.ntiles, c->cus * kb.wgpc));
    if (c->cfg.semantics == GPR_SEM_REFERENCE) return gpr_launch_mh_ref_wave(ab, kb.kc, grid, s);
    return gpr_launch_mh_config_wave(ab, kb.kc, kb.hi, c->cfg.proposal == GPR_PROP_MALA, kb.cls == 0, grid, s);
}

// All size-class launches of one block.  They touch disjoint chains and have their own
// tile counters and slabs, so they run concurrently: each persistent launch fills the
// GPU, and the next one's workgroups take the CUs that the previous one's tail leaves
// idle.  Forked from the context stream after the counter reset and joined back to it
// (also the form a graph capture records).
static hipError_t launch_buckets(gpr_ctx *c, const GprMHArgs &a) {
    const int nb = (int)c->buckets.size();
    if (nb == 1) return launch_bucket(c, a, 0, c->stream);
    hipError_t e = hipEventRecord(c->ev_fork, c->stream);
    for (int b = 0; b < nb && e == hipSuccess; b++) {
        e = hipStreamWaitEvent(c->side[b], c->ev_fork, 0);
        if (e == hipSuccess) e = launch_bucket(c, a, b, c->side[b]);
        if (e == hipSuccess) e = hipEventRecord(c->ev_join[b], c->side[b]);
    }
    for (int b = 0; b < nb && e == hipSuccess; b++) e = hipStreamWaitEvent(c->stream, c->ev_join[b], 0);
    return e;
}

// One MCMC block on the context stream.  on_device: step and block come from the
// device clock, which a final kernel advances (the form captured into graphs); else
// from the host counters, advanced here.
static int launch_block(gpr_ctx *c, int steps, int32_t block_seed, bool on_device = false) {
    GprMHArgs a = mh_args(c, steps, block_seed);
    if (on_device) a.clock = c->d_clock;
    const bool ref = c->cfg.semantics == GPR_SEM_REFERENCE;
    int sched = c->cfg.schedule;
    if (sched == GPR_SCHED_AUTO)
        sched = (ref && c->cfg.rng == GPR_RNG_LCG) ? GPR_SCHED_LANE
                                                   : (c->tuned_sched > 0 ? c->tuned_sched : GPR_SCHED_WAVE);
    if (c->cfg.rng == GPR_RNG_LCG && sched != GPR_SCHED_LANE)
        return fail(GPR_E_ARG, "the LCG stream is sequential per chain: it needs the lane schedule");
    if (sched == GPR_SCHED_LANE) {
        if (ref) HIP_TRY(gpr_launch_mh_ref_lane(a, c->stream));
        else {
            if (c->cfg.proposal == GPR_PROP_MALA)
                return fail(GPR_E_UNSUPPORTED, "MALA runs on the wave schedule only");
            int capmax = 1;
            for (const GprDataset &g : c->hds) capmax = std::max(capmax, g.cap);
            HIP_TRY(gpr_launch_mh_config_lane(a, capmax, c->stream));
        }
    } else {
        HIP_TRY(gpr_launch_fill_i32(c->d_tile_counter, (int)c->buckets.size(), 0, c->stream));
        HIP_TRY(launch_buckets(c, a));
    }
    c->sched_used = sched;
    if (on_device) {
        HIP_TRY(gpr_launch_clock_advance(c->d_clock, (uint32_t)steps, c->stream));
    } else {
        c->step += (uint32_t)steps;
        c->block++;
    }
    return GPR_OK;
}

extern "C" int gpr_run_block_async(gpr_ctx *c, int steps, int32_t block_seed) {
    if (!c || !c->have_state) return fail(GPR_E_STATE, "no chain state: call gpr_init_chains or gpr_write_state");
    if (steps < 1) return fail(GPR_E_ARG, "steps must be > 0 (main.cu:1028)");
    HIP_TRY(hipSetDevice(c->cfg.device));
    return launch_block(c, steps, block_seed);
}

extern "C" int gpr_run_block(gpr_ctx *c, int steps, int32_t block_seed) {
    if (!c || !c->have_state) return fail(GPR_E_STATE, "no chain state: call gpr_init_chains or gpr_write_state");
    if (steps < 1) return fail(GPR_E_ARG, "steps must be > 0 (main.cu:1028)");
    HIP_TRY(hipSetDevice(c->cfg.device));
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    int r = launch_block(c, steps, block_seed);
    if (r) return r;
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return GPR_OK;
}

static GprResampleArgs rs_args(gpr_ctx *c) {
    GprResampleArgs r{};
    r.ops = c->d_ops; r.cs = c->d_cs;
    r.len = c->cfg.semantics == GPR_SEM_CONFIG ? c->d_len : nullptr;
    r.prior = c->d_prior; r.ll = c->d_ll;
    r.nchains = c->N;
    r.cap_rows = c->rows;
    r.ds = c->d_ds;
    r.ndatasets = c->nd;
    r.good = c->d_good; r.ngood = c->d_ngood; r.map_chain = c->d_map_chain; r.map_post = c->d_map_post;
    r.cutoff = c->d_cutoff;
    r.seed = c->cfg.seed;
    r.block = c->block;
    r.nats = c->cfg.resample_nats;
    r.quantile = c->cfg.resample_quantile;
    r.semantics = c->cfg.semantics;
    r.opset = c->cfg.ops;
    r.nresampled = c->d_nres;
    return r;
}

static int launch_resample(gpr_ctx *c, bool on_device) {
    GprResampleArgs r = rs_args(c);
    if (on_device) r.clock = c->d_clock;
    HIP_TRY(gpr_launch_fill_i32(c->d_nres, 1, 0, c->stream));
    if (c->cfg.resample == GPR_RS_QUANTILE) {
        if (c->cfg.semantics != GPR_SEM_REFERENCE)
            return fail(GPR_E_ARG, "the quantile re-draw rule draws heap programs: reference semantics only");
        HIP_TRY(gpr_launch_resample_quantile(r, c->stream));
    } else {
        HIP_TRY(gpr_launch_resample(r, c->stream));
    }
    return GPR_OK;
}

extern "C" int gpr_resample(gpr_ctx *c) {
    if (!c || !c->have_state) return fail(GPR_E_STATE, "no chain state");
    if (c->cfg.resample == GPR_RS_NONE) return GPR_OK;
    HIP_TRY(hipSetDevice(c->cfg.device));
    const int r = launch_resample(c, false);
    if (r) return r;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GPR_OK;
}

// Capture nblocks x (MH block, clock advance, resampling) once per (nblocks, steps)
// and replay it; the device clock carries the step and block index across replays, so
// a replay continues the chains exactly as the same eager calls would.  Only kernels
// are recorded (counters are reset by fill kernels, never by memset nodes).
extern "C" int gpr_run_blocks_graph(gpr_ctx *c, int nblocks, int steps) {
    if (!c || !c->have_state) return fail(GPR_E_STATE, "no chain state: call gpr_init_chains or gpr_write_state");
    if (nblocks < 1 || steps < 1) return fail(GPR_E_ARG, "nblocks and steps must be > 0");
    if (c->cfg.rng == GPR_RNG_LCG)
        return fail(GPR_E_ARG, "LCG blocks are seeded from the host's rand(): run them with gpr_run_block");
    HIP_TRY(hipSetDevice(c->cfg.device));
    if (!c->gexec || c->g_nblocks != nblocks || c->g_steps != steps || c->gexec_stale) {
        c->gexec_stale = false;
        drop_graph(c);
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        int r = GPR_OK;
        for (int b = 0; b < nblocks && r == GPR_OK; b++) {
            r = launch_block(c, steps, 0, true);
            if (r == GPR_OK && c->cfg.resample != GPR_RS_NONE) r = launch_resample(c, true);
        }
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (r != GPR_OK) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        if (e != hipSuccess) return fail(GPR_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        c->graph = g;
        HIP_TRY(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        c->g_nblocks = nblocks;
        c->g_steps = steps;
    }
    c->h_clock[0] = c->step;
    c->h_clock[1] = c->block;
    HIP_TRY(hipMemcpyAsync(c->d_clock, c->h_clock, sizeof(c->h_clock), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms / nblocks;
    c->step += (uint32_t)(nblocks * steps);
    c->block += (uint32_t)nblocks;
    return GPR_OK;
}

// M = L L^T (FP64 sums in increasing index order, rounded to FP32) and L^-1 (FP64 forward
// substitution): the derived forms the kernels use, in the oracle's operation order.
static void metric_derive(GprMetric &m) {
#pragma clang fp contract(off)
    const int k = m.k, S = GPR_METRIC_MAX;
    for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) {
            double s = 0.0;
            for (int q = 0; q <= std::min(i, j); q++) s = s + (double)m.L[i * S + q] * (double)m.L[j * S + q];
            m.M[i * S + j] = (float)s;
        }
    for (int c = 0; c < k; c++)
        for (int i = 0; i < k; i++) {
            if (i < c) { m.Li[i * S + c] = 0.0; continue; }
            double s = i == c ? 1.0 : 0.0;
            for (int q = c; q < i; q++) s = s - (double)m.L[i * S + q] * m.Li[q * S + c];
            m.Li[i * S + c] = s / (double)m.L[i * S + i];
        }
}

extern "C" int gpr_set_metric(gpr_ctx *c, int dataset, int k, const int32_t *rows, const float *chol) {
    if (!c || !c->have_data) return fail(GPR_E_STATE, "gpr_set_data must come first");
    if (dataset < 0 || dataset >= c->nd) return fail(GPR_E_ARG, "dataset index out of range");
    if (c->cfg.semantics != GPR_SEM_CONFIG) return fail(GPR_E_ARG, "a proposal metric needs config semantics");
    if (k > 0 && !c->cfg.fixed_structure)
        return fail(GPR_E_ARG, "a proposal metric names token rows: it needs fixed_structure");
    if (k < 0 || k > GPR_METRIC_MAX) return fail(GPR_E_ARG, "metric size must be in [0, 16]");
    if (k > 0 && (!rows || !chol)) return fail(GPR_E_ARG, "null metric rows or factor");
    GprMetric m{};
    memset(m.slot, -1, sizeof(m.slot));
    m.k = k;
    for (int i = 0; i < k; i++) {
        if (rows[i] < 0 || rows[i] >= c->hds[dataset].cap) return fail(GPR_E_ARG, "metric row outside the cap");
        if (m.slot[rows[i]] >= 0) return fail(GPR_E_ARG, "metric rows must be distinct");
        m.rows[i] = rows[i];
        m.slot[rows[i]] = (int8_t)i;
        for (int j = 0; j < GPR_METRIC_MAX; j++) {
            const float v = j <= i && j < k ? chol[i * k + j] : 0.0f;
            if (!std::isfinite(v)) return fail(GPR_E_ARG, "metric factor must be finite");
            m.L[i * GPR_METRIC_MAX + j] = v;
        }
        if (!(m.L[i * GPR_METRIC_MAX + i] > 0.0f)) return fail(GPR_E_ARG, "metric factor needs a positive diagonal");
    }
    metric_derive(m);
    HIP_TRY(hipSetDevice(c->cfg.device));
    c->hmetric[dataset] = m;
    c->any_metric = false;
    for (const GprMetric &g : c->hmetric) c->any_metric = c->any_metric || g.k > 0;
    HIP_TRY(hipMemcpyAsync(c->d_metric + dataset, &c->hmetric[dataset], sizeof(GprMetric), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->gexec_stale = true;
    return GPR_OK;
}

// ---------------------------------------------------------------- autotuner
static int valid_tune(const gpr_tune &t) {
    return (t.n_class == 1 || t.n_class == 2 || t.n_class == 4) && (t.hi == 0 || t.hi == 1) &&
           (t.kc == 1 || t.kc == 2 || t.kc == 4) && (t.waves_per_wg == 1 || t.waves_per_wg == 2 || t.waves_per_wg == 4) &&
           t.wg_per_cu >= 1 && t.wg_per_cu <= 64;
}
static int class_index(int n_class) { return n_class == 1 ? 0 : (n_class == 2 ? 1 : 2); }

extern "C" int gpr_set_tuning(gpr_ctx *c, const gpr_tune *t, int n, int schedule) {
    if (!c || !c->have_data) return fail(GPR_E_STATE, "gpr_set_data must come first");
    if (n < 0 || (n > 0 && !t)) return fail(GPR_E_ARG, "null tuning table");
    if (schedule != GPR_SCHED_AUTO && schedule != GPR_SCHED_LANE && schedule != GPR_SCHED_WAVE)
        return fail(GPR_E_ARG, "bad schedule");
    for (int i = 0; i < n; i++)
        if (!valid_tune(t[i])) return fail(GPR_E_ARG, "tuning entry " + std::to_string(i) + " out of range");
    HIP_TRY(hipSetDevice(c->cfg.device));
    for (int i = 0; i < n; i++)
        c->ct[class_index(t[i].n_class)][t[i].hi] = ClassTune{t[i].kc, t[i].waves_per_wg, t[i].wg_per_cu, t[i].ms};
    c->tuned_sched = schedule == GPR_SCHED_AUTO ? -1 : schedule;
    build_tiles(c);
    c->gexec_stale = true;
    return upload_tiles(c);
}

extern "C" int gpr_get_tuning(const gpr_ctx *c, gpr_tune *out, int max_out, int *nout, int *schedule) {
    if (!c || !nout) return fail(GPR_E_ARG, "null argument");
    int k = 0;
    for (const KernelBucket &b : c->buckets) {
        if (out && k < max_out) {
            const ClassTune &t = c->ct[b.cls][b.hi];
            out[k] = gpr_tune{kClassN[b.cls], b.hi ? 1 : 0, t.kc, t.wpg, t.wgpc, t.ms};
        }
        k++;
    }
    *nout = k;
    if (schedule) *schedule = c->tuned_sched > 0 ? c->tuned_sched : GPR_SCHED_AUTO;
    return GPR_OK;
}

// Time every candidate shape of each size class present on the context's own chains:
// each candidate runs `steps` MH steps of that class from the same saved state, which
// is restored at the end (chains, counters and clock are as before the call).
extern "C" int gpr_autotune(gpr_ctx *c, int steps, gpr_tune *out, int max_out, int *nout) {
    if (!c || !c->have_state) return fail(GPR_E_STATE, "no chain state: call gpr_init_chains or gpr_write_state");
    if (steps < 1) return fail(GPR_E_ARG, "steps must be > 0");
    if (c->cfg.rng == GPR_RNG_LCG) return fail(GPR_E_ARG, "LCG chains run on the lane schedule only: nothing to tune");
    HIP_TRY(hipSetDevice(c->cfg.device));
    const size_t N = (size_t)c->N, R = (size_t)c->rows;
    uint8_t *s_ops = nullptr;
    float *s_cs = nullptr;
    int32_t *s_len = nullptr;
    double *s_pr = nullptr, *s_ll = nullptr;
    unsigned long long *s_cnt = nullptr;
    auto release = [&]() {
        void *ps[] = {s_ops, s_cs, s_len, s_pr, s_ll, s_cnt};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    };
#define TUNE_TRY(expr)                                                                         \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) { release(); return fail(GPR_E_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); } \
    } while (0)
    TUNE_TRY(hipMalloc((void **)&s_ops, R * N));
    TUNE_TRY(hipMalloc((void **)&s_cs, sizeof(float) * R * N));
    TUNE_TRY(hipMalloc((void **)&s_len, sizeof(int32_t) * N));
    TUNE_TRY(hipMalloc((void **)&s_pr, sizeof(double) * N));
    TUNE_TRY(hipMalloc((void **)&s_ll, sizeof(double) * N));
    TUNE_TRY(hipMalloc((void **)&s_cnt, sizeof(unsigned long long) * GPR_NCNT));
    auto copy_all = [&](bool save) -> hipError_t {
        hipError_t e = hipSuccess;
        auto cp = [&](void *dst, void *src, size_t bytes) {
            if (e == hipSuccess) e = hipMemcpyAsync(save ? dst : src, save ? src : dst, bytes, hipMemcpyDeviceToDevice, c->stream);
        };
        cp(s_ops, c->d_ops, R * N);
        cp(s_cs, c->d_cs, sizeof(float) * R * N);
        cp(s_len, c->d_len, sizeof(int32_t) * N);
        cp(s_pr, c->d_prior, sizeof(double) * N);
        cp(s_ll, c->d_ll, sizeof(double) * N);
        cp(s_cnt, c->d_counters, sizeof(unsigned long long) * GPR_NCNT);
        return e;
    };
    TUNE_TRY(copy_all(true));
    const GprMHArgs a = mh_args(c, steps, 0);
    const bool mala = c->cfg.proposal == GPR_PROP_MALA && c->cfg.semantics == GPR_SEM_CONFIG;
    double wave_total = 0.0;
    // the classes present, with their current shapes
    std::vector<std::pair<int, int>> classes;
    for (const KernelBucket &b : c->buckets) classes.push_back({b.cls, b.hi ? 1 : 0});
    for (auto [k, hi] : classes) {
        ClassTune best;
        best.ms = 1e30;
        for (int kc : {1, 2, 4})
            for (int wpg : {1, 2, 4})
                for (int wpc : {32, 16, 8}) {
                    if (c->cfg.semantics == GPR_SEM_REFERENCE && hi) continue;
                    if (mala && kc != 1) continue;      // MALA evaluates one point per lane
                    const ClassTune cand{kc, wpg, wpc / wpg, 0.0};
                    c->ct[k][hi] = cand;
                    build_tiles(c);
                    int bi = -1;
                    for (size_t i = 0; i < c->buckets.size(); i++)
                        if (c->buckets[i].cls == k && c->buckets[i].hi == (hi != 0)) bi = (int)i;
                    if (upload_tiles(c)) { release(); return GPR_E_HIP; }
                    TUNE_TRY(copy_all(false));
                    TUNE_TRY(gpr_launch_fill_i32(c->d_tile_counter, 6, 0, c->stream));
                    TUNE_TRY(hipEventRecord(c->ev0, c->stream));
                    TUNE_TRY(launch_bucket(c, a, bi, c->stream));
                    TUNE_TRY(hipEventRecord(c->ev1, c->stream));
                    TUNE_TRY(hipEventSynchronize(c->ev1));
                    float ms = 0.0f;
                    TUNE_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
                    // A later candidate (more points per lane, larger workgroups, fewer
                    // workgroups per CU) must be 1 % faster to replace the best: timings
                    // of a short run spread by about that much.  On a tie the class keeps
                    // more workgroups, its larger share of the CUs when the size classes
                    // of a block run concurrently (each is tuned alone).
                    if (ms < best.ms * 0.99) { best = cand; best.ms = ms; }
                }
        c->ct[k][hi] = best;
        wave_total += best.ms;
    }
    build_tiles(c);
    if (upload_tiles(c)) { release(); return GPR_E_HIP; }
    // schedule for GPR_SCHED_AUTO: the lane kernel over every chain against the tuned wave launches
    c->tuned_sched = GPR_SCHED_WAVE;
    if (c->cfg.schedule == GPR_SCHED_AUTO && !mala) {
        TUNE_TRY(copy_all(false));
        TUNE_TRY(hipEventRecord(c->ev0, c->stream));
        if (c->cfg.semantics == GPR_SEM_REFERENCE) {
            TUNE_TRY(gpr_launch_mh_ref_lane(a, c->stream));
        } else {
            int capmax = 1;
            for (const GprDataset &g : c->hds) capmax = std::max(capmax, g.cap);
            TUNE_TRY(gpr_launch_mh_config_lane(a, capmax, c->stream));
        }
        TUNE_TRY(hipEventRecord(c->ev1, c->stream));
        TUNE_TRY(hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        TUNE_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        if (ms < wave_total) c->tuned_sched = GPR_SCHED_LANE;
    }
    TUNE_TRY(copy_all(false));
    TUNE_TRY(hipStreamSynchronize(c->stream));
#undef TUNE_TRY
    release();
    c->gexec_stale = true;
    int k = 0;
    return gpr_get_tuning(c, out, max_out, nout ? nout : &k, nullptr);
}

extern "C" int gpr_get_stats(const gpr_ctx *c, gpr_stats *o) {
    if (!c || !o) return fail(GPR_E_ARG, "null argument");
    memset(o, 0, sizeof(*o));
    if (!c->have_data) return GPR_OK;
    unsigned long long h[GPR_NCNT];
    HIP_TRY(hipSetDevice(c->cfg.device));
    HIP_TRY(hipMemcpyAsync(h, c->d_counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    int32_t mc = -1;
    double mp = 0.0;
    HIP_TRY(hipMemcpyAsync(&mc, c->d_map_chain, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&mp, c->d_map_post, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    o->struct_proposed = (long long)h[GPR_CNT_SPROP];
    o->struct_accepted = (long long)h[GPR_CNT_SACC];
    o->const_proposed = (long long)h[GPR_CNT_CPROP];
    o->const_accepted = (long long)h[GPR_CNT_CACC];
    o->node_evals = (long long)h[GPR_CNT_EVALS];
    o->steps = (long long)h[GPR_CNT_STEPS];
    o->alg_flops = (long long)h[GPR_CNT_FLOPS];
    o->last_block_ms = c->last_ms;
    o->schedule_used = c->sched_used;
    o->map_chain = mc;
    o->map_posterior = mp;
    return GPR_OK;
}

extern "C" int gpr_reset_stats(gpr_ctx *c) {
    if (!c || !c->have_data) return GPR_OK;
    HIP_TRY(hipSetDevice(c->cfg.device));
    HIP_TRY(hipMemsetAsync(c->d_counters, 0, sizeof(unsigned long long) * GPR_NCNT, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GPR_OK;
}

extern "C" int gpr_sync(gpr_ctx *c) {
    if (!c) return fail(GPR_E_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GPR_OK;
}

extern "C" void *gpr_stream(gpr_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int gpr_eval_loglik(gpr_ctx *c, int dataset, const uint8_t *d_ops, const float *d_cs, const int32_t *d_len,
                               int nchains, int cap, int schedule, double *d_ll) {
    if (!c || !c->have_data) return fail(GPR_E_STATE, "gpr_set_data must come first");
    if (dataset < 0 || dataset >= c->nd) return fail(GPR_E_ARG, "dataset index out of range");
    if (!d_ops || !d_cs || !d_len || !d_ll || nchains < 1) return fail(GPR_E_ARG, "null or empty tape arrays");
    if (cap < 1 || cap > GPR_MAX_CAP) return fail(GPR_E_ARG, "cap must be in [1, 128]");
    HIP_TRY(hipSetDevice(c->cfg.device));
    GprEvalArgs a{};
    a.ops = d_ops; a.cs = d_cs; a.len = d_len; a.nchains = nchains; a.cap = cap;
    a.dx = c->d_x; a.dy = c->d_y; a.dw = c->d_w;
    a.d = c->hds[dataset];
    a.ll = d_ll;
    a.lnsym = log((double)c->cfg.nsym);
    if (schedule == GPR_SCHED_LANE) HIP_TRY(gpr_launch_eval_lane(a, c->stream));
    else HIP_TRY(gpr_launch_eval_wave(a, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GPR_OK;
}

extern "C" int gpr_grad_logpost(gpr_ctx *c, int dataset, const uint8_t *d_ops, const float *d_cs,
                                const int32_t *d_len, int nchains, int cap, float *d_grad, double *d_lp) {
    if (!c || !c->have_data) return fail(GPR_E_STATE, "gpr_set_data must come first");
    if (dataset < 0 || dataset >= c->nd) return fail(GPR_E_ARG, "dataset index out of range");
    if (!d_ops || !d_cs || !d_len || !d_grad || !d_lp || nchains < 1) return fail(GPR_E_ARG, "null or empty arrays");
    if (cap < 1 || cap > GPR_MAX_CAP) return fail(GPR_E_ARG, "cap must be in [1, 128]");
    HIP_TRY(hipSetDevice(c->cfg.device));
    GprEvalArgs a{};
    a.ops = d_ops; a.cs = d_cs; a.len = d_len; a.nchains = nchains; a.cap = cap;
    a.dx = c->d_x; a.dy = c->d_y; a.dw = c->d_w;
    a.d = c->hds[dataset];
    a.ll = d_lp;
    a.grad = d_grad;
    a.lnsym = log((double)c->cfg.nsym);
    HIP_TRY(gpr_launch_grad_wave(a, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GPR_OK;
}

// ---------------------------------------------------------------- no-allocation evaluation
// gpr_eval_loglik_noalloc_async keeps the dataset header and per-point weights, which
// gpr_set_data allocates and fills for gpr_eval_loglik, in this static device array:
// [0, 256) the GprDataset header, then GPR_NOALLOC_MAX_POINTS floats of weights.  Each
// device holds its own copy at its own address; host code uses only the address
// hipGetSymbolAddress returns for the current device, never the array's name.
#define GPR_NOALLOC_HDR_BYTES 256
static __device__ __align__(256) unsigned char scratch[GPR_NOALLOC_HDR_BYTES + sizeof(float) * GPR_NOALLOC_MAX_POINTS];
#define GPR_NOALLOC_MAX_DEVICES 64
static void *g_scratch_ptr[GPR_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// scratch's device address on `device` (current), looked up once per device
static int scratch_ptr(int device, unsigned char **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GPR_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = (unsigned char *)g_scratch_ptr[device];
    return GPR_OK;
}

extern "C" int gpr_eval_loglik_noalloc_async(int device, const float *d_x, const float *d_y, const float *d_sd, int n,
                                             const uint8_t *d_ops, const float *d_cs, const int32_t *d_len,
                                             int nchains, int cap, int schedule, double *d_ll, void *stream) {
    if (!d_x || !d_y || !d_sd || !d_ops || !d_cs || !d_len || !d_ll) return fail(GPR_E_ARG, "null device array");
    if (n < 1 || n > GPR_NOALLOC_MAX_POINTS)
        return fail(GPR_E_ARG, "n = " + std::to_string(n) + " points: the no-allocation entry takes 1 to " +
                                   std::to_string(GPR_NOALLOC_MAX_POINTS) + " (its static scratch size)");
    if (nchains < 1) return fail(GPR_E_ARG, "nchains must be >= 1");
    if (cap < 1 || cap > GPR_MAX_CAP) return fail(GPR_E_ARG, "cap must be in [1, 128]");
    if (schedule != GPR_SCHED_AUTO && schedule != GPR_SCHED_LANE && schedule != GPR_SCHED_WAVE)
        return fail(GPR_E_ARG, "bad schedule");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev || device >= GPR_NOALLOC_MAX_DEVICES)
        return fail(GPR_E_ARG, "invalid GPU device " + std::to_string(device));
    HIP_TRY(hipSetDevice(device));
    unsigned char *sp = nullptr;
    if (int r = scratch_ptr(device, &sp)) return r;
    GprDataset *hdr = (GprDataset *)sp;
    float *w = (float *)(sp + GPR_NOALLOC_HDR_BYTES);
    const hipStream_t s = (hipStream_t)stream;
    const int kc = kc_for(n);
    HIP_TRY(gpr_launch_dataset_prep(d_sd, n, nchains, cap, kc, hdr, w, s));
    GprEvalArgs a{};
    a.ops = d_ops; a.cs = d_cs; a.len = d_len; a.nchains = nchains; a.cap = cap;
    a.dx = d_x; a.dy = d_y; a.dw = w;
    a.d.n = n;
    a.d.kc = kc;
    a.d.cap = cap;
    a.d.nchains = nchains;
    a.dd = hdr;
    a.ll = d_ll;
    if (schedule == GPR_SCHED_LANE) HIP_TRY(gpr_launch_eval_lane(a, s));
    else HIP_TRY(gpr_launch_eval_wave(a, s));
    return GPR_OK;
}

// heap -> postfix of the reachable tree (host), main.cu:371-393 semantics
// (log 2 and the physical constants become constant tokens of their value, halving a
// c* token with c = 0.5)
static const int8_t kRefToTape[GPR_REF_NCODES] = {
    GPR_OP_ONE, GPR_OP_CONST, GPR_OP_INV, -1, -1, GPR_OP_ADD, GPR_OP_SUB, GPR_OP_RSUB, GPR_OP_CADD, GPR_OP_MUL,
    GPR_OP_CMUL, GPR_OP_DIV, GPR_OP_RDIV, GPR_OP_LOG, GPR_OP_EXP, GPR_OP_POW, GPR_OP_CPOW, GPR_OP_RPOW, GPR_OP_CRPOW,
    GPR_OP_SQRT, GPR_OP_SIN, GPR_OP_ASIN, GPR_OP_COS, GPR_OP_ACOS, GPR_OP_TAN, GPR_OP_ATAN, GPR_OP_ABS,
    GPR_OP_CONST, GPR_OP_CMUL, GPR_OP_CONST, GPR_OP_CONST, GPR_OP_CONST, GPR_OP_CONST, GPR_OP_CONST, GPR_OP_CONST,
    GPR_OP_CONST, GPR_OP_CONST, GPR_OP_CONST, GPR_OP_CONST, GPR_OP_CONST};

static int decode_heap(const uint8_t *hop, const float *hc, size_t stride, int slot, int pos, uint8_t *ops, float *cs) {
    if (slot > GPR_REF_SLOTS) { ops[pos] = GPR_OP_X; cs[pos] = 0.0f; return pos + 1; }
    const int o = hop[(size_t)(slot - 1) * stride];
    if (o >= GPR_REF_NCODES) return -1;
    if (o == GPR_REF_A) return decode_heap(hop, hc, stride, 2 * slot, pos, ops, cs);
    if (o == GPR_REF_B) return decode_heap(hop, hc, stride, 2 * slot + 1, pos, ops, cs);
    const int t = kRefToTape[o];
    const int ar = gpr_arity(t);
    if (ar >= 1) pos = decode_heap(hop, hc, stride, 2 * slot, pos, ops, cs);
    if (pos < 0) return pos;
    if (ar == 2) pos = decode_heap(hop, hc, stride, 2 * slot + 1, pos, ops, cs);
    if (pos < 0) return pos;
    ops[pos] = (uint8_t)t;
    cs[pos] = gpr_has_const(t) ? hc[(size_t)(slot - 1) * stride] : 0.0f;
    if (gpr_ref_is_physical(o)) cs[pos] = gpr_ref_physical_value(o);
    if (o == GPR_REF_HALF) cs[pos] = 0.5f;
    return pos + 1;
}

extern "C" int gpr_ref_decode(const uint8_t *hop, const float *hc, size_t stride, uint8_t *ops, float *cs) {
    if (!hop || !hc || !ops || !cs || stride < 1) return fail(GPR_E_ARG, "null argument");
    const int n = decode_heap(hop, hc, stride, 1, 0, ops, cs);
    if (n < 0) return fail(GPR_E_ARG, "bad heap op code");
    return n;
}

extern "C" int gpr_ref_op_set(int set, uint8_t *codes, int max_codes, const char **program_code) {
    if (set < 0 || set >= GPR_NUM_OPSETS) return fail(GPR_E_ARG, "unknown op set");
    const int n = gpr_opset_nops(set);
    if (codes) {
        if (max_codes < n) return fail(GPR_E_ARG, "codes holds fewer than NOPS entries");
        for (int i = 0; i < n; i++) codes[i] = (uint8_t)gpr_opset_code(set, i);
    }
    static const char *const kCode[GPR_NUM_OPSETS] = {GPR_REF_PROGRAM_CODE, "1CIab+-_#*@/|", "1CAB+-_#*@^",
                                                      "1CAB+-_#@", "B2epcghm:o.~Kab+-_*"};
    if (program_code) *program_code = kCode[set];
    return n;
}

// ---------------------------------------------------------------- test hooks
extern "C" int gpr_test_ref_decode(gpr_ctx *c, const uint8_t *d_hop, const float *d_hc, int nheaps, uint8_t *d_ops,
                                   float *d_cs, int32_t *d_meta) {
    if (!c || !d_hop || !d_hc || !d_ops || !d_cs || !d_meta || nheaps < 1) return fail(GPR_E_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(c->cfg.device));
    HIP_TRY(gpr_launch_test_ref_decode(d_hop, d_hc, nheaps, d_ops, d_cs, d_meta, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GPR_OK;
}

extern "C" int gpr_test_draws(gpr_ctx *c, const uint32_t *d_words, int n, uint64_t seed, float *d_out) {
    if (!c || !d_words || !d_out || n < 1) return fail(GPR_E_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(c->cfg.device));
    HIP_TRY(gpr_launch_test_draws(d_words, n, seed, d_out, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GPR_OK;
}

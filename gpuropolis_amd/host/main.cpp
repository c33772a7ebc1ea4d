// gpuropolis: the command-line run loop (main.cu:1010-1257) over the C ABI.
//
// Options keep the reference's names, defaults and meaning (main.cu:95-156); the
// options after --all are additions.  The run loop keeps the reference's order
// (main.cu:1156-1239): srand(seed); per block o: one MH block seeded by rand(), rows
// printed when o > burn && o % thin == 0, then resampling.  State stays on the device:
// it is copied to the host only for printing.
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include <chrono>
#include <string>
#include <vector>

#include "gpuropolis.h"
#include "io.hpp"

using namespace gpr_host;

namespace {

struct Options {
    // main.cu:96-109
    float sdscale = 1.0f;
    int N = 1024;
    int steps = 1000;
    int outer = 100;
    int thin = 1;
    int seed = -1;
    int burn = 0;
    int quiet = 0;
    int gpu = 0;
    int first_half = 0;
    int even_half = 0;
    int show_constants = 0;
    std::string in = "data.txt";
    std::string out = "out/";
    // additions
    int semantics = GPR_SEM_REFERENCE;
    int cap = 64;
    int dmax = GPR_MAX_DEPTH;
    int nsym = 10;
    int schedule = GPR_SCHED_AUTO;
    int proposal = GPR_PROP_RW;
    float sigma = -1.0f;
    float eps = -1.0f;
    int struct_every = 5;
    int fixed_structure = 0;
    int rng = GPR_RNG_PHILOX;
    int ops = GPR_OPS_DEFAULT;
    int resample = -1;
    double resample_nats = -1.0;
    float resample_quantile = -1.0f;
    int log = 1;
    std::string state_in, state_out;
    long tops = 0;
};

enum {
    OPT_SEMANTICS = 1000, OPT_CAP, OPT_DMAX, OPT_NSYM, OPT_SCHEDULE, OPT_PROPOSAL, OPT_SIGMA, OPT_EPS,
    OPT_STRUCT_EVERY, OPT_FIXED, OPT_RNG, OPT_RESAMPLE, OPT_NATS, OPT_QUANTILE, OPT_NOLOG, OPT_STATE_IN,
    OPT_STATE_OUT, OPT_OPS, OPT_TOPS, OPT_HELP
};

const struct option kLongOptions[] = {
    // main.cu:111-129
    {"in", required_argument, nullptr, 'd'},
    {"steps", required_argument, nullptr, 'i'},
    {"N", required_argument, nullptr, 'N'},
    {"out", required_argument, nullptr, 'O'},
    {"outer", required_argument, nullptr, 'o'},
    {"thin", required_argument, nullptr, 't'},
    {"seed", required_argument, nullptr, 's'},
    {"burn", required_argument, nullptr, 'b'},
    {"gpu", required_argument, nullptr, 'g'},
    {"sdscale", required_argument, nullptr, 'r'},
    {"first-half", no_argument, nullptr, 'f'},
    {"even-half", no_argument, nullptr, 'e'},
    {"show-constants", no_argument, nullptr, 'C'},
    {"quiet", no_argument, nullptr, 'q'},
    {"all", no_argument, nullptr, '_'},
    // additions
    {"semantics", required_argument, nullptr, OPT_SEMANTICS},
    {"cap", required_argument, nullptr, OPT_CAP},
    {"dmax", required_argument, nullptr, OPT_DMAX},
    {"nsym", required_argument, nullptr, OPT_NSYM},
    {"schedule", required_argument, nullptr, OPT_SCHEDULE},
    {"proposal", required_argument, nullptr, OPT_PROPOSAL},
    {"sigma", required_argument, nullptr, OPT_SIGMA},
    {"eps", required_argument, nullptr, OPT_EPS},
    {"struct-every", required_argument, nullptr, OPT_STRUCT_EVERY},
    {"fixed-structure", no_argument, nullptr, OPT_FIXED},
    {"rng", required_argument, nullptr, OPT_RNG},
    {"ops", required_argument, nullptr, OPT_OPS},
    {"resample", required_argument, nullptr, OPT_RESAMPLE},
    {"resample-nats", required_argument, nullptr, OPT_NATS},
    {"resample-quantile", required_argument, nullptr, OPT_QUANTILE},
    {"no-log", no_argument, nullptr, OPT_NOLOG},
    {"state-in", required_argument, nullptr, OPT_STATE_IN},
    {"state-out", required_argument, nullptr, OPT_STATE_OUT},
    {"tops", required_argument, nullptr, OPT_TOPS},
    {"help", no_argument, nullptr, OPT_HELP},
    {nullptr, 0, nullptr, 0}};

const char *kUsage =
    "usage: gpuropolis [--in data.txt] [--out out/] [--N 1024] [--steps 1000] [--outer 100]\n"
    "                  [--thin 1] [--burn 0] [--seed -1] [--gpu 0] [--sdscale 1]\n"
    "                  [--first-half] [--even-half] [--show-constants] [--quiet] [--all]\n"
    "  additions:      [--semantics reference|config] [--cap 64] [--dmax 16] [--nsym 10]\n"
    "                  [--schedule auto|lane|wave] [--proposal rw|mala] [--sigma S] [--eps E]\n"
    "                  [--struct-every 5] [--fixed-structure] [--rng philox|lcg]\n"
    "                  [--ops default|simplified|polynomial|linear|physical]\n"
    "                  [--resample none|map|quantile] [--resample-nats 100] [--resample-quantile 0.5]\n"
    "                  [--no-log] [--state-in FILE] [--state-out FILE] [--tops K]\n";

bool pick(const char *arg, std::initializer_list<std::pair<const char *, int>> names, int &out) {
    for (auto &p : names)
        if (!strcmp(arg, p.first)) {
            out = p.second;
            return true;
        }
    return false;
}

// process_arguments (main.cu:131-156): 0 on success, 1 on an unknown option or value
int parse(int argc, char **argv, Options &o) {
    int idx = 0, opt;
    optind = 1;
    while ((opt = getopt_long(argc, argv, "", kLongOptions, &idx)) != -1) {
        switch (opt) {
        case 'd': o.in = optarg; break;
        case 'i': o.steps = atoi(optarg); break;
        case 'N': o.N = atoi(optarg); break;
        case 'o': o.outer = atoi(optarg); break;
        case 'O': o.out = optarg; break;
        case 't': o.thin = atoi(optarg); break;
        case 'b': o.burn = atoi(optarg); break;
        case 'g': o.gpu = atoi(optarg); break;
        case 's': o.seed = atoi(optarg); break;
        case 'r': o.sdscale = (float)atof(optarg); break;
        case 'f': o.first_half = 1; break;
        case 'e': o.even_half = 1; break;
        case 'C': o.show_constants = 1; break;
        case 'q': o.quiet = 1; break;
        case '_': break;
        case OPT_SEMANTICS:
            if (!pick(optarg, {{"reference", GPR_SEM_REFERENCE}, {"config", GPR_SEM_CONFIG}}, o.semantics)) return 1;
            break;
        case OPT_CAP: o.cap = atoi(optarg); break;
        case OPT_DMAX: o.dmax = atoi(optarg); break;
        case OPT_NSYM: o.nsym = atoi(optarg); break;
        case OPT_SCHEDULE:
            if (!pick(optarg, {{"auto", GPR_SCHED_AUTO}, {"lane", GPR_SCHED_LANE}, {"wave", GPR_SCHED_WAVE}},
                      o.schedule))
                return 1;
            break;
        case OPT_PROPOSAL:
            if (!pick(optarg, {{"rw", GPR_PROP_RW}, {"mala", GPR_PROP_MALA}}, o.proposal)) return 1;
            break;
        case OPT_SIGMA: o.sigma = (float)atof(optarg); break;
        case OPT_EPS: o.eps = (float)atof(optarg); break;
        case OPT_STRUCT_EVERY: o.struct_every = atoi(optarg); break;
        case OPT_FIXED: o.fixed_structure = 1; break;
        case OPT_RNG:
            if (!pick(optarg, {{"philox", GPR_RNG_PHILOX}, {"lcg", GPR_RNG_LCG}}, o.rng)) return 1;
            break;
        case OPT_OPS:
            if (!pick(optarg, {{"default", GPR_OPS_DEFAULT}, {"simplified", GPR_OPS_SIMPLIFIED},
                               {"polynomial", GPR_OPS_POLYNOMIAL}, {"linear", GPR_OPS_LINEAR},
                               {"physical", GPR_OPS_PHYSICAL}}, o.ops))
                return 1;
            break;
        case OPT_RESAMPLE:
            if (!pick(optarg, {{"none", GPR_RS_NONE}, {"map", GPR_RS_MAP}, {"quantile", GPR_RS_QUANTILE}}, o.resample))
                return 1;
            break;
        case OPT_NATS: o.resample_nats = atof(optarg); break;
        case OPT_QUANTILE:
            o.resample_quantile = (float)atof(optarg);
            if (!(o.resample_quantile >= 0.0f && o.resample_quantile < 1.0f)) {
                fprintf(stderr, "--resample-quantile must be in [0, 1)\n");
                return 1;
            }
            break;
        case OPT_NOLOG: o.log = 0; break;
        case OPT_STATE_IN: o.state_in = optarg; break;
        case OPT_STATE_OUT: o.state_out = optarg; break;
        case OPT_TOPS: o.tops = atol(optarg); break;
        case OPT_HELP: fputs(kUsage, stdout); exit(0);
        default: return 1;
        }
    }
    return optind < argc ? 1 : 0;
}

// progress_bar (main.cu:997-1007)
void progress_bar(float pct) {
    const int width = 70;
    fprintf(stderr, "\r[");
    for (int p = 0; p < width; p++) fputc(p <= pct * width ? '=' : ' ', stderr);
    fprintf(stderr, "]");
}

const char *sem_name(int s) { return s == GPR_SEM_REFERENCE ? "reference" : "config"; }
const char *sched_name(int s) { return s == GPR_SCHED_LANE ? "lane" : s == GPR_SCHED_WAVE ? "wave" : "auto"; }
const char *rs_name(int s) { return s == GPR_RS_MAP ? "map" : s == GPR_RS_QUANTILE ? "quantile" : "none"; }

// The parameter and data header of samples.txt (main.cu:1087-1117).
void write_header(FILE *fp, const Options &o, int seed, int plen, int ncons, int gpu, const std::vector<Datum> &D) {
    fprintf(fp, "# -----------------------------------------------------------------\n");
    fprintf(fp, "# -- Parameters:\n");
    fprintf(fp, "# -----------------------------------------------------------------\n");
    fprintf(fp, "# \tInput data path: %s\n", o.in.c_str());
    fprintf(fp, "# \tOutput path: %s\n", o.out.c_str());
    fprintf(fp, "# \tMCMC Iterations (per block): %i\n", o.steps);
    fprintf(fp, "# \tBlocks: %i\n", o.outer);
    fprintf(fp, "# \tN chains: %i\n", o.N);
    fprintf(fp, "# \tSeed: %i\n", seed);
    fprintf(fp, "# \tMax program length: %i\n", plen);
    fprintf(fp, "# \tN Constants: %i\n", ncons);
    fprintf(fp, "# \tGPU: %i\n", gpu);
    fprintf(fp, "# \tNumber of data points: %i\n", (int)D.size());
    fprintf(fp, "#\n#\n");
    fprintf(fp, "# -----------------------------------------------------------------\n");
    fprintf(fp, "# -- Data:\n");
    fprintf(fp, "# -----------------------------------------------------------------\n");
    for (const Datum &d : D) fprintf(fp, "# \t%f\t%f\t%f\n", d.x, d.y, d.sd);
    fprintf(fp, "#\n#\n");
}

// One row per chain (main.cu:1173-1211).
void write_rows(FILE *fp, const Options &o, int block, int rows, const std::vector<uint8_t> &ops,
                const std::vector<float> &cs, const std::vector<int32_t> &len, const std::vector<double> &prior,
                const std::vector<double> &ll) {
    const int N = o.N;
    std::vector<uint8_t> P(rows);
    std::vector<float> C(rows);
    for (int h = 0; h < N; h++) {
        for (int i = 0; i < rows; i++) {
            P[i] = ops[(size_t)i * N + h];
            C[i] = cs[(size_t)i * N + h];
        }
        fprintf(fp, "%d\t%d\t", h, block);
        fprintf(fp, "%.3f\t%.3f\t%.3f\t", prior[h] + ll[h], prior[h], ll[h]);
        const bool ref = o.semantics == GPR_SEM_REFERENCE;
        const int L = ref ? rows : len[h];
        fprintf(fp, "%.1f\t", ref ? heap_length(P.data()) : tape_length(P.data(), L));
        const float deg = ref ? heap_degree(P.data(), C.data()) : tape_degree(P.data(), C.data(), L);
        if (deg == INFINITY) fprintf(fp, "NA\t");
        else fprintf(fp, "%.2f\t", deg);
        if (ref) {
            fprintf(fp, "\"%s\"\t\"%s\"\t\"%s\"\t\"%s\"\t", heap_string(P.data(), C.data(), "C").c_str(),
                    heap_string(P.data(), C.data(), "%.10f").c_str(), heap_latex(P.data(), C.data(), "%.2f").c_str(),
                    heap_code(P.data(), o.ops).c_str());
        } else {
            fprintf(fp, "\"%s\"\t\"%s\"\t\"%s\"\t\"%s\"\t", tape_string(P.data(), C.data(), L, "C").c_str(),
                    tape_string(P.data(), C.data(), L, "%.10f").c_str(),
                    tape_latex(P.data(), C.data(), L, "%.2f").c_str(), tape_code(P.data(), L).c_str());
        }
        if (o.show_constants)
            for (int c = 0; c < rows; c++) fprintf(fp, "%.10f\t", C[c]);
        fprintf(fp, "\n");
    }
}

// Binary chain state.  Header: magic, semantics, rows, N, then (current format) the
// seed, the Philox clock (next step, next resample block) and the number of blocks run;
// then ops [rows][N] u8, cs [rows][N] f32, len [N] i32.  Files of the older format (no
// seed, clock or block count) still load: they resume with a fresh clock.
constexpr int32_t kStateMagicV1 = 0x47505253, kStateMagic = 0x32525047;

struct StateClock {
    int32_t seed = 0;
    uint32_t step = 0, block = 0;
    int32_t blocks_done = 0;
    bool has_clock = false;
};

bool save_state(const std::string &path, int sem, int rows, int N, const StateClock &clk,
                const std::vector<uint8_t> &ops, const std::vector<float> &cs, const std::vector<int32_t> &len) {
    FILE *fp = fopen(path.c_str(), "wb");
    if (!fp) return false;
    const int32_t hdr[8] = {kStateMagic, sem, rows, N, clk.seed, (int32_t)clk.step, (int32_t)clk.block,
                            clk.blocks_done};
    bool ok = fwrite(hdr, sizeof hdr, 1, fp) == 1 && fwrite(ops.data(), 1, ops.size(), fp) == ops.size() &&
              fwrite(cs.data(), sizeof(float), cs.size(), fp) == cs.size() &&
              fwrite(len.data(), sizeof(int32_t), len.size(), fp) == len.size();
    return fclose(fp) == 0 && ok;
}

bool load_state(const std::string &path, int sem, int rows, int N, StateClock &clk, std::vector<uint8_t> &ops,
                std::vector<float> &cs, std::vector<int32_t> &len, std::string &err) {
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) {
        err = "cannot open state file " + path;
        return false;
    }
    int32_t hdr[8];
    bool ok = fread(hdr, sizeof(int32_t), 4, fp) == 4;
    if (ok && hdr[0] != kStateMagic && hdr[0] != kStateMagicV1) {
        err = "state file " + path + " is not a gpuropolis state file";
        fclose(fp);
        return false;
    }
    if (ok && (hdr[1] != sem || hdr[2] != rows || hdr[3] != N)) {
        err = "state file " + path + " does not match --semantics/--cap/--N";
        fclose(fp);
        return false;
    }
    clk = StateClock{};
    if (ok && hdr[0] == kStateMagic) {
        ok = fread(hdr + 4, sizeof(int32_t), 4, fp) == 4;
        clk.seed = hdr[4];
        clk.step = (uint32_t)hdr[5];
        clk.block = (uint32_t)hdr[6];
        clk.blocks_done = hdr[7];
        clk.has_clock = true;
        if (ok && clk.blocks_done < 0) {
            err = "state file " + path + " has a negative block count";
            fclose(fp);
            return false;
        }
    }
    ok = ok && fread(ops.data(), 1, ops.size(), fp) == ops.size() &&
         fread(cs.data(), sizeof(float), cs.size(), fp) == cs.size() &&
         fread(len.data(), sizeof(int32_t), len.size(), fp) == len.size();
    fclose(fp);
    if (!ok) err = "state file " + path + " is truncated";
    return ok;
}

int die(const char *what) {
    fprintf(stderr, "*** ERROR: %s: %s\n", what, gpr_last_error());
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    Options o;
    if (parse(argc, argv, o) == 1) {
        fprintf(stderr, "Invalid command line arguments.\n%s", kUsage);
        return 1;
    }
    // main.cu:1024-1031
    if (o.steps <= 0 || o.N <= 0 || o.burn < 0 || o.thin < 1 || o.outer < 0 || o.tops < 0) {
        fprintf(stderr, "*** ERROR: need --steps > 0, --N > 0, --burn >= 0, --thin >= 1, --outer >= 0, --tops >= 0\n");
        return 1;
    }
    const bool ref = o.semantics == GPR_SEM_REFERENCE;

    gpr_config cfg;
    gpr_config_default(&cfg);
    cfg.device = o.gpu;
    cfg.semantics = o.semantics;
    cfg.cap = o.cap;
    cfg.dmax = o.dmax;
    cfg.nsym = o.nsym;
    cfg.schedule = o.schedule;
    cfg.proposal = o.proposal;
    if (o.sigma > 0) cfg.sigma = o.sigma;
    if (o.eps > 0) cfg.eps = o.eps;
    cfg.struct_every = o.struct_every;
    cfg.fixed_structure = o.fixed_structure;
    cfg.rng = o.rng;
    cfg.ops = o.ops;
    cfg.resample = o.resample >= 0 ? o.resample : (ref ? GPR_RS_QUANTILE : GPR_RS_MAP);
    if (o.resample_nats >= 0) cfg.resample_nats = o.resample_nats;
    if (o.resample_quantile >= 0) cfg.resample_quantile = o.resample_quantile;

    const int rows = ref ? GPR_REF_SLOTS : o.cap;
    std::vector<uint8_t> ops((size_t)rows * o.N);
    std::vector<float> cs((size_t)rows * o.N);
    std::vector<int32_t> len(o.N);
    std::vector<double> prior(o.N), ll(o.N);
    std::string err;
    // A state file of the current format carries its run's seed and clock: the resumed
    // run continues that run's random streams and block numbering, whatever --seed says.
    StateClock clk;
    if (!o.state_in.empty() && !load_state(o.state_in, o.semantics, rows, o.N, clk, ops, cs, len, err)) {
        fprintf(stderr, "*** ERROR: %s\n", err.c_str());
        return 1;
    }

    // main.cu:1071-1073
    int seed = o.seed;
    if (seed == -1) seed = (int)time(nullptr);
    if (clk.has_clock) seed = clk.seed;
    cfg.seed = (uint64_t)(uint32_t)seed;

    std::vector<Datum> D;
    if (!load_data_file(o.in.c_str(), o.sdscale, o.first_half, o.even_half, D, err)) {
        fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    if (D.empty()) {
        fprintf(stderr, "*** ERROR: no data points in %s\n", o.in.c_str());
        return 1;
    }

    gpr_ctx *ctx = nullptr;
    if (gpr_create(&cfg, &ctx) != GPR_OK) return die("gpr_create");
    std::vector<float> x(D.size()), y(D.size()), sd(D.size());
    for (size_t i = 0; i < D.size(); i++) {
        x[i] = D[i].x;
        y[i] = D[i].y;
        sd[i] = D[i].sd;
    }
    const int32_t nd = (int32_t)D.size(), nch = o.N;
    if (gpr_set_data(ctx, 1, &nd, x.data(), y.data(), sd.data(), &nch, nullptr) != GPR_OK) return die("gpr_set_data");

    if (!o.state_in.empty()) {
        if (gpr_write_state(ctx, ops.data(), cs.data(), len.data()) != GPR_OK) return die("gpr_write_state");
        if (gpr_set_clock(ctx, clk.step, clk.block) != GPR_OK) return die("gpr_set_clock");
    } else if (gpr_init_chains(ctx) != GPR_OK) {
        return die("gpr_init_chains");
    }

    mkdir(o.out.c_str(), 0755);
    const std::string sample_path = o.out + "/samples.txt", log_path = o.out + "/log.txt";
    FILE *fp = fopen(sample_path.c_str(), "a");
    if (!fp) {
        fprintf(stderr, "*** ERROR: Cannot open file:\t%s\n", sample_path.c_str());
        return 1;
    }
    write_header(fp, o, seed, rows, rows, o.gpu, D);
    fclose(fp);

    FILE *lg = nullptr;
    if (o.log) {
        lg = fopen(log_path.c_str(), "a");
        if (!lg) {
            fprintf(stderr, "*** ERROR: Cannot open file:\t%s\n", log_path.c_str());
            return 1;
        }
        write_header(lg, o, seed, rows, rows, o.gpu, D);
        fprintf(lg, "# -- Run: %s, semantics %s, rng %s, proposal %s, resample %s, struct-every %d%s\n", gpr_version(),
                sem_name(cfg.semantics), cfg.rng == GPR_RNG_LCG ? "lcg" : "philox",
                cfg.proposal == GPR_PROP_MALA ? "mala" : "rw", rs_name(cfg.resample), cfg.struct_every,
                cfg.fixed_structure ? ", fixed structure" : "");
        fprintf(lg, "# block\tschedule\tdevice_ms\twall_ms\tsteps/s\tstruct_acc\tconst_acc\tmap_chain\tmap_posterior\n");
        fflush(lg);
    }

    srand(seed);  // main.cu:1107
    // A resumed run numbers its blocks on from the saved run's and skips the block
    // seeds that run drew, so burn-in, thinning and the LCG seeds continue as one run.
    const int b0 = clk.blocks_done;
    for (int b = 0; b < b0; b++) (void)rand();
    if (!o.quiet) progress_bar(0.0f);
    for (int b = b0; b < b0 + o.outer; b++) {
        gpr_reset_stats(ctx);
        const auto t0 = std::chrono::steady_clock::now();
        if (gpr_run_block(ctx, o.steps, rand()) != GPR_OK) return die("gpr_run_block");
        if (b > o.burn && b % o.thin == 0) {
            if (gpr_read_state(ctx, ops.data(), cs.data(), len.data(), prior.data(), ll.data()) != GPR_OK)
                return die("gpr_read_state");
            fp = fopen(sample_path.c_str(), "a");
            if (!fp) {
                fprintf(stderr, "*** ERROR: Cannot open file:\t%s\n", sample_path.c_str());
                return 1;
            }
            write_rows(fp, o, b, rows, ops, cs, len, prior, ll);
            fclose(fp);
        }
        if (cfg.resample != GPR_RS_NONE && gpr_resample(ctx) != GPR_OK) return die("gpr_resample");
        if (lg) {
            gpr_sync(ctx);
            const double wall =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            gpr_stats st;
            gpr_get_stats(ctx, &st);
            fprintf(lg, "%d\t%s\t%.3f\t%.3f\t%.4g\t%.4f\t%.4f\t%d\t%.3f\n", b, sched_name(st.schedule_used),
                    st.last_block_ms, wall, st.last_block_ms > 0 ? st.steps / (st.last_block_ms * 1e-3) : 0.0,
                    st.struct_proposed ? (double)st.struct_accepted / st.struct_proposed : 0.0,
                    st.const_proposed ? (double)st.const_accepted / st.const_proposed : 0.0, st.map_chain,
                    st.map_posterior);
            fflush(lg);
        }
        if (!o.quiet) progress_bar((float)(b - b0) / (float)o.outer);
    }
    if (!o.quiet) progress_bar(1.1f);
    if (!o.state_out.empty()) {
        if (gpr_read_state(ctx, ops.data(), cs.data(), len.data(), nullptr, nullptr) != GPR_OK)
            return die("gpr_read_state");
        StateClock out;
        out.seed = seed;
        out.blocks_done = b0 + o.outer;
        if (gpr_get_clock(ctx, &out.step, &out.block) != GPR_OK) return die("gpr_get_clock");
        if (!save_state(o.state_out, o.semantics, rows, o.N, out, ops, cs, len)) {
            fprintf(stderr, "*** ERROR: cannot write state file %s\n", o.state_out.c_str());
            return 1;
        }
    }
    if (lg) fclose(lg);
    gpr_destroy(ctx);
    // run-all.sh:45: sort -g -k3 samples.txt | tail -n K > tops.txt
    if (o.tops > 0 && !write_tops(sample_path.c_str(), (o.out + "/tops.txt").c_str(), o.tops, err)) {
        fprintf(stderr, "*** ERROR: %s\n", err.c_str());
        return 1;
    }
    fprintf(stderr, " Completed. \n");
    return 0;
}

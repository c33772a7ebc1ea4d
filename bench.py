"""Benchmark: the MH sampler on one MI355X, prints ONE JSON line.

Workload (BASELINE configuration "plain method, large"): one dataset of n=250 fp32
points, x ~ U[-5,5], y = a seeded random target expression (<= 12 nodes, constants
N(0, 2^2)) + noise of sd 1% of the y range; 131072 chains, postfix programs of at
most 64 tokens; a bench step is one MCMC block of 1000 MH steps on every chain
followed by MAP-relative resampling (chains > 100 nats below the MAP copy a good
chain), all on the device.  The config runs 20 blocks: --steps 20 by default.

value       whole-job MH steps per second (chains x steps / wall time of the K blocks)
roofline    the MH kernel: algorithmic FP32 ops per launch (n * (non-leaf tokens + 4)
            + constants, summed over every evaluated program, counted by the kernel)
            divided by the kernel's average launch time (HIP events on its stream),
            against the VALU lane-op issue peak (256 CU x 64 lanes x 2.4 GHz); the FP32
            datasheet peak (157.3 TFLOP/s, packed FMA) is reported beside it.  For the
            default configuration it also carries the HBM bytes per launch and the share
            of VALU and SALU issue slots in use, from rocprofv3 counters of the same
            launch (MEASURED_TRAFFIC, MEASURED_ISSUE) over the live launch time.
cpu_baseline the oracle's serial restatement of the same sampler (OpenMP over chains)
            on this host's cores, on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_LANE_OP_PEAK = 256 * 64 * 2.4e9      # lane-ops/s
FP32_DATASHEET_PEAK = 157.3e12            # FLOP/s, packed FMA
HBM_PEAK = 8.0e12

# HBM bytes per MH launch (FETCH_SIZE + WRITE_SIZE of rocprofv3 --pmc, scaled by the
# same counters on a plain copy of known size) for the default configuration, keyed by
# (schedule, chains, cap, MH steps per block, seed); the profiles/ summaries hold the runs.
MEASURED_TRAFFIC = {("wave", 131072, 64, 1000, 0): 267709984}
# VALU and SALU instructions per chain-step of the same launches (rocprofv3 --pmc
# SQ_INSTS_VALU, SQ_INSTS_SALU over chains x steps; the profiles/ summary holds the run).
# With the live launch time they give the share of issue slots each unit is busy: one
# wave64 instruction per SIMD every 4 cycles, 1024 SIMDs at 2.4 GHz.
MEASURED_ISSUE = {("wave", 131072, 64, 1000, 0): (297.9, 299.6)}

METRIC = "MH steps/sec + node×datapoint evals/sec (% VALU peak); ESS/sec MALA vs RW"


def cpu_baseline(ds, cap, seed, sample_chains, sample_steps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as O
    O.build()
    threads = int(os.environ.get("GPR_CPU_THREADS", min(16, os.cpu_count() or 1)))
    cfg = O.make_cfg(cap=cap, seed=seed)
    ops, cs, ln = O.cfg_init(cfg, sample_chains)
    x, y, sd = ds
    t0 = time.perf_counter()
    O.cfg_run_block(cfg, ops, cs, ln, x, y, sd, 0, sample_steps, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": sample_chains * sample_steps / dt, "unit": "MH steps/s", "cores": threads, "kind": "port",
            "sample": f"{sample_chains} chains x {sample_steps} MH steps of the same workload "
                      f"(oracle/oracle.c, OpenMP, {dt:.1f} s)"}


def step_split(args):
    """Cost per chain-step of the two step types: programs from two warm-up blocks of the
    large workload, then one block with structure moves off (constant steps only) and
    one with a structure move at every step, each on the default launch shape."""
    from gpuropolis_amd import lib as G
    from gpuropolis_amd import workloads as W
    ds, _ = W.dataset_L(args.seed)
    ctx = G.Context(cap=args.cap, seed=args.seed, schedule=G.SCHED_WAVE)
    ctx.set_data([ds], [args.chains])
    ctx.init_chains()
    for _ in range(2):
        ctx.run_block(args.mh_steps)
        ctx.resample()
    ops, cs, ln, _, _ = ctx.read_state()
    ctx.close()
    out = {"mean_tokens": float(ln.mean())}
    for name, kw in (("constant_steps", dict(fixed_structure=1)), ("structure_steps", dict(struct_every=1))):
        ctx = G.Context(cap=args.cap, seed=args.seed, schedule=G.SCHED_WAVE, **kw)
        ctx.set_data([ds], [args.chains])
        ctx.write_state(ops, cs, ln)
        ctx.run_block(args.mh_steps)
        ms = ctx.stats().last_block_ms
        out[name] = {"kernel_ms": ms, "ns_per_chain_step": ms * 1e6 / (args.chains * args.mh_steps)}
        ctx.close()
    print(json.dumps(out), flush=True)


def copy_calibration(mib):
    """5 launches of a plain copy kernel over mib MiB: the HBM counters of one launch
    should read mib MiB fetched and mib MiB written."""
    import torch
    a = torch.empty(mib << 18, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "bytes_per_launch": mib << 20}), flush=True)


def time_blocks(G, ds, args, sched, blocks, warmup, tune=False):
    ctx = G.Context(cap=args.cap, seed=args.seed, schedule=sched)
    ctx.set_data([ds], [args.chains])
    ctx.init_chains()
    for _ in range(warmup):
        ctx.run_block(args.mh_steps)
        ctx.resample()
    ctx.sync()
    table = None
    if tune:
        table, _ = ctx.autotune(args.tune_steps)
    ctx.reset_stats()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(blocks):
        ctx.run_block(args.mh_steps)          # HIP events around the MH launch on the context stream
        kernel_ms += ctx.stats().last_block_ms
        ctx.resample()
    ctx.sync()
    wall = time.perf_counter() - t0
    st = ctx.stats()
    ctx.close()
    return st, wall, kernel_ms, table


def ess_multichain(x):
    """Effective sample size of one scalar from x[chains, draws]: the chain-averaged
    autocovariance against the pooled variance (as split-R-hat ESS), truncated by Geyer's
    initial monotone sequence."""
    import numpy as np
    m, n = x.shape
    xc = x - x.mean(axis=1, keepdims=True)
    f = np.fft.rfft(xc, n=2 * n, axis=1)
    acov = np.fft.irfft(f * np.conj(f), axis=1)[:, :n].mean(axis=0) / n
    w = acov[0] * n / (n - 1)
    var_plus = (n - 1) / n * w + x.mean(axis=1).var(ddof=1)
    if not var_plus > 0:
        return float(m * n)
    rho = 1.0 - (w - acov * n / (n - 1)) / var_plus
    rho[0] = 1.0
    tau, prev = -1.0, np.inf
    for k in range(0, n - 1, 2):
        p = rho[k] + rho[k + 1]
        if p < 0:
            break
        p = min(p, prev)
        prev = p
        tau += 2 * p
    return float(m * n / max(tau, 1e-9))


def ess_rational(G, seed=0, nch=16384, thin=2, blocks=160):
    """The differentiation workload: the fixed rational function with 7 constants, n = 37.
    Warm-up adapts a proposal metric (Cholesky factor of the chains' cross-chain
    covariance, three rounds).  Then random-walk and MALA constant moves run with that
    metric, each at the step size giving it the most ESS per second among pilot values.
    ESS per second = the smallest ESS over the 7 constants (all chains pooled) per MH
    step x MH steps per second of device time, with the cost per MH step taken from a 1000-step block."""
    import numpy as np
    from gpuropolis_amd import workloads as W
    (x, y, sd), b = W.dataset_D(seed)
    rng = np.random.default_rng(seed + 100)
    O = np.zeros((32, nch), dtype=np.uint8)
    V = np.zeros((32, nch), dtype=np.float32)
    for c in range(nch):
        o, v = W.rational_tape(b * (1 + 0.002 * rng.normal(size=7)))
        O[:len(o), c] = o
        V[:len(o), c] = v
    L = np.full(nch, len(o), dtype=np.int32)
    rows = np.flatnonzero(o == G.OP["CONST"])

    def context(prop, step, cs, chol):
        kw = dict(eps=step) if prop == G.PROP_MALA else dict(sigma=step)
        ctx = G.Context(cap=32, seed=seed + 7, proposal=prop, fixed_structure=1, resample=G.RS_NONE,
                        schedule=G.SCHED_WAVE, **kw)
        ctx.set_data([(x, y, sd)], [nch])
        if chol is not None:
            ctx.set_metric(0, rows, chol)
        ctx.write_state(O, cs, L)
        return ctx

    # warm-up: identity metric, then twice the metric of the chains' spread
    cs, chol = V.copy(), None
    for step, nsteps in ((1e-3, 1000), (1.0, 500), (1.0, 500)):
        ctx = context(G.PROP_MALA, step, cs, chol)
        ctx.run_block(nsteps)
        cs = ctx.read_state()[1]
        ctx.close()
        cov = np.cov(cs[rows].astype(np.float64))
        chol = np.linalg.cholesky(cov + 1e-12 * np.eye(len(rows)) * np.trace(cov))

    def run(prop, step, nblocks, burn):
        ctx = context(prop, step, cs, chol)
        draws = []
        for i in range(nblocks):
            ctx.run_block(thin)
            if i >= burn:
                draws.append(ctx.read_state()[1][rows].astype(np.float64))
        st = ctx.stats()
        # cost per MH step from one long block (short blocks pay the per-launch start-up)
        ctx.reset_stats()
        ctx.run_block(1000)
        rate = ctx.stats().steps / (ctx.stats().last_block_ms / 1e3)
        ctx.close()
        D = np.array(draws)                                   # [draws, consts, chains]
        ess = [ess_multichain(D[:, j, :].T) for j in range(len(rows))]
        chain_steps = (nblocks - burn) * thin * nch
        return dict(step=step, ess_min=min(ess), ess_per_s=min(ess) / chain_steps * rate,
                    mh_steps_per_s=rate, ess_per_step=min(ess) / chain_steps,
                    acceptance=st.const_accepted / max(1, st.const_proposed),
                    mean=D.mean(axis=(0, 2)), sd=D.std(axis=(0, 2)), ess=np.array(ess))

    out = {}
    for name, prop, steps in (("rw", G.PROP_RW, [0.3, 0.5, 0.8, 1.2]), ("mala", G.PROP_MALA, [0.6, 0.9, 1.2, 1.6])):
        pilots = [run(prop, h, 60, 20) for h in steps]
        best = max(pilots, key=lambda r: r["ess_per_s"])
        out[name] = run(prop, best["step"], blocks, blocks // 4)
    # posterior means of the two samplers agree within 2 MC standard errors (sd / sqrt(ESS))
    se = np.sqrt(out["rw"]["sd"] ** 2 / out["rw"]["ess"] + out["mala"]["sd"] ** 2 / out["mala"]["ess"])
    z = np.abs(out["rw"]["mean"] - out["mala"]["mean"]) / se
    res = {"workload": f"differentiation: rational function, 7 constants, n=37, {nch} chains, proposal metric "
                       f"adapted in warm-up, {blocks} draws thinned by {thin} MH steps, first quarter discarded",
           "max_z_means": float(z.max())}
    for k in ("rw", "mala"):
        r = out[k]
        res[k] = {"step": r["step"], "acceptance": r["acceptance"], "ess_min": r["ess_min"],
                  "ess_per_step": r["ess_per_step"], "mh_steps_per_s": r["mh_steps_per_s"],
                  "ess_per_s": r["ess_per_s"]}
    res["mala_over_rw_ess_per_s"] = out["mala"]["ess_per_s"] / out["rw"]["ess_per_s"]
    return res


def batched(G, seed=0, blocks=3, mh_steps=1000):
    """The batched workload: 64 datasets (n log-uniform in [6, 250]), 4096 chains each,
    caps 32/64/128 round robin; blocks of MH steps + MAP resampling replayed as hipGraphs
    after autotuning."""
    from gpuropolis_amd import workloads as W
    dsets, caps, _ = W.datasets_B(seed)
    ctx = G.Context(cap=128, seed=seed, schedule=G.SCHED_WAVE)
    ctx.set_data(dsets, [4096] * len(dsets), caps)
    ctx.init_chains()
    ctx.run_blocks_graph(1, mh_steps)
    table, _ = ctx.autotune(50)
    ctx.run_blocks_graph(1, mh_steps)                       # capture with the tuned shapes
    ctx.reset_stats()
    ms = 0.0
    t0 = time.perf_counter()
    for _ in range(blocks):
        ctx.run_blocks_graph(1, mh_steps)
        ms += ctx.stats().last_block_ms
    wall = time.perf_counter() - t0
    st = ctx.stats()
    ctx.close()
    return {"workload": "batched: 64 datasets, n log-uniform [6, 250], 4096 chains each, caps 32/64/128, "
                        f"blocks of {mh_steps} MH steps + MAP-100 resample as hipGraph replays",
            "blocks": blocks, "mh_steps_per_s": st.steps / wall, "graph_ms_per_block": ms / blocks,
            "alg_tflops": st.alg_flops / (ms / 1e3) / 1e12,
            "frac_valu_peak": st.alg_flops / (ms / 1e3) / VALU_LANE_OP_PEAK, "launch_shape": table}


NIST_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "nist")


def nist_batch(G, seed=0, blocks=2, mh_steps=1000, chains=8192, tune_steps=50):
    """The NIST real-data batch: all 18 NIST StRD datasets (n 6 to 236, the reference's
    data-sources/NIST data.txt files) in one persistent launch, 8192 chains each (the
    reference's run-all.sh N), blocks of MH steps plus resampling.  Both semantics: the
    reference's heap-31 sampler with its median re-draw rule, and the postfix sampler
    (cap 64) with MAP-100 resampling."""
    import numpy as np
    names = sorted(os.listdir(NIST_DIR))
    dsets = []
    for nm in names:
        a = np.loadtxt(os.path.join(NIST_DIR, nm, "data.txt"), comments="#", ndmin=2)
        dsets.append(tuple(np.ascontiguousarray(a[:, k], dtype=np.float32) for k in range(3)))
    out = {"workload": f"NIST batch: {len(names)} NIST StRD datasets (n {min(len(d[0]) for d in dsets)} to "
                       f"{max(len(d[0]) for d in dsets)}), {chains} chains each, one persistent launch per "
                       f"size class, blocks of {mh_steps} MH steps + resampling", "blocks": blocks}
    for name, kw in (("reference", dict(semantics=G.SEM_REFERENCE, resample=G.RS_QUANTILE)),
                     ("config", dict(cap=64, resample=G.RS_MAP))):
        ctx = G.Context(seed=seed, schedule=G.SCHED_WAVE, **kw)
        ctx.set_data(dsets, [chains] * len(dsets))
        ctx.init_chains()
        ctx.run_block(mh_steps)
        ctx.resample()
        if tune_steps > 0:
            ctx.autotune(tune_steps)
        ctx.reset_stats()
        kms = 0.0
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(blocks):
            ctx.run_block(mh_steps)
            kms += ctx.stats().last_block_ms
            ctx.resample()
        ctx.sync()
        wall = time.perf_counter() - t0
        st = ctx.stats()
        ctx.close()
        out[name] = {"mh_steps_per_s": st.steps / wall, "kernel_ms_per_block": kms / blocks,
                     "node_evals_per_s": st.node_evals / wall,
                     "alg_tflops": st.alg_flops / (kms / 1e3) / 1e12}
    return out


def full_form(G, seed=0, blocks=1, mh_steps=200, chains=1 << 20):
    """The full form: 1,048,576 chains over 256 datasets (n log-uniform in [16, 1024], x on
    [-10, 10]), programs of at most 128 tokens, MALA constant moves and subtree structure
    moves in the batched persistent launch, MAP-100 resampling."""
    from gpuropolis_amd import workloads as W
    dsets, _ = W.datasets_F(seed)
    nd = len(dsets)
    per = [chains // nd] * nd
    ctx = G.Context(cap=128, seed=seed, schedule=G.SCHED_WAVE, proposal=G.PROP_MALA, eps=0.01)
    ctx.set_data(dsets, per)
    ctx.init_chains()
    ctx.run_block(mh_steps)
    ctx.resample()
    ctx.reset_stats()
    kms = 0.0
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(blocks):
        ctx.run_block(mh_steps)
        kms += ctx.stats().last_block_ms
        ctx.resample()
    ctx.sync()
    wall = time.perf_counter() - t0
    st = ctx.stats()
    ctx.close()
    return {"workload": f"full form: {sum(per)} chains, {nd} datasets (n log-uniform [16, 1024]), cap 128, "
                        f"MALA constants (eps 0.01) + subtree structure moves, MAP-100 resampling, "
                        f"blocks of {mh_steps} MH steps", "blocks": blocks,
            "mh_steps_per_s": st.steps / wall, "kernel_ms_per_block": kms / blocks,
            "node_evals_per_s": st.node_evals / wall, "alg_tflops": st.alg_flops / (kms / 1e3) / 1e12,
            "acceptance": {"structure": st.struct_accepted / max(1, st.struct_proposed),
                           "constants": st.const_accepted / max(1, st.const_proposed)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed MCMC blocks")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mh-steps", type=int, default=1000, help="MH steps per block")
    ap.add_argument("--chains", type=int, default=131072)
    ap.add_argument("--cap", type=int, default=64)
    ap.add_argument("--schedule", choices=["wave", "lane"], default="wave")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-chains", type=int, default=65536)
    ap.add_argument("--cpu-steps", type=int, default=1000)
    ap.add_argument("--other-schedule-blocks", type=int, default=2,
                    help="blocks timed on the other evaluator schedule for comparison (0: skip)")
    ap.add_argument("--batched-blocks", type=int, default=3, help="timed blocks of the batched workload (0: skip)")
    ap.add_argument("--no-ess", dest="ess", action="store_false", help="skip the MALA vs random-walk ESS run")
    ap.add_argument("--nist-blocks", type=int, default=2, help="timed blocks of the NIST batch (0: skip)")
    ap.add_argument("--full-blocks", type=int, default=1, help="timed blocks of the full form (0: skip)")
    ap.add_argument("--tune-steps", type=int, default=50,
                    help="autotune the wave launch shapes on this many MH steps after warm-up (0: defaults)")
    ap.add_argument("--step-split", action="store_true",
                    help="only time one block of constant steps and one of structure steps on the large "
                         "workload's programs (after two warm-up blocks) and print their cost per chain-step")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.step_split:
        step_split(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(datasets and chains shard across GPUs without communication, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from gpuropolis_amd import lib as G
    from gpuropolis_amd import workloads as W

    ds, _target = W.dataset_L(args.seed)
    sched = G.SCHED_WAVE if args.schedule == "wave" else G.SCHED_LANE
    st, wall, kernel_ms, table = time_blocks(G, ds, args, sched, args.steps, args.warmup,
                                             tune=args.tune_steps > 0 and args.schedule == "wave")
    mh_steps = st.steps
    value = mh_steps / wall
    evals_per_s = st.node_evals / wall
    flops_per_launch = st.alg_flops / args.steps
    avg_launch_s = kernel_ms / 1e3 / args.steps
    achieved = flops_per_launch / avg_launch_s
    kname = "k_mh_config_wave" if args.schedule == "wave" else "k_mh_config_lane"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "MH steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "plain method, large: 131072 chains, postfix cap 64, one dataset n=250, "
                               "blocks of 1000 MH steps + MAP-100 resample",
                   "chains": args.chains, "cap": args.cap, "n": 250, "mh_steps_per_block": args.mh_steps,
                   "schedule": args.schedule},
        "node_evals_per_s": evals_per_s,
        "launch_shape": table,
        "acceptance": {"structure": st.struct_accepted / max(1, st.struct_proposed),
                       "constants": st.const_accepted / max(1, st.const_proposed)},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_LANE_OP_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / VALU_LANE_OP_PEAK, "traffic": None,
                     "kernel": kname, "flops_per_launch": flops_per_launch,
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "peak_fp32_datasheet": FP32_DATASHEET_PEAK / 1e12,
                     "frac_fp32_datasheet": achieved / FP32_DATASHEET_PEAK},
    }
    issue = MEASURED_ISSUE.get((args.schedule, args.chains, args.cap, args.mh_steps, args.seed))
    if issue is not None:
        slots = 1024 * 2.4e9 / 4 * avg_launch_s
        chain_steps = args.chains * args.mh_steps
        out["roofline"]["issue_frac"] = {"valu": issue[0] * chain_steps / slots, "salu": issue[1] * chain_steps / slots}
    traffic = MEASURED_TRAFFIC.get((args.schedule, args.chains, args.cap, args.mh_steps, args.seed))
    if traffic is not None:
        out["roofline"]["traffic"] = traffic
        out["roofline"]["hbm_gbs"] = traffic / avg_launch_s / 1e9
        out["roofline"]["hbm_frac"] = traffic / avg_launch_s / HBM_PEAK
    if args.other_schedule_blocks > 0:
        other = "lane" if args.schedule == "wave" else "wave"
        o_st, o_wall, o_ms, _ = time_blocks(G, ds, args, G.SCHED_LANE if other == "lane" else G.SCHED_WAVE,
                                         args.other_schedule_blocks, 1)
        out["schedules"] = {
            args.schedule: {"mh_steps_per_s": value, "kernel_ms_per_block": avg_launch_s * 1e3},
            other: {"mh_steps_per_s": o_st.steps / o_wall, "blocks": args.other_schedule_blocks,
                    "kernel_ms_per_block": o_ms / args.other_schedule_blocks,
                    "alg_tflops": o_st.alg_flops / (o_ms / 1e3) / 1e12}}
    if args.batched_blocks > 0:
        out["batched"] = batched(G, args.seed, args.batched_blocks, args.mh_steps)
    if args.nist_blocks > 0:
        out["nist"] = nist_batch(G, args.seed, args.nist_blocks, args.mh_steps)
    if args.full_blocks > 0:
        out["full"] = full_form(G, args.seed, args.full_blocks)
    if args.ess:
        out["ess"] = ess_rational(G, args.seed)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ds, args.cap, args.seed, args.cpu_chains, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

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
            datasheet peak (157.3 TFLOP/s, packed FMA) is reported beside it.
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
MEASURED_TRAFFIC = {("wave", 131072, 64, 1000, 0): 176024608}

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


def time_blocks(G, ds, args, sched, blocks, warmup):
    ctx = G.Context(cap=args.cap, seed=args.seed, schedule=sched)
    ctx.set_data([ds], [args.chains])
    ctx.init_chains()
    for _ in range(warmup):
        ctx.run_block(args.mh_steps)
        ctx.resample()
    ctx.sync()
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
    return st, wall, kernel_ms


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
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
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
    st, wall, kernel_ms = time_blocks(G, ds, args, sched, args.steps, args.warmup)
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
        "acceptance": {"structure": st.struct_accepted / max(1, st.struct_proposed),
                       "constants": st.const_accepted / max(1, st.const_proposed)},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_LANE_OP_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / VALU_LANE_OP_PEAK, "traffic": None,
                     "kernel": kname, "flops_per_launch": flops_per_launch,
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "peak_fp32_datasheet": FP32_DATASHEET_PEAK / 1e12,
                     "frac_fp32_datasheet": achieved / FP32_DATASHEET_PEAK},
    }
    traffic = MEASURED_TRAFFIC.get((args.schedule, args.chains, args.cap, args.mh_steps, args.seed))
    if traffic is not None:
        out["roofline"]["traffic"] = traffic
        out["roofline"]["hbm_gbs"] = traffic / avg_launch_s / 1e9
        out["roofline"]["hbm_frac"] = traffic / avg_launch_s / HBM_PEAK
    if args.other_schedule_blocks > 0:
        other = "lane" if args.schedule == "wave" else "wave"
        o_st, o_wall, o_ms = time_blocks(G, ds, args, G.SCHED_LANE if other == "lane" else G.SCHED_WAVE,
                                         args.other_schedule_blocks, 1)
        out["schedules"] = {
            args.schedule: {"mh_steps_per_s": value, "kernel_ms_per_block": avg_launch_s * 1e3},
            other: {"mh_steps_per_s": o_st.steps / o_wall, "blocks": args.other_schedule_blocks,
                    "kernel_ms_per_block": o_ms / args.other_schedule_blocks,
                    "alg_tflops": o_st.alg_flops / (o_ms / 1e3) / 1e12}}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ds, args.cap, args.seed, args.cpu_chains, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

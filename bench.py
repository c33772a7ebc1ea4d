"""Benchmark: the MH sampler on one MI355X, prints ONE JSON line.

Headline workload, the full form: 1,048,576 chains over 256
datasets of uneven size (n log-uniform in [16, 1024], x ~ U[-10, 10], y = a seeded random
target expression of at most 24 nodes + noise of sd 1% of the y range), postfix programs
of at most 128 tokens, MALA constant moves (eps 0.01) and subtree structure moves in the
batched persistent launch, MAP-relative resampling (chains > 100 nats below their
dataset's MAP copy a good chain), all on the device.  A bench step is one MCMC block of
1000 MH steps on every chain followed by the resampling; --steps blocks are timed after
--warmup blocks.

value        whole-job MH steps per second (chains x steps / wall time of the K blocks)
roofline     the MH kernel of the headline (one persistent launch per block): algorithmic
             FP32 ops per launch (per evaluated program n * (non-leaf tokens + 4) +
             constants for the forward evaluation, plus n * (non-leaf tokens + 2 +
             constants) for each reverse sweep of a MALA gradient, counted by the kernel)
             over the kernel's average launch time (HIP events on the stream it runs on),
             against the VALU lane-op issue peak (256 CU x 64 lanes x 2.4 GHz); the FP32
             datasheet peak (157.3 TFLOP/s, packed FMA) is reported beside it.  HBM bytes
             per launch and the VALU/SALU issue shares come from rocprofv3 counters of the
             same launch (MEASURED_TRAFFIC, MEASURED_ISSUE), over the live launch time.
cpu_baseline the oracle's serial restatement of the same sampler (OpenMP over chains) on
             this host, on a bounded sample of the same workload; beside it the 1-thread
             rate and the reference semantics (the reference's heap-31 sampler on the 18
             NIST datasets) at both thread counts.
Sub-records: the plain large workload on both evaluator schedules, the batched workload,
the NIST batch in both semantics, and MALA against random walk on the differentiation
workload (ESS per second).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_LANE_OP_PEAK = 256 * 64 * 2.4e9      # lane-ops/s
FP32_DATASHEET_PEAK = 157.3e12            # FLOP/s, packed FMA
HBM_PEAK = 8.0e12

# HBM bytes per MH launch (FETCH_SIZE + WRITE_SIZE of rocprofv3 --pmc, scaled by the
# same counters on a plain copy of known size), keyed by (workload, chains, MH steps per
# block, seed); the profiles/ summaries hold the runs.  Valid for the kernel build they
# were measured on.  For the batched workload a "launch" is one block's six concurrent
# size-class launches.
MEASURED_TRAFFIC = {
    ("full", 1048576, 1000, 0): 1_323_373_952,
    ("large", 131072, 1000, 0): 163_308_256,
    ("batched", 262144, 1000, 0): 384_401_216,
}
# VALU and SALU instructions per chain-step of the same launches (rocprofv3 --pmc
# SQ_INSTS_VALU, SQ_INSTS_SALU over chains x steps).  With the live launch time they give
# the share of issue slots each unit is busy: one wave64 instruction per SIMD every 4
# cycles, 1024 SIMDs at 2.4 GHz.
MEASURED_ISSUE = {
    ("full", 1048576, 1000, 0): (1116.28, 758.48),
    ("large", 131072, 1000, 0): (281.12, 288.25),
    ("batched", 262144, 1000, 0): (319.92, 296.77),
}

METRIC = "MH steps/sec + node×datapoint evals/sec (% VALU peak); ESS/sec MALA vs RW"


def progress(name, value):
    """One line on stderr as each part of the run ends (the JSON line goes to stdout)."""
    print(f"bench.py: {name} done ({value:.4g})", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_threads():
    """Threads the CPU baseline uses: the host's share of cores for this process (the
    OpenMP thread count the environment grants, else the CPUs this process may run on)."""
    env = os.environ.get("GPR_CPU_THREADS") or os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(seed=0, scale=1.0):
    """The oracle (oracle/oracle.c, serial C with OpenMP over chains) on bounded samples:
    the full form's config sampler (MALA + subtree moves, cap 128) on 16 of its 256
    datasets (every 16th), and the reference sampler (heap-31 programs, the reference's
    moves) on the 18 NIST datasets, each with all granted threads and with one."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from gpuropolis_amd import workloads as W
    O.build()
    threads = host_threads()
    dsets, _ = W.datasets_F(seed)
    sel = list(range(0, len(dsets), 16))
    names = sorted(os.listdir(NIST_DIR))
    nist = [nist_data(nm) for nm in names]

    def run_config(nthreads, nch, steps):
        cfg = O.make_cfg(cap=128, seed=seed, proposal=1, eps=0.01)
        t0, tot = time.perf_counter(), 0
        for d in sel:
            x, y, sd = dsets[d]
            ops, cs, ln = O.cfg_init(cfg, nch, chain0=d * 4096)
            O.cfg_run_block(cfg, ops, cs, ln, x, y, sd, 0, steps, chain0=d * 4096, nthreads=nthreads)
            tot += nch * steps
        dt = time.perf_counter() - t0
        return tot / dt, dt, f"{len(sel)} of the full form's datasets (every 16th) x {nch} chains x {steps} MH steps " \
                             f"from the initial programs"

    def run_reference(nthreads, nch, steps):
        cfg = O.GpoRefCfg(0, seed, 0)
        t0, tot = time.perf_counter(), 0
        for d, (x, y, sd) in enumerate(nist):
            hop, hc = O.ref_init(seed, nch, chain0=d * 8192)
            O.ref_run_block(cfg, hop, hc, x, y, sd, 0, steps, chain0=d * 8192, nthreads=nthreads)
            tot += nch * steps
        dt = time.perf_counter() - t0
        return tot / dt, dt, f"the {len(nist)} NIST datasets x {nch} chains x {steps} MH steps from the initial programs"

    def sz(n):
        return max(8, int(n * scale))

    runs = {}
    for key, fn, nch_all, nch_one in (("config", run_config, 512, 64), ("reference", run_reference, 512, 40)):
        v, dt, what = fn(threads, sz(nch_all), 1000)
        v1, dt1, what1 = fn(1, sz(nch_one), 1000)
        runs[key] = {"all_threads": {"value": v, "threads": threads, "seconds": dt, "sample": what},
                     "one_thread": {"value": v1, "threads": 1, "seconds": dt1, "sample": what1}}
    c = runs["config"]["all_threads"]
    return {"value": c["value"], "unit": "MH steps/s", "cores": threads, "kind": "port",
            "sample": c["sample"] + f" (config sampler of the headline, {c['seconds']:.1f} s)",
            "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count(),
                     "affinity_cpus": len(os.sched_getaffinity(0)), "threads_granted": threads},
            "runs": runs}


def step_split(args):
    """Cost per chain-step of the step types, on the default launch shape: programs from
    warm-up blocks of a workload, then one block with structure moves off (constant
    steps only) and one with a structure move at every step.  --step-split large: the
    plain large workload (random walk); full: the full form (MALA), also with random-walk
    constant steps (sigma 0.01) on the same programs."""
    from gpuropolis_amd import lib as G
    from gpuropolis_amd import workloads as W
    if args.step_split == "large":
        dsets, chains, base = [W.dataset_L(args.seed)[0]], [args.chains], dict(cap=args.cap)
        warm = 2
    else:
        dsets = W.datasets_F(args.seed)[0]
        chains = [4096] * len(dsets)
        base = dict(cap=128, proposal=G.PROP_MALA, eps=0.01)
        warm = 1
    ctx = G.Context(seed=args.seed, schedule=G.SCHED_WAVE, **base)
    ctx.set_data(dsets, chains)
    ctx.init_chains()
    for _ in range(warm):
        ctx.run_block(args.mh_steps)
        ctx.resample()
    ops, cs, ln, _, _ = ctx.read_state()
    ctx.close()
    out = {"workload": args.step_split, "mean_tokens": float(ln.mean())}
    variants = [("constant_steps", dict(fixed_structure=1)), ("structure_steps", dict(struct_every=1))]
    if args.step_split == "full":
        variants.append(("constant_steps_random_walk", dict(fixed_structure=1, proposal=G.PROP_RW, sigma=0.01)))
    for name, kw in variants:
        ctx = G.Context(seed=args.seed, schedule=G.SCHED_WAVE, **{**base, **kw})
        ctx.set_data(dsets, chains)
        ctx.write_state(ops, cs, ln)
        ctx.run_block(args.mh_steps)
        st = ctx.stats()
        ms = st.last_block_ms
        out[name] = {"kernel_ms": ms, "ns_per_chain_step": ms * 1e6 / (sum(chains) * args.mh_steps),
                     "acceptance": (st.const_accepted + st.struct_accepted) / max(1, st.const_proposed +
                                                                                  st.struct_proposed),
                     "alg_flops": st.alg_flops, "node_evals": st.node_evals}
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
        return 0.0                       # no chain moved: no information
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


def ess_batched(G, seed=0, nch=512, thin=2, blocks=120):
    """MALA against random walk on the batched workload's datasets, each with its own
    target's structure fixed: the datasets whose target has 1 to 16 constants, nch chains
    each, all in one context (one persistent launch per size class).  Warm-up adapts a
    proposal metric per dataset (Cholesky factor of the chains' covariance of its
    constants, five rounds); each sampler then runs with each dataset at the step size
    that gave it the most ESS per step among five pilots.
    Per dataset, ESS = the smallest over its constants (all its chains pooled); ESS per
    second = ESS per chain-step x the MH steps per second of the whole batched launch
    (one 1000-step block).  Reported: the ratio of the summed ESS per second, and the
    median and geometric mean of the per-dataset ratios."""
    import numpy as np
    from gpuropolis_amd import workloads as W
    dsets, _, targets = W.datasets_B(seed)
    cap, CONST = 64, G.OP["CONST"]
    use = [d for d, (o, _) in enumerate(targets) if 1 <= int(np.sum(o == CONST)) <= 16 and len(o) <= cap]
    nd = len(use)
    N = nd * nch
    O = np.zeros((cap, N), dtype=np.uint8)
    V = np.zeros((cap, N), dtype=np.float32)
    L = np.zeros(N, dtype=np.int32)
    rng = np.random.default_rng(seed + 11)
    rows = []
    for k, d in enumerate(use):
        o, c = targets[d]
        r = np.flatnonzero(o == CONST)
        rows.append(r)
        sl = slice(k * nch, (k + 1) * nch)
        O[:len(o), sl] = o[:, None]
        V[:len(o), sl] = c[:, None]
        V[r, sl] += (1e-3 * (1 + np.abs(c[r]))[:, None] * rng.normal(size=(len(r), nch))).astype(np.float32)
        L[sl] = len(o)
    data = [dsets[d] for d in use]
    nrows = 1 + max(int(r.max()) for r in rows)           # the constants' rows (targets of at most 16 tokens)

    def context(prop, step, cs, chols, scale=None):
        # scale[k]: dataset k's step relative to the global one (its metric factor scaled;
        # for MALA the drift's metric scales with its square, as a step of eps * scale[k])
        kw = dict(eps=step) if prop == G.PROP_MALA else dict(sigma=step)
        ctx = G.Context(cap=cap, seed=seed + 7, proposal=prop, fixed_structure=1, resample=G.RS_NONE,
                        schedule=G.SCHED_WAVE, **kw)
        ctx.set_data(data, [nch] * nd)
        if chols is not None:
            for k in range(nd):
                ctx.set_metric(k, rows[k], chols[k] * (1.0 if scale is None else scale[k]))
        ctx.write_state(O, cs, L)
        return ctx

    def metric(cs):
        out = []
        for k in range(nd):
            x = cs[rows[k]][:, k * nch:(k + 1) * nch].astype(np.float64)
            cov = np.atleast_2d(np.cov(x))
            cov = cov + (1e-12 * np.trace(cov) + 1e-30) * np.eye(len(rows[k]))
            out.append(np.linalg.cholesky(cov))
        return out

    cs, chols = V.copy(), None
    for step, nsteps in ((1e-3, 1000), (1.0, 500), (1.0, 500), (1.0, 500), (1.0, 500)):
        ctx = context(G.PROP_MALA, step, cs, chols)
        ctx.run_block(nsteps)
        cs = ctx.read_state()[1]
        ctx.close()
        chols = metric(cs)

    def run(prop, step, nblocks, burn, scale=None):
        ctx = context(prop, step, cs, chols, scale)
        draws = []
        for i in range(nblocks):
            ctx.run_block(thin)
            if i >= burn:
                draws.append(ctx.read_state()[1][:nrows].copy())
        ll = ctx.read_state()[4]
        ctx.reset_stats()
        ctx.run_block(1000)
        rate = ctx.stats().steps / (ctx.stats().last_block_ms / 1e3)
        ctx.close()
        D = np.array(draws)                                   # [draws, nrows, N]
        chain_steps = (nblocks - burn) * thin * nch
        ess = np.zeros(nd)
        for k in range(nd):
            sl = slice(k * nch, (k + 1) * nch)
            ess[k] = min(ess_multichain(D[:, j, sl].astype(np.float64).T) for j in rows[k])
        valid = np.array([np.all(np.isfinite(ll[k * nch:(k + 1) * nch])) for k in range(nd)])
        return dict(step=step, ess_per_step=ess / chain_steps, rate=rate, valid=valid)

    out = {}
    for name, prop, steps in (("rw", G.PROP_RW, [0.4, 0.8, 1.6, 2.4, 3.2]),
                              ("mala", G.PROP_MALA, [0.6, 1.0, 1.4, 1.8, 2.4])):
        # per dataset the pilot step with the most ESS per step (one launch serves all, so
        # the cost per step is shared), then one run with each dataset at its own step
        pilots = [run(prop, h, 50, 15) for h in steps]
        eps = np.array([r["ess_per_step"] for r in pilots])
        scale = np.array(steps)[np.argmax(eps, axis=0)]
        out[name] = run(prop, 1.0, blocks, blocks // 4, scale)
        out[name]["scale"] = scale
    ok = out["rw"]["valid"] & out["mala"]["valid"]
    e_rw = out["rw"]["ess_per_step"][ok] * out["rw"]["rate"]
    e_ma = out["mala"]["ess_per_step"][ok] * out["mala"]["rate"]
    ratio = e_ma / e_rw
    nconst = np.array([len(r) for r in rows])[ok]
    res = {"workload": f"batched, fixed structures: the batched workload's {nd} datasets whose target has 1 to 16 "
                       f"constants, each target's structure fixed, {nch} chains per dataset in one context, proposal "
                       f"metric per dataset adapted in warm-up, {blocks} draws thinned by {thin} MH steps, first quarter "
                       f"discarded", "datasets_used": int(ok.sum()),
           "mala_over_rw_ess_per_s": float(e_ma.sum() / e_rw.sum()),
           "ratio_median": float(np.median(ratio)), "ratio_geomean": float(np.exp(np.mean(np.log(ratio))))}
    for name in ("rw", "mala"):
        r = out[name]
        res[name] = {"step_median": float(np.median(r["scale"][ok])), "mh_steps_per_s": r["rate"],
                     "ess_per_step_median": float(np.median(r["ess_per_step"][ok]))}
    for lo, hi in ((1, 2), (3, 4), (5, 16)):
        m = (nconst >= lo) & (nconst <= hi)
        if m.any():
            res[f"ratio_median_{lo}_to_{hi}_constants"] = float(np.median(ratio[m]))
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
    nch = 4096 * len(dsets)
    # the block's six size-class launches run concurrently: the roofline takes the graph
    # time of the block (HIP events around the replay on the context stream) as the launch
    return {"workload": "batched: 64 datasets, n log-uniform [6, 250], 4096 chains each, caps 32/64/128, "
                        f"blocks of {mh_steps} MH steps + MAP-100 resample as hipGraph replays",
            "blocks": blocks, "mh_steps_per_s": st.steps / wall, "graph_ms_per_block": ms / blocks,
            "alg_tflops": st.alg_flops / (ms / 1e3) / 1e12,
            "frac_valu_peak": st.alg_flops / (ms / 1e3) / VALU_LANE_OP_PEAK, "launch_shape": table,
            "roofline": roofline("k_mh_config_wave (six size-class launches per block)", st.alg_flops / blocks,
                                 ms / 1e3 / blocks, ("batched", nch, mh_steps, seed))}


NIST_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "nist")


def nist_data(name):
    import numpy as np
    a = np.loadtxt(os.path.join(NIST_DIR, name, "data.txt"), comments="#", ndmin=2)
    return tuple(np.ascontiguousarray(a[:, k], dtype=np.float32) for k in range(3))


def nist_batch(G, seed=0, blocks=2, mh_steps=1000, chains=8192, tune_steps=50):
    """The NIST real-data batch: all 18 NIST StRD datasets (n 6 to 236, the reference's
    data-sources/NIST data.txt files) in one persistent launch, 8192 chains each (the
    reference's run-all.sh N), blocks of MH steps plus resampling.  Both semantics: the
    reference's heap-31 sampler with its median re-draw rule, and the postfix sampler
    (cap 64) with MAP-100 resampling."""
    names = sorted(os.listdir(NIST_DIR))
    dsets = [nist_data(nm) for nm in names]
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
                     "alg_tflops": st.alg_flops / (kms / 1e3) / 1e12,
                     "frac_valu_peak": st.alg_flops / (kms / 1e3) / VALU_LANE_OP_PEAK}
    return out


def full_form(G, seed=0, blocks=5, warmup=1, mh_steps=1000, chains=1 << 20, tune_steps=20):
    """The full form: 1,048,576 chains over 256 datasets (n log-uniform in [16, 1024], x on
    [-10, 10]), programs of at most 128 tokens, MALA constant moves and subtree structure
    moves in the batched persistent launch, MAP-100 resampling.  Returns the stats, the
    wall time of the timed blocks, the summed device time of their MH launches (HIP events
    on the context stream, which the single persistent launch runs on), and the tuning."""
    from gpuropolis_amd import workloads as W
    dsets, _ = W.datasets_F(seed)
    nd = len(dsets)
    per = [chains // nd] * nd
    ctx = G.Context(cap=128, seed=seed, schedule=G.SCHED_WAVE, proposal=G.PROP_MALA, eps=0.01)
    ctx.set_data(dsets, per)
    ctx.init_chains()
    for _ in range(warmup):
        ctx.run_block(mh_steps)
        ctx.resample()
    table = None
    if tune_steps > 0:
        table, _ = ctx.autotune(tune_steps)
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
    return st, wall, kms, table, sum(per), nd


def roofline(kname, flops_per_launch, launch_s, key):
    achieved = flops_per_launch / launch_s
    r = {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_LANE_OP_PEAK / 1e12, "unit": "TFLOP/s",
         "frac": achieved / VALU_LANE_OP_PEAK, "traffic": None, "kernel": kname,
         "flops_per_launch": flops_per_launch, "avg_launch_ms": launch_s * 1e3,
         "peak_fp32_datasheet": FP32_DATASHEET_PEAK / 1e12, "frac_fp32_datasheet": achieved / FP32_DATASHEET_PEAK}
    issue = MEASURED_ISSUE.get(key)
    if issue is not None:
        slots = 1024 * 2.4e9 / 4 * launch_s
        chain_steps = key[1] * key[2]
        r["issue_frac"] = {"valu": issue[0] * chain_steps / slots, "salu": issue[1] * chain_steps / slots}
    traffic = MEASURED_TRAFFIC.get(key)
    if traffic is not None:
        r["traffic"] = traffic
        r["hbm_gbs"] = traffic / launch_s / 1e9
        r["hbm_frac"] = traffic / launch_s / HBM_PEAK
    return r


def large(G, args):
    """The plain large workload (131072 chains, cap 64, one dataset n = 250, random-walk
    constants): 20 blocks on the wave schedule, autotuned, with its roofline; then the lane
    schedule on a few blocks."""
    from gpuropolis_amd import workloads as W
    ds, _ = W.dataset_L(args.seed)
    st, wall, kernel_ms, table = time_blocks(G, ds, args, G.SCHED_WAVE, args.large_blocks, 1,
                                             tune=args.tune_steps > 0)
    launch_s = kernel_ms / 1e3 / args.large_blocks
    out = {"workload": f"plain method, large: {args.chains} chains, postfix cap 64, one dataset n=250, "
                       f"blocks of {args.mh_steps} MH steps + MAP-100 resample",
           "blocks": args.large_blocks, "mh_steps_per_s": st.steps / wall, "node_evals_per_s": st.node_evals / wall,
           "launch_shape": table,
           "roofline": roofline("k_mh_config_wave<1,false,false,false,false>", st.alg_flops / args.large_blocks, launch_s,
                                ("large", args.chains, args.mh_steps, args.seed))}
    if args.lane_blocks > 0:
        o_st, o_wall, o_ms, _ = time_blocks(G, ds, args, G.SCHED_LANE, args.lane_blocks, 1)
        out["schedules"] = {
            "wave": {"mh_steps_per_s": st.steps / wall, "kernel_ms_per_block": launch_s * 1e3},
            "lane": {"mh_steps_per_s": o_st.steps / o_wall, "blocks": args.lane_blocks,
                     "kernel_ms_per_block": o_ms / args.lane_blocks,
                     "alg_tflops": o_st.alg_flops / (o_ms / 1e3) / 1e12}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed MCMC blocks of the full form")
    ap.add_argument("--warmup", type=int, default=1, help="untimed blocks of the full form before tuning")
    ap.add_argument("--mh-steps", type=int, default=1000, help="MH steps per block")
    ap.add_argument("--chains", type=int, default=131072, help="chains of the plain large workload")
    ap.add_argument("--cap", type=int, default=64, help="token cap of the plain large workload")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-scale", type=float, default=1.0, help="scale of the CPU baseline's chain samples")
    ap.add_argument("--large-blocks", type=int, default=20, help="timed blocks of the plain large workload (0: skip)")
    ap.add_argument("--lane-blocks", type=int, default=1, help="blocks of the large workload on the lane schedule")
    ap.add_argument("--batched-blocks", type=int, default=3, help="timed blocks of the batched workload (0: skip)")
    ap.add_argument("--no-ess", dest="ess", action="store_false", help="skip the MALA vs random-walk ESS run")
    ap.add_argument("--nist-blocks", type=int, default=2, help="timed blocks of the NIST batch (0: skip)")
    ap.add_argument("--tune-steps", type=int, default=20,
                    help="autotune the wave launch shapes on this many MH steps after warm-up (0: defaults)")
    ap.add_argument("--only-headline", action="store_true", help="time the full form only (profiling runs)")
    ap.add_argument("--only", choices=["large", "batched", "nist", "ess", "ess_batched", "cpu"], default=None,
                    help="run one sub-record alone and print it (profiling runs)")
    ap.add_argument("--step-split", choices=["large", "full"], default=None,
                    help="only time one block of constant steps and one of structure steps on a workload's "
                         "programs (after warm-up blocks) and print their cost per chain-step")
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
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(datasets and chains shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)

    from gpuropolis_amd import lib as G

    if args.only:
        run = {"large": lambda: large(G, args),
               "batched": lambda: batched(G, args.seed, args.batched_blocks, args.mh_steps),
               "nist": lambda: nist_batch(G, args.seed, args.nist_blocks, args.mh_steps),
               "ess": lambda: ess_rational(G, args.seed),
               "ess_batched": lambda: ess_batched(G, args.seed),
               "cpu": lambda: cpu_baseline(args.seed, args.cpu_scale)}[args.only]
        print(json.dumps({args.only: run()}), flush=True)
        return

    st, wall, kms, table, nchains, nd = full_form(G, args.seed, args.steps, args.warmup, args.mh_steps,
                                                  tune_steps=args.tune_steps)
    launch_s = kms / 1e3 / args.steps
    out = {
        "metric": METRIC,
        "value": st.steps / wall,
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
        "config": {"workload": f"full form: {nchains} chains, {nd} datasets (n log-uniform [16, 1024], x on "
                               f"[-10, 10], targets <= 24 nodes + 1% noise), postfix cap 128, MALA constants "
                               f"(eps 0.01) + subtree structure moves, one persistent launch per block, blocks of "
                               f"{args.mh_steps} MH steps + MAP-100 resample",
                   "chains": nchains, "datasets": nd, "cap": 128, "mh_steps_per_block": args.mh_steps},
        "node_evals_per_s": st.node_evals / wall,
        "launch_shape": table,
        "acceptance": {"structure": st.struct_accepted / max(1, st.struct_proposed),
                       "constants": st.const_accepted / max(1, st.const_proposed)},
        "roofline": roofline("k_mh_config_wave<1,true,true,false,false>", st.alg_flops / args.steps, launch_s,
                             ("full", nchains, args.mh_steps, args.seed)),
    }
    progress("full form", out["value"])
    if not args.only_headline:
        if args.large_blocks > 0:
            out["large"] = large(G, args)
            progress("large", out["large"]["mh_steps_per_s"])
        if args.batched_blocks > 0:
            out["batched"] = batched(G, args.seed, args.batched_blocks, args.mh_steps)
            progress("batched", out["batched"]["mh_steps_per_s"])
        if args.nist_blocks > 0:
            out["nist"] = nist_batch(G, args.seed, args.nist_blocks, args.mh_steps)
            progress("nist", out["nist"]["config"]["mh_steps_per_s"])
        if args.ess:
            out["ess"] = ess_rational(G, args.seed)
            progress("ess", out["ess"]["mala_over_rw_ess_per_s"])
            out["ess_batched"] = ess_batched(G, args.seed)
            progress("ess_batched", out["ess_batched"]["mala_over_rw_ess_per_s"])
        if not args.no_cpu_baseline:
            out["cpu_baseline"] = cpu_baseline(args.seed, args.cpu_scale)
            progress("cpu_baseline", out["cpu_baseline"]["value"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

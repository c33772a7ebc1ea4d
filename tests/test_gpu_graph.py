"""MCMC blocks captured as hipGraphs: every replay (with eager reads of the state in
between) leaves exactly the state and counters the same eager gpr_run_block +
gpr_resample calls leave, for the config sampler (random walk and MALA) and the
reference sampler on the wave schedule.  The batched cases hold datasets of every size
class and both register forms, so a block is several launches forked onto their own
streams and joined (MALA: each two-register launch with its own token slabs)."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu

CASES = [("config-rw", dict(cap=64, seed=4, resample_nats=5.0)),
         ("config-mala", dict(cap=32, seed=6, proposal=1, eps=0.01, resample_nats=5.0)),
         ("reference", dict(semantics=1, seed=8, schedule=2, resample=2)),
         ("config-rw-batched", dict(cap=128, seed=5, resample_nats=5.0)),
         ("config-mala-batched", dict(cap=128, seed=7, proposal=1, eps=0.01, resample_nats=5.0))]


def set_data(ctx, name, nch):
    from gpuropolis_amd import workloads as W
    if name.endswith("batched"):
        dsets, caps, _ = W.datasets_B(0, nd=9)
        ctx.set_data(dsets, [nch // 9] * len(dsets), caps)
    else:
        ctx.set_data([W.dataset_S(0)], [nch])


def same_state(a, b):
    return all(np.array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8)) for u, v in zip(a, b))


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_graph_replays_match_eager(gpr, name, kw):
    nch, steps, reps = 1000, 40, 3
    eager = []
    with gpr.Context(**kw) as ctx:
        set_data(ctx, name, nch)
        ctx.init_chains()
        for _ in range(reps):
            ctx.run_block(steps)
            ctx.resample()
            eager.append(ctx.read_state())
        st_e = ctx.stats()
    with gpr.Context(**kw) as ctx:
        set_data(ctx, name, nch)
        ctx.init_chains()
        for r in range(reps):
            say(f"{name}: graph replay {r}")
            ctx.run_blocks_graph(1, steps)
            assert same_state(ctx.read_state(), eager[r]), (name, r)
        st_g = ctx.stats()
    for f in ("struct_proposed", "struct_accepted", "const_proposed", "const_accepted", "steps", "node_evals"):
        assert getattr(st_e, f) == getattr(st_g, f), f
    with gpr.Context(**kw) as ctx:
        set_data(ctx, name, nch)
        ctx.init_chains()
        say(f"{name}: one graph of {reps} blocks")
        ctx.run_blocks_graph(reps, steps)
        assert same_state(ctx.read_state(), eager[-1])
        ctx.run_block(steps)            # eager work after a replay continues the same chains
        ctx.resample()
        s1 = ctx.read_state()
    with gpr.Context(**kw) as ctx:
        set_data(ctx, name, nch)
        ctx.init_chains()
        for _ in range(reps + 1):
            ctx.run_block(steps)
            ctx.resample()
        assert same_state(ctx.read_state(), s1)


@pytest.mark.parametrize("name,kw", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_clock_resume_continues_the_streams(gpr, name, kw):
    """gpr_get_clock after eager blocks and graph replays; a new context given the saved
    chains and clock (gpr_write_state + gpr_set_clock) continues bit-identically."""
    from gpuropolis_amd import workloads as W
    nch, steps = 1000, 40
    with gpr.Context(**kw) as ctx:
        set_data(ctx, name, nch)
        ctx.init_chains()
        assert ctx.clock() == (0, 0)
        ctx.run_block(steps)
        ctx.resample()
        ctx.run_blocks_graph(1, steps)
        assert ctx.clock() == (2 * steps, 2)
        ops, cs, ln, _, _ = ctx.read_state()
        ctx.run_block(steps)
        ctx.resample()
        ctx.run_block(steps)
        ctx.resample()
        want = ctx.read_state()
        assert ctx.clock() == (4 * steps, 4)
    with gpr.Context(**kw) as ctx:
        set_data(ctx, name, nch)
        ctx.write_state(ops, cs, ln)
        ctx.set_clock(2 * steps, 2)
        say(f"{name}: resumed at step {2 * steps}")
        ctx.run_blocks_graph(2, steps)
        assert same_state(ctx.read_state(), want), name
        ctx.set_data([W.dataset_S(1)], [nch])
        assert ctx.clock() == (0, 0)

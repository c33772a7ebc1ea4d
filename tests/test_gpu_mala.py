"""MALA constant proposals in the wave MH kernel.

* Target invariance: on a one-constant program c*x with Gaussian data the posterior of
  c (Gaussian likelihood x Cauchy prior) is known by quadrature; MALA and random-walk
  chains both reproduce its mean and standard deviation (independent chains, so the
  standard error comes from the spread of per-chain means).  A correct sampler misses
  a 2-standard-error band in one run of 20 (over 20 seeds the oracle's z-scores here
  have mean 0.17 and sd 0.86), so the band is 3 standard errors; the same seed on the
  oracle gives the same estimate to a small fraction of a standard error.
* Same algorithm as the oracle: from the same start, device chains shadow the oracle's
  MALA chains (FP64 gradients there, FP32 adjoint sums here) on the rational workload.
"""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu


def toy():
    rng = np.random.default_rng(7)
    x = rng.uniform(0.0, 1.0, 20).astype(np.float32)
    y = (1.3 * x + rng.normal(0.0, 0.1, 20)).astype(np.float32)
    sd = np.full(20, 0.1, dtype=np.float32)
    return x, y, sd


def quadrature_mean_sd(x, y, sd):
    c = np.linspace(0.5, 2.1, 200001)
    xd, yd, s = (v.astype(np.float64) for v in (x, y, sd))
    lp = -np.sum((c[:, None] * xd[None, :] - yd[None, :]) ** 2 / (2 * s * s), axis=1) - np.log1p(c * c)
    w = np.exp(lp - lp.max())
    m = np.sum(w * c) / np.sum(w)
    return m, np.sqrt(np.sum(w * (c - m) ** 2) / np.sum(w))


@pytest.mark.parametrize("proposal", ["mala", "rw"])
def test_posterior_mean_matches_quadrature(oracle, gpr, proposal):
    x, y, sd = toy()
    m_true, s_true = quadrature_mean_sd(x, y, sd)
    nch, blocks, burn, steps = 4096, 40, 5, 10
    ops, cs = oracle.tape_arrays(["X", ("CONST", 0.7), "MUL"])
    O = np.zeros((32, nch), dtype=np.uint8)
    V = np.zeros((32, nch), dtype=np.float32)
    O[:3] = ops[:, None]
    V[:3] = cs[:, None]
    L = np.full(nch, 3, dtype=np.int32)
    prop = gpr.PROP_MALA if proposal == "mala" else gpr.PROP_RW
    with gpr.Context(cap=32, seed=3, proposal=prop, eps=0.04, sigma=0.06, fixed_structure=1,
                     resample=gpr.RS_NONE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(O, V, L)
        say(f"{proposal}: {nch} chains x {blocks} blocks x {steps} steps")
        samples = []
        for b in range(blocks):
            ctx.run_block(steps)
            if b >= burn:
                samples.append(ctx.read_state()[1][1].astype(np.float64))
        st = ctx.stats()
    S = np.array(samples)                      # [draws, chains]
    chain_means = S.mean(axis=0)
    m = chain_means.mean()
    se = chain_means.std(ddof=1) / np.sqrt(nch)
    acc = st.const_accepted / st.const_proposed
    say(f"{proposal}: mean {m:.6f} quadrature {m_true:.6f} se {se:.2e} sd {S.std():.5f} vs {s_true:.5f} "
        f"acceptance {acc:.3f}")
    assert abs(m - m_true) < 3 * se
    assert abs(S.std() / s_true - 1) < 0.03
    assert 0.2 < acc < 0.99
    # the oracle's chains from the same seed
    cfg = oracle.make_cfg(cap=32, proposal=1 if proposal == "mala" else 0, eps=0.04, sigma=0.06,
                          fixed_structure=1, seed=3)
    o_S = []
    for b in range(blocks):
        oracle.cfg_run_block(cfg, O, V, L, x, y, sd, b * steps, steps)
        if b >= burn:
            o_S.append(V[1].astype(np.float64).copy())
    m_o = np.array(o_S).mean()
    say(f"{proposal}: oracle mean {m_o:.6f}")
    assert abs(m - m_o) < 0.25 * se


def test_mala_shadows_oracle_on_rational_workload(oracle, gpr):
    from gpuropolis_amd.workloads import dataset_D, rational_tape
    (x, y, sd), b = dataset_D(0)
    rng = np.random.default_rng(4)
    nch, steps, eps = 64, 40, 1e-3
    O = np.zeros((32, nch), dtype=np.uint8)
    V = np.zeros((32, nch), dtype=np.float32)
    for c in range(nch):
        o, v = rational_tape(b * (1 + 0.002 * rng.normal(size=7)))
        O[:len(o), c] = o
        V[:len(o), c] = v
    L = np.full(nch, len(o), dtype=np.int32)
    cfg = oracle.make_cfg(cap=32, proposal=1, eps=eps, fixed_structure=1, seed=9)
    o_ops, o_cs, o_ln = O.copy(), V.copy(), L.copy()
    p_o, l_o, cnt = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, 0, steps)
    with gpr.Context(cap=32, seed=9, proposal=gpr.PROP_MALA, eps=eps, fixed_structure=1,
                     resample=gpr.RS_NONE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(O, V, L)
        say("mala shadowing: 64 chains x 40 steps, rational workload")
        ctx.run_block(steps)
        g_ops, g_cs, g_ln, g_pr, g_ll = ctx.read_state()
        st = ctx.stats()
    rel = np.max(np.abs(g_cs[:len(o)].astype(np.float64) - o_cs[:len(o)]) /
                 np.maximum(np.abs(o_cs[:len(o)].astype(np.float64)), 1e-3), axis=0)
    close = rel < 1e-3
    say("chains within 1e-3 of the oracle's MALA chain:", int(close.sum()), "of", nch,
        "acceptance gpu", st.const_accepted / st.const_proposed, "oracle", cnt[3] / max(cnt[2], 1))
    assert close.sum() >= int(0.9 * nch)
    assert cnt[3] > 0.3 * cnt[2]
    assert abs(st.const_accepted - cnt[3]) <= 0.1 * cnt[3] + 5


@pytest.mark.parametrize("cap,n", [(64, 32), (128, 70)])
def test_mala_structure_moves_follow_oracle(oracle, gpr, cap, n):
    """With a structure move at every step (struct_every = 1) the MALA kernel makes no
    constant move, and its structural proposals are evaluated by the stack-free forward
    sweep with a plan built per proposal: the trajectories follow the oracle's as the
    random-walk kernel's do (tapes over 64 tokens at cap 128)."""
    from gpu_util import check_paths, kappas, ll_agree
    rng = np.random.default_rng(cap + n)
    if n == 32:
        from gpuropolis_amd import workloads as W
        x, y, sd = W.dataset_S(0)
    else:
        x = rng.uniform(-5, 5, n).astype(np.float32)
        y = (1.5 * np.exp(-0.4 * x) + rng.normal(0, 0.05, n)).astype(np.float32)
        sd = np.full(n, 0.05, dtype=np.float32)
    nch, steps = 64, 200
    cfg = oracle.make_cfg(cap=cap, seed=77, proposal=1, eps=0.01, struct_every=1)
    ops, cs, ln = oracle.cfg_init(cfg, nch)
    o_ops, o_cs, o_ln = ops.copy(), cs.copy(), ln.copy()
    marg = np.zeros(nch)
    _, l1, cnt = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, 0, steps, margin=marg)
    p1 = np.array([oracle.tape_prior(o_ops[:o_ln[i], i], o_cs[:o_ln[i], i], 10) for i in range(nch)])
    with gpr.Context(cap=cap, seed=77, proposal=gpr.PROP_MALA, eps=0.01, struct_every=1,
                     resample=gpr.RS_NONE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(ops, cs, ln)
        say(f"mala structure moves cap={cap} n={n} steps={steps}")
        ctx.run_block(steps)
        g_ops, g_cs, g_ln, g_pr, g_ll = ctx.read_state()
        st = ctx.stats()
    same = (g_ln == o_ln) & np.all(g_ops == o_ops, axis=0) & np.all(g_cs.view(np.uint32) == o_cs.view(np.uint32), axis=0)
    say("longest tape", int(g_ln.max()))
    check_paths(same, marg, f"mala structure moves cap={cap} n={n}:")
    assert st.const_proposed == 0 and st.struct_proposed == nch * steps
    c = np.flatnonzero(same)
    assert np.array_equal(g_pr[c], p1[c])
    kap = kappas(oracle, o_ops, o_cs, o_ln, c, x, y, sd)
    ok, _ = ll_agree(g_ll[c], l1[c], n, kap, sd)
    assert ok.all()
    # a chain off the shared path may differ in every later accept
    assert abs(st.struct_accepted - cnt[1]) <= 3 + steps * (nch - same.sum())


def test_mala_mixed_moves_keep_state_consistent(oracle, gpr):
    """Structure moves and MALA constant moves mixed: after the block, every chain's
    stored prior is its final program's (bit for bit) and its stored LL matches the
    oracle's evaluation of that program."""
    from gpu_util import kappas, ll_agree
    from gpuropolis_amd import workloads as W
    x, y, sd = W.dataset_S(0)
    nch = 256
    with gpr.Context(cap=64, seed=31, proposal=gpr.PROP_MALA, eps=0.01, resample=gpr.RS_NONE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.init_chains()
        say("mala mixed moves: 256 chains x 300 steps")
        ctx.run_block(300)
        ops, cs, ln, pr, ll = ctx.read_state()
        st = ctx.stats()
    assert st.struct_accepted > 0 and st.const_accepted > 0
    p_ref = np.array([oracle.tape_prior(ops[:ln[i], i], cs[:ln[i], i], 10) for i in range(nch)])
    l_ref = np.array([oracle.tape_loglik(ops[:ln[i], i], cs[:ln[i], i], x, y, sd, mode=2) for i in range(nch)])
    assert np.array_equal(pr, p_ref)
    kap = kappas(oracle, ops, cs, ln, range(nch), x, y, sd)
    ok, _ = ll_agree(ll, l_ref, len(x), kap, sd)
    assert ok.all()


def test_mala_long_programs_shadow_oracle(oracle, gpr):
    """At cap 128 the MALA kernel keeps each wave's token slab in global memory: sums of
    c_j * x^(j mod 3 + 1) terms, 17 to 125 tokens in one launch, shadow the oracle's MALA
    chains."""
    rng = np.random.default_rng(17)
    x = rng.uniform(-1.0, 1.0, 90).astype(np.float32)
    y = (0.7 * x - 0.4 * x ** 2 + 0.2 * x ** 3 + rng.normal(0.0, 0.05, 90)).astype(np.float32)
    sd = np.full(90, 0.05, dtype=np.float32)
    nch, steps, eps, cap = 48, 40, 2e-3, 128
    tapes = []
    for c in range(nch):
        k = (3, 8, 12, 17, 21)[c % 5]
        toks = []
        for j in range(k):
            toks += ["X"] * (j % 3 + 1) + ["MUL"] * (j % 3) + [("CONST", float(rng.normal(0, 0.1))), "MUL"]
            toks += ["ADD"] if j else []
        tapes.append(oracle.tape_arrays(toks))
    O = np.zeros((cap, nch), dtype=np.uint8)
    V = np.zeros((cap, nch), dtype=np.float32)
    L = np.zeros(nch, dtype=np.int32)
    for c, (o, v) in enumerate(tapes):
        O[:len(o), c] = o
        V[:len(o), c] = v
        L[c] = len(o)
    assert L.max() > 100 and (L > 64).sum() >= nch // 2 and (L <= 64).sum() > 0
    cfg = oracle.make_cfg(cap=cap, proposal=1, eps=eps, fixed_structure=1, seed=23)
    o_ops, o_cs, o_ln = O.copy(), V.copy(), L.copy()
    _, _, cnt = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, 0, steps)
    with gpr.Context(cap=cap, seed=23, proposal=gpr.PROP_MALA, eps=eps, fixed_structure=1,
                     resample=gpr.RS_NONE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(O, V, L)
        say(f"mala long programs: {nch} chains, lengths {L.min()}..{L.max()}, {steps} steps")
        ctx.run_block(steps)
        g_cs = ctx.read_state()[1]
        st = ctx.stats()
    rel = np.max(np.abs(g_cs.astype(np.float64) - o_cs) / np.maximum(np.abs(o_cs.astype(np.float64)), 1e-3), axis=0)
    close = rel < 1e-3
    say("chains within 1e-3 of the oracle's:", int(close.sum()), "of", nch, "(over 64 tokens:",
        int(close[L > 64].sum()), "of", int((L > 64).sum()), ") acceptance gpu",
        st.const_accepted / st.const_proposed, "oracle", cnt[3] / max(cnt[2], 1))
    assert close[L > 64].sum() >= int(0.9 * (L > 64).sum())
    assert close.sum() >= int(0.9 * nch)
    assert cnt[3] > 0.2 * cnt[2]


@pytest.mark.parametrize("cap", [64, 128])
def test_mala_forward_mode_programs_shadow_oracle(oracle, gpr, cap):
    """Programs of one to seven constants (stack depth at most 5) take the forward-mode
    pass: one tangent for one constant, two for two, three for three, four in the
    two-plane stack for four, six in the three-plane stack for five or six, and two
    passes of four tangents for seven.  Sums of c_j f_j(x) over 150 points (three chunks), at cap 128 half of
    them padded past 64 tokens with (x - x) terms, shadow the oracle's MALA chains in
    every group."""
    rng = np.random.default_rng(29)
    n = 150
    x = rng.uniform(0.5, 2.0, n).astype(np.float32)
    y = (0.7 * x - 0.3 * x ** 2 + 0.5 / (x + 1) + 0.2 * np.exp(-x) + 0.1 * np.log(x)
         + 0.05 * x ** 3 - 0.02 * np.exp(x) + rng.normal(0.0, 0.05, n)).astype(np.float32)
    sd = np.full(n, 0.05, dtype=np.float32)
    terms = [["X"], ["X", "X", "MUL"], None, ["X", "NEG", "EXP"], ["X", "LOG"], ["X", "X", "X", "MUL", "MUL"],
             ["X", "EXP"]]
    nch, steps, eps = 70, 40, 2e-3
    O = np.zeros((cap, nch), dtype=np.uint8)
    V = np.zeros((cap, nch), dtype=np.float32)
    L = np.zeros(nch, dtype=np.int32)
    K = np.zeros(nch, dtype=np.int32)
    for c in range(nch):
        k = 1 + c % 7
        toks = []
        for j in range(k):
            cv = ("CONST", float(np.float32((0.7, -0.3, 0.5, 0.2, 0.1, 0.05, -0.02)[j] * (1 + 0.05 * rng.normal()))))
            if j == 2:
                toks += [cv, "X", "ONE", "ADD", "DIV"]          # c / (x + 1)
            else:
                toks += [cv] + terms[j] + ["MUL"]
            toks += ["ADD"] if j else []
        if cap == 128 and (c // 7) % 2 == 1:
            while len(toks) < 72:
                toks += ["X", "X", "SUB", "ADD"]
        o, v = oracle.tape_arrays(toks)
        O[:len(o), c] = o
        V[:len(o), c] = v
        L[c] = len(o)
        K[c] = k
    if cap == 128:
        assert (L > 64).sum() >= nch // 3
    cfg = oracle.make_cfg(cap=cap, proposal=1, eps=eps, fixed_structure=1, seed=31)
    o_ops, o_cs, o_ln = O.copy(), V.copy(), L.copy()
    _, _, cnt = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, 0, steps)
    with gpr.Context(cap=cap, seed=31, proposal=gpr.PROP_MALA, eps=eps, fixed_structure=1,
                     resample=gpr.RS_NONE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(O, V, L)
        say(f"mala forward-mode programs: cap {cap}, {nch} chains, 1-7 constants, lengths "
            f"{L.min()}..{L.max()}, {steps} steps")
        ctx.run_block(steps)
        g_cs = ctx.read_state()[1]
        st = ctx.stats()
    rel = np.max(np.abs(g_cs.astype(np.float64) - o_cs) / np.maximum(np.abs(o_cs.astype(np.float64)), 1e-3), axis=0)
    close = rel < 1e-3
    for k in range(1, 8):
        say(f"  {k} constants: {int(close[K == k].sum())} of {int((K == k).sum())} within 1e-3")
        assert close[K == k].sum() >= int(0.9 * (K == k).sum())
    say("acceptance gpu", st.const_accepted / st.const_proposed, "oracle", cnt[3] / max(cnt[2], 1))
    assert cnt[3] > 0.2 * cnt[2]
    assert abs(st.const_accepted - cnt[3]) <= 0.1 * cnt[3] + 5


def test_mala_deep_stack_programs_shadow_oracle(oracle, gpr):
    """Stack depth at the forward-mode pass's limits and one past them: every leaf pushed
    before the first operator, so a program of D leaves has depth D.  (1, 2 or 3
    constants, depth 9), (4, depth 8) and (5 or 6, depth 6) take one forward-mode pass at
    its deepest stack; (5, depth 7), (6, depth 8), (7, depth 5) and (8, depth 8) take two
    passes of four tangents; (3, depth 10), (4 or 5, depth 9), (8, depth 9) and (9, depth
    5) take the adjoint sweeps.  Each group fits its own dataset of 150 points (three
    chunks), made from its program at the starting constants' centre, and its chains
    shadow the oracle's."""
    rng = np.random.default_rng(53)
    n, per, steps, eps, cap = 150, 7, 40, 2e-3, 64
    groups = [(1, 9), (2, 9), (3, 9), (4, 8), (5, 6), (6, 6), (5, 7), (6, 8), (7, 5), (8, 8), (3, 10), (4, 9),
              (5, 9), (8, 9), (9, 5)]
    x = rng.uniform(0.5, 1.5, n).astype(np.float32)
    sd = np.full(n, 0.05, dtype=np.float32)

    def program(k, depth, cvals):
        # leaves: the first m constants folded by + and * in turn (at least one x leaf
        # follows), the other constants, then x and exp(-x) in turn; operators + and * in
        # turn from the innermost
        m = max(1, k - depth + 2)
        toks = [("CONST", float(cvals[0]))]
        for j in range(1, m):
            toks += [("CONST", float(cvals[j])), "ADD" if j % 2 else "MUL"]
        toks += [("CONST", float(cvals[j])) for j in range(m, k)]
        for j in range(depth - 1 - k + m):
            toks += ["X"] if j % 2 == 0 else ["X", "NEG", "EXP"]
        toks += ["ADD" if j % 2 == 0 else "MUL" for j in range(depth - 1)]
        return toks

    O = np.zeros((cap, per * len(groups)), dtype=np.uint8)
    V = np.zeros((cap, per * len(groups)), dtype=np.float32)
    L = np.zeros(per * len(groups), dtype=np.int32)
    data = []
    for g, (k, depth) in enumerate(groups):
        centre = rng.uniform(0.3, 0.8, k) * rng.choice([-1.0, 1.0], k)
        o, v = oracle.tape_arrays(program(k, depth, centre.astype(np.float32)))
        assert oracle.lib().gpo_tape_depth(oracle.u8(o), len(o)) == depth
        f = np.array([oracle.eval_tape(o, v, xi) for xi in x])
        data.append((x, (f + rng.normal(0.0, 0.05, n)).astype(np.float32), sd))
        for c in range(g * per, (g + 1) * per):
            o, v = oracle.tape_arrays(program(k, depth, (centre * (1 + 0.005 * rng.normal(size=k))).astype(np.float32)))
            O[:len(o), c] = o
            V[:len(o), c] = v
            L[c] = len(o)
    cfg = oracle.make_cfg(cap=cap, proposal=1, eps=eps, fixed_structure=1, seed=59)
    o_cs = V.copy()
    cnt = np.zeros(4, dtype=np.int64)
    for g in range(len(groups)):
        s = slice(g * per, (g + 1) * per)
        o_ops, cs_g, o_ln = O[:, s].copy(), V[:, s].copy(), L[s].copy()
        _, _, c_g = oracle.cfg_run_block(cfg, o_ops, cs_g, o_ln, *data[g], 0, steps, chain0=g * per)
        o_cs[:, s] = cs_g
        cnt += c_g
    with gpr.Context(cap=cap, seed=59, proposal=gpr.PROP_MALA, eps=eps, fixed_structure=1,
                     resample=gpr.RS_NONE) as ctx:
        ctx.set_data(data, [per] * len(groups))
        ctx.write_state(O, V, L)
        say(f"mala deep stacks: {len(groups)} groups (constants, depth) {groups}, {per} chains each, {steps} steps")
        ctx.run_block(steps)
        g_cs = ctx.read_state()[1]
        st = ctx.stats()
    rel = np.max(np.abs(g_cs.astype(np.float64) - o_cs) / np.maximum(np.abs(o_cs.astype(np.float64)), 1e-3), axis=0)
    close = rel < 1e-3
    for g, kd in enumerate(groups):
        s = slice(g * per, (g + 1) * per)
        say(f"  {kd}: {int(close[s].sum())} of {per} within 1e-3")
        assert close[s].sum() >= per - 1
    say("acceptance gpu", st.const_accepted / st.const_proposed, "oracle", cnt[3] / max(cnt[2], 1))
    assert cnt[3] > 0.2 * cnt[2]
    assert abs(st.const_accepted - cnt[3]) <= 0.1 * cnt[3] + 5


def test_mala_zero_tangent_meets_infinite_partial(oracle, gpr):
    """c * x^(1/2) with x = 0 among the data: at x = 0 the partial of x^(1/2) in x is
    infinite and its partial in the exponent is NaN, but neither operand carries a
    tangent of c, so the forward-mode gradient must stay finite (a zero tangent times any
    partial is zero) and the chains move as the oracle's do."""
    rng = np.random.default_rng(41)
    n = 80
    x = np.concatenate([[0.0, 0.0], rng.uniform(0.0, 2.0, n - 2)]).astype(np.float32)
    y = (1.7 * np.sqrt(x) + rng.normal(0.0, 0.05, n)).astype(np.float32)
    sd = np.full(n, 0.05, dtype=np.float32)
    nch, steps, eps, cap = 32, 40, 2e-3, 64
    O = np.zeros((cap, nch), dtype=np.uint8)
    V = np.zeros((cap, nch), dtype=np.float32)
    L = np.zeros(nch, dtype=np.int32)
    for c in range(nch):
        o, v = oracle.tape_arrays([("CONST", float(np.float32(1.7 * (1 + 0.05 * rng.normal())))), "X", "ONE", "ONE",
                                   "ONE", "ADD", "DIV", "POW", "MUL"])
        O[:len(o), c] = o
        V[:len(o), c] = v
        L[c] = len(o)
    cfg = oracle.make_cfg(cap=cap, proposal=1, eps=eps, fixed_structure=1, seed=43)
    o_ops, o_cs, o_ln = O.copy(), V.copy(), L.copy()
    _, _, cnt = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, 0, steps)
    with gpr.Context(cap=cap, seed=43, proposal=gpr.PROP_MALA, eps=eps, fixed_structure=1,
                     resample=gpr.RS_NONE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(O, V, L)
        say(f"mala zero tangent x infinite partial: {nch} chains, {steps} steps")
        ctx.run_block(steps)
        g_cs = ctx.read_state()[1]
        st = ctx.stats()
    rel = np.abs(g_cs[0].astype(np.float64) - o_cs[0]) / np.abs(o_cs[0].astype(np.float64))
    say("chains within 1e-3 of the oracle's:", int((rel < 1e-3).sum()), "of", nch, "acceptance gpu",
        st.const_accepted / st.const_proposed, "oracle", cnt[3] / max(cnt[2], 1))
    assert cnt[3] > 0.3 * cnt[2]
    assert (rel < 1e-3).sum() >= int(0.9 * nch)
    assert abs(st.const_accepted - cnt[3]) <= 0.1 * cnt[3] + 5

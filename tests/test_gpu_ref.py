"""Reference-semantics kernels (heap-31 programs, the reference's proposals and median
resampling) against the oracle's restatement of MH_simple_kernel (main.cu:679-813).

Every draw is bit-identical (Philox substreams, or the reference LCG seeded
block_seed + chain), so chains follow the oracle's trajectories: identical heaps and
constants, identical priors; likelihoods within tolerance (device vs host libm in the
programs' transcendentals).  A chain may leave the shared path only where the oracle's
trajectory met an accept decision within rounding of its threshold
(gpu_util.check_paths).
"""
import numpy as np
import pytest

from gpu_util import check_paths, kappa, ll_agree, say

pytestmark = pytest.mark.gpu


def data(name):
    import os
    if name == "S":
        from gpuropolis_amd import workloads as W
        return W.dataset_S(0)
    a = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "nist", name, "data.txt"))
    return a[:, 0].astype(np.float32), a[:, 1].astype(np.float32), a[:, 2].astype(np.float32)


def test_ref_init_matches_oracle(oracle, gpr):
    x, y, sd = data("S")
    hop, hc = oracle.ref_init(21, 3000)
    with gpr.Context(semantics=gpr.SEM_REFERENCE, seed=21) as ctx:
        ctx.set_data([(x, y, sd)], [3000])
        say("ref init parity")
        ctx.init_chains()
        g_hop, g_hc, _, _, _ = ctx.read_state()
    assert np.array_equal(g_hop, hop)
    assert np.array_equal(g_hc.view(np.uint32), hc.view(np.uint32))


# Kirby2 and Hahn1 (n = 151, 236) span several 64-point chunks, so the wave kernel's
# early rejection (the LL bound from the pre-drawn accept u) takes part
CASES = [("lane", 0, "S", 200), ("lane", 1, "S", 200), ("wave", 0, "S", 200), ("wave", 0, "BoxBOD", 300),
         ("wave", 0, "Thurber", 150), ("lane", 1, "Misra1a", 100), ("wave", 0, "Kirby2", 120),
         ("wave", 0, "Hahn1", 80)]


@pytest.mark.parametrize("sched,rng,dname,steps", CASES)
def test_ref_trajectories_match_oracle(oracle, gpr, sched, rng, dname, steps):
    x, y, sd = data(dname)
    nch = 64
    hop, hc = oracle.ref_init(5, nch)
    o_hop, o_hc = hop.copy(), hc.copy()
    cfg = oracle.GpoRefCfg(rng, 99)
    m0, m1 = np.zeros(nch), np.zeros(nch)
    p0, l0 = oracle.ref_run_block(cfg, o_hop, o_hc, x, y, sd, 0, steps, random_seed=4321, margin=m0)
    p1, l1 = oracle.ref_run_block(cfg, o_hop, o_hc, x, y, sd, steps, steps, random_seed=8765, margin=m1)
    sch = gpr.SCHED_LANE if sched == "lane" else gpr.SCHED_WAVE
    with gpr.Context(semantics=gpr.SEM_REFERENCE, seed=99, rng=rng, schedule=sch) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(hop, hc)
        say(f"ref trajectories sched={sched} rng={rng} data={dname} steps={steps}")
        ctx.run_block(steps, 4321)
        ctx.run_block(steps, 8765)
        g_hop, g_hc, _, g_pr, g_ll = ctx.read_state()
        st = ctx.stats()
    if sched == "wave" and len(x) > 64:
        # the lane kernel evaluates every point of every proposal; on the same chains the
        # wave kernel stops most proposals after their first chunk
        with gpr.Context(semantics=gpr.SEM_REFERENCE, seed=99, rng=rng, schedule=gpr.SCHED_LANE) as lctx:
            lctx.set_data([(x, y, sd)], [nch])
            lctx.write_state(hop, hc)
            lctx.run_block(steps, 4321)
            lctx.run_block(steps, 8765)
            lst = lctx.stats()
        say("evaluated token x points, wave / lane:", st.node_evals, lst.node_evals)
        assert st.node_evals < 0.8 * lst.node_evals
    same = np.all(g_hop == o_hop, axis=0) & np.all(g_hc.view(np.uint32) == o_hc.view(np.uint32), axis=0)
    check_paths(same, np.minimum(m0, m1), f"{sched} rng={rng} {dname}:")
    c = np.flatnonzero(same)
    assert np.array_equal(g_pr[c], p1[c])
    # programs whose LL is sensitive to one-ulp differences of the device libm get that
    # sensitivity (kappa) added to the tolerance
    kap = np.array([kappa(oracle, *oracle.ref_decode(o_hop[:, i].copy(), o_hc[:, i].copy())[:2], x, y, sd)
                    for i in c])
    good, strict = ll_agree(g_ll[c], l1[c], len(x), kap, sd)
    say("LL checks:", int(good.sum()), "of", len(c), "agree,", int(strict.sum()), "within 1e-5 of |LL|")
    assert good.all(), (g_ll[c][~good], l1[c][~good], kap[~good])


def test_ref_quantile_resample_matches_oracle(oracle, gpr):
    x, y, sd = data("S")
    nch = 2001
    hop, hc = oracle.ref_init(8, nch)
    with gpr.Context(semantics=gpr.SEM_REFERENCE, seed=8, resample=gpr.RS_QUANTILE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(hop, hc)
        ctx.run_block(20, 1)
        g_hop, g_hc, _, g_pr, g_ll = ctx.read_state()
        say("ref quantile resample")
        ctx.resample()
        r_hop, r_hc, _, _, _ = ctx.read_state()
    o_hop, o_hc = g_hop.copy(), g_hc.copy()
    cutoff = np.zeros(1)
    nre = oracle.lib().gpo_ref_resample(8, 1, nch, 0, oracle.u8(o_hop), oracle.f32(o_hc), oracle.f64(g_pr),
                                        oracle.f64(g_ll), oracle.f64(cutoff), 0)
    assert nre > 0
    assert np.array_equal(r_hop, o_hop)
    assert np.array_equal(r_hc.view(np.uint32), o_hc.view(np.uint32))


# ------------------------------------------------------------------------ op sets
@pytest.mark.parametrize("ops", [1, 2, 3, 4])
def test_ref_op_set_init_and_resample_match_oracle(oracle, gpr, ops):
    """Initial draws and the median rule's re-draws of each op set (main.cu:51-86) are
    the oracle's, bit for bit."""
    x, y, sd = data("S")
    nch = 1001
    hop, hc = oracle.ref_init(31, nch, ops=ops)
    with gpr.Context(semantics=gpr.SEM_REFERENCE, seed=31, ops=ops, resample=gpr.RS_QUANTILE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        say(f"op set {ops}: init and quantile resample")
        ctx.init_chains()
        i_hop, i_hc, _, _, _ = ctx.read_state()
        ctx.run_block(10, 1)
        g_hop, g_hc, _, g_pr, g_ll = ctx.read_state()
        ctx.resample()
        r_hop, r_hc, _, _, _ = ctx.read_state()
    assert np.array_equal(i_hop, hop)
    assert np.array_equal(i_hc.view(np.uint32), hc.view(np.uint32))
    o_hop, o_hc = g_hop.copy(), g_hc.copy()
    cutoff = np.zeros(1)
    nre = oracle.lib().gpo_ref_resample(31, 1, nch, 0, oracle.u8(o_hop), oracle.f32(o_hc), oracle.f64(g_pr),
                                        oracle.f64(g_ll), oracle.f64(cutoff), ops)
    assert nre > 0
    assert np.array_equal(r_hop, o_hop)
    assert np.array_equal(r_hc.view(np.uint32), o_hc.view(np.uint32))


@pytest.mark.parametrize("sched,ops,dname", [("wave", 2, "Misra1a"), ("wave", 4, "Misra1a"), ("lane", 4, "Misra1a"),
                                             ("wave", 3, "Misra1a"), ("lane", 1, "Misra1a"), ("wave", 4, "Kirby2")])
def test_ref_op_set_trajectories_match_oracle(oracle, gpr, sched, ops, dname):
    """The sampler with each op set follows the oracle's trajectories: moves draw the op
    set's codes, the prior uses its NOPS, and the physical constants and halving
    evaluate as main.cu:429-441 on both schedules (Kirby2, n = 151: with the wave
    kernel's early rejection)."""
    x, y, sd = data(dname)
    nch, steps = 64, 200
    hop, hc = oracle.ref_init(6, nch, ops=ops)
    o_hop, o_hc = hop.copy(), hc.copy()
    cfg = oracle.GpoRefCfg(0, 77, ops)
    m0, m1 = np.zeros(nch), np.zeros(nch)
    _, _ = oracle.ref_run_block(cfg, o_hop, o_hc, x, y, sd, 0, steps, margin=m0)
    p1, l1 = oracle.ref_run_block(cfg, o_hop, o_hc, x, y, sd, steps, steps, margin=m1)
    sch = gpr.SCHED_LANE if sched == "lane" else gpr.SCHED_WAVE
    with gpr.Context(semantics=gpr.SEM_REFERENCE, seed=77, ops=ops, schedule=sch) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(hop, hc)
        say(f"op set {ops} trajectories sched={sched} data={dname}")
        ctx.run_block(steps)
        ctx.run_block(steps)
        g_hop, g_hc, _, g_pr, g_ll = ctx.read_state()
    same = np.all(g_hop == o_hop, axis=0) & np.all(g_hc.view(np.uint32) == o_hc.view(np.uint32), axis=0)
    check_paths(same, np.minimum(m0, m1), f"op set {ops} {sched} {dname}:")
    c = np.flatnonzero(same)
    bad = c[g_pr[c] != p1[c]]
    for i in bad[:4]:
        h = o_hop[:, i].copy()
        say("prior mismatch chain", i, g_pr[i], p1[i], g_pr[i] - p1[i], "heap", list(h),
            "length", oracle.lib().gpo_ref_fold_length(oracle.u8(h), 1))
    assert len(bad) == 0
    kap = np.array([kappa(oracle, *oracle.ref_decode(o_hop[:, i].copy(), o_hc[:, i].copy())[:2], x, y, sd)
                    for i in c])
    ok, _ = ll_agree(g_ll[c], l1[c], len(x), kap, sd)
    assert ok.all()
    codes, _ = oracle.opset(ops)
    assert set(np.unique(g_hop)) <= set(codes)



def test_heap_arrays_in_the_references_types(oracle, gpr):
    """gpr_write_heap_i32 / gpr_read_heap_i32 take main.cu's own arrays (int op codes,
    main.cu:45, 1123-1130): the state they write is the uint8 state, a block run from it
    equals the block run from the same state written as uint8, and codes outside the op
    table are rejected."""
    x, y, sd = data("S")
    nch = 256
    hop, hc = oracle.ref_init(12, nch)
    out = []
    for form in ("u8", "i32"):
        with gpr.Context(semantics=gpr.SEM_REFERENCE, seed=12, schedule=gpr.SCHED_WAVE) as ctx:
            ctx.set_data([(x, y, sd)], [nch])
            if form == "u8":
                ctx.write_state(hop, hc)
            else:
                ctx.write_heap_i32(hop.astype(np.int32), hc)
                g_hop, g_hc, _, _, _ = ctx.read_state()
                assert np.array_equal(g_hop, hop) and np.array_equal(g_hc.view(np.uint32), hc.view(np.uint32))
                bad = hop.astype(np.int32)
                bad[5, 7] = gpr.ROP["KB"] + 1
                with pytest.raises(gpr.GprError, match="outside"):
                    ctx.write_heap_i32(bad, hc)
            say(f"heap arrays as {form}: one block")
            ctx.run_block(50, 7)
            out.append(ctx.read_heap_i32())
    for u, v in zip(out[0], out[1]):
        assert np.array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))
    P, Cs, pr, ll = out[1]
    assert P.dtype == np.int32 and P.min() >= 0 and P.max() < len(gpr.REF_OPS)

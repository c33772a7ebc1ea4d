"""The oracle's restatement of the reference sampler (MH_simple_kernel, main.cu:679-813)
pinned without a build of the reference.

The reference needs nvcc, so no golden trajectories of it exist.  Its moves are pinned
here in two ways.

* Closed forms.  An invalid current posterior forces every proposal through
  (main.cu:706-708, 730, 792), so with sd = 0 data (every log-likelihood NaN) the chain
  is the proposal itself:
  * the root structure proposal redraws all 31 slots from U{0..NOPS-1} of the op set;
  * the constant proposal is the mixture of main.cu:751-765: integers sgn*G - 1 (G
    geometric, P(G = k) = 0.9 * 0.1^(k-1)) with probability 0.15, else N(0, 1).
  With valid data, changes of a single leaf-level slot occur with frequency
  q * P(u < expf(a)), where q is the chance that a structure proposal makes exactly that
  change and a the acceptance of main.cu:727; changes of dead slots leave the posterior
  and the node count alone, so a = 0 and they are always accepted (quirk 7).  A program
  with no counted node (pure x) has forward = -log 0 = +inf and never moves (quirk 6).
* A second transcription.  `replay_chain` restates main.cu:683-806 in Python from the
  reference source, on primitives the other oracle tests pin (Philox and LCG streams,
  the spec logf / expf / sin 2 pi, the folds, prior and likelihood).  It keeps the
  reference's arithmetic types and the forward / backward terms carried across steps
  (quirk 1).  The oracle's C chains must equal it bit for bit.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import stats

import oracle_lib as O

HEAP = 31
F32 = np.float32
# the values of random_float (main.cu:230-232): float(k) / 1e6 in double, then float
GRID = (np.arange(1000000, dtype=np.float64) / 1e6).astype(F32)


def lib():
    return O.lib()


def data(valid=True):
    x = np.linspace(0.25, 2.0, 16).astype(F32)
    y = (1.5 * x + 0.3 + 0.05 * np.sin(7 * x)).astype(F32)
    sd = np.full(16, 1.0 if valid else 0.0, dtype=F32)
    return x, y, sd


def post_of(h, c, x, y, sd, ops=0):
    L = lib()
    h = np.ascontiguousarray(h, dtype=np.uint8)
    c = np.ascontiguousarray(c, dtype=F32)
    pr = L.gpo_ref_prior_ops(O.u8(h), O.f32(c), 1, ops)
    ll = L.gpo_ref_likelihood(O.u8(h), O.f32(c), 1, O.f32(x), O.f32(y), O.f32(sd), len(x))
    return pr, ll


# ------------------------------------------------------------------ closed forms
@pytest.mark.parametrize("ops", [0, 4])
def test_root_structure_proposal_is_uniform_over_the_op_set(ops):
    """Invalid current posterior: n = 0 (main.cu:706-708), every slot is a descendant of
    the root and is redrawn from U{0..NOPS-1} (main.cu:718), ABS included (quirk 9)."""
    codes, _ = O.opset(ops)
    nch = 60000
    hop, hc = O.ref_init(3, nch, ops=ops)
    x, y, sd = data(valid=False)
    O.ref_run_block(O.GpoRefCfg(0, 11, ops), hop, hc, x, y, sd, 0, 1)
    assert set(np.unique(hop)) == set(codes)
    k = len(codes)
    chi = 0.0
    for i in range(HEAP):
        cnt = np.array([(hop[i] == c).sum() for c in codes])
        chi += ((cnt - nch / k) ** 2 / (nch / k)).sum()
    dof = HEAP * (k - 1)
    assert stats.chi2.sf(chi, dof) > 1e-6, (chi, dof)
    # slots are drawn independently: pairs of neighbouring slots
    same = (hop[:-1] == hop[1:]).mean()
    assert abs(same - 1.0 / k) < 5 * math.sqrt((1.0 / k) * (1 - 1.0 / k) / (nch * (HEAP - 1)))


def test_constant_proposal_is_the_reference_mixture():
    """After step 1 (a constant step) of a forced-accept chain every constant is a fresh
    draw of main.cu:751-765: P(integer) = 0.15, integer values sgn*G - 1 with P(sgn = -1)
    = 1/2 and P(G = k) = 0.9 * 0.1^(k-1) on the random_float grid (so -1 never occurs,
    quirk 3), and N(0, 1) otherwise."""
    nch = 40000
    hop, hc = O.ref_init(5, nch)
    x, y, sd = data(valid=False)
    O.ref_run_block(O.GpoRefCfg(0, 21), hop, hc, x, y, sd, 0, 2)
    v = hc.ravel().astype(np.float64)
    fin = np.isfinite(v)
    # u = 0 of random_float (probability 1e-6 per draw) gives inf or NaN (quirk 10)
    assert (~fin).sum() < 20
    v = v[fin]
    n = len(v)
    isint = v == np.round(v)
    p = isint.mean()
    assert abs(p - 0.15) < 5 * math.sqrt(0.15 * 0.85 / n), p
    iv = v[isint]
    assert not np.any(iv == -1.0)
    for val, prob in ((0, 0.5 * 0.9), (1, 0.5 * 0.09), (2, 0.5 * 0.009), (-2, 0.5 * 0.9), (-3, 0.5 * 0.09),
                      (-4, 0.5 * 0.009)):
        got = (iv == val).mean()
        assert abs(got - prob) < 5 * math.sqrt(prob * (1 - prob) / len(iv)) + 1e-5, (val, got, prob)
    z = v[~isint]
    ks = stats.kstest(z, "norm")
    assert ks.pvalue > 1e-6, ks
    assert abs(z.mean()) < 5 / math.sqrt(len(z)) and abs(z.var() - 1) < 5 * math.sqrt(2 / len(z))


@pytest.mark.parametrize("fill", ["A", "B"])
def test_pure_x_program_never_moves(fill):
    """count_nodes of x is 0, forward = -log 0 = +inf (main.cu:702), so every acceptance
    of the launch is -inf or NaN: a valid chain at x never moves (quirk 6), and the stale
    forward term also blocks every constant move of the launch (quirk 1)."""
    nch = 256
    hop = np.full((HEAP, nch), O.ROP[fill], dtype=np.uint8)
    hc = np.random.default_rng(0).normal(0, 1, (HEAP, nch)).astype(F32)
    h0, c0 = hop.copy(), hc.copy()
    x, y, sd = data()
    prior, ll = O.ref_run_block(O.GpoRefCfg(0, 4), hop, hc, x, y, sd, 0, 50)
    assert np.array_equal(hop, h0) and np.array_equal(hc.view(np.uint32), c0.view(np.uint32))
    pr, l0 = post_of(h0[:, 0], c0[:, 0], x, y, sd)
    assert np.isfinite(pr + l0) and ll[0] == l0 and prior[0] == pr


def _count_heap():
    """Root chain A, A, A to heap slot 8 (1-based) = CPLUS (C = 0.5) over slot 16 = ONE:
    f = 1.5, two counted nodes.  Slot 16 is the one live leaf-level slot; turning it into
    A or B (f = x + 0.5) halves the node count, so the -log count terms move those
    acceptances by log 2."""
    R = O.ROP
    h = np.full(HEAP, R["ONE"], dtype=np.uint8)
    c = np.full(HEAP, 0.5, dtype=F32)
    for k, v in {1: "A", 2: "A", 4: "A", 8: "CPLUS", 16: "ONE"}.items():
        h[k - 1] = R[v]
    x = np.linspace(0.25, 2.0, 16).astype(F32)
    y = (1.5 + 0.05 * np.sin(7 * x)).astype(F32)
    return h, c, (x, y, np.full(16, 1.2, dtype=F32))


def test_single_slot_changes_follow_the_acceptance_law():
    """One structure step from a fixed valid state.  A change of leaf-level slot k to code
    v alone comes from a proposal at k (1/31 * 1/27) or at an ancestor whose other redrawn
    slots kept their codes (1/31 * 27^-(2^(d+1)-1)), and is kept with probability
    P(u < expf(a)), u on the grid {0, 1e-6, ..}, a = post' - post - log c' + log c
    (main.cu:702-730).  Observed counts over all 16 x 26 cells fit these expectations;
    dead slots (a = 0) are always kept."""
    L = lib()
    h0, c0, (x, y, sd) = _count_heap()
    pr0, ll0 = post_of(h0, c0, x, y, sd)
    p0 = pr0 + ll0
    cnt0 = L.gpo_ref_fold_count(O.u8(h0), 1)
    assert np.isfinite(p0) and cnt0 == 2
    nch = 1000000
    hop = np.repeat(h0[:, None], nch, axis=1)
    hc = np.repeat(c0[:, None], nch, axis=1)
    O.ref_run_block(O.GpoRefCfg(0, 99), hop, hc, x, y, sd, 0, 1)
    diff = hop != h0[:, None]
    one = diff.sum(axis=0) == 1
    slot = np.argmax(diff, axis=0)
    del diff
    q = sum(27.0 ** -(2 ** (d + 1) - 1) for d in range(5)) / 31
    chi, cells, dead_seen, low_obs, low_exp, halved = 0.0, 0, 0, 0, 0.0, 0
    for k in range(15, HEAP):
        for v in range(27):
            if v == h0[k]:
                continue
            h = h0.copy()
            h[k] = v
            pr, ll = post_of(h, c0, x, y, sd)
            cnt = L.gpo_ref_fold_count(O.u8(h), 1)
            fwd = -float(L.gpo_logf(float(cnt0)))
            bwd = -float(L.gpo_logf(float(cnt)))
            a = (pr + ll) - p0 + bwd - fwd
            if not np.isfinite(pr + ll):
                pk = 0.0
            elif a > 0:
                pk = 1.0
            else:
                t = F32(L.gpo_expf(float(F32(a))))
                pk = np.count_nonzero(GRID < t) / 1e6
            obs = int(np.count_nonzero(one & (slot == k) & (hop[k] == v)))
            exp = nch * q * pk
            if k != 15:
                assert a == 0.0 and pk == 1.0
                dead_seen += obs > 0
            elif cnt == 1:
                halved += 0.2 < pk < 0.8       # cells whose outcome the count term doubles
            if exp >= 20:
                assert abs(obs - exp) < 5 * math.sqrt(exp), (k, v, obs, exp)
            if exp >= 5:
                chi += (obs - exp) ** 2 / exp
                cells += 1
            else:
                low_obs += obs
                low_exp += exp
    chi += (low_obs - low_exp) ** 2 / max(low_exp, 1.0)
    assert halved == 2 and dead_seen >= 15 * 26 - 5
    assert cells > 300 and stats.chi2.sf(chi, cells) > 1e-6, (chi, cells)


# ------------------------------------------------------------------ second transcription
class _Src:
    """rand31 of the reference (main.cu:213-216): the LCG, or a Philox substream's word >> 1."""

    def __init__(self, lcg=None, sub=None):
        self.rx = lcg
        self.st = None
        if sub is not None:
            self.st = O.GpoStream()
            lib().gpo_stream_init(C.byref(self.st), *sub)

    def rand31(self):
        if self.st is not None:
            return lib().gpo_stream_next(C.byref(self.st)) >> 1
        return lib().gpo_lcg_next(C.byref(self.rx))


def _random_int(n, s):                                   # main.cu:218-228
    div = 0x7FFFFFFF // (n + 1)
    while True:
        r = s.rand31() // div
        if r < n:
            return r


def _random_float(s):                                    # main.cu:230-232
    return F32(float(F32(_random_int(1000000, s))) / 1000000.0)


def _random_normal(s):                                   # main.cu:237-241, spec logf / sin 2 pi
    u = _random_float(s)
    v = _random_float(s)
    return F32(np.sqrt(F32(-2.0) * F32(lib().gpo_logf(u)))) * F32(lib().gpo_sin2pi(v))


LOG01 = F32(math.log(float(F32(0.1))))                  # logf(PROPOSE_GEOM_R)
LOG015 = F32(math.log(float(F32(0.15))))                # logf(PROPOSE_INT)
LOG1M015 = math.log(1.0 - float(F32(0.15)))             # log(1.0 - PROPOSE_INT)
LOG1M01 = math.log(1.0 - float(F32(0.1)))
LOG2PI = F32(math.log(float(F32(2.0 * 3.141592653589))))


def _random_geometric(s):                                # main.cu:251-254
    return F32(np.ceil(F32(lib().gpo_logf(_random_float(s))) / LOG01))


def _lgeom(n):                                           # main.cu:307-309 at r = 0.1
    return float(F32(F32(n - F32(1.0)) * LOG01)) + LOG1M01


def _lnormal(x):                                         # main.cu:299-301 at s = 1
    return -float(F32(x * x)) / 2.0 - 0.5 * (float(LOG2PI) + 0.0)


def _is_int(x):                                          # main.cu:192
    return bool(np.isfinite(x)) and abs(float(F32(x - F32(np.floor(abs(x) + 0.5)) * np.sign(x)))) < 1e-6


def _logplusexp(a, b):                                   # main.cu:190, float logf / expf
    L = lib()
    m = max(a, b)
    return float(L.gpo_logf(F32(F32(L.gpo_expf(float(F32(a - m)))) + F32(L.gpo_expf(float(F32(b - m))))))) + m


def replay_chain(cfg, chain, step0, steps, random_seed, P, C_, x, y, sd):
    L = lib()
    nops = L.gpo_opset_nops(cfg.ops)
    P = P.copy()
    Cs = C_.copy()

    def evaluate():
        return post_of(P, Cs, x, y, sd, cfg.ops)

    cur_prior, cur_ll = evaluate()
    cur = cur_prior + cur_ll
    lcg = C.c_uint32((random_seed + chain) & 0xFFFFFFFF)
    fwd = bwd = 0.0

    def src(step, purpose, slot=0):
        return _Src(lcg=lcg) if cfg.rng == 1 else _Src(sub=(cfg.seed, chain, step, purpose, slot))

    for i in range(steps):
        step = step0 + i
        if i % 5 == 0:
            fwd = -float(L.gpo_logf(float(L.gpo_ref_fold_count(O.u8(P), 1))))
            s = src(step, O.P_STRUCT)
            n = 0 if not np.isfinite(cur) else _random_int(HEAP, s)
            old = P.copy()
            for k in range(n, HEAP):
                if L.gpo_ref_is_descendant(n, k):
                    P[k] = L.gpo_opset_code(cfg.ops, _random_int(nops, s))
            bwd = -float(L.gpo_logf(float(L.gpo_ref_fold_count(O.u8(P), 1))))
            restore = lambda: P.__setitem__(slice(None), old)  # noqa: E731
        else:
            oldc = Cs.copy()
            for c in range(HEAP):
                s = src(step, O.P_CPROP, c)
                if _random_float(s) < F32(1.0):
                    if _random_float(s) < F32(0.15):
                        sgn = F32(-1.0) if _random_float(s) < 0.5 else F32(1.0)
                        Cs[c] = F32(sgn * _random_geometric(s)) - F32(1.0)
                        fwd += (_lgeom(Cs[c] + F32(1.0)) + (0.0 if Cs[c] == 0.0 else math.log(2.0))) + float(LOG015)
                    else:
                        Cs[c] = F32(1.0) * _random_normal(s)
                        fwd += _lnormal(Cs[c]) + LOG1M015
                    if _is_int(oldc[c]):
                        bwd += _logplusexp((_lgeom(oldc[c] + F32(1.0)) + (0.0 if Cs[c] == 0.0 else math.log(2.0)))
                                           + float(LOG015), _lnormal(oldc[c]) + LOG1M015)
                    else:
                        bwd += _lnormal(oldc[c]) + LOG1M015
                else:
                    raise AssertionError("random_float < 1.0 always holds")
            restore = lambda: Cs.__setitem__(slice(None), oldc)  # noqa: E731
        pp, pl = evaluate()
        ppost = pp + pl
        acc = ppost - cur + bwd - fwd
        s = src(step, O.P_ACCEPT)
        if (np.isfinite(ppost) and (acc > 0.0 or _random_float(s) < F32(L.gpo_expf(float(F32(acc)))))) \
                or not np.isfinite(cur):
            cur, cur_prior, cur_ll = ppost, pp, pl
        else:
            restore()
    return P, Cs, cur_prior, cur_ll


@pytest.mark.parametrize("rng,ops", [(0, 0), (1, 0), (0, 4)])
def test_second_transcription_reproduces_the_oracle(rng, ops):
    """Eleven steps (three structure steps, eight constant steps, forward / backward
    carried between them) of 160 chains: identical programs, constants, priors and
    likelihoods.  A fifth of the initial constants are small integers, so both backward
    forms of main.cu:768-775 occur."""
    L = lib()
    nch, steps, rseed = 160, 11, 4321
    hop, hc = O.ref_init(7, nch, ops=ops)
    rng_ = np.random.default_rng(1)
    ints = rng_.random(hc.shape) < 0.2
    hc[ints] = rng_.integers(-3, 4, ints.sum()).astype(F32)
    x, y, sd = data()
    h0, c0 = hop.copy(), hc.copy()
    cfg = O.GpoRefCfg(rng, 17, ops)
    prior, ll = O.ref_run_block(cfg, hop, hc, x, y, sd, 5, steps, random_seed=rseed)
    moved = 0
    for ch in range(nch):
        P, Cs, pr, l = replay_chain(cfg, ch, 5, steps, rseed, h0[:, ch], c0[:, ch], x, y, sd)
        assert np.array_equal(P, hop[:, ch]), ch
        assert np.array_equal(Cs.view(np.uint32), hc[:, ch].view(np.uint32)), ch
        assert pr == prior[ch] and (l == ll[ch] or (math.isnan(l) and math.isnan(ll[ch]))), ch
        moved += not (np.array_equal(P, h0[:, ch]) and np.array_equal(Cs, c0[:, ch]))
    assert moved > nch // 4
    assert L.gpo_opset_nops(ops) > 0

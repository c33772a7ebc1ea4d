"""Pins of the oracle's program layer (CPU only).

heap fold (program_fold, main.cu:371-393)  vs an independent recursive evaluation
decode heap -> postfix                      vs the fold, bit-exact, all 40 heap codes
length / count (main.cu:477-608)            vs closed forms on the decoded tape
is_descendant (main.cu:660-677)             vs brute-force subtree enumeration
NIST StRD certified models                  residual sums of squares of the certified fits
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from nist_models import MODELS

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ARITH = ["ONE", "CONST", "INV", "A", "B", "PLUS", "MINUS", "RMINUS", "CPLUS", "TIMES", "CTIMES", "DIV", "RDIV",
         "SQRT", "ABS"]


def rec_eval(hop, hc, i, x):
    """Recursive evaluation of heap slot i (1-based) with IEEE float32 arithmetic."""
    f = np.float32
    if i > 31:
        return f(x)
    o = O.REF_OPS[hop[i - 1]]
    C = f(hc[i - 1])
    a = lambda: rec_eval(hop, hc, 2 * i, x)       # noqa: E731
    b = lambda: rec_eval(hop, hc, 2 * i + 1, x)   # noqa: E731
    with np.errstate(all="ignore"):
        return {
            "ONE": lambda: f(1), "CONST": lambda: C, "INV": lambda: f(1) / a(), "A": a, "B": b,
            "PLUS": lambda: a() + b(), "MINUS": lambda: a() - b(), "RMINUS": lambda: b() - a(),
            "CPLUS": lambda: a() + C, "TIMES": lambda: a() * b(), "CTIMES": lambda: C * a(),
            "DIV": lambda: a() / b(), "RDIV": lambda: b() / a(), "SQRT": lambda: np.sqrt(a()),
            "ABS": lambda: np.abs(a()),
        }[o]()


def random_heap(rng, ops):
    idx = np.array([O.ROP[o] for o in ops])
    hop = rng.choice(idx, 31).astype(np.uint8)
    hc = (rng.normal(0, 2, 31)).astype(np.float32)
    return hop, hc


def test_fold_matches_recursive_evaluation():
    rng = np.random.default_rng(0)
    L = O.lib()
    for _ in range(400):
        hop, hc = random_heap(rng, ARITH)
        for x in (0.37, -2.5, 3.0):
            got = L.gpo_ref_fold_eval(O.u8(hop), O.f32(hc), 1, x)
            ref = rec_eval(hop, hc, 1, x)
            assert (np.isnan(got) and np.isnan(ref)) or np.float32(got) == ref


def test_decode_matches_fold_all_ops():
    rng = np.random.default_rng(1)
    L = O.lib()
    for _ in range(1000):
        hop, hc = random_heap(rng, O.REF_OPS)
        ops, cs, slot = O.ref_decode(hop, hc)
        assert O.lib().gpo_tape_depth(O.u8(ops), len(ops)) <= 6
        for x in (0.7, -1.3):
            a = L.gpo_ref_fold_eval(O.u8(hop), O.f32(hc), 1, x)
            b = O.eval_tape(ops, cs, x)
            assert (np.isnan(a) and np.isnan(b)) or np.float32(a) == np.float32(b)


def test_length_and_count_closed_forms():
    rng = np.random.default_rng(2)
    L = O.lib()
    # constant-operand heap ops cost 2 (CONST_LENGTH_PRIOR + 1); halving decodes to a c* token
    # but costs 1 + a, so C-op tokens are counted by their heap op
    cops = {O.ROP[n] for n in ("CPLUS", "CTIMES", "CPOW", "CRPOW")}
    for _ in range(2000):
        hop, hc = random_heap(rng, O.REF_OPS)
        ops, _, slot = O.ref_decode(hop, hc)
        n_tok, n_x = len(ops), int(np.sum(ops == O.OP["X"]))
        n_c = int(sum(int(hop[s]) in cops for s in slot if s != 255))
        assert L.gpo_ref_fold_length(O.u8(hop), 1) == n_tok + n_c
        assert L.gpo_ref_fold_count(O.u8(hop), 1) == n_tok - n_x


def test_count_zero_for_pure_x():
    hop = np.full(31, O.ROP["A"], dtype=np.uint8)     # every slot passes its left child: f = x
    assert O.lib().gpo_ref_fold_count(O.u8(hop), 1) == 0
    assert O.lib().gpo_ref_fold_length(O.u8(hop), 1) == 1.0


def test_degree_of_polynomials():
    L = O.lib()
    hop = np.full(31, O.ROP["A"], dtype=np.uint8)
    hc = np.zeros(31, dtype=np.float32)
    hop[0] = O.ROP["TIMES"]                            # x * x
    assert L.gpo_ref_fold_degree(O.u8(hop), O.f32(hc), 1) == 2.0
    hop[0] = O.ROP["PLUS"]                             # x + x
    assert L.gpo_ref_fold_degree(O.u8(hop), O.f32(hc), 1) == 1.0
    hop[0] = O.ROP["EXP"]
    assert L.gpo_ref_fold_degree(O.u8(hop), O.f32(hc), 1) == np.inf


def test_is_descendant_brute_force():
    L = O.lib()
    for n in range(31):
        sub, frontier = {n + 1}, [n + 1]
        while frontier:
            j = frontier.pop()
            for c in (2 * j, 2 * j + 1):
                if c <= 31:
                    sub.add(c)
                    frontier.append(c)
        for k in range(31):
            assert L.gpo_ref_is_descendant(n, k) == ((k + 1) in sub)


def test_tape_depth_and_subtree():
    ops, _ = O.tape_arrays(["X", ("CONST", 1.0), "ADD", "X", "MUL", "EXP"])
    L = O.lib()
    assert L.gpo_tape_depth(O.u8(ops), len(ops)) == 2
    assert L.gpo_subtree_start(O.u8(ops), 5) == 0
    assert L.gpo_subtree_start(O.u8(ops), 2) == 0
    assert L.gpo_subtree_start(O.u8(ops), 3) == 3
    bad, _ = O.tape_arrays(["X", "ADD"])
    assert L.gpo_tape_depth(O.u8(bad), 2) == -1


# --------------------------------------------------------------- NIST certified fits
def nist(name):
    a = np.loadtxt(os.path.join(GOLDEN, "nist", name, "data.txt"))
    return a[:, 0], a[:, 1], a[:, 2]


@pytest.mark.parametrize("name", sorted(MODELS))
def test_nist_certified_rss(name):
    cert = json.load(open(os.path.join(GOLDEN, "nist_certified.json")))[name]
    b = [cert["params"][f"b{i + 1}"] for i in range(len(cert["params"]))]
    # certified data from the .dat file (the reference's data.txt copies are rounded or
    # edited for some datasets); sd = the certified residual SD, as data.txt carries it
    x, y = np.array(cert["x"]), np.array(cert["y"])
    sd = np.full(len(x), cert["residual_sd"])
    f, tape = MODELS[name]
    assert len(x) == cert["n"]
    assert np.allclose(nist(name)[2], cert["residual_sd"], rtol=3e-3)   # data.txt sd column, 3 digits
    # the NIST formula with the certified parameters reproduces the certified RSS
    rss_exact = np.sum((y - f(b, x)) ** 2)
    assert abs(rss_exact - cert["rss"]) <= 1e-7 * cert["rss"], (rss_exact, cert["rss"])
    # the oracle's tape evaluator + Gaussian LL, on float32 data and constants, against the
    # same formula evaluated in FP64 on identically rounded inputs
    if tape is None:
        from gpuropolis_amd.workloads import rational_tape
        ops, cs = rational_tape(np.array(b))
    else:
        ops, cs = O.tape_arrays(tape(b))
    x32, y32, sd32 = (v.astype(np.float32) for v in (x, y, sd))
    bf = [float(np.float32(v)) for v in b]
    with np.errstate(all="ignore"):
        fr = f(bf, x32.astype(np.float64))
    s = float(sd32[0])
    rss_ref = np.sum((y32.astype(np.float64) - fr) ** 2)
    ll = O.tape_loglik(ops, cs, x32, y32, sd32, mode=1)
    rss_oracle = -2 * s * s * (ll + len(x) * np.log(s) + 0.5 * len(x) * np.log(2 * np.pi))
    assert abs(rss_oracle - rss_ref) <= 1e-6 * rss_ref + 1e-9, (rss_oracle, rss_ref)


def test_reference_likelihood_matches_tape_likelihood():
    """compute_likelihood (main.cu:454-468, fold + lnormalpdf) vs decode + tape LL."""
    rng = np.random.default_rng(7)
    L = O.lib()
    x = rng.uniform(-2, 2, 40).astype(np.float32)
    y = np.sin(x).astype(np.float32)
    sd = np.full(40, 0.3, dtype=np.float32)
    n_valid = 0
    for _ in range(500):
        hop, hc = random_heap(rng, O.REF_OPS)
        a = L.gpo_ref_likelihood(O.u8(hop), O.f32(hc), 1, O.f32(x), O.f32(y), O.f32(sd), 40)
        ops, cs, _ = O.ref_decode(hop, hc)
        b = O.tape_loglik(ops, cs, x, y, sd, 0)
        if not np.isfinite(a) and np.isfinite(b):
            # quirk of lnormalpdf (main.cu:300): (f-y)^2 is squared in float and overflows
            # to inf for |f-y| > 1.8e19, where the FP64 tape likelihood stays finite
            assert b < -1e30
            continue
        assert np.isfinite(a) == np.isfinite(b)
        if np.isfinite(a):
            n_valid += 1
            assert abs(a - b) <= 1e-6 * max(abs(b), 40)
    assert n_valid > 50

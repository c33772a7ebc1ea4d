"""The evaluation stack bound GPR_MAX_DEPTH = 16 restricts the config sampler's target:
programs whose postfix stack exceeds 16 entries get no posterior mass (the proposals
that would reach them are rejected, in the oracle as on the device).  This bounds what
the restriction removes: under the description-length prior (weight nsym^-len with
nsym = 10; a constant's Cauchy density integrates to 1), the programs of at most 128
tokens deeper than 16 hold 2e-7 of the prior mass.

The mass per length and depth bound comes from a dynamic program over stack heights
(leaves x and const push, the 3 unary ops keep the height, the 5 binary ops pop one),
pinned against brute-force enumeration with the oracle's own depth and prior functions
on every program of up to 5 tokens."""
import itertools
import math

import numpy as np

import oracle_lib as O

LEAVES = ["X", "CONST"]
UNARY = ["EXP", "LOG", "NEG"]
BINARY = ["ADD", "SUB", "MUL", "DIV", "POW"]
SYMS = LEAVES + UNARY + BINARY


def dp_mass(max_depth, cap, nsym=10):
    """Prior mass (without the constants' densities) of programs of each length 1..cap
    whose stack height never exceeds max_depth."""
    w = 1.0 / nsym
    f = {0: 1.0}
    out = []
    for _ in range(cap):
        g = {}
        for h, v in f.items():
            if h + 1 <= max_depth:
                g[h + 1] = g.get(h + 1, 0.0) + v * len(LEAVES) * w
            if h >= 1:
                g[h] = g.get(h, 0.0) + v * len(UNARY) * w
            if h >= 2:
                g[h - 1] = g.get(h - 1, 0.0) + v * len(BINARY) * w
        f = g
        out.append(f.get(1, 0.0))
    return np.array(out)


def test_dp_matches_enumeration_with_the_oracle():
    lib = O.lib()
    # the oracle's Cauchy term of a constant valued 0 (float logs), taken out below
    c0 = O.tape_prior(np.array([O.OP["CONST"]], np.uint8), np.zeros(1, np.float32), 10) - \
        O.tape_prior(np.array([O.OP["X"]], np.uint8), np.zeros(1, np.float32), 10)
    for L in range(1, 6):
        tot = {1: 0.0, 2: 0.0, 3: 0.0, 16: 0.0}
        for toks in itertools.product(SYMS, repeat=L):
            ops = np.array([O.OP[t] for t in toks], dtype=np.uint8)
            depth = lib.gpo_tape_depth(O.u8(ops), L)
            if depth < 0:                    # malformed postfix
                continue
            # the description-length part of the oracle's prior
            lp = O.tape_prior(ops, np.zeros(L, dtype=np.float32), 10)
            wgt = math.exp(lp - int(np.sum(ops == O.OP["CONST"])) * c0)
            for D in tot:
                if depth <= D:
                    tot[D] += wgt
        for D, t in tot.items():
            assert abs(t - dp_mass(D, L)[L - 1]) <= 1e-12 * max(t, 1e-300), (L, D, t)


def test_depth_bound_removes_negligible_prior_mass():
    full = dp_mass(10 ** 6, 128).sum()
    bounded = dp_mass(16, 128).sum()
    frac = 1.0 - bounded / full
    assert 0.0 < frac < 1e-6, frac
    # the same bound for the full form's programs at the depths that occur in practice
    assert 1.0 - dp_mass(8, 128).sum() / full < 1e-3

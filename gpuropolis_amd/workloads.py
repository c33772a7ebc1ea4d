"""Seeded synthetic workloads for the five benchmark configurations.

S  one dataset, n=32, x~U[0,10], y = 2.5 exp(-0.3x) + 1 + N(0, 0.05^2), sd 0.05;
   1024 chains, cap 32, 10 blocks x 1000 steps, seed 0.
L  one dataset, n=250, x~U[-5,5], y = random target (<= 12 nodes, constants
   N(0, 2^2)) + noise of sd 1% of the y range; 131072 chains, cap 64, 20 x 1000.
B  64 datasets, n log-uniform in [6, 250], targets <= 16 nodes + 1% noise; 4096
   chains each, caps 32/64/128 round-robin; 20 x 1000.
D  rational (b1+b2x+b3x^2+b4x^3)/(1+b5x+b6x^2+b7x^3), n=37, x~U[-3,2], b~N(0,1)
   with denominator roots outside [-3,2], noise sd 0.01; 16384 chains.
F  256 datasets, n log-uniform in [16, 1024], x~U[-10,10], targets <= 24 nodes +
   1% noise; 1048576 chains, cap 128.

Targets are postfix programs over the config grammar {+,-,*,/,pow,exp,log,neg,x,c},
evaluated in FP64 by numpy.
"""
import math

import numpy as np

from .lib import OP

BIN = ["ADD", "SUB", "MUL", "DIV", "POW"]
UN = ["EXP", "LOG", "NEG"]


def eval_postfix64(ops, cs, x):
    x = np.asarray(x, dtype=np.float64)
    st = []
    with np.errstate(all="ignore"):
        for o, c in zip(ops, cs):
            o = int(o)
            if o == OP["X"]:
                st.append(x)
            elif o == OP["CONST"]:
                st.append(np.full_like(x, float(c)))
            elif o == OP["ONE"]:
                st.append(np.ones_like(x))
            elif o in (OP["ADD"], OP["SUB"], OP["MUL"], OP["DIV"], OP["POW"]):
                b = st.pop()
                a = st.pop()
                st.append({OP["ADD"]: a + b, OP["SUB"]: a - b, OP["MUL"]: a * b, OP["DIV"]: a / b,
                           OP["POW"]: np.power(a, b)}[o])
            elif o == OP["EXP"]:
                st.append(np.exp(st.pop()))
            elif o == OP["LOG"]:
                st.append(np.log(st.pop()))
            elif o == OP["NEG"]:
                st.append(-st.pop())
            else:
                raise ValueError(o)
    return st[0]


def random_target(rng, max_nodes, x, const_sd=2.0, min_nodes=5):
    """Random postfix target over the config grammar: finite on x, not constant."""
    for _ in range(10000):
        n_target = int(rng.integers(min_nodes, max_nodes + 1))
        # grow a random tree of about n_target nodes in prefix order, then reverse
        ops, cs = [], []
        need = 1
        while need > 0:
            room = n_target - len(ops)
            if room <= need:
                kind = "leaf"
            else:
                kind = rng.choice(["leaf", "bin", "un"], p=[0.3, 0.5, 0.2])
            if kind == "leaf":
                if rng.random() < 0.5:
                    ops.append(OP["X"]); cs.append(0.0)
                else:
                    ops.append(OP["CONST"]); cs.append(float(np.float32(rng.normal(0.0, const_sd))))
                need -= 1
            elif kind == "bin":
                ops.append(OP[str(rng.choice(BIN))]); cs.append(0.0)
                need += 1
            else:
                ops.append(OP[str(rng.choice(UN))]); cs.append(0.0)
        ops = ops[::-1]
        cs = cs[::-1]
        if len(ops) > max_nodes or len(ops) < min_nodes:
            continue
        y = eval_postfix64(ops, cs, x)
        if not np.all(np.isfinite(y)):
            continue
        if np.ptp(y) < 1e-3 or np.max(np.abs(y)) > 1e6:
            continue
        if OP["X"] not in ops:
            continue
        return np.array(ops, dtype=np.uint8), np.array(cs, dtype=np.float32), y
    raise RuntimeError("no finite target found")


def dataset_S(seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 10.0, 32)
    y = 2.5 * np.exp(-0.3 * x) + 1.0 + rng.normal(0.0, 0.05, 32)
    sd = np.full(32, 0.05)
    return x.astype(np.float32), y.astype(np.float32), sd.astype(np.float32)


def dataset_random(rng, n, max_nodes, lo, hi):
    x = rng.uniform(lo, hi, n)
    ops, cs, y = random_target(rng, max_nodes, x)
    noise = 0.01 * float(np.ptp(y))
    y = y + rng.normal(0.0, noise, n)
    sd = np.full(n, noise)
    return (x.astype(np.float32), y.astype(np.float32), sd.astype(np.float32)), (ops, cs)


def dataset_L(seed=0):
    rng = np.random.default_rng(seed + 1)
    return dataset_random(rng, 250, 12, -5.0, 5.0)


def datasets_B(seed=0, nd=64):
    rng = np.random.default_rng(seed + 2)
    out, targets = [], []
    for d in range(nd):
        n = int(round(math.exp(rng.uniform(math.log(6), math.log(250)))))
        ds, tg = dataset_random(rng, n, 16, -5.0, 5.0)
        out.append(ds)
        targets.append(tg)
    caps = [(32, 64, 128)[d % 3] for d in range(nd)]
    return out, caps, targets


def datasets_F(seed=0, nd=256):
    rng = np.random.default_rng(seed + 3)
    out, targets = [], []
    for d in range(nd):
        n = int(round(math.exp(rng.uniform(math.log(16), math.log(1024)))))
        ds, tg = dataset_random(rng, n, 24, -10.0, 10.0)
        out.append(ds)
        targets.append(tg)
    return out, targets


RATIONAL_TOKENS = None


def rational_tape(b):
    """(b1 + b2 x + b3 x^2 + b4 x^3) / (1 + b5 x + b6 x^2 + b7 x^3) as a postfix tape (Horner)."""
    def c(v):
        return ("CONST", float(np.float32(v)))
    # numerator: ((b4 x + b3) x + b2) x + b1
    num = [c(b[3]), "X", "MUL", c(b[2]), "ADD", "X", "MUL", c(b[1]), "ADD", "X", "MUL", c(b[0]), "ADD"]
    # denominator: ((b7 x + b6) x + b5) x + 1
    den = [c(b[6]), "X", "MUL", c(b[5]), "ADD", "X", "MUL", c(b[4]), "ADD", "X", "MUL", "ONE", "ADD"]
    toks = num + den + ["DIV"]
    ops = np.array([OP[t[0]] if isinstance(t, tuple) else OP[t] for t in toks], dtype=np.uint8)
    cs = np.array([t[1] if isinstance(t, tuple) else 0.0 for t in toks], dtype=np.float32)
    return ops, cs


def dataset_D(seed=0):
    rng = np.random.default_rng(seed + 4)
    while True:
        b = rng.normal(0.0, 1.0, 7)
        roots = np.roots([b[6], b[5], b[4], 1.0])
        real = roots[np.abs(roots.imag) < 1e-9].real
        if np.all((real < -3.0) | (real > 2.0)):
            break
    x = rng.uniform(-3.0, 2.0, 37)
    f = (b[0] + b[1] * x + b[2] * x ** 2 + b[3] * x ** 3) / (1 + b[4] * x + b[5] * x ** 2 + b[6] * x ** 3)
    y = f + rng.normal(0.0, 0.01, 37)
    sd = np.full(37, 0.01)
    return (x.astype(np.float32), y.astype(np.float32), sd.astype(np.float32)), b.astype(np.float32)

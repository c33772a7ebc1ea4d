"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests and the
CPU-baseline leg of bench.py.  The oracle is the checker, never the product."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

# unified op codes (oracle.h / include/gpuropolis.h)
OPS = ["X", "CONST", "ONE", "ADD", "SUB", "MUL", "DIV", "POW", "RSUB", "RDIV", "RPOW", "NEG", "EXP",
       "LOG", "INV", "SQRT", "SIN", "ASIN", "COS", "ACOS", "TAN", "ATAN", "ABS", "CADD", "CMUL",
       "CPOW", "CRPOW"]
OP = {n: i for i, n in enumerate(OPS)}
# reference heap op codes, main.cu:53
REF_OPS = ["ONE", "CONST", "INV", "A", "B", "PLUS", "MINUS", "RMINUS", "CPLUS", "TIMES", "CTIMES", "DIV",
           "RDIV", "LOG", "EXP", "POW", "CPOW", "RPOW", "CRPOW", "SQRT", "SIN", "ASIN", "COS", "ACOS",
           "TAN", "ATAN", "ABS",
           # the further ops of the PHYSICAL_CONSTANTS op set (main.cu:82)
           "LOGOF2", "HALF", "E", "PI", "CLIGHT", "G", "HBAR", "MU0", "EL", "MP", "ME", "E0", "KB"]
ROP = {n: i for i, n in enumerate(REF_OPS)}
P_STRUCT, P_CPROP, P_ACCEPT, P_INIT, P_RESAMPLE = 1, 2, 3, 4, 5

_u8p = C.POINTER(C.c_uint8)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)
_i64p = C.POINTER(C.c_long)


class GpoStream(C.Structure):
    _fields_ = [("key", C.c_uint32 * 2), ("ctr", C.c_uint32 * 4), ("buf", C.c_uint32 * 4), ("idx", C.c_int)]


class GpoCfg(C.Structure):
    _fields_ = [("cap", C.c_int), ("dmax", C.c_int), ("nsym", C.c_int), ("proposal", C.c_int),
                ("sigma", C.c_float), ("eps", C.c_float), ("struct_every", C.c_int),
                ("fixed_structure", C.c_int), ("seed", C.c_uint64), ("mk", C.c_int), ("mrows", C.c_int * 16),
                ("mL", C.c_float * 256)]


class GpoRefCfg(C.Structure):
    _fields_ = [("rng", C.c_int), ("seed", C.c_uint64), ("ops", C.c_int)]


class GpoChainStats(C.Structure):
    _fields_ = [("prior", C.c_double), ("ll", C.c_double), ("n_struct_prop", C.c_long),
                ("n_struct_acc", C.c_long), ("n_const_prop", C.c_long), ("n_const_acc", C.c_long),
                ("min_margin", C.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR], stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    sig = {
        "gpo_philox4x32_10": (None, [_u32p, _u32p, _u32p]),
        "gpo_stream_init": (None, [C.POINTER(GpoStream), C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
        "gpo_stream_next": (C.c_uint32, [C.POINTER(GpoStream)]),
        "gpo_lcg_next": (C.c_uint32, [_u32p]),
        "gpo_logf": (C.c_float, [C.c_float]),
        "gpo_expf": (C.c_float, [C.c_float]),
        "gpo_sin2pi": (C.c_float, [C.c_float]),
        "gpo_normal_icdf": (C.c_float, [C.c_float]),
        "gpo_u01": (C.c_float, [C.c_uint32]),
        "gpo_ref_fold_eval": (C.c_float, [_u8p, _f32p, C.c_size_t, C.c_float]),
        "gpo_ref_fold_length": (C.c_float, [_u8p, C.c_size_t]),
        "gpo_ref_fold_count": (C.c_int, [_u8p, C.c_size_t]),
        "gpo_ref_fold_degree": (C.c_float, [_u8p, _f32p, C.c_size_t]),
        "gpo_ref_is_descendant": (C.c_int, [C.c_int, C.c_int]),
        "gpo_ref_decode": (C.c_int, [_u8p, _f32p, C.c_size_t, _u8p, _f32p, _u8p]),
        "gpo_ref_lnormalpdf": (C.c_double, [C.c_float, C.c_float]),
        "gpo_ref_lcauchypdf": (C.c_double, [C.c_float, C.c_float]),
        "gpo_ref_lgeometricpdf": (C.c_double, [C.c_float, C.c_float]),
        "gpo_ref_likelihood": (C.c_double, [_u8p, _f32p, C.c_size_t, _f32p, _f32p, _f32p, C.c_int]),
        "gpo_ref_prior": (C.c_double, [_u8p, _f32p, C.c_size_t]),
        "gpo_arity": (C.c_int, [C.c_int]),
        "gpo_dl_cost": (C.c_int, [C.c_int]),
        "gpo_tape_depth": (C.c_int, [_u8p, C.c_int]),
        "gpo_subtree_start": (C.c_int, [_u8p, C.c_int]),
        "gpo_eval_tape_f32": (C.c_float, [_u8p, _f32p, C.c_int, C.c_float]),
        "gpo_eval_tape_f64": (C.c_double, [_u8p, _f32p, C.c_int, C.c_double]),
        "gpo_tape_loglik": (C.c_double, [_u8p, _f32p, C.c_int, _f32p, _f32p, _f32p, C.c_int, C.c_int]),
        "gpo_tape_prior": (C.c_double, [_u8p, _f32p, C.c_int, C.c_double, C.c_int]),
        "gpo_tape_grad": (C.c_int, [_u8p, _f32p, C.c_int, _f32p, _f32p, _f32p, C.c_int, _f64p, _f64p]),
        "gpo_tape_logpost_f64": (C.c_double, [_u8p, _f64p, C.c_int, _f32p, _f32p, _f32p, C.c_int]),
        "gpo_cfg_gen_table": (None, [_u8p, _f64p]),
        "gpo_cfg_init_chain": (C.c_int, [C.POINTER(GpoCfg), C.c_uint32, _u8p, _f32p]),
        "gpo_cfg_run_chain": (None, [C.POINTER(GpoCfg), C.c_uint32, C.c_uint64, C.c_int, _u8p, _f32p,
                                     C.POINTER(C.c_int), _f32p, _f32p, _f32p, C.c_int,
                                     C.POINTER(GpoChainStats)]),
        "gpo_cfg_run_block": (None, [C.POINTER(GpoCfg), C.c_int, C.c_uint32, C.c_uint64, C.c_int, _u8p, _f32p,
                                     _i32p, _f32p, _f32p, _f32p, C.c_int, _f64p, _f64p, _i64p, C.c_int]),
        "gpo_cfg_run_block_m": (None, [C.POINTER(GpoCfg), C.c_int, C.c_uint32, C.c_uint64, C.c_int, _u8p, _f32p,
                                       _i32p, _f32p, _f32p, _f32p, C.c_int, _f64p, _f64p, _i64p, C.c_int, _f64p]),
        "gpo_ref_run_block_m": (None, [C.POINTER(GpoRefCfg), C.c_int, C.c_uint32, C.c_uint64, C.c_int, C.c_int32,
                                       _u8p, _f32p, _f32p, _f32p, _f32p, C.c_int, _f64p, _f64p, C.c_int, _f64p]),
        "gpo_cfg_resample": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, C.c_double,
                                       _u8p, _f32p, _i32p, _f64p, _f64p, _i32p]),
        "gpo_ref_run_block": (None, [C.POINTER(GpoRefCfg), C.c_int, C.c_uint32, C.c_uint64, C.c_int, C.c_int32,
                                     _u8p, _f32p, _f32p, _f32p, _f32p, C.c_int, _f64p, _f64p, C.c_int]),
        "gpo_ref_draws": (None, [C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_int, C.c_int, _f64p]),
        "gpo_ref_stream_draws": (None, [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_int, C.c_int, _f64p]),
        "gpo_ref_init": (None, [C.c_uint64, C.c_int, C.c_uint32, _u8p, _f32p, C.c_int]),
        "gpo_ref_resample": (C.c_int, [C.c_uint64, C.c_uint32, C.c_int, C.c_uint32, _u8p, _f32p, _f64p, _f64p,
                                       _f64p, C.c_int]),
        "gpo_opset_nops": (C.c_int, [C.c_int]),
        "gpo_opset_code": (C.c_int, [C.c_int, C.c_int]),
        "gpo_opset_program_code": (C.c_char_p, [C.c_int]),
        "gpo_ref_prior_ops": (C.c_double, [_u8p, _f32p, C.c_size_t, C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def ptr(a, t):
    return a.ctypes.data_as(t)


def u8(a):
    return ptr(a, _u8p)


def f32(a):
    return ptr(a, _f32p)


def f64(a):
    return ptr(a, _f64p)


def i32(a):
    return ptr(a, _i32p)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().gpo_philox4x32_10(ptr(c, _u32p), ptr(k, _u32p), ptr(out, _u32p))
    return out


def stream(seed, chain, step, purpose, slot, count):
    s = GpoStream()
    lib().gpo_stream_init(C.byref(s), seed, chain, step, purpose, slot)
    return np.array([lib().gpo_stream_next(C.byref(s)) for _ in range(count)], dtype=np.uint32)


def tape_arrays(tokens):
    """tokens: list of op names or (name, const) -> (ops u8, cs f32)"""
    ops, cs = [], []
    for t in tokens:
        if isinstance(t, tuple):
            ops.append(OP[t[0]]); cs.append(t[1])
        else:
            ops.append(OP[t]); cs.append(0.0)
    return np.array(ops, dtype=np.uint8), np.array(cs, dtype=np.float32)


def eval_tape(ops, cs, x, f64=False):
    ops = np.ascontiguousarray(ops, dtype=np.uint8)
    cs = np.ascontiguousarray(cs, dtype=np.float32)
    if f64:
        return lib().gpo_eval_tape_f64(u8(ops), f32(cs), len(ops), float(x))
    return lib().gpo_eval_tape_f32(u8(ops), f32(cs), len(ops), float(x))


def tape_loglik(ops, cs, x, y, sd, mode=0):
    ops = np.ascontiguousarray(ops, dtype=np.uint8)
    cs = np.ascontiguousarray(cs, dtype=np.float32)
    x, y, sd = (np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, sd))
    return lib().gpo_tape_loglik(u8(ops), f32(cs), len(ops), f32(x), f32(y), f32(sd), len(x), mode)


def tape_prior(ops, cs, nsym, cauchy_slots=0):
    ops = np.ascontiguousarray(ops, dtype=np.uint8)
    cs = np.ascontiguousarray(cs, dtype=np.float32)
    return lib().gpo_tape_prior(u8(ops), f32(cs), len(ops), float(np.log(nsym)), cauchy_slots)


def tape_grad(ops, cs, x, y, sd):
    ops = np.ascontiguousarray(ops, dtype=np.uint8)
    cs = np.ascontiguousarray(cs, dtype=np.float32)
    x, y, sd = (np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, sd))
    g = np.zeros(len(ops), dtype=np.float64)
    lp = np.zeros(1, dtype=np.float64)
    rc = lib().gpo_tape_grad(u8(ops), f32(cs), len(ops), f32(x), f32(y), f32(sd), len(x), f64(g), f64(lp))
    return rc, g, float(lp[0])


def tape_logpost_f64(ops, cs64, x, y, sd):
    ops = np.ascontiguousarray(ops, dtype=np.uint8)
    cs64 = np.ascontiguousarray(cs64, dtype=np.float64)
    x, y, sd = (np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, sd))
    return lib().gpo_tape_logpost_f64(u8(ops), f64(cs64), len(ops), f32(x), f32(y), f32(sd), len(x))


def central_diff_grad(ops, cs, x, y, sd, rel=1e-6):
    """FP64 central differences of tape_logpost_f64 at the constant tokens; also the
    same at half the step, so callers can see how well the difference has converged."""
    c64 = np.asarray(cs, dtype=np.float64)
    g, g2 = np.zeros(len(ops)), np.zeros(len(ops))
    for t in range(len(ops)):
        if not (ops[t] == OP["CONST"] or ops[t] >= OP["CADD"]):
            continue
        for out, h in ((g, rel * max(1.0, abs(c64[t]))), (g2, 0.5 * rel * max(1.0, abs(c64[t])))):
            cp, cm = c64.copy(), c64.copy()
            cp[t] += h
            cm[t] -= h
            out[t] = (tape_logpost_f64(ops, cp, x, y, sd) - tape_logpost_f64(ops, cm, x, y, sd)) / (2 * h)
    return g, g2


def ref_decode(hop, hc):
    hop = np.ascontiguousarray(hop, dtype=np.uint8)
    hc = np.ascontiguousarray(hc, dtype=np.float32)
    ops = np.zeros(64, dtype=np.uint8)
    cs = np.zeros(64, dtype=np.float32)
    slot = np.zeros(64, dtype=np.uint8)
    n = lib().gpo_ref_decode(u8(hop), f32(hc), 1, u8(ops), f32(cs), u8(slot))
    return ops[:n].copy(), cs[:n].copy(), slot[:n].copy()


def make_cfg(cap=32, dmax=16, nsym=10, proposal=0, sigma=0.1, eps=0.05, struct_every=5,
             fixed_structure=0, seed=0, metric=None):
    """metric: None or (rows, L) with L a K x K lower-triangular factor (K <= 16)."""
    cfg = GpoCfg(cap, dmax, nsym, proposal, sigma, eps, struct_every, fixed_structure, seed)
    if metric is not None:
        rows, L = metric
        cfg.mk = len(rows)
        for i, r in enumerate(rows):
            cfg.mrows[i] = int(r)
            for j in range(i + 1):
                cfg.mL[i * 16 + j] = float(np.float32(L[i][j]))
    return cfg


def cfg_init(cfg, nchains, chain0=0):
    ops = np.zeros((cfg.cap, nchains), dtype=np.uint8)
    cs = np.zeros((cfg.cap, nchains), dtype=np.float32)
    ln = np.zeros(nchains, dtype=np.int32)
    o = np.zeros(256, dtype=np.uint8)
    v = np.zeros(256, dtype=np.float32)
    for c in range(nchains):
        L = lib().gpo_cfg_init_chain(C.byref(cfg), chain0 + c, u8(o), f32(v))
        ln[c] = L
        ops[:L, c] = o[:L]
        cs[:L, c] = v[:L]
    return ops, cs, ln


def cfg_run_block(cfg, ops, cs, ln, x, y, sd, step0, steps, chain0=0, nthreads=0, margin=None):
    """margin: optional float64 [nchains] receiving each chain's accept margin (the
    smallest |a - log u| over the likelihood scale among its decisions)."""
    nch = ln.shape[0]
    prior = np.zeros(nch, dtype=np.float64)
    ll = np.zeros(nch, dtype=np.float64)
    cnt = np.zeros(4, dtype=np.int64)
    x, y, sd = (np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, sd))
    lib().gpo_cfg_run_block_m(C.byref(cfg), nch, chain0, step0, steps, u8(ops), f32(cs), i32(ln),
                              f32(x), f32(y), f32(sd), len(x), f64(prior), f64(ll), ptr(cnt, _i64p), nthreads,
                              None if margin is None else f64(margin))
    return prior, ll, cnt


def ref_run_block(refcfg, hop, hc, x, y, sd, step0, steps, random_seed=0, chain0=0, nthreads=0, margin=None):
    nch = hop.shape[1]
    prior = np.zeros(nch, dtype=np.float64)
    ll = np.zeros(nch, dtype=np.float64)
    x, y, sd = (np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, sd))
    lib().gpo_ref_run_block_m(C.byref(refcfg), nch, chain0, step0, steps, random_seed, u8(hop), f32(hc),
                              f32(x), f32(y), f32(sd), len(x), f64(prior), f64(ll), nthreads,
                              None if margin is None else f64(margin))
    return prior, ll


def ref_init(seed, nchains, chain0=0, ops=0):
    hop = np.zeros((31, nchains), dtype=np.uint8)
    hc = np.zeros((31, nchains), dtype=np.float32)
    lib().gpo_ref_init(seed, nchains, chain0, u8(hop), f32(hc), ops)
    return hop, hc


def opset(ops):
    """(heap codes in enum order, PROGRAM_CODE) of op set `ops`."""
    n = lib().gpo_opset_nops(ops)
    return [lib().gpo_opset_code(ops, i) for i in range(n)], lib().gpo_opset_program_code(ops).decode()


def ref_draws(kind, n, rng=0, seed=0, lcg_seed=0, arg=0):
    out = np.zeros(n, dtype=np.float64)
    lib().gpo_ref_draws(kind, rng, seed, lcg_seed, n, arg, f64(out))
    return out


def ref_stream_draws(kind, n, seed, chain, step, purpose, slot=0, arg=0):
    """n draws of a reference transform from the Philox substream (chain, step, purpose, slot)."""
    out = np.zeros(n, dtype=np.float64)
    lib().gpo_ref_stream_draws(kind, seed, chain, step, purpose, slot, n, arg, f64(out))
    return out

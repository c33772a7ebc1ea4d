"""ctypes binding of libgpuropolis.so (the C ABI in include/gpuropolis.h).

The library is the product; this module only marshals numpy arrays and device
pointers.  It raises if the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# GPR_LIB names another build of the library (A/B measurements of build variants)
LIB_PATH = os.environ.get("GPR_LIB") or os.path.join(PKG, "libgpuropolis.so")

OPS = ["X", "CONST", "ONE", "ADD", "SUB", "MUL", "DIV", "POW", "RSUB", "RDIV", "RPOW", "NEG", "EXP",
       "LOG", "INV", "SQRT", "SIN", "ASIN", "COS", "ACOS", "TAN", "ATAN", "ABS", "CADD", "CMUL",
       "CPOW", "CRPOW"]
OP = {n: i for i, n in enumerate(OPS)}
REF_OPS = ["ONE", "CONST", "INV", "A", "B", "PLUS", "MINUS", "RMINUS", "CPLUS", "TIMES", "CTIMES", "DIV",
           "RDIV", "LOG", "EXP", "POW", "CPOW", "RPOW", "CRPOW", "SQRT", "SIN", "ASIN", "COS", "ACOS",
           "TAN", "ATAN", "ABS", "LOGOF2", "HALF", "E", "PI", "CLIGHT", "G", "HBAR", "MU0", "EL", "MP", "ME",
           "E0", "KB"]
ROP = {n: i for i, n in enumerate(REF_OPS)}

SEM_CONFIG, SEM_REFERENCE = 0, 1
SCHED_AUTO, SCHED_LANE, SCHED_WAVE = 0, 1, 2
PROP_RW, PROP_MALA = 0, 1
RNG_PHILOX, RNG_LCG = 0, 1
RS_NONE, RS_MAP, RS_QUANTILE = 0, 1, 2
OPS_DEFAULT, OPS_SIMPLIFIED, OPS_POLYNOMIAL, OPS_LINEAR, OPS_PHYSICAL = 0, 1, 2, 3, 4
MAX_CAP, MAX_DEPTH = 128, 16

EXPORTS = ["gpr_config_default", "gpr_create", "gpr_destroy", "gpr_last_error", "gpr_version", "gpr_set_data",
           "gpr_init_chains", "gpr_write_state", "gpr_read_state", "gpr_run_block", "gpr_run_block_async",
           "gpr_resample", "gpr_run_blocks_graph", "gpr_get_stats", "gpr_reset_stats", "gpr_sync", "gpr_stream",
           "gpr_eval_loglik", "gpr_grad_logpost", "gpr_ref_decode", "gpr_test_ref_decode", "gpr_test_draws",
           "gpr_autotune", "gpr_set_tuning", "gpr_get_tuning", "gpr_set_metric", "gpr_ref_op_set",
           "gpr_get_clock", "gpr_set_clock", "gpr_eval_loglik_noalloc_async", "gpr_write_heap_i32",
           "gpr_read_heap_i32"]
NOALLOC_MAX_POINTS = 65536


class Config(C.Structure):
    _fields_ = [("device", C.c_int), ("semantics", C.c_int), ("cap", C.c_int), ("dmax", C.c_int),
                ("nsym", C.c_int), ("proposal", C.c_int), ("sigma", C.c_float), ("eps", C.c_float),
                ("struct_every", C.c_int), ("fixed_structure", C.c_int), ("schedule", C.c_int), ("rng", C.c_int),
                ("resample", C.c_int), ("resample_nats", C.c_double), ("resample_quantile", C.c_float),
                ("seed", C.c_uint64), ("ops", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("struct_proposed", C.c_longlong), ("struct_accepted", C.c_longlong),
                ("const_proposed", C.c_longlong), ("const_accepted", C.c_longlong), ("steps", C.c_longlong),
                ("node_evals", C.c_longlong), ("alg_flops", C.c_longlong), ("last_block_ms", C.c_double), ("schedule_used", C.c_int),
                ("map_chain", C.c_int), ("map_posterior", C.c_double)]


class Tune(C.Structure):
    _fields_ = [("n_class", C.c_int), ("hi", C.c_int), ("kc", C.c_int), ("waves_per_wg", C.c_int),
                ("wg_per_cu", C.c_int), ("ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class GprError(RuntimeError):
    pass


_lib = None


def load():
    """Load the in-tree library; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GprError(f"{LIB_PATH} is not built: run `python -m gpuropolis_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sigs = {
        "gpr_config_default": (C.c_int, [C.POINTER(Config)]),
        "gpr_create": (C.c_int, [C.POINTER(Config), C.POINTER(vp)]),
        "gpr_destroy": (C.c_int, [vp]),
        "gpr_last_error": (C.c_char_p, []),
        "gpr_version": (C.c_char_p, []),
        "gpr_set_data": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp, vp]),
        "gpr_init_chains": (C.c_int, [vp]),
        "gpr_write_state": (C.c_int, [vp, vp, vp, vp]),
        "gpr_read_state": (C.c_int, [vp, vp, vp, vp, vp, vp]),
        "gpr_get_clock": (C.c_int, [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "gpr_set_clock": (C.c_int, [vp, C.c_uint32, C.c_uint32]),
        "gpr_run_block": (C.c_int, [vp, C.c_int, C.c_int32]),
        "gpr_run_block_async": (C.c_int, [vp, C.c_int, C.c_int32]),
        "gpr_resample": (C.c_int, [vp]),
        "gpr_run_blocks_graph": (C.c_int, [vp, C.c_int, C.c_int]),
        "gpr_get_stats": (C.c_int, [vp, C.POINTER(Stats)]),
        "gpr_reset_stats": (C.c_int, [vp]),
        "gpr_sync": (C.c_int, [vp]),
        "gpr_stream": (vp, [vp]),
        "gpr_eval_loglik": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_int, vp]),
        "gpr_grad_logpost": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_int, C.c_int, vp, vp]),
        "gpr_autotune": (C.c_int, [vp, C.c_int, C.POINTER(Tune), C.c_int, C.POINTER(C.c_int)]),
        "gpr_set_metric": (C.c_int, [vp, C.c_int, C.c_int, vp, vp]),
        "gpr_set_tuning": (C.c_int, [vp, C.POINTER(Tune), C.c_int, C.c_int]),
        "gpr_get_tuning": (C.c_int, [vp, C.POINTER(Tune), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "gpr_ref_decode": (C.c_int, [vp, vp, C.c_size_t, vp, vp]),
        "gpr_ref_op_set": (C.c_int, [C.c_int, vp, C.c_int, C.POINTER(C.c_char_p)]),
        "gpr_test_ref_decode": (C.c_int, [vp, vp, vp, C.c_int, vp, vp, vp]),
        "gpr_test_draws": (C.c_int, [vp, vp, C.c_int, C.c_uint64, vp]),
        "gpr_write_heap_i32": (C.c_int, [vp, vp, vp]),
        "gpr_read_heap_i32": (C.c_int, [vp, vp, vp, vp, vp]),
        "gpr_eval_loglik_noalloc_async": (C.c_int, [C.c_int, vp, vp, vp, C.c_int, vp, vp, vp, C.c_int, C.c_int,
                                                    C.c_int, vp, vp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc < 0:
        raise GprError(f"gpuropolis error {rc}: {load().gpr_last_error().decode()}")
    return rc


def _p(a):
    return None if a is None else a.ctypes.data


def default_config(**kw):
    cfg = Config()
    check(load().gpr_config_default(C.byref(cfg)))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    return cfg


class Context:
    """One sampler context: datasets + chain state resident on one GPU."""

    def __init__(self, cfg=None, **kw):
        self.cfg = cfg if cfg is not None else default_config(**kw)
        h = C.c_void_p()
        check(load().gpr_create(C.byref(self.cfg), C.byref(h)))
        self.h = h
        self.n = None
        self.chains = None

    def close(self):
        if self.h:
            load().gpr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def rows(self):
        return 31 if self.cfg.semantics == SEM_REFERENCE else self.cfg.cap

    def set_data(self, datasets, chains, caps=None):
        """datasets: list of (x, y, sd) arrays; chains: list of chain counts."""
        n = np.array([len(d[0]) for d in datasets], dtype=np.int32)
        x = np.ascontiguousarray(np.concatenate([d[0] for d in datasets]), dtype=np.float32)
        y = np.ascontiguousarray(np.concatenate([d[1] for d in datasets]), dtype=np.float32)
        sd = np.ascontiguousarray(np.concatenate([d[2] for d in datasets]), dtype=np.float32)
        ch = np.array(chains, dtype=np.int32)
        cp = None if caps is None else np.array(caps, dtype=np.int32)
        check(load().gpr_set_data(self.h, len(n), _p(n), _p(x), _p(y), _p(sd), _p(ch), _p(cp)))
        self.n = n
        self.chains = ch
        self.N = int(ch.sum())

    def init_chains(self):
        check(load().gpr_init_chains(self.h))

    def write_state(self, ops, cs, length=None):
        ops = np.ascontiguousarray(ops, dtype=np.uint8)
        cs = np.ascontiguousarray(cs, dtype=np.float32)
        ln = None if length is None else np.ascontiguousarray(length, dtype=np.int32)
        check(load().gpr_write_state(self.h, _p(ops), _p(cs), _p(ln)))

    def read_state(self):
        ops = np.zeros((self.rows, self.N), dtype=np.uint8)
        cs = np.zeros((self.rows, self.N), dtype=np.float32)
        ln = np.zeros(self.N, dtype=np.int32)
        pr = np.zeros(self.N, dtype=np.float64)
        ll = np.zeros(self.N, dtype=np.float64)
        check(load().gpr_read_state(self.h, _p(ops), _p(cs), _p(ln), _p(pr), _p(ll)))
        return ops, cs, ln, pr, ll

    def write_heap_i32(self, P, Cs):
        """Reference semantics: heap op codes as int32 [31][N] (main.cu's op type) and constants."""
        P = np.ascontiguousarray(P, dtype=np.int32)
        Cs = np.ascontiguousarray(Cs, dtype=np.float32)
        check(load().gpr_write_heap_i32(self.h, _p(P), _p(Cs)))

    def read_heap_i32(self):
        P = np.zeros((31, self.N), dtype=np.int32)
        Cs = np.zeros((31, self.N), dtype=np.float32)
        pr = np.zeros(self.N, dtype=np.float64)
        ll = np.zeros(self.N, dtype=np.float64)
        check(load().gpr_read_heap_i32(self.h, _p(P), _p(Cs), _p(pr), _p(ll)))
        return P, Cs, pr, ll

    def clock(self):
        """(step, block): where the next block's and the next resample's Philox streams start."""
        st, bl = C.c_uint32(), C.c_uint32()
        check(load().gpr_get_clock(self.h, C.byref(st), C.byref(bl)))
        return st.value, bl.value

    def set_clock(self, step, block):
        check(load().gpr_set_clock(self.h, step, block))

    def run_block(self, steps, block_seed=0):
        check(load().gpr_run_block(self.h, steps, block_seed))

    def run_block_async(self, steps, block_seed=0):
        check(load().gpr_run_block_async(self.h, steps, block_seed))

    def resample(self):
        check(load().gpr_resample(self.h))

    def run_blocks_graph(self, nblocks, steps):
        check(load().gpr_run_blocks_graph(self.h, nblocks, steps))

    def stats(self):
        s = Stats()
        check(load().gpr_get_stats(self.h, C.byref(s)))
        return s

    def reset_stats(self):
        check(load().gpr_reset_stats(self.h))

    def sync(self):
        check(load().gpr_sync(self.h))

    def stream(self):
        return load().gpr_stream(self.h)

    def eval_loglik_dev(self, dataset, d_ops, d_cs, d_len, nchains, cap, schedule, d_ll):
        """All pointers are device addresses (ints)."""
        check(load().gpr_eval_loglik(self.h, dataset, d_ops, d_cs, d_len, nchains, cap, schedule, d_ll))

    def test_ref_decode_dev(self, d_hop, d_hc, nheaps, d_ops, d_cs, d_meta):
        check(load().gpr_test_ref_decode(self.h, d_hop, d_hc, nheaps, d_ops, d_cs, d_meta))

    def test_draws_dev(self, d_words, n, seed, d_out):
        check(load().gpr_test_draws(self.h, d_words, n, seed, d_out))

    def autotune(self, steps=50):
        """Tune the launch shape of every size class on the current chains; returns the
        table as a list of dicts and the schedule picked for GPR_SCHED_AUTO."""
        out = (Tune * 6)()
        n = C.c_int(0)
        check(load().gpr_autotune(self.h, steps, out, 6, C.byref(n)))
        return self.tuning()

    def tuning(self):
        out = (Tune * 6)()
        n, sched = C.c_int(0), C.c_int(0)
        check(load().gpr_get_tuning(self.h, out, 6, C.byref(n), C.byref(sched)))
        return [out[i].as_dict() for i in range(n.value)], sched.value

    def set_tuning(self, entries, schedule=SCHED_AUTO):
        arr = (Tune * max(1, len(entries)))()
        for i, e in enumerate(entries):
            for k, v in e.items():
                setattr(arr[i], k, v)
        check(load().gpr_set_tuning(self.h, arr, len(entries), schedule))

    def set_metric(self, dataset, rows, chol):
        """Proposal metric of dataset's constant moves: token rows and the lower Cholesky
        factor of the proposal covariance (rows=[] removes it)."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        L = np.ascontiguousarray(np.tril(np.asarray(chol, dtype=np.float64)), dtype=np.float32) if len(r) else None
        check(load().gpr_set_metric(self.h, dataset, len(r), _p(r) if len(r) else None, _p(L)))

    def grad_logpost_dev(self, dataset, d_ops, d_cs, d_len, nchains, cap, d_grad, d_lp):
        check(load().gpr_grad_logpost(self.h, dataset, d_ops, d_cs, d_len, nchains, cap, d_grad, d_lp))


def eval_loglik_noalloc_async(device, d_x, d_y, d_sd, n, d_ops, d_cs, d_len, nchains, cap, schedule, d_ll,
                              stream=None):
    """gpr_eval_loglik_noalloc_async: every array is a device address (int); the work is
    enqueued on `stream` (a hipStream_t as int, None for the null stream)."""
    check(load().gpr_eval_loglik_noalloc_async(device, d_x, d_y, d_sd, n, d_ops, d_cs, d_len, nchains, cap,
                                               schedule, d_ll, stream))


def ref_op_set(ops):
    """(heap codes of op set `ops` in its enum order, its PROGRAM_CODE)."""
    codes = np.zeros(32, dtype=np.uint8)
    pc = C.c_char_p()
    n = check(load().gpr_ref_op_set(ops, _p(codes), 32, C.byref(pc)))
    return [int(c) for c in codes[:n]], pc.value.decode()


def ref_decode(hop, hc):
    hop = np.ascontiguousarray(hop, dtype=np.uint8)
    hc = np.ascontiguousarray(hc, dtype=np.float32)
    ops = np.zeros(64, dtype=np.uint8)
    cs = np.zeros(64, dtype=np.float32)
    n = check(load().gpr_ref_decode(_p(hop), _p(hc), 1, _p(ops), _p(cs)))
    return ops[:n].copy(), cs[:n].copy()

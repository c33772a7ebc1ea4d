"""The C-ABI library on the CPU: it loads, exports every function include/gpuropolis.h
declares, its host-only entries agree with the oracle, and calls that need a GPU fail
with an error code and message instead of crashing (no compute calls here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "gpuropolis.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gpr_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(gpr):
    L = gpr.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(gpr.EXPORTS) <= set(names)


def test_op_tables_agree_with_header():
    txt = open(os.path.join(ROOT, "include", "gpuropolis.h")).read()
    ops = re.search(r"GPR_OP_X = 0,(.*?)GPR_NUM_OPS", txt, re.S).group(1)
    names = ["X"] + re.findall(r"GPR_OP_([A-Z]+)", ops)
    from gpuropolis_amd.lib import OPS
    assert names == OPS == O.OPS
    code = re.search(r'GPR_REF_PROGRAM_CODE "([^"]+)"', txt).group(1)
    assert len(code) == 27                     # strlen(PROGRAM_CODE) == NOPS, main.cu:1026


def test_config_defaults(gpr):
    cfg = gpr.default_config()
    assert cfg.struct_every == 5 and abs(cfg.sigma - 0.1) < 1e-7 and cfg.resample_nats == 100.0
    assert abs(cfg.resample_quantile - 0.5) < 1e-7 and cfg.dmax == 16 and cfg.nsym == 10


def test_ref_decode_matches_oracle(gpr):
    rng = np.random.default_rng(0)
    for _ in range(300):
        hop = rng.integers(0, 27, 31).astype(np.uint8)
        hc = rng.normal(0, 3, 31).astype(np.float32)
        a_ops, a_cs = gpr.ref_decode(hop, hc)
        b_ops, b_cs, _ = O.ref_decode(hop, hc)
        assert np.array_equal(a_ops, b_ops) and np.array_equal(a_cs.view(np.uint32), b_cs.view(np.uint32))


def test_errors_are_reported_not_fatal(gpr):
    L = gpr.load()
    cfg = gpr.default_config(cap=500)
    h = C.c_void_p()
    rc = L.gpr_create(C.byref(cfg), C.byref(h))
    assert rc == -1 and b"cap" in L.gpr_last_error()
    cfg = gpr.default_config()
    rc = L.gpr_create(C.byref(cfg), C.byref(h))
    if rc != 0:                                # no GPU in this container
        assert L.gpr_last_error()
    else:
        L.gpr_destroy(h)
    bad = np.full(31, 200, dtype=np.uint8)
    assert L.gpr_ref_decode(bad.ctypes.data, np.zeros(31, np.float32).ctypes.data, 1,
                            np.zeros(64, np.uint8).ctypes.data, np.zeros(64, np.float32).ctypes.data) == -1


def test_op_sets_need_reference_semantics(gpr):
    """argument checks before any device call"""
    with pytest.raises(gpr.GprError, match="reference semantics only"):
        gpr.Context(semantics=gpr.SEM_CONFIG, ops=2)
    with pytest.raises(gpr.GprError, match="op set"):
        gpr.Context(semantics=gpr.SEM_REFERENCE, ops=7)


def test_autograd_wrapper_checks_tapes_on_the_host():
    """The gradient wrapper rejects tapes the adjoint kernel cannot take: lengths outside
    [1, cap], unknown op codes, malformed postfix, stacks deeper than 16."""
    import torch
    from gpuropolis_amd import lib as G
    from gpuropolis_amd.autograd import check_tapes
    X, C, ADD, EXP = G.OP["X"], G.OP["CONST"], G.OP["ADD"], G.OP["EXP"]

    def tapes(progs, cap=8):
        ops = torch.zeros((cap, len(progs)), dtype=torch.uint8)
        for i, p in enumerate(progs):
            ops[:len(p), i] = torch.tensor(p, dtype=torch.uint8)
        return ops, torch.tensor([len(p) for p in progs], dtype=torch.int32)

    check_tapes(*tapes([[X], [X, C, ADD, EXP], [C, X, X, ADD, ADD]]))
    for bad in ([X, ADD], [X, X], [ADD], [X, C, ADD, ADD], [X, 200]):
        with pytest.raises(G.GprError):
            check_tapes(*tapes([[X], bad]))
    ops, ln = tapes([[X, C, ADD]])
    for L in (0, 9):
        with pytest.raises(G.GprError):
            check_tapes(ops, torch.tensor([L], dtype=torch.int32))
    deep = [X] * 17 + [ADD] * 16
    check_tapes(*tapes([[X] * 16 + [ADD] * 15], cap=40))
    with pytest.raises(G.GprError):
        check_tapes(*tapes([deep], cap=40))


def test_noalloc_eval_rejects_sizes_before_any_device_call(gpr):
    """The no-allocation evaluator checks its sizes first: n beyond its static scratch, or
    below 1, and bad caps are errors with a message, with no device touched."""
    L = gpr.load()
    fake = 0x1000     # never dereferenced: the argument checks come first
    for n, cap in ((gpr.NOALLOC_MAX_POINTS + 1, 32), (0, 32), (100, 0), (100, 129)):
        rc = L.gpr_eval_loglik_noalloc_async(0, fake, fake, fake, n, fake, fake, fake, 4, cap, gpr.SCHED_WAVE, fake,
                                             None)
        assert rc == -1, (n, cap)
    L.gpr_eval_loglik_noalloc_async(0, fake, fake, fake, gpr.NOALLOC_MAX_POINTS + 1, fake, fake, fake, 4, 32,
                                    gpr.SCHED_WAVE, fake, None)
    assert b"static scratch" in L.gpr_last_error()

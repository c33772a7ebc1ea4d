"""Op sets of the reference (main.cu:51-86) as run-time presets (CPU pins).

* Tables.  The heap codes of each op set in its enum order are restated below from the
  enum lists of main.cu:53, 59, 65, 72 and 82, with each PROGRAM_CODE (one character per
  op: main.cu:1026 asserts strlen(PROGRAM_CODE) == NOPS).  The oracle and the library
  must hold the same tables.
* Physical constants and halving.  The fold returns the float of each double literal of
  main.cu:429-441, and log(2.0) and x/2.0, checked against numpy.  The decoded tape
  evaluates to the fold.
* Prior.  log(NOPS + 1) of the op set.  The default set keeps its value, and the other
  sets differ from it by -length * (log(NOPS + 1) - log 28).
* Sampler.  Initial draws use the first NOPS - 1 codes of the set (main.cu:1133), with
  uniform frequencies.  Structure moves draw all NOPS (main.cu:718).
"""
import math

import numpy as np
import pytest

import oracle_lib as O

ENUMS = {
    0: ["ONE", "CONST", "INV", "A", "B", "PLUS", "MINUS", "RMINUS", "CPLUS", "TIMES", "CTIMES", "DIV", "RDIV",
        "LOG", "EXP", "POW", "CPOW", "RPOW", "CRPOW", "SQRT", "SIN", "ASIN", "COS", "ACOS", "TAN", "ATAN", "ABS"],
    1: ["ONE", "CONST", "INV", "A", "B", "PLUS", "MINUS", "RMINUS", "CPLUS", "TIMES", "CTIMES", "DIV", "RDIV"],
    2: ["ONE", "CONST", "A", "B", "PLUS", "MINUS", "RMINUS", "CPLUS", "TIMES", "CTIMES", "CPOW"],
    3: ["ONE", "CONST", "A", "B", "PLUS", "MINUS", "RMINUS", "CPLUS", "CTIMES"],
    4: ["LOGOF2", "HALF", "E", "PI", "CLIGHT", "G", "HBAR", "MU0", "EL", "MP", "ME", "E0", "KB", "A", "B",
        "PLUS", "MINUS", "RMINUS", "TIMES"],
}
PROGRAM_CODE = {0: "1CIab+-_#*@/|LE^pVPRSsCcTtA", 1: "1CIab+-_#*@/|", 2: "1CAB+-_#*@^", 3: "1CAB+-_#@",
                4: "B2epcghm:o.~Kab+-_*"}
# main.cu:429-441, double literals (E is the elementary charge itself, the rest are logs)
PHYSICAL = {"E": 1.6021766208e-19, "PI": 1.14472988584915, "CLIGHT": 19.5186009865452,
            "G": -23.4302046558627, "HBAR": -78.2347583540371, "MU0": -13.5870714043359,
            "EL": -43.2777536741305, "MP": -61.6554051167679, "ME": -69.1708328401338,
            "E0": -25.4501305691713, "KB": -52.6369038080531, "LOGOF2": math.log(2.0)}


@pytest.mark.parametrize("ops", sorted(ENUMS))
def test_op_set_tables(ops):
    codes, pc = O.opset(ops)
    assert codes == [O.ROP[n] for n in ENUMS[ops]]
    assert pc == PROGRAM_CODE[ops] and len(pc) == len(codes)


def test_library_op_set_tables_match_oracle():
    from gpuropolis_amd import lib as G
    for ops in ENUMS:
        assert G.ref_op_set(ops) == O.opset(ops)


def test_physical_constants_and_halving_fold_and_decode():
    L = O.lib()
    for name, v in PHYSICAL.items():
        hop = np.full(31, O.ROP["A"], dtype=np.uint8)
        hop[0] = O.ROP[name]
        hc = np.full(31, 7.0, dtype=np.float32)      # heap constants are not used by these ops
        got = L.gpo_ref_fold_eval(O.u8(hop), O.f32(hc), 1, 0.3)
        assert np.float32(got) == np.float32(v), name
        ops, cs, _ = O.ref_decode(hop, hc)
        assert len(ops) == 1 and np.float32(O.eval_tape(ops, cs, 0.3)) == np.float32(v)
        assert L.gpo_ref_fold_length(O.u8(hop), 1) == 1.0 and L.gpo_ref_fold_count(O.u8(hop), 1) == 1
        assert L.gpo_ref_fold_degree(O.u8(hop), O.f32(hc), 1) == 0.0
    hop = np.full(31, O.ROP["A"], dtype=np.uint8)
    hop[0] = O.ROP["HALF"]
    hc = np.zeros(31, dtype=np.float32)
    for x in (0.3, -7.25, 1e-39, 3e38):
        got = L.gpo_ref_fold_eval(O.u8(hop), O.f32(hc), 1, x)
        assert np.float32(got) == np.float32(np.float64(np.float32(x)) / 2.0)
        ops, cs, _ = O.ref_decode(hop, hc)
        assert np.float32(O.eval_tape(ops, cs, x)) == np.float32(got)
    assert L.gpo_ref_fold_length(O.u8(hop), 1) == 2.0       # 1 + a, a = x
    assert L.gpo_ref_fold_degree(O.u8(hop), O.f32(hc), 1) == 1.0


def test_prior_uses_the_op_sets_nops():
    rng = np.random.default_rng(3)
    L = O.lib()
    for _ in range(200):
        hop = rng.integers(0, len(O.REF_OPS), 31).astype(np.uint8)
        hc = rng.normal(0, 2, 31).astype(np.float32)
        length = L.gpo_ref_fold_length(O.u8(hop), 1)
        p0 = L.gpo_ref_prior(O.u8(hop), O.f32(hc), 1)
        assert L.gpo_ref_prior_ops(O.u8(hop), O.f32(hc), 1, 0) == p0
        if not np.isfinite(p0):
            continue
        for ops in (1, 2, 3, 4):
            n = len(ENUMS[ops])
            d = L.gpo_ref_prior_ops(O.u8(hop), O.f32(hc), 1, ops) - p0
            assert abs(d - (-length * (math.log(n + 1) - math.log(28)))) < 1e-9 * max(1.0, abs(p0))


@pytest.mark.parametrize("ops", sorted(ENUMS))
def test_init_draws_use_the_op_set(ops):
    codes = [O.ROP[n] for n in ENUMS[ops]]
    hop, _ = O.ref_init(17, 4000, ops=ops)
    allowed = codes[:-1]                                   # random_int(NOPS-1), main.cu:1133
    assert set(np.unique(hop)) <= set(allowed)
    counts = np.array([(hop == c).sum() for c in allowed])
    expect = hop.size / len(allowed)
    assert np.all(np.abs(counts - expect) < 6 * np.sqrt(expect))


@pytest.mark.parametrize("ops", [1, 2, 3, 4])
def test_structure_moves_draw_the_whole_op_set(ops):
    from gpuropolis_amd.workloads import dataset_S
    x, y, sd = dataset_S(0)
    codes = [O.ROP[n] for n in ENUMS[ops]]
    hop, hc = O.ref_init(3, 256, ops=ops)
    cfg = O.GpoRefCfg(0, 9, ops)
    prior, ll = O.ref_run_block(cfg, hop, hc, x, y, sd, 0, 300)
    assert set(np.unique(hop)) <= set(codes)
    assert np.isin(codes[-1], hop)                         # the last code appears only through moves
    for c in range(0, 256, 17):
        assert prior[c] == O.lib().gpo_ref_prior_ops(O.u8(np.ascontiguousarray(hop[:, c])),
                                                      O.f32(np.ascontiguousarray(hc[:, c])), 1, ops)

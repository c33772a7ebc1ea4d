"""The CLI's host side on the CPU: the data loader (load_data_file, main.cu:319-363),
the samples.txt display functions (main.cu:615-654, 818-976) and the command line's
error paths (main.cu:131-156, 1017-1031).

Pins: the printed expressions, read back as arithmetic, are the same functions as the
oracle's evaluators; lengths and degrees equal the oracle's folds; the templates of
string_dispatch / latex_dispatch are known answers on one-operator heaps; the loader's
trimming is checked on files whose kept rows are known."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def io(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("shim") / "libhostio.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "include"),
                           os.path.join(HERE, "host_io_shim.cpp"),
                           os.path.join(ROOT, "gpuropolis_amd", "host", "io.cpp"), "-o", out])
    L = C.CDLL(out)
    L.shim_heap_length.restype = C.c_float
    L.shim_heap_degree.restype = C.c_float
    L.shim_tape_length.restype = C.c_float
    L.shim_tape_degree.restype = C.c_float
    return L


def u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def f32(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def heap_text(io, P, Cs, which, fmt=b"C"):
    buf = C.create_string_buffer(1 << 20)
    n = io.shim_heap(u8(P), f32(Cs), which, fmt, buf, len(buf))
    assert n >= 0
    return buf.value.decode()


def tape_text(io, ops, cs, which, fmt=b"C"):
    buf = C.create_string_buffer(1 << 20)
    n = io.shim_tape(u8(ops), f32(cs), len(ops), which, fmt, buf, len(buf))
    assert n >= 0
    return buf.value.decode()


def load(io, path, sdscale=1.0, first=0, even=0):
    xyz = np.zeros(3 * 4096, dtype=np.float32)
    err = C.create_string_buffer(4096)
    n = io.shim_load(str(path).encode(), C.c_float(sdscale), first, even, f32(xyz), 4096, err, len(err))
    if n < 0:
        return None, err.value.decode()
    return xyz[:3 * n].reshape(n, 3), ""


# ------------------------------------------------------------------------ loader
def test_loader_reads_golden_data_like_loadtxt(io):
    for name in sorted(os.listdir(os.path.join(HERE, "golden", "nist"))):
        p = os.path.join(HERE, "golden", "nist", name, "data.txt")
        d, err = load(io, p)
        assert d is not None, err
        ref = np.loadtxt(p, comments="#").astype(np.float32)
        assert np.array_equal(d, ref)


def test_loader_comments_blank_lines_and_scientific(io, tmp_path):
    p = tmp_path / "d.txt"
    p.write_text("# header\n1\t2\t3\n\n   \t\n#x\n1e-3\t-2.5E+2\t4e0\n0.5 0.25 1\n")
    d, err = load(io, p, sdscale=2.0)
    assert d is not None, err
    assert np.array_equal(d, np.array([[1, 2, 6], [1e-3, -250, 8], [0.5, 0.25, 2]], dtype=np.float32))


@pytest.mark.parametrize("n", [0, 1, 2, 7, 8])
def test_loader_first_and_even_half(io, tmp_path, n):
    p = tmp_path / "d.txt"
    p.write_text("".join(f"{i}\t{10 * i}\t1\n" for i in range(n)))
    d, _ = load(io, p, first=1)
    assert list(d[:, 0]) == list(range(n // 2))                    # keep rows [0, n/2)
    d, _ = load(io, p, even=1)
    assert list(d[:, 0]) == list(range(1, n, 2))                   # erase every other row from the end
    d, _ = load(io, p, first=1, even=1)
    assert list(d[:, 0]) == list(range(1, n // 2, 2))


def test_loader_errors(io, tmp_path):
    d, err = load(io, tmp_path / "missing.txt")
    assert d is None and err.startswith("*** ERROR: Cannot open file:")
    p = tmp_path / "bad.txt"
    p.write_text("1\t2\t3\nnot a number\n")
    d, err = load(io, p)
    assert d is None and err.startswith("*** ERROR IN PARSING INPUT\tnot a number")
    p.write_text("1\t2\t0\n")
    d, err = load(io, p)                                            # assert((data_t)sd > 0), main.cu:336
    assert d is None and "sd must be positive" in err


# ------------------------------------------------------------------------ heap display
NP = {"log": "np.log", "exp": "np.exp", "sqrt": "np.sqrt", "sin": "np.sin", "cos": "np.cos", "tan": "np.tan",
      "asin": "np.arcsin", "acos": "np.arccos", "atan": "np.arctan", "abs": "np.abs"}


def as_python(expr):
    import re
    expr = expr.replace("^", "**")
    return re.sub(r"\b(asin|acos|atan|log|exp|sqrt|sin|cos|tan|abs)\(", lambda m: NP[m.group(1)] + "(", expr)


def eval_expr(expr, x):
    with np.errstate(all="ignore"):
        v = eval(as_python(expr), {"np": np, "x": x, "F": np.float64})
    return np.broadcast_to(np.asarray(v, dtype=np.float64), x.shape)


def test_heap_strings_evaluate_to_the_reference_tree(io):
    """displaystring with exact constants, read back as arithmetic, equals the FP64
    evaluation of the reachable tree (program_fold semantics) at every x."""
    rng = np.random.default_rng(3)
    xs = np.linspace(-2.5, 3.5, 13)
    n_ok = 0
    for it in range(400):
        P = rng.integers(0, 27, 31).astype(np.uint8)
        if it % 2:          # favour pass-through and binary ops so trees grow deep
            P = rng.choice(np.array([3, 4, 5, 6, 7, 9, 11, 12, 8, 10, 1, 15], dtype=np.uint8), 31)
        Cs = (rng.normal(0, 2, 31)).astype(np.float32)
        s = heap_text(io, P, Cs, 0, b"F(%.17g)")     # numpy scalars, parenthesised ("-c^x" would read as -(c^x))
        got = eval_expr(s, xs)
        ops, cs, _ = O.ref_decode(P, Cs)
        want = np.array([O.eval_tape(ops, cs, float(x), f64=True) for x in xs])
        ok = np.isclose(got, want, rtol=1e-9, atol=1e-300, equal_nan=True) | (np.isinf(got) & np.isinf(want))
        assert ok.all(), (s, got[~ok], want[~ok])
        n_ok += 1
    assert n_ok == 400


def test_heap_length_degree_code_match_oracle_folds(io):
    rng = np.random.default_rng(4)
    code = "1CIab+-_#*@/|LE^pVPRSsCcTtA"
    for _ in range(1000):
        P = rng.integers(0, 27, 31).astype(np.uint8)
        if _ % 2:
            P = rng.integers(0, len(O.REF_OPS), 31).astype(np.uint8)   # with the PHYSICAL_CONSTANTS ops
        Cs = rng.choice(np.array([0.0, 1.0, 2.0, 3.0, -1.0, 0.5], dtype=np.float32), 31)
        assert io.shim_heap_length(u8(P)) == O.lib().gpo_ref_fold_length(u8(P), 1)
        a = io.shim_heap_degree(u8(P), f32(Cs))
        b = O.lib().gpo_ref_fold_degree(u8(P), f32(Cs), 1)
        assert a == b or (np.isnan(a) and np.isnan(b))
        assert heap_text(io, P, Cs, 2) == "".join(code[p] if p < 27 else "?" for p in P)


def test_heap_code_column_per_op_set(io):
    """the code column prints each op set's PROGRAM_CODE character of the op's enum value;
    heap codes outside the op set print '?'"""
    rng = np.random.default_rng(6)
    for ops in range(5):
        codes, pc = O.opset(ops)
        for _ in range(50):
            P = rng.choice(np.array(codes, dtype=np.uint8), 31)
            P[rng.integers(0, 31)] = [c for c in range(len(O.REF_OPS)) if c not in codes][0]
            buf = C.create_string_buffer(64)
            n = io.shim_heap_code(u8(P), ops, buf, 64)
            want = "".join(pc[codes.index(p)] if p in codes else "?" for p in P)
            assert buf.value.decode()[:n] == want


def one_op_heap(op, c_root=2.5, c_left=1.5):
    """root = op, slot 2 = CONST (so a = the constant), slot 3 = A (b = x), the rest
    pass x up through A"""
    P = np.full(31, O.ROP["A"], dtype=np.uint8)
    P[0] = O.ROP[op]
    P[1] = O.ROP["CONST"]
    Cs = np.zeros(31, dtype=np.float32)
    Cs[0], Cs[1] = c_root, c_left
    return P, Cs


# string_dispatch / latex_dispatch (main.cu:821-873, 901-953) with a = "1.50", b = "x", C = "2.50"
KNOWN = {
    "ONE": ("1", "1"), "CONST": ("2.50", "2.50"), "INV": ("(1/1.50)", "1/1.50"), "A": ("1.50", "1.50"),
    "B": ("x", "x"), "PLUS": ("(1.50+x)", "1.50+x"), "CPLUS": ("(1.50+2.50)", "1.50+2.50"),
    "MINUS": ("(1.50-x)", "1.50-x"), "RMINUS": ("(x-1.50)", "x-1.50"), "TIMES": ("(x*1.50)", "x 1.50"),
    "CTIMES": ("(1.50*2.50)", "1.50 \\cdot 2.50"), "DIV": ("(1.50/x)", "\\frac{1.50}{x}"),
    "RDIV": ("(x/1.50)", "\\frac{x}{1.50}"), "LOG": ("log(1.50)", "\\log(1.50)"), "EXP": ("exp(1.50)", "e^{1.50}"),
    "POW": ("(1.50^x)", "1.50^{x}"), "CPOW": ("(1.50^2.50)", "1.50^{2.50}"), "RPOW": ("(x^1.50)", "x^{1.50}"),
    "CRPOW": ("(2.50^1.50)", "2.50^{1.50}"), "SQRT": ("sqrt(1.50)", "\\sqrt{1.50}"),
    "SIN": ("sin(1.50)", "sin(1.50)"), "ASIN": ("asin(1.50)", "asin(1.50)"), "COS": ("cos(1.50)", "cos(1.50)"),
    "ACOS": ("acos(1.50)", "acos(1.50)"), "TAN": ("tan(1.50)", "tan(1.50)"), "ATAN": ("atan(1.50)", "atan(1.50)"),
    "ABS": ("abs(1.50)", "\\mid 1.50\\mid "),
    # PHYSICAL_CONSTANTS ops (main.cu:855-868, 935-948); the LaTeX of log 2 is unbalanced there
    "E": ("E", "e"), "PI": ("PI", "\\pi"), "CLIGHT": ("CLIGHT", "c"), "G": ("G", "G"), "HBAR": ("HBAR", "\\bar{h}"),
    "MU0": ("MU0", "\\mu_0"), "EL": ("EL", "e_l"), "MP": ("MP", "m_p"), "ME": ("ME", "m_e"),
    "E0": ("E0", "\\epsilon_0"), "KB": ("KB", "k_B"), "LOGOF2": ("log(2.0)", "\\log 2)"),
    "HALF": ("(1.50/2.0)", "(1.50/2.0)"),
}


@pytest.mark.parametrize("op", sorted(KNOWN))
def test_heap_display_templates(io, op):
    P, Cs = one_op_heap(op)
    s, tex = KNOWN[op]
    assert heap_text(io, P, Cs, 0, b"%.2f") == s
    assert heap_text(io, P, Cs, 1, b"%.2f") == tex
    assert heap_text(io, P, Cs, 0, b"C") == s.replace("1.50", "C").replace("2.50", "C")


def test_heap_strings_have_no_length_limit(io):
    """the reference copies children through 100-byte buffers (main.cu:876); long
    programs print in full here"""
    P = np.full(31, O.ROP["PLUS"], dtype=np.uint8)
    Cs = np.zeros(31, dtype=np.float32)
    s = heap_text(io, P, Cs, 0, b"%.10f")
    assert s.count("x") == 32 and s.count("+") == 31
    assert np.allclose(eval_expr(s, np.array([1.25])), 32 * 1.25)


# ------------------------------------------------------------------------ postfix display
def test_tape_strings_evaluate_to_the_tape(io):
    cfg = O.make_cfg(cap=64, seed=11)
    ops, cs, ln = O.cfg_init(cfg, 300)
    xs = np.linspace(-1.5, 2.5, 11)
    for c in range(300):
        o, v = ops[:ln[c], c].copy(), cs[:ln[c], c].copy()
        s = tape_text(io, o, v, 0, b"F(%.17g)")
        got = eval_expr(s, xs)
        want = np.array([O.eval_tape(o, v, float(x), f64=True) for x in xs])
        ok = np.isclose(got, want, rtol=1e-9, atol=1e-300, equal_nan=True) | (np.isinf(got) & np.isinf(want))
        assert ok.all(), (s, got[~ok], want[~ok])
        assert io.shim_tape_length(u8(o), len(o)) == sum(O.lib().gpo_dl_cost(int(t)) for t in o)
        assert len(tape_text(io, o, v, 2)) == len(o)


TAPE_DEGREES = [
    (["X"], 1.0), ([("CONST", 2.0)], 0.0), (["X", "X", "MUL"], 2.0),
    (["X", "X", "MUL", ("CONST", 3.0), "ADD"], 2.0), (["X", ("CONST", 3.0), "POW"], 3.0),
    (["X", ("CONST", 2.5), "POW"], float("inf")), (["X", "X", "POW"], float("inf")),
    (["X", "EXP"], float("inf")), (["X", "X", "MUL", ("CONST", 4.0), "DIV"], 2.0),
    (["X", "X", "DIV"], float("inf")), (["X", "NEG", "X", "X", "MUL", "MUL"], 3.0),
    (["X", "X", "SUB", "X", "ADD"], 1.0),
]


@pytest.mark.parametrize("tokens,deg", TAPE_DEGREES)
def test_tape_degree_known_answers(io, tokens, deg):
    ops, cs = O.tape_arrays(tokens)
    assert io.shim_tape_degree(u8(ops), f32(cs), len(ops)) == deg


def test_tape_latex_known_answers(io):
    ops, cs = O.tape_arrays(["X", ("CONST", 2.0), "POW", "X", "DIV", "NEG", "EXP"])
    assert tape_text(io, ops, cs, 1, b"%.2f") == "e^{-\\frac{x^{2.00}}{x}}"
    assert tape_text(io, ops, cs, 0, b"C") == "exp((-((x^C)/x)))"


# ------------------------------------------------------------------------ command line
@pytest.fixture(scope="module")
def cli():
    from gpuropolis_amd import build
    if not os.path.exists(build.CLI):
        if not os.path.exists(build.LIB):
            build.build_lib()
        build.build_cli()
    return build.CLI


def run_cli(cli, *args):
    return subprocess.run([cli, *args], capture_output=True, text=True, timeout=60)


def test_cli_rejects_bad_arguments(cli, tmp_path):
    r = run_cli(cli, "--bogus")
    assert r.returncode == 1 and "Invalid command line arguments." in r.stderr
    r = run_cli(cli, "--schedule", "nonsense")
    assert r.returncode == 1 and "Invalid command line arguments." in r.stderr
    r = run_cli(cli, "--steps", "0")
    assert r.returncode == 1
    r = run_cli(cli, "--thin", "0")
    assert r.returncode == 1
    for q in ("1", "1.5", "-0.1", "nan"):
        r = run_cli(cli, "--resample-quantile", q)
        assert r.returncode == 1 and "--resample-quantile must be in [0, 1)" in r.stderr, q


def test_cli_reports_data_errors_before_touching_a_device(cli, tmp_path):
    r = run_cli(cli, "--in", str(tmp_path / "nope.txt"), "--out", str(tmp_path / "o"))
    assert r.returncode == 1 and "*** ERROR: Cannot open file:" in r.stderr
    p = tmp_path / "bad.txt"
    p.write_text("1 2 3\nx y z\n")
    r = run_cli(cli, "--in", str(p), "--out", str(tmp_path / "o"))
    assert r.returncode == 1 and "*** ERROR IN PARSING INPUT" in r.stderr
    assert not (tmp_path / "o" / "samples.txt").exists()


def test_cli_reports_state_file_errors_before_touching_a_device(cli, tmp_path):
    data = tmp_path / "d.txt"
    data.write_text("1 2 1\n2 3 1\n")
    common = ["--in", str(data), "--out", str(tmp_path / "o"), "--N", "8"]
    r = run_cli(cli, *common, "--state-in", str(tmp_path / "missing.bin"))
    assert r.returncode == 1 and "cannot open state file" in r.stderr
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"\0" * 64)
    r = run_cli(cli, *common, "--state-in", str(bad))
    assert r.returncode == 1 and "not a gpuropolis state file" in r.stderr
    other = tmp_path / "other.bin"   # current-format header for N = 16, reference semantics
    other.write_bytes(np.array([0x32525047, 1, 31, 16, 1, 0, 0, 0], np.int32).tobytes())
    r = run_cli(cli, *common, "--state-in", str(other))
    assert r.returncode == 1 and "does not match" in r.stderr
    short = tmp_path / "short.bin"
    short.write_bytes(np.array([0x32525047, 1, 31, 8, 1, 0, 0, 0], np.int32).tobytes() + b"\1" * 31)
    r = run_cli(cli, *common, "--state-in", str(short))
    assert r.returncode == 1 and "truncated" in r.stderr
    r = run_cli(cli, *common, "--tops", "-1")
    assert r.returncode == 1


# ------------------------------------------------------------------------ tops.txt
def _tops(io, src, out, k):
    err = C.create_string_buffer(4096)
    assert io.shim_tops(str(src).encode(), str(out).encode(), C.c_long(k), err, len(err)) == 0, err.value
    return out.read_bytes()


def _sort_tail(src, k):
    env = dict(os.environ, LC_ALL="C")
    s = subprocess.run(["sort", "-g", "-k3", str(src)], capture_output=True, check=True, env=env).stdout
    lines = s.splitlines(keepends=True)
    return b"".join(lines[len(lines) - k:]) if k else b""


def _samples_like(rng, nrows):
    """Header, datum lines and rows shaped as samples.txt, with ties, infinities and
    lines too short to have a third field."""
    out = ["# ----", "# -- Parameters:", "# \tInput data path: d.txt", "#", "#"]
    out += ["# \t%f\t%f\t%f" % (x, 3 * x - 1, 0.5) for x in rng.normal(0, 30, 12)]
    post = np.round(rng.normal(-40, 25, nrows), 1)
    post[rng.integers(0, nrows, nrows // 5)] = post[0]          # ties on the key
    for i, p in enumerate(post):
        body = "%.3f\t%.3f\t%.3f\t%.1f\tNA\t\"x\"\t\"x\"\t\"x\"\t\"%s\"\t" % (p, p / 3, p - p / 3, 3.0, "ab"[i % 2] * 5)
        out.append("%d\t%d\t%s" % (i % 97, i // 97, body))
    out += ["%d\t%d\t-inf\t-inf\t0.000\t" % (5, 1), "%d\t%d\tinf\t1.000\t1.000\t" % (6, 1),
            "%d\t%d\tnan\t1\t1\t" % (7, 1), "%d\t%d\tnan\t1\t1\t" % (2, 1), "3\t1\t-nan\t1", "2\t1\t-nan\t1", "only two", "", "  7\t8\t  1.5e1 z",
            "1 2 3.25e-1", "1\t2\t+0x1p3", "1\t2\t-0.000", "1\t2\t0.000"]
    rng.shuffle(out)
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("seed", [0, 1])
def test_tops_matches_sort_g_k3_tail(io, tmp_path, seed):
    """tops.txt equals `LC_ALL=C sort -g -k3 samples.txt | tail -n K` (run-all.sh:45)."""
    rng = np.random.default_rng(seed)
    src = tmp_path / "samples.txt"
    src.write_text(_samples_like(rng, 3000))
    nlines = len(src.read_bytes().splitlines())
    for k in (0, 1, 7, 100, 2999, nlines, nlines + 50):
        assert _tops(io, src, tmp_path / "tops.txt", k) == _sort_tail(src, min(k, nlines)), k


def test_tops_without_trailing_newline_and_errors(io, tmp_path):
    src = tmp_path / "s.txt"
    src.write_bytes(b"1\t1\t-2.5\ta\n1\t2\t9\tb\n1\t3\t4")
    assert _tops(io, src, tmp_path / "t.txt", 2) == _sort_tail(src, 2) == b"1\t3\t4\n1\t2\t9\tb\n"
    err = C.create_string_buffer(4096)
    assert io.shim_tops(str(tmp_path / "none.txt").encode(), str(tmp_path / "t.txt").encode(), C.c_long(3), err,
                        len(err)) == -1
    assert b"cannot open" in err.value

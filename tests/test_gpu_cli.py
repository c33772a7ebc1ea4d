"""The gpuropolis CLI end to end on the device: samples.txt header and rows in the
reference's format (main.cu:1087-1117, 1173-1211), printed likelihoods consistent with
the printed programs, runs replayable from --seed, log.txt written, state files
round-trip."""
import os
import re
import subprocess

import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "golden", "nist", "Misra1a", "data.txt")
CODE = "1CIab+-_#*@/|LE^pVPRSsCcTtA"


@pytest.fixture(scope="module")
def cli():
    from gpuropolis_amd import build
    assert os.path.exists(build.CLI), "gpuropolis CLI is not built"
    return build.CLI


def run(cli, out, *args):
    cmd = [cli, "--in", DATA, "--out", str(out), "--quiet", *args]
    say("cli:", " ".join(cmd[1:]))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return open(os.path.join(out, "samples.txt")).read()


def rows_of(text):
    return [ln for ln in text.splitlines() if not ln.startswith("#")]


NUM = re.compile(r"(?<![\w.)])-?\d+(?:\.\d+)?(?:[eE][-+]?\d+)?")
FN = {"asin": "np.arcsin", "acos": "np.arccos", "atan": "np.arctan", "log": "np.log", "exp": "np.exp",
      "sqrt": "np.sqrt", "sin": "np.sin", "cos": "np.cos", "tan": "np.tan", "abs": "np.abs"}


# PHYSICAL_CONSTANTS names of the string column and their values (main.cu:429-439)
PHYS = {"CLIGHT": 19.5186009865452, "HBAR": -78.2347583540371, "MU0": -13.5870714043359,
        "E0": -25.4501305691713, "KB": -52.6369038080531, "EL": -43.2777536741305, "MP": -61.6554051167679,
        "ME": -69.1708328401338, "PI": 1.14472988584915, "E": 1.6021766208e-19, "G": -23.4302046558627}


def eval_expr(expr, x):
    e = NUM.sub(lambda m: "F(" + m.group(0) + ")", expr).replace("^", "**")
    e = re.sub(r"\b(asin|acos|atan|log|exp|sqrt|sin|cos|tan|abs)\(", lambda m: FN[m.group(1)] + "(", e)
    with np.errstate(all="ignore"):
        v = eval(e, {"np": np, "x": x, "F": np.float64, "nan": np.nan, "inf": np.inf, **PHYS})
    return np.broadcast_to(np.asarray(v, dtype=np.float64), x.shape)


def loglik(f, y, sd):
    with np.errstate(all="ignore"):
        r = f - y
        return float(np.sum(-(r * r) / (2 * sd * sd) - 0.5 * (np.log(2 * np.pi) + 2 * np.log(sd))))


def check_rows(rows, N, blocks, nconst, d, ref, code=CODE):
    assert len(rows) == N * len(blocks)
    x, y, sd = (d[:, i].astype(np.float64) for i in range(3))
    agree = total = 0
    for k, row in enumerate(rows):
        f = row.split("\t")
        assert len(f) == 11 + nconst + 1 and f[-1] == ""
        assert int(f[0]) == k % N and int(f[1]) == blocks[k // N]
        post, prior, ll = float(f[2]), float(f[3]), float(f[4])
        if np.isfinite(post):
            assert abs(post - (prior + ll)) <= 0.0015 + 1e-9 * abs(post)
        assert f[6] == "NA" or re.fullmatch(r"-?(\d+\.\d\d|nan|inf)", f[6])
        for s in f[7:11]:
            assert s[0] == '"' and s[-1] == '"'
        if ref:
            assert len(f[10]) == 33 and all(ch in code for ch in f[10][1:-1])
        for c in f[11:11 + nconst]:
            assert re.fullmatch(r"-?\d+\.\d{10}", c) or c in ("nan", "-nan", "inf", "-inf")
        if np.isfinite(ll):
            total += 1
            mine = loglik(eval_expr(f[8][1:-1], x), y, sd)
            agree += np.isfinite(mine) and abs(mine - ll) <= max(1e-3 * abs(ll), 0.02)
    say("rows whose printed program reproduces the printed likelihood:", agree, "of", total)
    assert total > 0 and agree >= 0.9 * total


def header_lines(out, N, steps, outer, seed, plen, d):
    h = ["# -----------------------------------------------------------------", "# -- Parameters:",
         "# -----------------------------------------------------------------", f"# \tInput data path: {DATA}",
         f"# \tOutput path: {out}", f"# \tMCMC Iterations (per block): {steps}", f"# \tBlocks: {outer}",
         f"# \tN chains: {N}", f"# \tSeed: {seed}", f"# \tMax program length: {plen}", f"# \tN Constants: {plen}",
         "# \tGPU: 0", f"# \tNumber of data points: {len(d)}", "#", "#",
         "# -----------------------------------------------------------------", "# -- Data:",
         "# -----------------------------------------------------------------"]
    h += ["# \t%f\t%f\t%f" % tuple(map(float, r)) for r in d]
    return h + ["#", "#"]


def test_cli_reference_semantics(cli, tmp_path):
    d = np.loadtxt(DATA).astype(np.float32)
    out = tmp_path / "a"
    text = run(cli, out, "--N", "256", "--steps", "100", "--outer", "4", "--seed", "7", "--show-constants")
    lines = text.splitlines()
    hdr = header_lines(str(out), 256, 100, 4, 7, 31, d)
    assert lines[:len(hdr)] == hdr
    check_rows(rows_of(text), 256, [1, 2, 3], 31, d, ref=True)
    log = open(os.path.join(out, "log.txt")).read().splitlines()
    assert log[:len(hdr)] == hdr
    assert len([ln for ln in log if not ln.startswith("#")]) == 4
    # replayable from the seed; --burn and --thin select blocks o > burn && o % thin == 0
    text2 = run(cli, tmp_path / "b", "--N", "256", "--steps", "100", "--outer", "4", "--seed", "7",
                "--show-constants")
    assert rows_of(text2) == rows_of(text)
    text3 = run(cli, tmp_path / "c", "--N", "64", "--steps", "50", "--outer", "7", "--seed", "7", "--burn", "2",
                "--thin", "2")
    assert sorted({int(r.split("\t")[1]) for r in rows_of(text3)}) == [4, 6]
    # samples.txt is appended to, never truncated
    text4 = run(cli, tmp_path / "c", "--N", "64", "--steps", "50", "--outer", "2", "--seed", "8")
    assert text4.startswith(text3) and len(rows_of(text4)) == len(rows_of(text3)) + 64


def test_cli_config_semantics_and_state_files(cli, tmp_path):
    d = np.loadtxt(DATA).astype(np.float32)
    st = str(tmp_path / "state.bin")
    text = run(cli, tmp_path / "a", "--semantics", "config", "--cap", "32", "--N", "256", "--steps", "200",
               "--outer", "3", "--seed", "5", "--show-constants", "--state-out", st)
    check_rows(rows_of(text), 256, [1, 2], 32, d, ref=False)
    assert os.path.getsize(st) == 32 + 256 * 32 * 5 + 256 * 4
    # the resumed run numbers its blocks on from the saved run's three
    text2 = run(cli, tmp_path / "b", "--semantics", "config", "--cap", "32", "--N", "256", "--steps", "200",
                "--outer", "2", "--seed", "6", "--state-in", st)
    assert sorted({int(r.split("\t")[1]) for r in rows_of(text2)}) == [3, 4] and len(rows_of(text2)) == 512
    r = subprocess.run([cli, "--in", DATA, "--out", str(tmp_path / "c"), "--quiet", "--semantics", "config",
                        "--cap", "64", "--N", "256", "--state-in", st], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "does not match" in r.stderr


def test_cli_op_sets(cli, tmp_path):
    """--ops chooses the reference's op set (main.cu:51-86): the code column prints the op
    set's PROGRAM_CODE, and the printed programs (physical constants by name) reproduce
    the printed likelihoods."""
    d = np.loadtxt(DATA).astype(np.float32)
    for name, code in (("physical", "B2epcghm:o.~Kab+-_*"), ("polynomial", "1CAB+-_#*@^")):
        text = run(cli, tmp_path / name, "--ops", name, "--N", "256", "--steps", "100", "--outer", "3",
                   "--seed", "3", "--show-constants")
        rows = rows_of(text)
        check_rows(rows, 256, [1, 2], 31, d, ref=True, code=code)
        used = {ch for r in rows for ch in r.split("\t")[10][1:-1]}
        say(name, "codes printed:", "".join(sorted(used)))
        assert len(used) >= len(code) - 2


@pytest.mark.parametrize("args", [
    ("--N", "256", "--steps", "100"),                                                   # heap-31, quantile rule
    ("--N", "256", "--steps", "100", "--rng", "lcg", "--schedule", "lane"),             # the reference's LCG
    ("--semantics", "config", "--cap", "32", "--N", "256", "--steps", "100"),           # postfix, MAP rule
    ("--semantics", "config", "--cap", "32", "--N", "128", "--steps", "60", "--proposal", "mala"),
], ids=["ref", "ref-lcg", "config", "config-mala"])
def test_cli_resume_is_bit_identical_to_one_run(cli, tmp_path, args):
    """Four blocks in one run, and two runs of two blocks joined by --state-out/--state-in,
    print the same rows and leave the same state file: the state file carries the seed,
    the Philox clock and the block count, and the resumed run skips the saved run's
    block seeds."""
    one, half, end = (str(tmp_path / f) for f in ("one.bin", "half.bin", "end.bin"))
    text = run(cli, tmp_path / "one", *args, "--outer", "4", "--seed", "11", "--show-constants", "--state-out", one)
    run(cli, tmp_path / "two", *args, "--outer", "2", "--seed", "11", "--show-constants", "--state-out", half)
    text2 = run(cli, tmp_path / "two", *args, "--outer", "2", "--seed", "99", "--show-constants",
                "--state-in", half, "--state-out", end)
    rows, rows2 = rows_of(text), rows_of(text2)
    say("resume", args, "rows", len(rows), len(rows2))
    assert len(rows) == 3 * int(args[args.index("--N") + 1])
    assert rows2 == rows
    assert open(one, "rb").read() == open(end, "rb").read()
    assert "# \tSeed: 11" in text2.split("# \tSeed: 11", 1)[1]   # the resumed header prints the saved seed


def test_cli_tops_file_is_sort_g_k3_tail(cli, tmp_path):
    """--tops K writes tops.txt as run-all.sh:45 does: sort -g -k3 samples.txt | tail -n K."""
    run(cli, tmp_path / "t", "--N", "256", "--steps", "50", "--outer", "4", "--seed", "3", "--tops", "300")
    got = open(tmp_path / "t" / "tops.txt", "rb").read()
    s = subprocess.run(["sort", "-g", "-k3", str(tmp_path / "t" / "samples.txt")], capture_output=True, check=True,
                       env=dict(os.environ, LC_ALL="C")).stdout.splitlines(keepends=True)
    assert got == b"".join(s[-300:])
    keys = [float(ln.split(b"\t")[2]) for ln in got.splitlines() if not ln.startswith(b"#")]
    assert keys == sorted(keys)

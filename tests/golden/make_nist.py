"""Copy the reference's NIST StRD fixtures into tests/golden/nist/ (data files) and
extract their certified values into tests/golden/nist_certified.json.

Reads data-sources/NIST/*/{data.txt,*.dat} under /root/reference (the reference's
own fixtures; NIST data are US-government works).  Run once in the build container:
    python tests/golden/make_nist.py
"""
import json
import os
import re
import shutil

SRC = "/root/reference/data-sources/NIST"
HERE = os.path.dirname(os.path.abspath(__file__))


def certified(datfile):
    txt = open(datfile, encoding="latin-1").read().replace("\r", "")
    params = {}
    for m in re.finditer(r"^\s*(b\d+)\s*=\s*\S+\s+\S+\s+([-+0-9.Ee]+)\s+([-+0-9.Ee]+)", txt, re.M):
        params[m.group(1)] = float(m.group(2))
    rss = float(re.search(r"Residual Sum of Squares:\s*([-+0-9.Ee]+)", txt).group(1))
    rsd = float(re.search(r"Residual Standard Deviation:\s*([-+0-9.Ee]+)", txt).group(1))
    nobs = int(re.search(r"Number of Observations:\s*(\d+)", txt).group(1))
    # data section: "Data:  y  x" header, then one "y x" pair per line
    lines = txt.split("\n")
    start = max(i for i, l in enumerate(lines) if re.match(r"^Data:\s+y\s+x\s*$", l)) + 1
    ys, xs = [], []
    for l in lines[start:]:
        parts = l.split()
        if len(parts) == 2:
            ys.append(float(parts[0]))
            xs.append(float(parts[1]))
    return {"params": params, "rss": rss, "residual_sd": rsd, "n": nobs, "x": xs, "y": ys}


def main():
    out = {}
    for name in sorted(os.listdir(SRC)):
        d = os.path.join(SRC, name)
        if not os.path.isdir(d):
            continue
        os.makedirs(os.path.join(HERE, "nist", name), exist_ok=True)
        shutil.copyfile(os.path.join(d, "data.txt"), os.path.join(HERE, "nist", name, "data.txt"))
        dats = [f for f in os.listdir(d) if f.endswith(".dat")]
        if dats:
            out[name] = certified(os.path.join(d, dats[0]))
    json.dump(out, open(os.path.join(HERE, "nist_certified.json"), "w"), indent=1, sort_keys=True)
    print("datasets:", len(os.listdir(os.path.join(HERE, "nist"))), "certified:", sorted(out))


if __name__ == "__main__":
    main()

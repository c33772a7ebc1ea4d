"""Write tests/golden/philox_kat.json: Philox4x32-10 outputs computed by torch's
ATen/core/PhiloxRNGEngine.h (an installed source header), compiled on the host.

torch's engine(seed, subsequence, offset) uses key = seed and
counter = {offset_lo, offset_hi, subseq_lo, subseq_hi}; the first four draws are
one Philox4x32-10 block."""
import json, os, subprocess, tempfile
import torch

CASES = [(0, 0, 0), (1, 0, 0), (0x243F6A8885A308D3, 0x13198A2E03707344, 7),
         (0xFFFFFFFFFFFFFFFF, 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFFFFFFFFFF), (42, 123456789, 98765),
         (0xA4093822299F31D0, 5, 1)]
SRC = r'''
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
#include <cstdlib>
int main(int argc, char** argv) {
  for (int i = 1; i + 2 < argc; i += 3) {
    unsigned long long s = strtoull(argv[i], 0, 0), q = strtoull(argv[i+1], 0, 0), o = strtoull(argv[i+2], 0, 0);
    at::Philox4_32 e(s, q, o);
    unsigned a = e(), b = e(), c = e(), d = e();
    printf("%u %u %u %u\n", a, b, c, d);
  }
}
'''

def main():
    inc = os.path.join(os.path.dirname(torch.__file__), "include")
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "k.cpp"), "w").write(SRC)
        exe = os.path.join(d, "k")
        subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", inc, os.path.join(d, "k.cpp"), "-o", exe])
        args = [str(v) for c in CASES for v in c]
        out = subprocess.check_output([exe] + args, text=True).split("\n")
    rows = []
    for (s, q, o), line in zip(CASES, out):
        w = [int(v) for v in line.split()]
        rows.append({"key": [s & 0xFFFFFFFF, s >> 32],
                     "ctr": [o & 0xFFFFFFFF, o >> 32, q & 0xFFFFFFFF, q >> 32], "out": w})
    path = os.path.join(os.path.dirname(__file__), "philox_kat.json")
    json.dump({"source": "torch ATen/core/PhiloxRNGEngine.h " + torch.__version__, "vectors": rows},
              open(path, "w"), indent=1)
    print("wrote", path)

if __name__ == "__main__":
    main()

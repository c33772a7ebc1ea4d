"""Build libgpuropolis.so (HIP, gfx950) and the gpuropolis CLI in-tree.

    python -m gpuropolis_amd.build            # library + CLI
    python -m gpuropolis_amd.build --lib      # library only

hipcc cross-compiles for gfx950 without a GPU.  Every HIP translation unit gets
-Xclang -target-feature -Xclang -packed-fp32-ops so no packed-FP32 instructions are
emitted.
"""
import argparse
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "libgpuropolis.so")
CLI = os.path.join(PKG, "bin", "gpuropolis")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GPR_ARCH", "gfx950")

HIP_SOURCES = ["gpr_kernels.hip", "gpr_ref.hip", "gpr_api.hip"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics", "-Wall", "-Wno-unused-parameter", "-Wno-unused-function",
             "-I", os.path.join(ROOT, "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib(verbose=False, extra=()):
    objs = []
    os.makedirs(os.path.join(PKG, "_obj"), exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(PKG, "_obj", src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_cli(verbose=False):
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cpp")] if os.path.isdir(HOST) else []
    if not srcs:
        return None
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp,
          "-L", PKG, "-lgpuropolis", "-Wl,-rpath,$ORIGIN/..", "-lm"], verbose)
    os.replace(tmp, CLI)
    return CLI


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args(argv)
    build_lib(args.verbose)
    if not args.lib:
        build_cli(args.verbose)


if __name__ == "__main__":
    sys.exit(main())

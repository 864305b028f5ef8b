"""Builds libclothsim.so (the HIP product) in-tree with hipcc for gfx950.

    python -m clothsimulator_amd.build [--jobs N]

Every translation unit is compiled with -ffp-contract=off (the bit-exact parts
depend on it) and with packed FP32 instructions disabled.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libclothsim.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC, "-x", "hip", "--offload-arch=gfx950"] + COMMON + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    objs = []
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, log in ex.map(_compile, srcs):
            noisy = [l for l in log.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if rc != 0:
                failed.append((src, log))
            elif verbose and noisy:
                print("\n".join(noisy), file=sys.stderr)
            objs.append(obj)
    if failed:
        for src, log in failed:
            print("=== %s\n%s" % (src, log), file=sys.stderr)
        raise RuntimeError("hipcc failed on %d file(s)" % len(failed))
    tmp = LIB + ".tmp"
    # --no-undefined: a symbol missing from the objects fails here, not at load time
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()

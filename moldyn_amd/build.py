"""Build libmoldyn_hip.so (gfx950) in-tree with hipcc.

python -m moldyn_amd.build   or   moldyn_amd.build.build()
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmoldyn_hip.so")
SOURCES = ["md_kernels.hip", "md_mlp.hip", "md_mlp_train.hip", "md_api.hip", "md_train_api.hip", "md_frames.hip", "md_init.hip",
           "md_noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else "hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # the FP64 paths restate Rust arithmetic, which never contracts a*b+c
    "-ffp-contract=off",
    # packed FP32 forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{msgs}")
    return out, msgs


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, msgs in results:
        if verbose and msgs.strip():
            print(msgs, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    shutil.move(tmp, LIB)
    return LIB


def build_host_checks(out_dir):
    """Host-only build of every source under AddressSanitizer and
    UndefinedBehaviorSanitizer, linked with csrc/host_checks.cpp into
    out_dir/host_checks (no device code is compiled, so the binary runs
    without a GPU).  Returns the binary's path."""
    os.makedirs(out_dir, exist_ok=True)
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all"]
    # the device code is built unoptimised: only the host code runs
    base = [a for a in FLAGS if a != "-O3"] + ["-O1", "-g", "-fno-omit-frame-pointer", "-Xarch_device", "-O0"]

    def one(src):
        out = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
        r = subprocess.run([HIPCC] + base + san + ["-c", os.path.join(CSRC, src), "-o", out],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host-only build of {src} failed:\n{r.stdout}{r.stderr}")
        return out

    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(one, SOURCES))
    drv = os.path.join(out_dir, "host_checks.o")
    r = subprocess.run([HIPCC, "-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-fPIC"] + san +
                       ["-I" + os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, "host_checks.cpp"),
                        "-o", drv], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host_checks.cpp failed:\n" + r.stdout + r.stderr)
    exe = os.path.join(out_dir, "host_checks")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950"] + san + ["-o", exe, drv] + objs + ["-lpthread"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host_checks link failed:\n" + r.stdout + r.stderr)
    return exe


if __name__ == "__main__":
    print(build(verbose=True))

"""Build the gfx950 HIP library libfivo2d.so in-tree (csrc/), with hipcc.

Every HIP translation unit is compiled with the packed-FP32 target feature off
(all arithmetic here is FP64; the flag keeps hipcc from emitting packed FP32
instructions in any helper code) and with multiply-add fusion limited to single
expressions (-ffp-contract=on). Usage: python build.py [--jobs N]
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(CSRC, "libfivo2d.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["fivo_kernels.hip", "fivo_batch_kernels.hip", "fivo_relax.hip", "fivo_api.hip", "fivo_batch.hip",
           "fivo_mlp.hip", "fivo_freezeout.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fno-gpu-rdc",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # multiply-adds fuse only within one source expression, so the same body
         # rounds the same way in every kernel it is inlined into
         "-ffp-contract=on",
         "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]


def _compile(src):
    obj = os.path.join(CSRC, src.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs=5, verbose=False):
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        res = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in res]
    if verbose:
        for _, err in res:
            if err.strip():
                print(err)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(OUT + ".tmp", OUT)
    for o in objs:
        os.remove(o)
    return OUT


if __name__ == "__main__":
    j = int(sys.argv[sys.argv.index("--jobs") + 1]) if "--jobs" in sys.argv else 3
    print(build(j, verbose=True))

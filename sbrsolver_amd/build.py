"""Builds libsbr_amd.so (gfx950) in-tree with hipcc.

All translation units are compiled with -ffp-contract=off: the trace's discrete decisions must come
out of the same IEEE operation sequence as the reference's FP64 code, so no multiply-add may be fused.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libsbr_amd.so")
SOURCES = [
    os.path.join(HERE, "csrc", "engine.hip"),
    os.path.join(HERE, "csrc", "ingest.hip"),
    os.path.join(HERE, "csrc", "capi.cpp"),
    os.path.join(HERE, "csrc", "serve.cpp"),
    os.path.join(HERE, "csrc", "host", "preprocessor.cpp"),
    os.path.join(HERE, "csrc", "host", "rtsolver.cpp"),
    os.path.join(HERE, "csrc", "host", "ray.cpp"),
    os.path.join(HERE, "csrc", "host", "fieldcompute.cpp"),
]
DEMO = os.path.join(HERE, "sbr_demo")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
]


def build(verbose=False, jobs=4):
    from concurrent.futures import ThreadPoolExecutor
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(HERE, "csrc", os.path.basename(src) + ".o")
        cmds.append([HIPCC] + FLAGS + ["-x", "hip", "-c", src, "-o", obj])
        objs.append(obj)

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, cmds))
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    # The reference's main() against the library (links libsbr_amd.so from its own directory).
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                           os.path.join(HERE, "csrc", "host", "sbr_demo.cpp"), "-o", DEMO + ".tmp", "-L", HERE,
                           "-lsbr_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(DEMO + ".tmp", DEMO)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)

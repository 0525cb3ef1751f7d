"""Build libhqc_amd.so from hqc_amd/csrc/*.hip for gfx950 (hipcc, in-tree).

    python -m hqc_amd.build          # compile (always) and report the path
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhqc_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep hipcc from emitting packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def build(verbose=False):
    """Compile every kernel source into one shared library; returns its path."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    cmds, objs = [], []
    for src in sources():
        obj = os.path.join(HERE, "_build", os.path.basename(src) + ".o")
        cmds.append([HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-c", src, "-o", obj])
        objs.append(obj)

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    # one hipcc per source, a few at a time (each takes seconds to a minute)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(compile_one, cmds))
    tmp = LIB + ".tmp"
    # rocBLAS for the SCF driver's batched GEMMs (scf.hip)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + \
        ["-L/opt/rocm/lib", "-lrocblas"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

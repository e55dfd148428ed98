"""Build recipe for the engine library (gfx950, in-tree).

``python -m coldcore_amd.build`` compiles ``csrc/dct_engine.hip`` with hipcc
into ``coldcore_amd/libcoldcore_covering.so``, the drop-in for the
reference's ``build/plugins/libcoldcore_covering.so``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
#: translation units: the ABI (walks, reductions, host logic) and one per
#: sweep-kernel family, compiled in parallel and linked into one library
UNITS = ("dct_engine.hip", "sweep_hamming.hip", "sweep_lee.hip",
         "sweep_linf.hip", "sweep_round.hip", "sweep_round_q67.hip",
         "sweep_round_q810.hip", "sweep_packed_q2.hip",
         "sweep_packed_q3.hip", "sweep_packed_q4.hip", "sweep_packed_q5.hip",
         "mlp.hip", "batch.hip", "noalloc.hip")
LIB = os.path.join(HERE, "libcoldcore_covering.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-fvisibility=hidden",
    # packed fp32 ops are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=True, jobs=None):
    from concurrent.futures import ThreadPoolExecutor
    odir = os.path.join(HERE, "_obj")
    os.makedirs(odir, exist_ok=True)
    objs = [os.path.join(odir, u.replace(".hip", ".o")) for u in UNITS]
    cmds = [[HIPCC, *FLAGS, "-c", "-o", o, os.path.join(CSRC, u)]
            for u, o in zip(UNITS, objs)]
    if verbose:
        for c in cmds:
            print(" ".join(c), flush=True)
    jobs = jobs or min(len(cmds), os.cpu_count() or 1, 16)
    with ThreadPoolExecutor(jobs) as ex:
        for r in ex.map(lambda c: subprocess.run(c, check=True), cmds):
            pass
    tmp = LIB + ".tmp"
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)

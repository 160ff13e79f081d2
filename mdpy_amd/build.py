"""Build libmdx.so in-tree with hipcc for gfx950 (no CMake, no torch extension machinery).

    python -m mdpy_amd.build            # or mdpy_amd.build.build()

Flags: -O3, --offload-arch=gfx950, and packed-FP32 instruction generation disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``) so the compiler never emits
v_pk_{add,mul,fma}_f32.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmdx.so")
SOURCES = ["mdx_api.hip", "mdx_nlist.hip", "mdx_nonbonded.hip", "mdx_pme.hip", "mdx_bonded.hip", "mdx_mlp.hip",
           "mdx_noalloc.hip", "mdx_analyse.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-command-line-argument"]


def build(verbose=False, extra=()):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [HIPCC] + FLAGS + list(extra) + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmdx.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

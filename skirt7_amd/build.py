"""Builds libskirt7_amd.so for gfx950 in-tree (skirt7_amd/lib/).

All HIP code is compiled with -ffp-contract=off, so the FP64 walks repeat the reference's
arithmetic without fused multiply-adds, and with packed FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "csrc", f) for f in ("sg_kernels.hip", "sg_fit.hip", "sg_treebuild.hip", "sg_build.cpp", "sg_voronoi.cpp", "sg_io.cpp", "sg_ga.cpp")]
OUT = os.path.join(HERE, "lib", "libskirt7_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
         "-Wno-unused-command-line-argument"]


def build(verbose=False, out=OUT, defines=()):
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-D" + d for d in defines] + ["-o", tmp] + SRC + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libskirt7_amd.so")
    if verbose:
        sys.stderr.write("\n".join(l for l in (r.stdout + r.stderr).splitlines() if "error" in l or "warning: " in l
                                   and "packed-fp32" not in l))
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))

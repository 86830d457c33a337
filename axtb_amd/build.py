"""Build of the HIP engine: ``libaxtb_hip.so`` in this package directory, for gfx950.

One translation unit (csrc/engine.hip includes the kernels) compiled by hipcc and linked with
rocBLAS (large plain FP64 products). Packed FP32 instructions are disabled for all device code
(``-packed-fp32-ops``).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build_engine(verbose=False):
    out = os.path.join(HERE, "libaxtb_hip.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, os.path.join(CSRC, "engine.hip"), "-o", tmp, "-lrocblas"]
    r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_engine(verbose=True))

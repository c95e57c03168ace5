"""Build libgtn_hip.so (the gfx950 HIP kernels behind the C ABI of include/gtn_hip.h) in-tree.

Usage:  python -m grassmanntn_amd.build     (or grassmanntn_amd.build.build() from Python)
Every .hip file under csrc/ is compiled with hipcc for gfx950 and linked into
grassmanntn_amd/libgtn_hip.so.  Compiles every time it is called.
"""
import os
import subprocess
import sys
import concurrent.futures as cf

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgtn_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
         # packed FP32 instructions are kept out of all device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # matrix-core accumulators stay in the VGPRs they are used from (no per-iteration copies
         # between accumulation and vector registers in the GEMM loops)
         "-Xarch_device", "-mllvm=-amdgpu-mfma-vgpr-form",
         "-Wno-unused-result"]


def _compile(src, obj):
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, err))
    return obj


def build(verbose=True):
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, s.replace(".hip", ".o")) for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(_compile, [os.path.join(CSRC, s) for s in srcs], objs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)

"""Build the CPU (AddressSanitizer) version of libgtn_hip.so from the unmodified HIP sources --
TEST INFRASTRUCTURE ONLY.  Output: tests/cpu_hip/_build/libgtn_cpu.so (not committed)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "grassmanntn_amd", "csrc")
OUT_DIR = os.path.join(HERE, "_build")
OUT = os.path.join(OUT_DIR, "libgtn_cpu.so")
CXX = "/opt/rocm/llvm/bin/clang++"
FLAGS = ["-std=c++20", "-O1", "-g", "-fPIC", "-fsanitize=address", "-shared-libasan", "-fno-omit-frame-pointer",
         "-Wno-unused-result", "-Wno-ignored-attributes", "-I", os.path.join(HERE, "include")]


def build():
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]
    srcs.append(os.path.join(HERE, "emu_runtime.cpp"))
    for s in srcs:
        o = os.path.join(OUT_DIR, os.path.basename(s) + ".o")
        cmd = [CXX] + FLAGS + ["-x", "c++", "-c", s, "-o", o]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError("CPU build of %s failed:\n%s" % (s, r.stderr[-4000:]))
        objs.append(o)
    r = subprocess.run([CXX, "-shared", "-fsanitize=address", "-shared-libasan", "-o", OUT] + objs + ["-lpthread"],
                       capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr)
    return OUT


def asan_runtime():
    return subprocess.run([CXX, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                          text=True).stdout.strip()


if __name__ == "__main__":
    print(build())
    print(asan_runtime())
    sys.exit(0)

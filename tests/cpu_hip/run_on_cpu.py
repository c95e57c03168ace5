"""Run test functions of a GPU test module (default tests/test_gpu_core.py) on the CPU against the
AddressSanitizer build of the library's HIP sources -- TEST INFRASTRUCTURE ONLY.

  LD_PRELOAD=$(clang++ -print-file-name=libclang_rt.asan-x86_64.so) ASAN_OPTIONS=detect_leaks=0 \\
      python tests/cpu_hip/run_on_cpu.py [--module=test_gpu_block] [--exclude=substr] [name-substring ...]
"""
import inspect
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

from kernel_emulator import emulate  # noqa: E402
import importlib  # noqa: E402


def combos(f):
    marks = [m for m in getattr(f, "pytestmark", []) if m.name == "parametrize"]
    out = [{}]
    for m in marks:
        names = [a.strip() for a in m.args[0].split(",")]
        new = []
        for c in out:
            for v in m.args[1]:
                v = v if len(names) > 1 else (v,)
                d = dict(c)
                d.update(dict(zip(names, v)))
                new.append(d)
        out = new
    return out


def main(args):
    mods = [a[len("--module="):] for a in args if a.startswith("--module=")]
    T = importlib.import_module(mods[0] if mods else "test_gpu_core")
    args = [a for a in args if not a.startswith("--module=")]
    filters = [a for a in args if not a.startswith("--exclude=")]
    excludes = [a[len("--exclude="):] for a in args if a.startswith("--exclude=")]
    lib = os.path.join(HERE, "_build", "libgtn_cpu.so")
    fails = 0
    with emulate(real_lib_path=lib):
        names = sorted((n for n in dir(T) if n.startswith("test_")),
                       key=lambda n: inspect.getsourcelines(getattr(T, n))[1])
        for n in names:
            if filters and not any(f in n for f in filters):
                continue
            f = getattr(T, n)
            for c in combos(f):
                if any(x in str(c) for x in excludes):
                    continue
                print("RUN ", n, str(c)[:90], flush=True)
                try:
                    f(**c)
                    print("PASS", n, flush=True)
                except Exception as e:  # report and continue
                    fails += 1
                    print("FAIL", n, type(e).__name__, str(e)[:400], flush=True)
    print("failures:", fails)
    return fails


if __name__ == "__main__":
    sys.exit(1 if main(sys.argv[1:]) else 0)

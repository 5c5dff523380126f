"""Builds the CPU emulation of libsprgpu (tests/emu/_build/libsprgpu_emu.so):
the library's own kernel and runtime sources compiled by g++ with
AddressSanitizer and UBSan against include/hip/hip_runtime.h, a CPU stand-in
for the HIP subset they use (each block's lanes run as threads).  Test
infrastructure only."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libsprgpu_emu.so")
CXX = ["g++", "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-ffp-contract=off", "-fPIC", "-w", "-pthread"]


def sources():
    src = os.path.join(ROOT, "core_amd", "csrc")
    deps = [os.path.join(src, f) for f in os.listdir(src) if f.endswith((".hip", ".cpp", ".h"))]
    deps += [os.path.join(HERE, f) for f in ("emu.cpp", "prep.py", "build.py")]
    deps += [os.path.join(HERE, "include", "hip", "hip_runtime.h"), os.path.join(ROOT, "include", "sprgpu.h")]
    return deps


def stale():
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    return any(os.path.getmtime(f) > t for f in sources())


def build(force=False):
    if not force and not stale():
        return LIB
    subprocess.check_call([sys.executable, os.path.join(HERE, "prep.py")])
    objdir = os.path.join(OUT, "obj")
    os.makedirs(objdir, exist_ok=True)
    srcdir = os.path.join(OUT, "src")
    units = [os.path.join(srcdir, f) for f in sorted(os.listdir(srcdir)) if f.endswith(".cpp")]
    units.append(os.path.join(HERE, "emu.cpp"))
    inc = ["-I", os.path.join(HERE, "include"), "-I", srcdir, "-I", os.path.join(ROOT, "include")]

    def compile_one(u):
        o = os.path.join(objdir, os.path.basename(u)[:-4] + ".o")
        subprocess.check_call(CXX + inc + ["-c", u, "-o", o])
        return o

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, units))
    subprocess.check_call(CXX + ["-shared", "-o", LIB] + objs)
    return LIB


if __name__ == "__main__":
    print(build(force="-f" in sys.argv))

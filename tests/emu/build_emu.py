"""CPU build of colb_amd/csrc against the HIP emulation in tests/emu (g++, AddressSanitizer).
TEST INFRASTRUCTURE ONLY: the result is loaded by tests that check kernel indexing on the CPU."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "colb_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libabd_emu.so")
FLAGS = ["-std=c++20", "-DABD_EMULATION", "-O1", "-g", "-fPIC", "-fno-omit-frame-pointer", "-fsanitize=address",
         "-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "include"), "-pthread",
         "-Wno-unused-result", "-Wno-attributes"]


SOURCES = ("abd_kernels.hip", "abd_broad.hip", "abd_sort.hip", "abd_mlp.hip", "abd_direct.hip", "abd_api.hip")


def build(asan=True, jobs=4):
    os.makedirs(OUT, exist_ok=True)
    flags = FLAGS if asan else [f for f in FLAGS if f != "-fsanitize=address"]

    def compile_one(src):
        obj = os.path.join(OUT, src.replace(".hip", ".o"))
        r = subprocess.run(["g++", *flags, "-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj],
                           capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(r.stderr[-5000:])
        return obj

    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    r = subprocess.run(["g++", *flags, "-shared", "-o", LIB, *objs], capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr[-3000:])
    return LIB


def build_sort_driver():
    """Unit test of the radix sort and the multi-level scan (tests/emu/sort_driver.cpp)."""
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, "sort_driver")
    flags = [f for f in FLAGS if f != "-fsanitize=address"] + ["-I", HERE]
    r = subprocess.run(["g++", *flags, "-x", "c++", os.path.join(CSRC, "abd_sort.hip"), "-x", "c++",
                        os.path.join(HERE, "sort_driver.cpp"), "-o", exe], capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr[-3000:])
    return exe


def asan_runtime():
    return subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()


if __name__ == "__main__":
    print(build())

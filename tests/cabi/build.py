"""Compile abi_check.cpp with g++ against include/yoxel.h, linked to the
in-tree libyoxel.so (run path relative to the binary). Output:
tests/cabi/abi_check."""
import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
LIBDIR = ROOT / "yoxel-voxel_amd" / "lib"
BIN = HERE / "abi_check"


def build():
    rel = LIBDIR.relative_to(ROOT)
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-I", str(ROOT / "include"), str(HERE / "abi_check.cpp"),
           "-L", str(LIBDIR), "-lyoxel", f"-Wl,-rpath,$ORIGIN/../../{rel}", "-o", str(BIN)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"abi_check build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return BIN


if __name__ == "__main__":
    print(build())

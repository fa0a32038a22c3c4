"""Build libyoxel.so for gfx950 in-tree (hipcc; no JIT cache).

Run as a script or call build(). Every source is compiled on each call.
"""
import os
import pathlib
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libyoxel.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no fused multiply-add, so products and sums round as written
# (the traversal is bit-exact against the CPU restatement).
# Correctly rounded fp32 division and square root on the device.
# -packed-fp32-ops off: no v_pk_*_f32 generation.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"), "-I", str(CSRC),
    "-Wno-unused-command-line-argument",
]
ARCH = ["--offload-arch=gfx950"]


def _sources():
    return sorted(CSRC.glob("*.cpp")) + sorted(CSRC.glob("*.hip"))


def _compile(src: pathlib.Path, objdir: pathlib.Path) -> pathlib.Path:
    obj = objdir / (src.name + ".o")
    cmd = [HIPCC] + ARCH + COMMON + ["-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC] + COMMON + ["-x", "c++", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> pathlib.Path:
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, objdir), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC] + ARCH + ["-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs] + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

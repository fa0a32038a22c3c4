"""Host-side sanitizer run (AddressSanitizer + UndefinedBehaviorSanitizer).

  python tests/sanitize/run.py [--fuzz N]

1. libyoxel.so rebuilt with the sanitizers on its host code only (every hipcc
   flag behind -Xarch_host; device code is compiled as usual), into
   tests/sanitize/_build/. The CPU test suite (-m "not gpu") runs against it
   (YOXEL_LIB) with clang's ASan runtime preloaded.
2. The oracle's C restatement rebuilt with gcc's sanitizers; the CPU suite runs
   against it (YORACLE_LIB) with gcc's runtimes preloaded.
3. A bounded fuzz of yx_scene_load_vox / yx_scene_from_nodes on the sanitized
   library: truncated files, header node counts past the file, child ids out
   of range, cycles, random flags; each case in order, in one process that a
   sanitizer report ends.
Any report fails the run (halt on the first error); leak checking is off
(the Python interpreter itself holds allocations until exit).
"""
import argparse
import glob
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
CSRC = ROOT / "yoxel-voxel_amd" / "csrc"
BUILD = HERE / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
COMMON = ["-O1", "-g", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", str(ROOT / "include"), "-I", str(CSRC),
          "-Wno-unused-command-line-argument"]


def host_san():
    out = []
    for f in SAN:
        out += ["-Xarch_host", f]
    return out


def run(cmd, **kw):
    r = subprocess.run(cmd, **kw)
    if r.returncode != 0:
        raise SystemExit(f"failed ({r.returncode}): {' '.join(map(str, cmd))}")
    return r


def build_library():
    BUILD.mkdir(exist_ok=True)
    objs = []
    for src in sorted(CSRC.glob("*.cpp")) + sorted(CSRC.glob("*.hip")):
        obj = BUILD / (src.name + ".o")
        if src.suffix == ".hip":
            cmd = [HIPCC, "--offload-arch=gfx950"] + COMMON + host_san() + ["-c", str(src), "-o", str(obj)]
        else:
            cmd = [HIPCC] + COMMON + host_san() + ["-x", "c++", "-c", str(src), "-o", str(obj)]
        print("compile", src.name, flush=True)
        run(cmd)
        objs.append(str(obj))
    lib = BUILD / "libyoxel.so"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + host_san() + ["-o", str(lib)] + objs + ["-lpthread"])
    return lib


def build_oracle():
    BUILD.mkdir(exist_ok=True)
    lib = BUILD / "liboracle.so"
    srcs = sorted(glob.glob(str(ROOT / "oracle" / "*.c")))
    run(["gcc", "-O1", "-g", "-std=c99", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared"] + SAN +
        ["-o", str(lib)] + srcs + ["-lm"])
    return lib


def clang_asan_runtime():
    libs = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not libs:
        raise SystemExit("clang's ASan runtime not found under /opt/rocm/llvm")
    return libs[-1]


def gcc_runtimes():
    return [subprocess.run(["gcc", f"-print-file-name={n}"], capture_output=True, text=True).stdout.strip()
            for n in ("libasan.so", "libubsan.so")]


def sanitizer_env(preload, **extra):
    env = dict(os.environ)
    env["LD_PRELOAD"] = ":".join(preload)
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env.update(extra)
    return env


FUZZ = r'''
import ctypes, os, struct, sys
import numpy as np
sys.path.insert(0, os.environ["YX_ROOT"])
import __graft_entry__ as E
yx = E.load_package()
L = yx._lib.lib()
rng = np.random.default_rng(int(os.environ.get("FUZZ_SEED", "1")))
n_cases = int(os.environ["FUZZ_N"])
base = yx.SVOData.build_synthetic(depth=5, num_spheres=6, radius_min=2.0, radius_max=6.0, heightfield=True, seed=3)
nodes0 = base.nodes().copy()
root0 = base.GetRoot()
path = os.path.join(os.environ["FUZZ_DIR"], "case.vox")
outcomes = {}
for i in range(n_cases):
    nodes = nodes0.copy()
    root = root0
    kind = i % 6
    hdr_size = nodes.shape[0]
    data = None
    if kind == 0:  # truncated anywhere
        blob = struct.pack("<4I", root, 0, 0, hdr_size) + nodes.tobytes()
        data = blob[: int(rng.integers(0, len(blob)))]
    elif kind == 1:  # header claims more (or far more) nodes than the file holds
        hdr_size = int(rng.choice([hdr_size + 1, hdr_size * 2, 0xFFFFFFFF, int(rng.integers(0, 1 << 31))]))
    elif kind == 2:  # child ids out of range
        for _ in range(int(rng.integers(1, 5))):
            r, s = int(rng.integers(0, nodes.shape[0])), int(rng.integers(0, 8))
            if not (nodes[r, 0] >> s) & 1:
                nodes[r, 2 + s] = int(rng.integers(nodes.shape[0], 0x80000000))
    elif kind == 3:  # a cycle: a child slot pointing at an ancestor or itself
        r, s = int(rng.integers(0, nodes.shape[0])), int(rng.integers(0, 8))
        nodes[r, 0] &= ~np.uint32(1 << s)
        nodes[r, 0] &= ~np.uint32(1 << (8 + s))
        nodes[r, 2 + s] = int(rng.choice([root, r]))
    elif kind == 4:  # random flags and children
        r = int(rng.integers(0, nodes.shape[0]))
        nodes[r] = rng.integers(0, 1 << 32, size=10, dtype=np.uint64).astype(np.uint32)
    else:  # bad root
        root = int(rng.choice([nodes.shape[0], 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF]))
    if data is None:
        data = struct.pack("<4I", root, 0, 0, hdr_size) + nodes.tobytes()
    with open(path, "wb") as f:
        f.write(data)
    h = ctypes.c_void_p()
    rc = L.yx_scene_load_vox(path.encode(), ctypes.byref(h))
    if rc == 0:
        info = yx._lib.SceneInfo()
        L.yx_scene_get_info(h, ctypes.byref(info))
        ptr, nn, rt = ctypes.c_void_p(), ctypes.c_uint32(), ctypes.c_uint32()
        L.yx_scene_nodes(h, ctypes.byref(ptr), ctypes.byref(nn), ctypes.byref(rt))
        if nn.value:  # read every word the scene hands out
            words = np.ctypeslib.as_array((ctypes.c_uint32 * (nn.value * 10)).from_address(ptr.value))
            int(words.sum())
        filled = ctypes.c_uint64()
        L.yx_scene_fill_interior(h, ctypes.byref(filled))
        L.yx_scene_free(h)
    outcomes[rc] = outcomes.get(rc, 0) + 1
print("fuzz cases", n_cases, "status counts", dict(sorted(outcomes.items())), flush=True)
'''


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fuzz", type=int, default=1200, help="fuzz cases")
    ap.add_argument("--tests", default="tests", help="pytest target")
    args = ap.parse_args()
    lib = build_library()
    olib = build_oracle()
    env_a = sanitizer_env([clang_asan_runtime()], YOXEL_LIB=str(lib))
    print("pass 1: CPU tests against the host-sanitized library", flush=True)
    run([sys.executable, "-m", "pytest", "-x", "-q", "-s", "-m", "not gpu", "-p", "no:cacheprovider", args.tests],
        cwd=ROOT, env=env_a)
    print("pass 2: CPU tests against the sanitized oracle", flush=True)
    env_b = sanitizer_env(gcc_runtimes(), YORACLE_LIB=str(olib))
    run([sys.executable, "-m", "pytest", "-x", "-q", "-s", "-m", "not gpu", "-p", "no:cacheprovider", args.tests],
        cwd=ROOT, env=env_b)
    print(f"pass 3: {args.fuzz} .vox fuzz cases against the host-sanitized library", flush=True)
    fuzz_dir = BUILD / "fuzz"
    fuzz_dir.mkdir(exist_ok=True)
    env_c = dict(env_a, YX_ROOT=str(ROOT), FUZZ_N=str(args.fuzz), FUZZ_DIR=str(fuzz_dir))
    run([sys.executable, "-c", FUZZ], cwd=ROOT, env=env_c)
    print("sanitizer run clean", flush=True)


if __name__ == "__main__":
    main()

"""Build recipe for libgrid2grid_amd.so (HIP, gfx950) and the test oracle.

The library is compiled in-tree with hipcc so the built .so travels with the
source tree. `python -m grid2grid_amd.build` builds everything.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgrid2grid_amd.so")

# (source, object name, defines): the gather's variants are compiled once per
# (workgroup size, feature level) family so that they build in parallel
SOURCES = [("g2g_kernels.hip", "g2g_kernels", []), ("g2g_host.cpp", "g2g_host", []),
           ("g2g_serve.cpp", "g2g_serve", []), ("g2g_rank_pairs.hip", "g2g_rank_pairs", []),
           ("g2g_cache.cpp", "g2g_cache", []), ("g2g_noalloc.hip", "g2g_noalloc", [])] + [
    ("g2g_gather_variants.hip", f"g2g_gather_{th}_{feat}", [f"-DG2G_THREADS={th}", f"-DG2G_FEAT={feat}"])
    for th in (256, 512) for feat in (0, 1, 2)]

HIPCC_FLAGS = [
    "-std=c++17",
    "-O3",
    "--offload-arch=gfx950",
    # code object v5: loadable by the HIP runtime bundled with PyTorch too
    "-mcode-object-version=5",
    "-fPIC",
    "-shared",
    # no FMA contraction anywhere: the scaled path must round like the reference
    "-ffp-contract=off",
    # packed-FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # atomics stay as written: the gather's work counter is read after the
    # work it overlaps, not right after it is issued
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def _compile(src, obj, defines, verbose):
    cmd = [hipcc()] + [f for f in HIPCC_FLAGS if f != "-shared"] + defines + [
        "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
    noise = [l for l in (r.stderr or "").splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n" + "\n".join(noise))
    if verbose and noise:
        print("\n".join(noise), file=sys.stderr)


def build_library(verbose=False, incremental=False):
    """Every source compiled to an object (in parallel, under _build/), then
    linked. incremental: keep objects newer than their source and every header
    of csrc/ and include/ (for development; the default rebuilds all)."""
    from concurrent.futures import ThreadPoolExecutor
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    headers += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
                if f.endswith(".h")]
    newest_header = max((os.path.getmtime(h) for h in headers), default=0)
    jobs = []
    objs = []
    for name, oname, defines in SOURCES:
        src = os.path.join(CSRC, name)
        obj = os.path.join(obj_dir, oname + ".o")
        objs.append(obj)
        fresh = (incremental and os.path.exists(obj) and os.path.getmtime(obj) > os.path.getmtime(src)
                 and os.path.getmtime(obj) > newest_header)
        if not fresh:
            jobs.append((src, obj, defines))
    with ThreadPoolExecutor(max_workers=max(1, min(len(jobs), 8))) as ex:
        for f in [ex.submit(_compile, src, obj, defines, verbose) for src, obj, defines in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-lpthread", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    return LIB


HOST_REPLAYS = {"gather_replay.cpp"}


def build_cpp_tests(verbose=False):
    """C++ programs under tests/cpp written against the header-only facade
    (include/grid2grid), linked to the library; built into tests/cpp/bin."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "bin")
    os.makedirs(out_dir, exist_ok=True)
    built = []
    for f in sorted(os.listdir(src_dir)):
        if not f.endswith(".cpp"):
            continue
        if f in HOST_REPLAYS:  # host-only replays of the kernels' index arithmetic: shared objects
            out = os.path.join(out_dir, "lib" + f[:-4] + ".so")
            cmd = [hipcc(), "-std=c++17", "-O2", "--offload-arch=gfx950", "-fPIC", "-shared",
                   "-I" + os.path.join(ROOT, "include"),
                   "-I" + CSRC, os.path.join(src_dir, f), "-o", out]
        else:
            out = os.path.join(out_dir, f[:-4])
            cmd = [hipcc(), "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
                   os.path.join(src_dir, f), "-L" + HERE, "-lgrid2grid_amd",
                   "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../../../grid2grid_amd", "-o", out]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {f} failed:\n{r.stderr}")
        built.append(out)
    return built


def build_oracle(verbose=False):
    """Compile the C oracle and, when the reference tree is present, the
    reference harness (oracle/_ref). Test infrastructure only."""
    cmd = ["make", "-s", "-C", os.path.join(ROOT, "oracle")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    if verbose and (r.stdout or r.stderr):
        print(r.stdout + r.stderr)


def main():
    build_library(verbose=True)
    build_cpp_tests(verbose=True)
    build_oracle(verbose=True)
    print("built", LIB)


if __name__ == "__main__":
    main()

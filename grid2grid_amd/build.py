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

SOURCES = ["g2g_kernels.hip", "g2g_host.cpp", "g2g_serve.cpp", "g2g_rank_pairs.hip"]

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


def build_library(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + [
        "-I" + os.path.join(ROOT, "include"),
        "-I" + CSRC,
    ] + srcs + ["-lpthread", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
    noise = [l for l in (r.stderr or "").splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(noise))
    if verbose and noise:
        print("\n".join(noise), file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


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

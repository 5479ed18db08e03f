"""Build the gfx950 device library thoth_amd/lib/libthoth_mcl.so.

    python -m thoth_amd.build

Compiles the HIP kernels with hipcc for gfx950 only (FP64 kernels with
-ffp-contract=off, FP32 kernels with contraction), the host C++ with g++, and
links everything with hipcc.  Packed-FP32 instructions are disabled on every
device compile.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_obj"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libthoth_mcl.so"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
JK = PKG / "data" / "new-joe-kuo-6.21201.4096.txt"

DEVICE_FLAGS = ["-O3", "-std=c++20", "--offload-arch=gfx950", "-fPIC", "-fvisibility=hidden",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-I%s" % CSRC, "-I%s" % (ROOT / "include")]
HOST_FLAGS = ["-O2", "-std=c++2b", "-fPIC", "-Wall", "-Wextra", "-fvisibility=hidden", "-pthread", "-D__HIP_PLATFORM_AMD__",
              "-I%s" % os.path.join(ROCM, "include"), "-I%s" % CSRC, "-I%s" % (ROOT / "include"),
              '-DTHOTH_JK_PATH="%s"' % JK]

# (source, compiler, extra flags)
UNITS = [
    ("mcl_sweep_f64.hip", "hipcc", ["-ffp-contract=off"]),
    ("mcl_sweep_f32.hip", "hipcc", []),
    ("mcl_sweep_f32_uniform.hip", "hipcc", []),
    # five-equity uniform books: ILP-first scheduling (same occupancy, better
    # interleaving of the equities' steps; the one-equity kernels lose
    # occupancy under it and stay in the unit above)
    ("mcl_sweep_f32_uniform5.hip", "hipcc", ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]),
    ("mcl_util.hip", "hipcc", []),
    ("mcl_cv.hip", "hipcc", []),
    ("mcl_sweep_waves.hip", "hipcc", []),
    ("mcl_pathwise.hip", "hipcc", ["-ffp-contract=off"]),
    ("mcl_pathwise_f32.hip", "hipcc", []),
    ("mcl_noalloc.hip", "hipcc", []),
    ("mcl_prng_f64.hip", "hipcc", ["-ffp-contract=off"]),
    ("mcl_prng_f32.hip", "hipcc", []),
    ("mcl_prng.cpp", "g++", []),
    ("mcl_lsm.hip", "hipcc", ["-ffp-contract=off"]),
    ("mcl_gpu.cpp", "g++", []),
    ("mcl_books.cpp", "g++", []),
    ("mcl_market.cpp", "g++", []),
    ("mcl_host.cpp", "g++", []),
    ("mcl_cv_host.cpp", "g++", []),
    ("mcl_grid.cpp", "g++", []),
    ("mcl_split.cpp", "g++", []),
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    OBJ.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    jobs = []
    for src, cc, extra in UNITS:
        obj = OBJ / (src.rsplit(".", 1)[0] + ".o")
        if cc == "hipcc":
            cmd = [HIPCC] + DEVICE_FLAGS + extra + ["-c", str(CSRC / src), "-o", str(obj)]
        else:
            cmd = ["g++"] + HOST_FLAGS + extra + ["-c", str(CSRC / src), "-o", str(obj)]
        jobs.append((cmd, obj))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", str(tmp)] + [str(o) for _, o in jobs]
         + ["-L%s" % os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,%s" % os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

"""Build librskx.so in-tree with hipcc for gfx950.

Each source is compiled to its own object in parallel (they share no device
code: every kernel is launched from the file that defines it), then the
objects are linked into the shared library.
"""
import os
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["api.hip", "factor_kernels.hip", "sweep_kernels.hip", "tree_kernels.hip", "tree.cpp",
           "refine_kernels.hip", "zfactor_kernels.hip", "zsweep_kernels.hip", "mlp_kernels.hip",
           "elim_mfma.hip", "big_kernels.hip"]
LIB = os.path.join(HERE, "librskx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# the largest boxes' dense factorizations and products run through rocSOLVER /
# rocBLAS (the same sonames torch loads, so one copy of each per process)
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


# per-source flags: the register-resident elimination unrolls nested loops
# over up to 3 x 24 register slots; above LLVM's default size limit for
# "#pragma unroll" a partly unrolled loop would index the slot array
# dynamically and move it to scratch memory
EXTRA = {"factor_kernels.hip": ["-mllvm", "-pragma-unroll-threshold=100000"]}


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)


def build(verbose=False, src_dir=None, out=None):
    """Compile the sources of src_dir (default flam_amd/csrc) into out
    (default flam_amd/librskx.so)."""
    src_dir = src_dir or os.path.join(HERE, "csrc")
    out = out or LIB
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + src_dir]
    with tempfile.TemporaryDirectory(prefix="rskx_build_") as tmpd:
        jobs = []
        for s in SOURCES:
            obj = os.path.join(tmpd, os.path.splitext(s)[0] + ".o")
            jobs.append(([HIPCC, *FLAGS, *EXTRA.get(s, []), *inc, "-c", os.path.join(src_dir, s),
                          "-o", obj], obj))
        if verbose:
            for cmd, _ in jobs:
                print(" ".join(cmd))
        with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
            list(ex.map(_run, [c for c, _ in jobs]))
        tmp = out + ".tmp"
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *[o for _, o in jobs], *LIBS, "-o", tmp])
        os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))

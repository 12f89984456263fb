"""Build the gfx950 HIP library ppkmhd_amd/libppkmhd_amd.so with hipcc (in-tree).

    python -m ppkmhd_amd.build [--jobs N] [--verbose]

Kernel files are compiled once per arithmetic variant: ``fast`` (the compiler may contract
a*b+c into FMA; repeated denominators are inverted once) and ``strict`` (-ffp-contract=off,
every division as written: bit-identical to the reference's double arithmetic).  Every HIP
compile disables packed FP32 instructions.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libppkmhd_amd.so")
ARCH = "gfx950"


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the library)")


COMMON = [
    "-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-lambda-capture", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]

# (source, object, extra flags)
UNITS = [
    ("step_staged.hip", "step_staged_fast.o", ["-DPPK_NS=fast", "-ffp-contract=fast"]),
    ("step_staged.hip", "step_staged_strict.o",
     ["-DPPK_NS=strict", "-DPPK_EXACT_DIV", "-ffp-contract=off"]),
    ("fused.hip", "fused_fast.o", ["-DPPK_NS=fast", "-ffp-contract=fast"]),
    ("fused.hip", "fused_strict.o", ["-DPPK_NS=strict", "-DPPK_EXACT_DIV", "-ffp-contract=off"]),
    ("march3d.hip", "march3d_fast.o", ["-DPPK_NS=fast", "-ffp-contract=fast"]),
    ("march3d.hip", "march3d_strict.o",
     ["-DPPK_NS=strict", "-DPPK_EXACT_DIV", "-ffp-contract=off"]),
    ("batch.hip", "batch_fast.o", ["-DPPK_NS=fast", "-ffp-contract=fast"]),
    ("batch.hip", "batch_strict.o", ["-DPPK_NS=strict", "-DPPK_EXACT_DIV", "-ffp-contract=off"]),
    # the network's accumulators in VGPRs (no copies out of AGPRs before each ReLU)
    ("surrogate.hip", "surrogate_fast.o",
     ["-DPPK_NS=fast", "-ffp-contract=fast", "-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("surrogate.hip", "surrogate_strict.o",
     ["-DPPK_NS=strict", "-DPPK_EXACT_DIV", "-ffp-contract=off", "-mllvm",
      "-amdgpu-mfma-vgpr-form"]),
    ("hydro.hip", "hydro_fast.o", ["-DPPK_NS=fast", "-ffp-contract=fast"]),
    ("hydro.hip", "hydro_strict.o", ["-DPPK_NS=strict", "-DPPK_EXACT_DIV", "-ffp-contract=off"]),
    ("mlp.hip", "mlp.o", []),
    ("diag.hip", "diag.o", []),
    ("noalloc.hip", "noalloc.o", ["-ffp-contract=off"]),
    ("capi.cpp", "capi.o", ["--offload-host-only"]),
    ("config.cpp", "config.o", []),
    ("ics.cpp", "ics.o", ["-ffp-contract=off"]),
    ("ics_device.hip", "ics_device.o", ["-ffp-contract=off"]),
    ("io_vtk.cpp", "io_vtk.o", []),
    ("io_restart.cpp", "io_restart.o", ["--offload-host-only"]),
    ("batch_capi.cpp", "batch_capi.o", ["--offload-host-only"]),
    ("mlp_capi.cpp", "mlp_capi.o", ["--offload-host-only"]),
    ("hydro_capi.cpp", "hydro_capi.o", ["--offload-host-only", "-ffp-contract=off"]),
]


def _units():
    return list(UNITS)


def _compile(unit, verbose=False):
    src, obj, extra = unit
    cmd = [hipcc()] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD, obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs=None, verbose=False):
    """Compile every HIP/C++ unit for gfx950 and link the shared library.  Always rebuilds."""
    os.makedirs(BUILD, exist_ok=True)
    units = _units()
    jobs = jobs or min(len(units), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(u, verbose), units))
    objs = [os.path.join(BUILD, o) for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    sys.exit(main())

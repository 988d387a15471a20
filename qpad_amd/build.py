"""Build libqpad_hip.so (gfx950) in-tree with hipcc.

python -m qpad_amd.build  [--jobs N]

Every .hip source under qpad_amd/csrc is compiled to an object in
qpad_amd/_build and linked into qpad_amd/libqpad_hip.so.  Packed FP32
instructions are disabled (-packed-fp32-ops) for all device code.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqpad_hip.so")
ARCH = os.environ.get("QPAD_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-I", os.path.join(HERE, "..", "include")]


def _compile(src, force=False):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    deps = [src] + glob.glob(os.path.join(CSRC, "*.h")) + [os.path.join(HERE, "..", "include", "qpad_hip.h")]
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
        return obj, ""
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs=4, verbose=False, force=False):
    """Compile (every source when force, else the out-of-date ones) and link."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s_: _compile(s_, force), srcs))
    objs = [o for o, _ in results]
    if verbose:
        for o, w in results:
            if w.strip():
                print(w, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_diag()
    build_fortran_driver()
    return LIB


DIAG_SRC = os.path.join(CSRC, "qpad_diag.cpp")
DIAG_LIB = os.path.join(HERE, "libqpad_diag.so")
HDF5_ROOT = os.environ.get("QPAD_HDF5_ROOT", "/opt/conda")


def build_diag():
    """libqpad_diag.so: QPAD's HDF5 diagnostics (host C++ on HDF5 1.10 and
    libqpad_hip's C ABI).  Skipped, with a message, where no HDF5 is
    installed; the core library does not depend on it."""
    hdr = os.path.join(HDF5_ROOT, "include", "hdf5.h")
    if not os.path.exists(hdr):
        print(f"libqpad_diag.so not built: no HDF5 headers at {hdr}", file=sys.stderr)
        return None
    hlib = os.path.join(HDF5_ROOT, "lib")
    tmp = DIAG_LIB + ".tmp"
    # libstdc++ and libgcc linked in, and a RUNPATH (direct dependencies
    # only): the HDF5 prefix also holds an older libstdc++, which must not
    # shadow the system one that the ROCm runtime needs
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-I", os.path.join(HDF5_ROOT, "include"),
           DIAG_SRC, "-o", tmp, "-static-libstdc++", "-static-libgcc", "-L", HERE, "-lqpad_hip", "-L", hlib,
           "-lhdf5", "-Wl,--enable-new-dtags", f"-Wl,-rpath,$ORIGIN:{hlib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libqpad_diag.so build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, DIAG_LIB)
    return DIAG_LIB


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "qpad_slice_driver")


def build_fortran_driver():
    """qpad_amd/fortran/qpad_slice_driver: the Fortran module and a program
    that drives a 3D step through it in QPAD's order (amdflang).  Skipped,
    with a message, where amdflang is missing."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        print("qpad_slice_driver not built: no amdflang", file=sys.stderr)
        return None
    work = os.path.join(BUILD, "fortran")
    os.makedirs(work, exist_ok=True)
    mod_o = os.path.join(work, "qpad_hip_mod.o")
    tmp = DRIVER + ".tmp"
    for cmd in ([fc, "-O2", "-c", os.path.join(FORTRAN, "qpad_hip_mod.f03"), "-o", mod_o, "-J", work],
                [fc, "-O2", "-I", work, os.path.join(FORTRAN, "qpad_slice_driver.f03"), mod_o, "-L", HERE,
                 "-lqpad_hip", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--force", action="store_true", help="recompile every source")
    args = ap.parse_args()
    print(build(args.jobs, args.verbose, args.force))

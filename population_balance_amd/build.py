"""Build liblesgo_hip.so (gfx950) in-tree from csrc/*.hip with hipcc.

    python -m population_balance_amd.build [--jobs N]
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblesgo_hip.so")
ARCH = os.environ.get("LP_OFFLOAD_ARCH", "gfx950")
SOURCES = ["lp_fft.hip", "lp_lag.hip", "lp_convec.hip", "lp_api.hip", "lp_noalloc.hip"]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed-FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


FORTRAN = os.path.join(HERE, "fortran")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


def build_fortran(verbose=False):
    """Compile the ISO_C_BINDING module, the reference-signature shims
    (lesgo_dropin.f90) and the two Fortran hosts that use them with amdflang (the
    reference's language).  Skipped (returns None) when no Fortran compiler exists."""
    if not os.path.exists(FC):
        return None
    mod_dir = os.path.join(OBJ, "fortran")
    os.makedirs(mod_dir, exist_ok=True)
    outs = []
    for prog, mods in (("drop_in_demo", ["lesgo_amd_iface.f90"]),
                       ("dropin_check", ["lesgo_amd_iface.f90", "lesgo_dropin.f90"])):
        out = os.path.join(FORTRAN, prog)
        cmd = [FC, "-O2", f"-J{mod_dir}", *[os.path.join(FORTRAN, m) for m in mods],
               os.path.join(FORTRAN, prog + ".f90"), "-L" + HERE, "-llesgo_hip",
               "-Wl,-rpath," + HERE, "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {prog}:\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"built {out}")
        outs.append(out)
    return outs


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs or len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)

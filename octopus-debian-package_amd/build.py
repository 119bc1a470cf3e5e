"""Build liboctopus_hip.so (HIP, gfx950) in-tree.

    python octopus-debian-package_amd/build.py [--jobs N]

Every .hip file of csrc/ is compiled with hipcc for gfx950 and linked into
lib/liboctopus_hip.so.  Floating-point contraction is off for host and device code:
the kernels spell out every fused multiply-add with fma(), so results do not depend on
the compiler's contraction choices.  The Fortran iso_c_binding module (fortran/octhip_m.F90)
and its host driver are then compiled with amdflang into lib/octhip_driver.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIBNAME = "liboctopus_hip.so"
OBJDIR = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("OCTHIP_FC", "/opt/rocm/bin/amdflang")
FDIR = PKG / "fortran"
DRIVER = "octhip_driver"
ARCH = "gfx950"

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    noise = "is not a recognized feature"  # the host pass does not know the gfx feature flag
    msgs = [ln for ln in (r.stdout + r.stderr).splitlines() if ln.strip() and noise not in ln]
    if msgs:
        sys.stderr.write("\n".join(msgs) + "\n")


def build(jobs: int = 8, verbose: bool = False) -> Path:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [OBJDIR / (s.stem + ".o") for s in srcs]
    cmds = [[HIPCC, *COMMON, "-c", str(s), "-o", str(o)] for s, o in zip(srcs, objs)]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    # roctx: the library's ranges (profiling regions, C-ABI entry points) for rocprofv3 --marker-trace
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-lrocprofiler-sdk-roctx"])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    build_fortran(verbose)
    return out


def build_fortran(verbose: bool = False) -> Path | None:
    """amdflang: the bind(C) module and the driver program linked to liboctopus_hip.so."""
    if shutil.which(FC) is None and not Path(FC).exists():
        if verbose:
            print(f"{FC} not found: Fortran binding not built")
        return None
    mod = OBJDIR / "fmod"
    mod.mkdir(parents=True, exist_ok=True)
    exe = LIBDIR / DRIVER
    tmp = LIBDIR / (DRIVER + ".tmp")
    _run([FC, "-O2", f"-J{mod}", "-o", str(tmp), str(FDIR / "octhip_m.F90"), str(FDIR / "octhip_driver.F90"),
          f"-L{LIBDIR}", "-loctopus_hip", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, exe)
    if verbose:
        print(f"built {exe}")
    return exe


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    build(args.jobs, verbose=True)

"""Build libmeltgpu.so (the gfx950 HIP library) in-tree with hipcc, and the
Fortran binding module with amdflang when it is installed.

The library is compiled with -ffp-contract=off so that every expression rounds
exactly as the reference's Fortran (built without FMA) does, and with packed
FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libmeltgpu.so"
OBJDIR = PKG / "_build"
SOURCES = ["api.hip", "euler.hip", "poisson.hip", "ibm.hip", "lagr.hip", "diag.hip", "hit.hip", "remesh.hip"]
ARCH = os.environ.get("MELT_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def _source_stamp() -> float:
    files = [CSRC / f for f in SOURCES] + list(CSRC.glob("*.h")) + [ROOT / "include" / "melt_ibm.h"]
    return max(f.stat().st_mtime for f in files)


def build_library(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    OBJDIR.mkdir(exist_ok=True)
    stamp = _source_stamp()
    jobs = []
    for src in SOURCES:
        obj = OBJDIR / (Path(src).stem + ".o")
        jobs.append([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=min(len(jobs), 8)) as ex:
        for fut in [ex.submit(_run, j) for j in jobs]:
            fut.result()
    # every object must see the same headers (the context struct is shared)
    if _source_stamp() != stamp:
        raise RuntimeError("sources changed while the library was compiling; build again")
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *[str(OBJDIR / (Path(s).stem + ".o")) for s in SOURCES],
          "-o", str(tmp)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_fortran(verbose: bool = False) -> Path | None:
    """Compile the iso_c_binding module (the reference-named drop-ins)."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if Path("/opt/rocm/bin/amdflang").exists() else None)
    if not fc:
        return None
    src = PKG / "fortran" / "melt_ibm_f.f90"
    out = OBJDIR / "melt_ibm_f.o"
    OBJDIR.mkdir(exist_ok=True)
    _run([fc, "-c", "-fdefault-real-8", "-fdefault-double-8", "-O2", "-fPIC", f"-J{OBJDIR}", str(src), "-o", str(out)])
    if verbose:
        print(f"built {out}")
    return out


def build_fortran_program(src: Path, out: Path, verbose: bool = False) -> Path | None:
    """Link a Fortran program that uses the binding module against the library
    (run-time search path relative to the program, so the tree can move)."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if Path("/opt/rocm/bin/amdflang").exists() else None)
    mod = build_fortran()
    if not fc or mod is None:
        return None
    rel = os.path.relpath(PKG, out.parent)
    _run([fc, "-O2", "-fdefault-real-8", "-fdefault-double-8", f"-I{OBJDIR}", str(src), str(mod), f"-L{PKG}", "-lmeltgpu",
          f"-Wl,-rpath,$ORIGIN/{rel}", "-Wl,-rpath,/opt/rocm/lib", "-o", str(out)])
    if verbose:
        print(f"built {out}")
    return out


def main() -> None:
    build_library(verbose=True)
    build_fortran(verbose=True)


if __name__ == "__main__":
    sys.exit(main())

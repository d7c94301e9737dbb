"""ORACLE -- test infrastructure only.

Loads the C FP64 restatement of the melt_IBM substep (liborc.so, built by
oracle/Makefile) and, where present, the reference's own tridiag_periodic.f90
compiled under oracle/_ref.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py use this package; the product never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "liborc.so"
REF_TRIDIAG = HERE / "_ref" / "libref_tridiag.so"
REFERENCE = Path("/root/reference")


def build(with_ref: bool = True) -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "liborc.so"], check=True)
    if with_ref and (REFERENCE / "tridiag_periodic.f90").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", f"REFDIR={REFERENCE}"], check=True)


def load() -> C.CDLL:
    if not LIB.exists():
        build(with_ref=False)
    lib = C.CDLL(str(LIB))
    lib.orc_det_sin_export.restype = C.c_double
    lib.orc_det_sin_export.argtypes = [C.c_double]
    lib.orc_inv4_export.restype = C.c_int
    lib.orc_inv4_export.argtypes = [C.c_void_p]
    return lib


def load_ref_tridiag():
    """The reference's tridiag_periodic (Fortran, by-reference args) or None."""
    if not REF_TRIDIAG.exists():
        return None
    lib = C.CDLL(str(REF_TRIDIAG))
    fn = lib.tridiag_periodic_
    fn.restype = None
    fn.argtypes = [C.c_void_p] * 4 + [C.POINTER(C.c_int)] * 3
    return fn


def context(cfg):
    """A MeltContext driving the oracle library through the same ABI."""
    from melt_ibm_amd.api import MeltContext
    return MeltContext(cfg, lib=load(), prefix="orc_")

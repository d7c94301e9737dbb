"""The C ABI library and its host mirror, without a GPU: the library loads,
exports every entry point include/melt_ibm.h declares, and the Python binding
matches the header's structs."""
import ctypes as C
import re
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "melt_ibm.h").read_text()
    return sorted(set(re.findall(r"\b(melt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from melt_ibm_amd.api import load_library
    lib = load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert len(declared_symbols()) >= 45


def test_oracle_mirrors_the_abi():
    import oracle
    lib = oracle.load()
    # device-side services, and the precision form's controls (the oracle is the
    # FP64 reference path, it has no FP32 form to refine)
    skip = {"melt_device_ok", "melt_timing_enable", "melt_timing_select", "melt_timing_reset", "melt_timing_count", "melt_timing_name", "melt_timing_read",
            "melt_set_refinement", "melt_get_refinement_stats", "melt_graph_enable", "melt_stream",
            "melt_set_tuning", "melt_upload_field_device", "melt_divgck_noalloc",
            "melt_snapshot_size", "melt_snapshot_save", "melt_snapshot_load"}
    missing = [s for s in declared_symbols() if s not in skip and not hasattr(lib, "orc_" + s[len("melt_"):])]
    assert not missing, missing


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of melt_config / melt_body / melt_layout agree with what a C
    compiler lays out from the header (size and every field offset)."""
    from melt_ibm_amd.api import Body, _Config, _Layout
    prog = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/melt_ibm.h"', "int main(void){"]
    for cname, py in (("melt_config", _Config), ("melt_body", Body), ("melt_layout", _Layout)):
        prog.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            prog.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    prog.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(prog))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n") if line)
    for cname, py in (("melt_config", _Config), ("melt_body", Body), ("melt_layout", _Layout)):
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"


def test_timing_slot_names_without_device():
    from melt_ibm_amd.api import load_library
    lib = load_library()
    lib.melt_timing_name.restype = C.c_char_p
    names = [lib.melt_timing_name(i).decode() for i in range(lib.melt_timing_count())]
    assert "explicit_xsolve" in names and "fft_z_solve" in names and "substep" in names


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "not supported" in r.stdout


def test_fortran_binding_compiles():
    from melt_ibm_amd import build
    if build.build_fortran() is None:
        pytest.skip("amdflang not installed")

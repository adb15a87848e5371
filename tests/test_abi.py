"""The C-ABI library loads and exports every entry point include/bbpd.h declares (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

from bb_pd_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "bbpd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bbpd_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "bbpd_calculatedeformedlength" in names and "bbpd_solver_create" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_host_only_calls_work_without_gpu():
    L = _lib.lib()
    assert L.bbpd_version() >= 100
    assert L.bbpd_device_count() >= 0
    spec = _lib.BeamSpec(4, 4, 4, 1.0, 3.0, 0, 0.0, 0.0, 1.0)     # hf^2 = 9: ambiguous, rejected
    n, b = C.c_int64(), C.c_int64()
    assert L.bbpd_beam_count(C.byref(spec), C.byref(n), C.byref(b)) == -1
    assert b"integer" in L.bbpd_last_error()


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors of the header's structs have the C layout."""
    import subprocess
    structs = {"bbpd_problem": _lib.Problem, "bbpd_beam_spec": _lib.BeamSpec, "bbpd_observables": _lib.Observables,
               "bbpd_batch_spec": _lib.BatchSpec, "bbpd_region_status": _lib.RegionStatus}
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "bbpd.h"', "int main(void){"]
    for cname, py in structs.items():
        src.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            src.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(HEADER.parent), str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for line in filter(None, out):
        cname, field, val = line.split()
        py = structs[cname]
        got = C.sizeof(py) if field == "sizeof" else getattr(py, field).offset
        assert got == int(val), (cname, field, got, val)


def test_one_hip_runtime_per_process():
    """libbbpd_hip binds to the HIP/HSA runtime torch already loaded (two runtimes in one process
    leave the second without devices)."""
    _lib.lib()
    maps = Path("/proc/self/maps").read_text()
    hip = set(re.findall(r"\S*libamdhip64\S*", maps))
    hsa = set(re.findall(r"\S*libhsa-runtime64\S*", maps))
    assert len(hip) == 1 and len(hsa) == 1, (hip, hsa)

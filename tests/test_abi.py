"""CPU tests of the C ABI boundary: the library loads, exports every function
declared in include/phj_gpu.h, and the Python mirror's struct layouts equal the
C compiler's. No compute calls (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

import phj_gpu as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "phj_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:phj_status|uint32_t|const char \*)\s*(phj_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = G.load_library()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(G.exported_symbols())
    assert lib.phj_abi_version() == 1


def test_symbols_are_exported_from_the_shared_object():
    out = subprocess.run(["nm", "-D", "--defined-only", G.library_path()], capture_output=True, text=True).stdout
    syms = set(re.findall(r"\bT (phj_\w+)", out))
    assert set(declared_functions()) <= syms


STRUCTS = {
    "phj_column": G._Column, "phj_relation": G._Relation, "phj_tuning": G._Tuning, "phj_join_desc": G._JoinDesc,
    "phj_phase_timing": G._Phase, "phj_join_result": G._JoinResult, "phj_partitioned": G._Partitioned,
    "phj_checksum": G._Checksum, "phj_gen_spec": G._GenSpec, "phj_row_ref": G._RowRef,
}


def test_struct_layouts_match_the_c_header():
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "phj_gpu.h"', "int main(void){"]
    for cname, py in STRUCTS.items():
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    src.append('printf("phj_ht_cell %zu\\n", sizeof(phj_ht_cell)); return 0; }')
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "layout.c")
        exe = os.path.join(td, "layout")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), "-o", exe, c], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.split() for l in lines if l)
    for cname, py in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)
    assert int(got["phj_ht_cell"]) == 16  # src/HashTable.h:35-40


def test_context_without_device_reports_no_device():
    lib = G.load_library()
    p = C.c_void_p()
    rc = lib.phj_ctx_create(0, C.byref(p))
    if rc == G.OK:  # a HIP device is visible here
        lib.phj_ctx_destroy(p)
        pytest.skip("a HIP device is present")
    assert rc == G.ERR_NO_DEVICE


def test_null_context_is_rejected():
    lib = G.load_library()
    assert lib.phj_synchronize(None) == G.ERR_INVALID_ARGUMENT
    assert lib.phj_join(None, None, None) == G.ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("parts,passes,want", [(256, 1, [8]), (256, 2, [4, 4]), (1024, 3, [4, 3, 3]), (2, 1, [1])])
def test_radix_config_from_partitions_follows_cli(parts, passes, want):
    assert G.RadixConfig.from_partitions(parts, passes).pass_bits == want


def test_radix_config_rejects_too_many_passes():
    with pytest.raises(ValueError):
        G.RadixConfig.from_partitions(4, 3)

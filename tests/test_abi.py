"""The C-ABI library loads and exports every function include/dml.h declares
(no device compute: these run without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "dml.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|int64_t|const char \*)\s*(dml_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    import dml
    L = ctypes.CDLL(dml.LIB_PATH)
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(dml.EXPORTS) == names


def test_strerror_is_host_only():
    import dml
    L = dml.lib()
    assert L.dml_strerror(0) == b"ok"
    assert L.dml_strerror(-4) == b"particle out of tessellation"


def test_config_struct_matches_header_layout():
    import dml
    # dml_config: 3+1+1+9+9+3+3+3+3 doubles, 6 int32, u64, double, 4 int32
    assert ctypes.sizeof(dml.Config) == 8 * 35 + 4 * 6 + 8 + 8 + 4 * 4


def test_config_layout_equals_c_compiler(tmp_path):
    """the ctypes mirror against gcc's own layout of include/dml.h"""
    import dml
    import subprocess
    src = tmp_path / "lay.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "dml.h"\n'
                   'int main(void) { printf("%zu %zu %zu %zu\\n", sizeof(dml_config), '
                   'offsetof(dml_config, seed), offsetof(dml_config, nadj), '
                   'offsetof(dml_config, gcmc_mode)); return 0; }\n')
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, timeout=60)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    C = dml.Config
    assert got == [ctypes.sizeof(C), C.seed.offset, C.nadj.offset, C.gcmc_mode.offset]

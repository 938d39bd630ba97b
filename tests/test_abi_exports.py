"""The C-ABI libraries load without a GPU and export every function that a
header under include/ declares (the drop-in boundary: a StarPU application
built against these headers links against these libraries)."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBS = ["libstarpu_amd_kernels.so", "libstarpu_amd.so", "libstarpu_amd_chol.so"]


def declared_functions(path):
    """Names of the function prototypes in a C header: comments, preprocessor
    lines and brace bodies removed, then every statement of the form
    `<type words and stars> name(<params>)`; typedefs and struct members
    (function pointers are written `(*name)(...)`) do not match."""
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    src = re.sub(r"//[^\n]*", " ", src)
    # the C++ guards (extern "C" { ... }) sit in #ifdef __cplusplus blocks
    src = re.sub(r"#\s*ifdef\s+__cplusplus.*?#\s*endif", " ", src, flags=re.S)
    src = "\n".join(l for l in src.splitlines() if not l.lstrip().startswith("#"))
    # drop struct/enum/union bodies
    while True:
        new = re.sub(r"\{[^{}]*\}", ";", src)
        if new == src:
            break
        src = new
    names = []
    for stmt in src.split(";"):
        stmt = " ".join(stmt.split())
        if not stmt or stmt.startswith("typedef"):
            continue
        m = re.match(r"^(?:extern\s+)?[\w\s\*]+?\b([A-Za-z_]\w*)\s*\((.*)\)$", stmt)
        if m:
            names.append(m.group(1))
    return names


HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def test_headers_declare_functions():
    allnames = [n for h in HEADERS for n in declared_functions(h)]
    # spot checks that the parser sees the boundary of the reference's API
    for want in ("starpu_init", "starpu_task_insert", "starpu_matrix_data_register", "starpu_data_map_filters",
                 "starpu_hip_get_local_stream", "starpu_hash_crc32c_be_ptr", "sa_dgemm_nt", "chol_factor"):
        assert want in allnames, want
    assert len(allnames) > 100


@pytest.mark.parametrize("header", [os.path.basename(h) for h in HEADERS])
def test_every_declared_function_is_exported(header):
    libs = []
    for name in LIBS:
        path = os.path.join(ROOT, "starpu_amd", name)
        assert os.path.exists(path), f"{path} not built"
        libs.append(C.CDLL(path, mode=C.RTLD_GLOBAL))
    missing = []
    for fn in declared_functions(os.path.join(ROOT, "include", header)):
        if not any(hasattr(L, fn) for L in libs):
            missing.append(fn)
    assert not missing, f"{header}: declared but not exported: {missing}"

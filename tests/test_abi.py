"""The C-ABI library loads, exports every function include/*.h declares, and
the ctypes mirrors of its structs have the C layout (no GPU needed)."""
import ctypes as C
import re
import subprocess
import tempfile
from pathlib import Path

import pytest

from monkeygl_amd import _capi

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = []
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(mk_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_headers_declare_functions():
    names = declared_functions()
    assert "mk_render" in names and "mk_copy_volume_data" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_symbol_list_matches_header():
    assert sorted(_capi.SYMBOLS) == declared_functions()


def test_struct_layout_matches_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "monkeygl_hip.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu\n", sizeof(mk_render_params), offsetof(mk_render_params, alpha_wwwl),
           offsetof(mk_render_params, step_scale), sizeof(mk_plane_params), offsetof(mk_plane_params, half_num));
    printf("%zu %zu %zu %zu %zu\n", sizeof(mk_batch_item), offsetof(mk_batch_item, render),
           offsetof(mk_batch_item, plane), offsetof(mk_batch_item, cpr_sampling),
           offsetof(mk_batch_item, cpr_dirs_offset));
    return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "layout.c"
        c.write_text(src)
        exe = Path(d) / "layout"
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(c), "-o", str(exe)], check=True)
        got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    R, Pp, B = _capi.RenderParams, _capi.PlaneParams, _capi.BatchItem
    assert got == [C.sizeof(R), R.alpha_wwwl.offset, R.step_scale.offset, C.sizeof(Pp), Pp.half_num.offset,
                   C.sizeof(B), B.render.offset, B.plane.offset, B.cpr_sampling.offset, B.cpr_dirs_offset.offset]


def test_no_gpu_calls_fail_cleanly_without_device():
    """Without a GPU, context creation reports an error instead of crashing."""
    lib = _capi.load()
    if lib.mk_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_capi.MkError):
        _capi.Context(0)

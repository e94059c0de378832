"""CPU checks of the C-ABI library: it loads and exports every symbol that
include/retrofire_amd.h declares (no device calls)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "retrofire_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rf_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from retrofire_amd import _ffi
    L = _ffi.lib()
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if getattr(L, n, None) is None]
    assert missing == []
    assert set(names) == set(_ffi.SIGNATURES), "ctypes signatures and header disagree"


def test_status_strings():
    from retrofire_amd import _ffi
    L = _ffi.lib()
    for s in [0, -1, -2, -3, -4, -5, -6, -7, -8, -9]:
        assert len(L.rf_status_str(s)) >= 2


def test_struct_layouts_match_header():
    """The ctypes mirrors have the sizes the C compiler gives the header's
    structs (compiled with the host compiler)."""
    import ctypes as C
    import subprocess
    import tempfile
    from retrofire_amd import _ffi
    code = r'''
#include <stdio.h>
#include "retrofire_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(rf_ctx), sizeof(rf_stats), sizeof(rf_target),
         sizeof(rf_draw), sizeof(rf_mesh_draw), sizeof(rf_frame), sizeof(rf_request), sizeof(rf_result),
         sizeof(rf_tune_entry));
  return 0;
}'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(code)
        exe = os.path.join(d, "s")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        sizes = [int(x) for x in subprocess.check_output([exe]).split()]
    ours = [C.sizeof(t) for t in (_ffi.RfCtx, _ffi.RfStats, _ffi.RfTarget, _ffi.RfDraw, _ffi.RfMeshDraw,
                                  _ffi.RfFrame, _ffi.RfRequest, _ffi.RfResult, _ffi.RfTuneEntry)]
    assert ours == sizes


def test_no_device_without_gpu():
    """Opening a device fails cleanly (no crash) where no GPU is visible."""
    import ctypes as C
    import retrofire_amd as rf
    from retrofire_amd import _ffi
    h = C.c_void_p()
    st = _ffi.lib().rf_device_open(0, C.byref(h))
    if st == 0:
        pytest.skip("a GPU is visible")
    assert st == -4


def test_launch_primitive_limit_is_refused_before_any_device_work():
    """A draw of 2^29 primitives or more cannot be indexed by the rasterizer's
    32-bit (index << 3 | fan) entries: rf_render refuses it with
    RF_E_UNSUPPORTED while checking its arguments, before it touches the
    device (the handle below is a zeroed stand-in that is never read)."""
    import ctypes as C
    import numpy as np
    import retrofire_amd as rf
    from retrofire_amd import _ffi
    from retrofire_amd import mat as M
    L = _ffi.lib()
    fake_dev = C.create_string_buffer(4096)
    tris = np.zeros(3, np.uint32)
    pos = np.zeros(3, np.float32)
    col = np.zeros((4, 4, 4), np.uint8)
    keep = []
    d = rf.make_draw(tris.reshape(1, 3), pos.reshape(1, 3), rf.shader.Flat(), M.identity(),
                     M.viewport(0, 4, 4, 0), keep)
    t = rf.make_target(rf.Colorbuf(col))
    c = rf.Context().to_c()
    for n, want in ((1 << 29, -5), ((1 << 29) + 12345, -5), (1 << 40, -5)):
        d.n_tris = n
        assert L.rf_render(C.cast(fake_dev, C.c_void_p), C.byref(d), C.byref(t), C.byref(c), None) == want
    assert not col.any()

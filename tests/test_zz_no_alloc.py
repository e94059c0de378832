"""rf_render_noalloc (the plain call with every device buffer carved from the
library's static 64 MiB array, no device allocation) against rf_render: the
same colour bytes, depth bits and Stats on two calls in a row, the second with
a smaller input; oversized calls refused before any device work."""
import numpy as np
import pytest

import retrofire_amd as rf
from retrofire_amd import scenes as S

pytestmark = pytest.mark.gpu


def _both(sc, clear, seed):
    rng = np.random.default_rng(seed)
    col = rng.integers(0, 256, size=(sc.height, sc.width, 4), dtype=np.uint8)
    dep = rng.uniform(0.0, 0.01, size=(sc.height, sc.width)).astype(np.float32)
    out = []
    for fn in (rf.render, rf.render_noalloc):
        c, d = col.copy(), dep.copy()
        ctx = rf.Context(clear=clear)
        shd = rf.shader.TexRepeatPot(sc.tex) if sc.kind == 3 else rf.shader.Gouraud()
        fn(sc.tris, (sc.pos, sc.attr), shd, sc.mvp, sc.viewport, rf.Framebuf(c, d), ctx)
        s = ctx.stats
        out.append((c, d, [s.calls, s.prims_i, s.prims_o, s.verts_i, s.verts_o, s.frags_i, s.frags_o]))
    return out


def test_noalloc_equals_render_twice_second_smaller():
    dev = rf.default_device()
    big, small = S.crates1080(), S.sphere640()
    held = None
    for k, (sc, clear) in enumerate(((big, True), (small, False))):
        (c0, d0, s0), (c1, d1, s1) = _both(sc, clear, seed=k)
        if held is None:
            held = dev.bytes  # after the first rf_render: what the library holds
        assert np.array_equal(c0, c1), f"call {k} ({sc.name}) colour"
        assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32)), f"call {k} ({sc.name}) depth"
        assert s0 == s1 and s1[6] > 10000, f"call {k} stats {s0} {s1}"
    # the no-alloc calls added nothing to the library's device allocations
    # (the second rf_render of the smaller scene fits what the first grew)
    assert dev.bytes == held


def test_noalloc_refuses_what_does_not_fit():
    sc = S.sphere640()
    col = np.zeros((4096, 4096, 4), np.uint8)  # 64 MiB of colour alone, plus depth
    dep = np.zeros((4096, 4096), np.float32)
    with pytest.raises(rf.RenderError) as e:
        rf.render_noalloc(sc.tris, (sc.pos, sc.attr), rf.shader.Gouraud(), sc.mvp, sc.viewport,
                          rf.Framebuf(col, dep), rf.Context())
    assert e.value.status == -9
    assert not col.any()
    with pytest.raises(rf.RenderError) as e:
        rf.render_noalloc(sc.tris, (sc.pos, sc.attr), rf.shader.Gouraud(), sc.mvp, sc.viewport,
                          rf.Framebuf(col[:480, :640], dep[:480, :640]), rf.Context(depth_sort="back_to_front"))
    assert e.value.status == -5

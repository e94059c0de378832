"""The binning margin bounds the drift of the reference's serially stepped
edges (ScanlineIter::next, raster.rs:84-113): the scanlines tri_fill makes for
a triangle (computed by the oracle) never leave the vertex x range widened by
the margin the setup kernel gives that triangle. Cases include long, nearly
vertical edges far from x = 0 whose slope sits just above half an ulp of x,
where every row step rounds up by the same amount and the stepped edge runs
pixels past its end vertex (more than the former fixed 2-pixel margin)."""
import ctypes as C

import numpy as np

import oracle_lib as O


def f32(x):
    return np.float32(x)


def bin_margin(xs, rows):
    """The setup kernel's margin (bin_margin, rf_kernels.hip) in float32."""
    xmin, xmax = f32(min(xs)), f32(max(xs))
    X = f32(max(abs(xmin), abs(xmax)) + f32(1.0))
    ulp = np.uint32((np.array(X, np.float32).view(np.uint32) & np.uint32(0x7F800000)) - np.uint32(23 << 23)).view(
        np.float32)
    D = f32(f32(rows) * ulp) + f32(f32(xmax - xmin) * f32(2.0 ** -20))
    return 2 + int(np.floor(f32(D + f32(0.5))))


def spans(v):
    """tri_fill's scanlines (y, x0, x1) of a triangle of (x, y) vertices."""
    L = O.lib()
    vv = np.zeros(12, np.float32)
    for k in range(3):
        vv[4 * k:4 * k + 4] = (v[k][0], v[k][1], 0.5, 0.0)
    max_sl = 40000
    sl = np.zeros(4 * max_sl, np.int64)
    fr = np.zeros(2, np.float32)
    nf = C.c_int64()
    n = L.oracle_tri_fill_f32(vv.ctypes.data_as(C.POINTER(C.c_float)), sl.ctypes.data_as(C.POINTER(C.c_int64)),
                              max_sl, fr.ctypes.data_as(C.POINTER(C.c_float)), 0, C.byref(nf))
    if n < 0:  # fragments are not asked for: only the scanlines matter
        n = max_sl
    return sl[: 4 * n].reshape(-1, 4)


def check(v):
    xs = [f32(p[0]) for p in v]
    ys = [f32(p[1]) for p in v]
    rows = np.floor(f32(max(ys) + f32(0.5))) - np.floor(f32(min(ys) + f32(0.5)))
    m = bin_margin(xs, rows)
    c0, c1 = int(np.floor(min(xs))) - m, int(np.floor(max(xs))) + m
    worst = 0.0
    for y, x0, x1, cnt in spans(v):
        if x1 > x0:
            assert x0 >= max(c0, 0) and x1 - 1 <= c1, (v, m, x0, x1)
            worst = max(worst, float(x1 - 1) - float(max(xs)), float(min(xs)) - float(x0))
    return worst


def test_margin_bounds_drift_of_long_steep_edges():
    rng = np.random.default_rng(11)
    beyond = 0
    for i in range(120):
        X = float(rng.uniform(1000.0, 16300.0))
        rows = int(rng.integers(200, 16300))
        ulp = 2.0 ** (np.floor(np.log2(X)) - 23)
        dl = float(rng.uniform(0.5, 1.5)) * ulp * float(rng.choice([0.5, 1.5, 2.5, 1.0, 3.5]))
        dl *= float(rng.choice([-1.0, 1.0]))
        y0 = float(rng.uniform(0.0, 1.0))
        top = (X, y0)
        bot = (X + dl * rows, y0 + rows)
        side = (X - float(rng.uniform(1.0, 40.0)), y0 + rows * float(rng.uniform(0.2, 1.0)))
        v = [top, bot, side] if i % 2 else [top, side, bot]
        worst = check(v)
        beyond += worst > 1.5
    assert beyond >= 5, f"only {beyond} cases drift more than 1.5 px past their vertices"


def test_margin_random_triangles():
    rng = np.random.default_rng(12)
    for _ in range(300):
        c = rng.uniform(0, 2000, size=2)
        v = [tuple(c + rng.uniform(-300, 300, size=2)) for _ in range(3)]
        check(v)

"""Host-side matrix constructors, in the reference's f32 operation order.

These build the `mvp` and `to_screen` matrices handed to `render()`; the
reference builds them on the host too (core/src/math/mat.rs). Every scalar is a
numpy float32 and every operation is one IEEE binary32 operation in the order
the Rust code performs it, so the matrices are bit-identical to the
reference's for the same inputs (trigonometric inputs aside, which come from
the platform libm in both cases).

Matrices are row-major 4x4 float32 arrays, applied as row . column
(mat.rs:929-970).
"""
from __future__ import annotations

import math

import numpy as np

F = np.float32


def dot(a, b) -> np.float32:
    """`dot` (core/src/math/vec.rs:231-238): a left fold starting from +0.0."""
    acc = F(0.0)
    for x, y in zip(a, b):
        acc = F(acc + F(F(x) * F(y)))
    return acc


def mat(rows) -> np.ndarray:
    return np.array(rows, dtype=np.float32).reshape(4, 4)


def identity() -> np.ndarray:
    return np.eye(4, dtype=np.float32)


def compose(outer: np.ndarray, inner: np.ndarray) -> np.ndarray:
    """`outer.compose(inner)` (mat.rs:268-285): applying the result equals
    applying `inner`, then `outer`. Entry [j][i] = dot(outer row j, inner
    column i)."""
    out = np.zeros((4, 4), dtype=np.float32)
    for j in range(4):
        for i in range(4):
            out[j, i] = dot(outer[j, :], inner[:, i])
    return out


def then(first: np.ndarray, second: np.ndarray) -> np.ndarray:
    """`first.then(second)` (mat.rs:292-299) == `second.compose(first)`."""
    return compose(second, first)


def translate3(x: float, y: float, z: float) -> np.ndarray:
    """translate3 (mat.rs:1078-1085)."""
    return mat([[1, 0, 0, F(x)], [0, 1, 0, F(y)], [0, 0, 1, F(z)], [0, 0, 0, 1]])


def scale3(x: float, y: float, z: float) -> np.ndarray:
    """scale3: diagonal scaling of points."""
    return mat([[F(x), 0, 0, 0], [0, F(y), 0, 0], [0, 0, F(z), 0], [0, 0, 0, 1]])


def rotate_y(turns: float) -> np.ndarray:
    """rotate_y (mat.rs:1167-1175) for an angle given in turns; sin/cos are
    evaluated in float32 from the radian value (Angle::sin_cos)."""
    rads = F(F(turns) * F(2.0 * math.pi))
    s, c = F(np.sin(rads)), F(np.cos(rads))
    return mat([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]])


def perspective(focal_ratio: float, aspect_ratio: float, near: float, far: float) -> np.ndarray:
    """perspective (mat.rs:1255-1279)."""
    near, far = F(near), F(far)
    assert focal_ratio > 0 and aspect_ratio > 0 and near > 0 and far > near
    e00 = F(focal_ratio)
    e11 = F(e00 * F(aspect_ratio))
    e22 = F(F(far + near) / F(far - near))
    e23 = F(F(F(F(2.0) * far) * near) / F(near - far))
    return mat([[e00, 0, 0, 0], [0, e11, 0, 0], [0, 0, e22, e23], [0, 0, 1, 0]])


def viewport(x0: int, y0: int, x1: int, y1: int) -> np.ndarray:
    """viewport(pt2(x0, y0)..pt2(x1, y1)) (mat.rs:1304-1315)."""
    x0, y0, x1, y1 = F(x0), F(y0), F(x1), F(y1)
    dx = F(F(x1 - x0) / F(2.0))
    dy = F(F(y1 - y0) / F(2.0))
    return mat([[dx, 0, 0, F(x0 + dx)], [0, dy, 0, F(y0 + dy)], [0, 0, 1, 0], [0, 0, 0, 1]])


def fov_diagonal_focal_ratio(degrees: float, aspect_ratio: float) -> np.float32:
    """Fov::Diagonal(a).focal_ratio(aspect) (cam.rs:120-143)."""
    a = F(F(degrees) * F(math.pi / 180.0))
    ratio = F(F(1.0) / F(np.tan(F(a / F(2.0)))))
    ar = F(aspect_ratio)
    diag = F(np.sqrt(F(F(1.0) + F(F(F(1.0) / ar) / ar))))
    return F(ratio * diag)


def apply_pt(m: np.ndarray, p) -> np.ndarray:
    """Mat4::apply for a point (mat.rs:945-949)."""
    h = [F(p[0]), F(p[1]), F(p[2]), F(1.0)]
    return np.array([dot(m[0], h), dot(m[1], h), dot(m[2], h)], dtype=np.float32)


def apply_proj(m: np.ndarray, p) -> np.ndarray:
    """ProjMat3::apply (mat.rs:965-969)."""
    h = [F(p[0]), F(p[1]), F(p[2]), F(1.0)]
    return np.array([dot(m[r], h) for r in range(4)], dtype=np.float32)

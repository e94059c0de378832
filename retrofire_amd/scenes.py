"""Seeded synthetic scenes for the render path's workloads.

The generators follow the reference's own: the Xorshift64 generator and the
Uniform<f32> distribution bit for bit (core/src/math/rand.rs:118-327, pinned
by its doctests in tests/test_scenes.py), and the mesh builders structurally
(the lathe sphere geom/src/solids/lathe.rs:98-190, the box
geom/src/solids/platonic.rs:132-241, the icosphere geom/src/solids.rs:25-77).
Each scene is generated once; the oracle and the GPU consume the same arrays.

Workloads (seed = Xorshift64::DEFAULT_SEED + k for workload k):
  sphere640     640x480 Gouraud UV-sphere, 64 x 32, random vertex colours
  crates1080    1920x1080, 1000 textured unit cubes on a 100 x 100 field
  icospheres1080  1920x1080, 10,000 icospheres (subdivision 2, r = 0.05)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import mat as M

F = np.float32
DEFAULT_SEED = 378682147834061
MASK64 = (1 << 64) - 1


class Xorshift64:
    """Xorshift64 (rand.rs:118-183)."""

    def __init__(self, seed: int = DEFAULT_SEED):
        assert seed != 0
        self.x = seed & MASK64

    def next_bits(self) -> int:
        x = self.x
        x ^= (x << 13) & MASK64
        x ^= x >> 7
        x ^= (x << 17) & MASK64
        self.x = x
        return x

    def uniform_f32(self, start: float, end: float) -> np.float32:
        """Uniform<f32> (rand.rs:320-327): a random f32 in [1, 2) from 23 bits,
        minus one, scaled."""
        mant = self.next_bits() >> 41
        unit = F(np.uint32((127 << 23) | mant).view(np.float32)) - F(1.0)
        return F(F(unit * F(F(end) - F(start))) + F(start))

    def uniform_u32(self, start: int, end: int) -> int:
        """Uniform<u32> (rand.rs:258-262)."""
        bits = self.next_bits() & 0xFFFFFFFF
        return bits % (end - start) + start

    def uniform_usize(self, start: int, end: int) -> int:
        """Uniform<usize> (rand.rs:286-290)."""
        return self.next_bits() % (end - start) + start

    def point_in_unit_ball(self) -> np.ndarray:
        """VectorsInUnitBall (rand.rs:543-553): rejection sampling."""
        while True:
            v = np.array([self.uniform_f32(-1.0, 1.0) for _ in range(3)], dtype=np.float32)
            if F(F(F(v[0] * v[0]) + F(v[1] * v[1])) + F(v[2] * v[2])) <= F(1.0):
                return v


@dataclass
class Scene:
    name: str
    width: int
    height: int
    kind: int                   # shader kind
    pos: np.ndarray             # (n, 3) float32
    attr: np.ndarray | None     # (n, k) float32
    tris: np.ndarray            # (m, 3) uint32
    mvp: np.ndarray
    viewport: np.ndarray
    tex: np.ndarray | None = None   # (h, w, 4) uint8
    flat: tuple = (255, 255, 255, 255)
    objects: int = 1
    tris_per_object: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def n_verts(self):
        return self.pos.shape[0]

    @property
    def n_tris(self):
        return self.tris.shape[0]


# ------------------------------------------------------------------ meshes

def lathe_sphere(sectors: int, segments: int, radius: float = 1.0):
    """Sphere{sectors, segments, radius} via the lathe (lathe.rs:98-190,
    240-261): (segments + 1) rings of (sectors + 1) vertices, 2 * sectors *
    segments faces."""
    secs, segs = sectors, segments
    vps = segs + 1
    # 0.0.vary_to(1.0, vps): t stepped by 1/(vps-1)
    step = F(F(1.0) / F(vps - 1))
    ts, t = [], F(0.0)
    for _ in range(vps):
        ts.append(t)
        t = F(t + step)
    rot = F(2.0 * math.pi) / F(secs)
    c, s = F(math.cos(rot)), F(math.sin(rot))
    pos = []
    for t in ts:
        a = F(F(-0.25) + F(F(F(0.25) - F(-0.25)) * t))  # (-0.25).lerp(&0.25, t) turns
        ang = F(a * F(2.0 * math.pi))
        nx, ny = F(math.cos(ang)), F(math.sin(ang))
        px, py = F(nx * F(radius)), F(ny * F(radius))
        x, z = px, F(0.0)
        for _u in range(secs + 1):
            pos.append((x, py, z))
            x, z = F(F(c * x) - F(s * z)), F(F(s * x) + F(c * z))
    faces = []
    n = secs + 1
    for j in range(1, vps):
        for i in range(1, n):
            p, q = (j - 1) * n + i - 1, (j - 1) * n + i
            r, s_ = j * n + i - 1, j * n + i
            faces.append((p, q, s_))
            faces.append((p, s_, r))
    return np.array(pos, dtype=np.float32), np.array(faces, dtype=np.uint32)


_BOX_COORDS = [(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)]
_BOX_UV = [(0, 0), (1, 0), (0, 1), (1, 1)]
_BOX_VERTS = [
    (0b011, 0), (0b010, 1), (0b001, 2), (0b000, 3),
    (0b110, 0), (0b111, 1), (0b100, 2), (0b101, 3),
    (0b000, 0), (0b100, 1), (0b001, 2), (0b101, 3),
    (0b011, 0), (0b111, 1), (0b010, 2), (0b110, 3),
    (0b010, 0), (0b110, 1), (0b000, 2), (0b100, 3),
    (0b111, 0), (0b011, 1), (0b101, 2), (0b001, 3),
]
_BOX_FACES = [(0, 1, 3), (0, 3, 2), (4, 5, 7), (4, 7, 6), (8, 9, 11), (8, 11, 10),
              (12, 13, 15), (12, 15, 14), (16, 17, 19), (16, 19, 18), (20, 21, 23), (20, 23, 22)]


def cube(side: float = 1.0):
    """Cube{side_len} as Mesh<TexCoord> (platonic.rs:132-241): 24 vertices
    with per-face UVs, 12 triangles."""
    d = F(F(side) / F(2.0))
    lo, hi = -d, d
    pos, uv = [], []
    for ci, ui in _BOX_VERTS:
        c = _BOX_COORDS[ci]
        pos.append(tuple(F(lo + F(F(hi - lo) * F(c[k]))) for k in range(3)))
        uv.append(_BOX_UV[ui])
    return np.array(pos, np.float32), np.array(uv, np.float32), np.array(_BOX_FACES, np.uint32)


PHI = F((1.0 + math.sqrt(5.0)) / 2.0)
_ICO_COORDS = [(-PHI, 0, -1), (-PHI, 0, 1), (PHI, 0, -1), (PHI, 0, 1),
               (-1, -PHI, 0), (1, -PHI, 0), (-1, PHI, 0), (1, PHI, 0),
               (0, -1, -PHI), (0, 1, -PHI), (0, -1, PHI), (0, 1, PHI)]
_ICO_FACES = [(0, 4, 1), (0, 1, 6), (2, 3, 5), (2, 7, 3), (4, 8, 5), (4, 5, 10), (6, 7, 9), (6, 11, 7),
              (8, 0, 9), (8, 9, 2), (10, 11, 1), (10, 3, 11), (0, 8, 4), (1, 4, 10), (0, 6, 9),
              (1, 11, 6), (2, 5, 8), (3, 10, 5), (2, 9, 7), (3, 7, 11)]


def icosphere(radius: float, depth: int):
    """Icosphere(radius, depth) (solids.rs:25-77): recursive 4-way subdivision
    with midpoints memoised per directed edge, vertices normalised and
    scaled."""
    coords = [np.array(c, dtype=np.float32) for c in _ICO_COORDS]
    memo = {}
    faces = []

    def mid(i, j):
        if (i, j) not in memo:
            a, b = coords[i], coords[j]
            coords.append((a + (b - a) * F(0.5)).astype(np.float32))
            memo[(i, j)] = len(coords) - 1
        return memo[(i, j)]

    def rec(d, i, j, k):
        if d == 0:
            faces.append((i, j, k))
            return
        ij, ik, jk = mid(i, j), mid(i, k), mid(j, k)
        rec(d - 1, i, ij, ik)
        rec(d - 1, j, jk, ij)
        rec(d - 1, k, ik, jk)
        rec(d - 1, ij, jk, ik)

    for f in _ICO_FACES:
        rec(depth, *f)
    out = []
    for p in coords:
        ln = F(np.sqrt(F(F(F(p[0] * p[0]) + F(p[1] * p[1])) + F(p[2] * p[2]))))
        out.append([F(F(p[k] / ln) * F(radius)) for k in range(3)])
    return np.array(out, np.float32), np.array(faces, np.uint32)


def transform_points(m: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Mat4::apply over many points (mat.rs:945-949), vectorised in float32
    with the dot's left-fold order."""
    p = p.astype(np.float32)
    out = np.empty_like(p)
    for r in range(3):
        acc = np.zeros(p.shape[0], np.float32)
        acc = (acc + m[r, 0] * p[:, 0]).astype(np.float32)
        acc = (acc + m[r, 1] * p[:, 1]).astype(np.float32)
        acc = (acc + m[r, 2] * p[:, 2]).astype(np.float32)
        acc = (acc + m[r, 3] * F(1.0)).astype(np.float32)
        out[:, r] = acc
    return out


# ------------------------------------------------------------------ scenes

def sphere640(seed_offset: int = 0) -> Scene:
    """One UV-sphere (64 x 32) with random vertex colours, Gouraud, 640x480,
    camera at distance 3, 60 degree horizontal fov."""
    rng = Xorshift64(DEFAULT_SEED + 0 + seed_offset)
    pos, tris = lathe_sphere(64, 32, 1.0)
    cols = np.array([[rng.uniform_f32(0.0, 1.0) for _ in range(3)] for _ in range(pos.shape[0])], np.float32)
    w, h = 640, 480
    sqrt3 = F(1.7320508)  # SQRT_3 (math.rs:157) = 1/tan(30 deg)
    mvp = M.then(M.translate3(0.0, 0.0, 3.0), M.perspective(sqrt3, F(w) / F(h), 0.1, 100.0))
    # y-down viewport: the lathe winds its outside clockwise in y-up screen
    # space, so this orientation keeps the near hemisphere front-facing
    vp = M.viewport(0, 0, w, h)
    return Scene("sphere640", w, h, 1, pos, cols, tris, mvp, vp, objects=1, tris_per_object=tris.shape[0])


def crate_texture(rng: Xorshift64, n: int = 256) -> np.ndarray:
    """n x n RGBA8: an 8 x 8 checker (32-texel squares) XOR per-channel noise."""
    y, x = np.mgrid[0:n, 0:n]
    base = np.where(((x // 32) ^ (y // 32)) & 1, 0xC0, 0x30).astype(np.uint8)
    noise = np.array([rng.next_bits() for _ in range(n * n)], dtype=np.uint64).reshape(n, n)
    tex = np.empty((n, n, 4), np.uint8)
    for c in range(3):
        tex[..., c] = base ^ ((noise >> np.uint64(8 * c)) & np.uint64(0x3F)).astype(np.uint8)
    tex[..., 3] = 255
    return tex


def crates_world(n_crates: int = 1000, seed: int = DEFAULT_SEED + 1):
    """World-space vertices of `n_crates` unit cubes at uniform xz in a
    100 x 100 field, y = 0, random yaw; one merged Mesh<TexCoord>."""
    rng = Xorshift64(seed)
    cpos, cuv, ctri = cube(1.0)
    pos, uv, tris = [], [], []
    for i in range(n_crates):
        x = rng.uniform_f32(-50.0, 50.0)
        z = rng.uniform_f32(-50.0, 50.0)
        yaw = rng.uniform_f32(0.0, 1.0)
        model = M.then(M.rotate_y(float(yaw)), M.translate3(x, 0.0, z))
        pos.append(transform_points(model, cpos))
        uv.append(cuv)
        tris.append(ctri + np.uint32(24 * i))
    tex = crate_texture(rng)
    return np.concatenate(pos), np.concatenate(uv), np.concatenate(tris), tex


def crates_objects(n_crates: int = 1000, seed: int = DEFAULT_SEED + 1):
    """The crates of crates_world as separate objects, as the crates demo
    keeps them (crates.rs:114-133, Obj scene.rs:14-18): one cube Mesh<TexCoord>
    and a model-to-world matrix per crate."""
    rng = Xorshift64(seed)
    cpos, cuv, ctri = cube(1.0)
    models = []
    for i in range(n_crates):
        x = rng.uniform_f32(-50.0, 50.0)
        z = rng.uniform_f32(-50.0, 50.0)
        yaw = rng.uniform_f32(0.0, 1.0)
        models.append(M.then(M.rotate_y(float(yaw)), M.translate3(x, 0.0, z)))
    tex = crate_texture(rng)
    return cpos, cuv, ctri, models, tex


def first_person(x: float, y: float, z: float, yaw_turns: float = 0.0) -> np.ndarray:
    """World-to-view of a level first-person camera at (x, y, z) turned by
    `yaw_turns` about +y (heading +z at yaw 0): the inverse camera transform
    (cam.rs:404-418)."""
    return M.then(M.translate3(-x, -y, -z), M.rotate_y(-yaw_turns))


def crates1080(n_crates: int = 1000) -> Scene:
    """1000 textured crates, SamplerRepeatPot, first-person camera at height
    1.7 (near-plane clipping), diagonal fov 90 degrees, 1920x1080."""
    pos, uv, tris, tex = crates_world(n_crates)
    w, h = 1920, 1080
    aspect = F(w) / F(h)
    proj = M.perspective(M.fov_diagonal_focal_ratio(90.0, aspect), aspect, 0.1, 1000.0)
    mvp = M.then(first_person(0.0, 1.7, 0.0), proj)
    vp = M.viewport(0, h, w, 0)
    return Scene("crates1080", w, h, 3, pos, uv, tris, mvp, vp, tex=tex, objects=n_crates,
                 tris_per_object=12)


def icospheres_world(n: int = 10000, seed: int = DEFAULT_SEED + 2):
    rng = Xorshift64(seed)
    ipos, itri = icosphere(0.05, 2)
    centers = np.empty((n, 3), np.float32)
    cols = np.empty((n, 3), np.float32)
    for i in range(n):
        centers[i] = rng.point_in_unit_ball()
        cols[i] = [rng.uniform_f32(0.0, 1.0) for _ in range(3)]
    nv = ipos.shape[0]
    pos = (ipos[None, :, :] + centers[:, None, :]).astype(np.float32).reshape(-1, 3)
    attr = np.repeat(cols, nv, axis=0).astype(np.float32)
    tris = (itri[None, :, :] + (np.arange(n, dtype=np.uint32) * np.uint32(nv))[:, None, None]).reshape(-1, 3)
    return pos, attr, tris.astype(np.uint32), nv, itri.shape[0]


def icospheres1080(n: int = 10000) -> Scene:
    """10,000 icospheres with one random colour each (constant per-vertex
    Color3f, bit-identical to flat shading), camera at z = 2.5, 1920x1080."""
    pos, attr, tris, nv, nt = icospheres_world(n)
    w, h = 1920, 1080
    sqrt3 = F(1.7320508)
    mvp = M.then(M.translate3(0.0, 0.0, 2.5), M.perspective(sqrt3, F(w) / F(h), 0.1, 100.0))
    vp = M.viewport(0, h, w, 0)
    return Scene("icospheres1080", w, h, 1, pos, attr, tris, mvp, vp, objects=n, tris_per_object=nt)


SCENES = {"sphere640": sphere640, "crates1080": crates1080, "icospheres1080": icospheres1080}


# ---------------------------------------------------------------------------
# Batched and served workloads: frames and requests over the scenes above.
# ---------------------------------------------------------------------------
RESOLUTIONS = [(256, 256), (512, 512), (1280, 720), (1920, 1080), (2048, 2048)]


def crates_mvp(w: int, h: int, yaw: float) -> np.ndarray:
    """First person at (0, 1.7, 0) turned by `yaw` (turns), diagonal fov 90
    degrees, near 0.1, far 1000, aspect w / h."""
    aspect = F(w) / F(h)
    proj = M.perspective(M.fov_diagonal_focal_ratio(90.0, aspect), aspect, 0.1, 1000.0)
    return M.then(first_person(0.0, 1.7, 0.0, yaw), proj)


def spheres_mvp(w: int, h: int, orbit: float) -> np.ndarray:
    """Orbit of `orbit` turns about +y at distance 2.5 from the cloud's centre,
    horizontal fov 60 degrees, near 0.1, far 100, aspect w / h."""
    view = M.then(M.rotate_y(orbit), M.translate3(0.0, 0.0, 2.5))
    return M.then(view, M.perspective(F(1.7320508), F(w) / F(h), 0.1, 100.0))


def batch_frames(n: int, crate_tris: int, sphere_tris: int, tris_per_sphere: int, seed: int = 0):
    """The batched workload's frames: n frames, each a resolution drawn
    uniformly from RESOLUTIONS and a seeded mix of the crates (10-100 % of the
    crates, 12 triangles each) and the icospheres (10-100 % of the spheres)
    with random camera poses. Returns dicts of w, h, vp, mvp_c, mvp_i and the
    triangle counts nc, ni of the two draws (prefixes of the merged meshes)."""
    rng = Xorshift64(DEFAULT_SEED + 3 + seed)
    out = []
    for _ in range(n):
        w, h = RESOLUTIONS[rng.uniform_usize(0, 5)]
        fc = float(rng.uniform_f32(0.1, 1.0))
        fi = float(rng.uniform_f32(0.1, 1.0))
        yaw = float(rng.uniform_f32(0.0, 1.0))
        orbit = float(rng.uniform_f32(0.0, 1.0))
        nc = max(1, int(fc * crate_tris / 12)) * 12
        ni = max(1, int(fi * sphere_tris / tris_per_sphere)) * tris_per_sphere
        out.append(dict(w=w, h=h, vp=M.viewport(0, h, w, 0), mvp_c=crates_mvp(w, h, yaw),
                        mvp_i=spheres_mvp(w, h, orbit), nc=nc, ni=ni))
    return out


def serve_meshes():
    """The served workload's 16 resident meshes: 8 crate fields of 125, 250,
    ..., 1000 crates (textured) and 8 icosphere clouds of 250, 500, ..., 2000
    spheres (one colour each). Returns [(kind, pos, attr, tris)] and the crate
    texture; kind is "crates" or "spheres"."""
    out = []
    tex = None
    for k in range(8):
        pos, uv, tris, t = crates_world(125 * (k + 1), seed=DEFAULT_SEED + 100 + k)
        tex = t if tex is None else tex
        out.append(("crates", pos, uv, tris))
    for k in range(8):
        pos, attr, tris, _, _ = icospheres_world(250 * (k + 1), seed=DEFAULT_SEED + 200 + k)
        out.append(("spheres", pos, attr, tris))
    return out, tex


def serve_stream(rate: float, n: int, seed: int = 5):
    """An open-loop request stream at `rate` requests/s: n requests of
    (arrival time s, width, height, mesh index 0..15, pose in turns), with
    exponential gaps, resolutions from RESOLUTIONS and meshes uniform over
    serve_meshes()."""
    rng = Xorshift64(DEFAULT_SEED + seed)
    t, out = 0.0, []
    for _ in range(n):
        t += -math.log(max(1e-7, 1.0 - float(rng.uniform_f32(0.0, 1.0)))) / rate
        w, h = RESOLUTIONS[rng.uniform_usize(0, 5)]
        mi = rng.uniform_usize(0, 16)
        pose = float(rng.uniform_f32(0.0, 1.0))
        out.append((t, w, h, mi, pose))
    return out

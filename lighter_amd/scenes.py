"""Seeded synthetic scenes for the baking workloads, and a builder that issues
the corresponding ``ltr_*`` calls on any library exporting lighter's API.

Scenes are produced with numpy's PCG64 generator (seed = 0x4C540000 + config
index by default); libc ``rand`` is never used here, because the baker itself
draws its AO rotation offsets from libc ``rand`` (lighter.cpp:819,1300).

Geometry is laid out as planar rectangular "charts" (room walls, box faces,
terrain patches), each tessellated into quads and packed into the instance's
UV2 atlas with a shelf packer.  Triangles wind counter-clockwise around the
shading normal.
"""
from __future__ import annotations

import ctypes as C
import functools
import dataclasses as dc
import math
from typing import Callable, Optional

import numpy as np

from . import api

SEED_BASE = 0x4C540000


@dc.dataclass
class Part:
    positions: np.ndarray  # (V,3) float32
    normals: np.ndarray  # (V,3) float32
    tex1: np.ndarray  # (V,2) float32
    tex2: np.ndarray  # (V,2) float32
    indices: np.ndarray  # (I,) uint32
    shadow: int = 1


@dc.dataclass
class Instance:
    matrix: np.ndarray  # (4,4) float32, ltr_MAT4 layout (translation in row 3)
    importance: float = 1.0
    shadow: int = 1
    ident: str = ""


@dc.dataclass
class Mesh:
    ident: str
    parts: list
    instances: list


@dc.dataclass
class SceneSpec:
    name: str
    meshes: list
    lights: list  # list of dicts of LightInfo fields
    config: dict = dc.field(default_factory=dict)
    lm_size: Optional[int] = None  # every instance gets lm_size x lm_size
    albedo: Optional[dict] = None  # part_id -> rgb, applied through sample_fn
    emission: Optional[dict] = None  # part_id -> rgb, applied through sample_fn
    samples: list = dc.field(default_factory=list)  # external samples: (pos, nrm)
    sample_cb: Optional[Callable] = None  # sample_fn body: called with the ltr_SampleRequest, after albedo/emission

    @property
    def num_triangles(self) -> int:
        n = 0
        for m in self.meshes:
            tri = sum(len(p.indices) // 3 for p in m.parts)
            n += tri * len(m.instances)
        return n


IDENTITY = np.eye(4, dtype=np.float32)


def normalize_rows_f32(v):
    """Vec3::Normalized (lighter_int.hpp:404-415) of each row in float32:
    v * (1 / sqrt((x x + y y) + z z)), zero rows stay zero.  Int_RDGenLinks
    normalises the patch normals this way (lighter.cpp:683)."""
    v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
    l2 = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]
    out = np.zeros_like(v)
    nz = l2 != 0
    inv = np.float32(1.0) / np.sqrt(l2[nz])
    out[nz] = v[nz] * inv[:, None]
    return out


# --------------------------------------------------------------------------
# charts and atlas packing
# --------------------------------------------------------------------------
@dc.dataclass
class Chart:
    origin: np.ndarray  # corner
    du: np.ndarray  # edge along chart u
    dv: np.ndarray  # edge along chart v
    nu: int
    nv: int
    part: int
    height: Optional[np.ndarray] = None  # (nv+1, nu+1) displacement along +z (terrain)

    @functools.cached_property
    def size_m(self):
        return float(np.linalg.norm(self.du)), float(np.linalg.norm(self.dv))


@dc.dataclass
class QuadSet:
    """Many flat 1 x 1 charts of one part (box faces), as arrays: corner,
    u edge and v edge of each quad (float64, (n, 3))."""
    origin: np.ndarray
    du: np.ndarray
    dv: np.ndarray
    part: int

    def __len__(self):
        return len(self.origin)


def _shelf_pack(wh, W, H, pad):
    """Shelf-pack rectangles wh (n, 2) in pixels, tallest first (ties in
    input order); returns the corners (n, 2) or None.  Rows fill left to
    right while pad + sum(w + pad) <= W; a row is as tall as its first
    (tallest) rectangle; rows stack from y = pad with pad between them."""
    n = len(wh)
    order = np.argsort(-wh[:, 1], kind="stable")
    w = wh[order, 0]
    h = wh[order, 1]
    cw = np.cumsum(w + pad)
    out = np.empty((n, 2), np.int64)
    start, y, base = 0, pad, 0
    while start < n:
        end = int(np.searchsorted(cw, base + W - pad, side="right"))
        if end <= start or y + h[start] + pad > H:
            return None
        out[order[start:end], 0] = pad + cw[start:end] - (w[start:end] + pad) - base
        out[order[start:end], 1] = y
        y += int(h[start]) + pad
        base = int(cw[end - 1])
        start = end
    return out


def _chart_metric(charts, quads=()):
    m = [(c.size_m[0], c.size_m[1]) for c in charts]
    arr = [np.array(m, np.float64).reshape(-1, 2)]
    for q in quads:
        arr.append(np.stack([np.sqrt((q.du * q.du).sum(-1)), np.sqrt((q.dv * q.dv).sum(-1))], -1))
    return np.concatenate(arr)


def pack_charts(charts, W, H, pad=2, quads=()):
    """Pick the largest texels-per-metre scale at which all charts (then the
    quads of each QuadSet) fit: 40 bisection steps over [0, 4096]; each
    chart gets max(1, int(size_m * scale)) pixels per side."""
    lo, hi = 0.0, 4096.0
    best = None
    metric = _chart_metric(charts, quads)
    for _ in range(40):
        s = 0.5 * (lo + hi)
        sizes = np.maximum(1, (metric * s).astype(np.int64))
        pos = _shelf_pack(sizes, W, H, pad)
        if pos is None:
            hi = s
        else:
            lo = s
            best = (pos, sizes)
    assert best is not None, "charts do not fit the atlas"
    return best


def _quad_arrays(O, DU, DV, rect, W, H):
    """Vertices and indices of many 1x1 charts at once, in the per-chart
    layout: vertices (a, b) = (0,0), (1,0), (0,1), (1,1), triangles
    (v00, v10, v11), (v11, v01, v00).  rect: (x0, y0, w, h) per quad."""
    n = len(O)
    A = np.array([0.0, 1.0, 0.0, 1.0])
    B = np.array([0.0, 0.0, 1.0, 1.0])
    P = O[:, None, :] + A[None, :, None] * DU[:, None, :] + B[None, :, None] * DV[:, None, :]
    nrm = np.cross(DU, DV)
    nrm = nrm / np.sqrt((nrm * nrm).sum(-1, keepdims=True))
    N = np.broadcast_to(nrm[:, None, :], P.shape)
    T1 = np.broadcast_to(np.stack([A, B], -1)[None], (n, 4, 2))
    T2 = np.stack([(rect[:, 0:1] + A[None] * rect[:, 2:3]) / W, (rect[:, 1:2] + B[None] * rect[:, 3:4]) / H], -1)
    idx = (np.array([0, 1, 3, 3, 2, 0])[None, :] + 4 * np.arange(n)[:, None]).ravel()
    verts = (P.reshape(-1, 3), N.reshape(-1, 3), T1.reshape(-1, 2), T2.reshape(-1, 2))
    return verts, idx


def _quad_batch(lst, W, H):
    """_quad_arrays over (chart, x0, y0, w, h) tuples of 1x1 charts."""
    O = np.array([c.origin for c, *_ in lst], np.float64)
    DU = np.array([c.du for c, *_ in lst], np.float64)
    DV = np.array([c.dv for c, *_ in lst], np.float64)
    rect = np.array([(x0, y0, wp, hp) for _, x0, y0, wp, hp in lst], np.float64)
    return _quad_arrays(O, DU, DV, rect, W, H)


def charts_to_part_arrays(charts, W, H, pad=2, quads=()):
    """Tessellate charts (and the quads of each QuadSet, packed after them)
    and assign UV2 atlas coordinates; returns per-part arrays."""
    pos_rect, sizes = pack_charts(charts, W, H, pad, quads)
    by_part = {}
    k = len(charts)
    for q in quads:
        assert all(c.part != q.part for c in charts), "a QuadSet's part holds no other chart"
        rect = np.concatenate([pos_rect[k:k + len(q)], sizes[k:k + len(q)]], 1).astype(np.float64)
        by_part.setdefault(q.part, []).append(_quad_arrays(q.origin, q.du, q.dv, rect, W, H))
        k += len(q)
    # parts made only of single flat quads (boxes) are tessellated in one batch
    quad_parts = {}
    for c in charts:
        q = c.nu == 1 and c.nv == 1 and c.height is None
        quad_parts[c.part] = quad_parts.get(c.part, True) and q
    batch = {}
    for c, (x0, y0), (wp, hp) in zip(charts, pos_rect, sizes):
        if quad_parts[c.part]:
            batch.setdefault(c.part, []).append((c, x0, y0, wp, hp))
            continue
        a = np.linspace(0.0, 1.0, c.nu + 1, dtype=np.float64)
        b = np.linspace(0.0, 1.0, c.nv + 1, dtype=np.float64)
        A, B = np.meshgrid(a, b)  # (nv+1, nu+1)
        P = c.origin[None, None, :] + A[..., None] * c.du[None, None, :] + B[..., None] * c.dv[None, None, :]
        if c.height is not None:
            P = P.copy()
            P[..., 2] += c.height
        n = np.cross(c.du, c.dv)
        n = n / np.linalg.norm(n)
        if c.height is not None:
            # terrain normals from central differences of the displaced grid
            gz_u = np.gradient(P[..., 2], axis=1) / (np.linalg.norm(c.du) / c.nu)
            gz_v = np.gradient(P[..., 2], axis=0) / (np.linalg.norm(c.dv) / c.nv)
            N = np.stack([-gz_u, -gz_v, np.ones_like(gz_u)], -1)
            N /= np.linalg.norm(N, axis=-1, keepdims=True)
        else:
            N = np.broadcast_to(n, P.shape)
        T1 = np.stack([A, B], -1)
        T2 = np.stack([(x0 + A * wp) / W, (y0 + B * hp) / H], -1)
        verts = (P.reshape(-1, 3), N.reshape(-1, 3), T1.reshape(-1, 2), T2.reshape(-1, 2))
        iu, iv = np.meshgrid(np.arange(c.nu), np.arange(c.nv))
        v00 = (iv * (c.nu + 1) + iu).ravel()
        v10 = v00 + 1
        v01 = v00 + (c.nu + 1)
        v11 = v01 + 1
        idx = np.stack([v00, v10, v11, v11, v01, v00], -1).ravel()
        lst = by_part.setdefault(c.part, [])
        lst.append((verts, idx))
    for pid, lst in batch.items():
        by_part[pid] = [_quad_batch(lst, W, H)]
    parts = {}
    for pid, lst in by_part.items():
        P, N, T1, T2, I = [], [], [], [], []
        base = 0
        for (p, n, t1, t2), idx in lst:
            P.append(p); N.append(n); T1.append(t1); T2.append(t2)
            I.append(idx + base)
            base += len(p)
        parts[pid] = Part(
            np.ascontiguousarray(np.concatenate(P), dtype=np.float32),
            np.ascontiguousarray(np.concatenate(N), dtype=np.float32),
            np.ascontiguousarray(np.concatenate(T1), dtype=np.float32),
            np.ascontiguousarray(np.concatenate(T2), dtype=np.float32),
            np.ascontiguousarray(np.concatenate(I), dtype=np.uint32),
        )
    return [parts[k] for k in sorted(parts)]


def box_charts(lo, hi, part, with_bottom=True):
    """Six outward-facing faces of an axis-aligned box, one quad each."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    dx, dy, dz = hi - lo
    X = np.array([dx, 0, 0]); Y = np.array([0, dy, 0]); Z = np.array([0, 0, dz])
    ch = [
        Chart(lo + Z, X, Y, 1, 1, part),  # top, +z
        Chart(lo, X, Z, 1, 1, part),  # -y side
        Chart(lo + Y + X, -X, Z, 1, 1, part),  # +y side
        Chart(lo + Y, -Y, Z, 1, 1, part),  # -x side
        Chart(lo + X, Y, Z, 1, 1, part),  # +x side
    ]
    if with_bottom:
        ch.append(Chart(lo + Y, X, -Y, 1, 1, part))  # bottom, -z
    return ch


def box_quads(lo, hi, part, with_bottom=True):
    """box_charts of many boxes at once (lo, hi: (n, 3)): the faces of each
    box in box_charts' order, box after box, as one QuadSet."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    d = hi - lo
    z = np.zeros(len(lo))
    X = np.stack([d[:, 0], z, z], 1)
    Y = np.stack([z, d[:, 1], z], 1)
    Z = np.stack([z, z, d[:, 2]], 1)
    faces = [(lo + Z, X, Y), (lo, X, Z), (lo + Y + X, -X, Z), (lo + Y, -Y, Z), (lo + X, Y, Z)]
    if with_bottom:
        faces.append((lo + Y, X, -Y))
    O = np.stack([f[0] for f in faces], 1).reshape(-1, 3)
    DU = np.stack([f[1] for f in faces], 1).reshape(-1, 3)
    DV = np.stack([f[2] for f in faces], 1).reshape(-1, 3)
    return QuadSet(O, DU, DV, part)


def room_charts(size=(10.0, 10.0, 4.0), tess=8, part=0):
    """Six inward-facing walls of a room [0,sx]x[0,sy]x[0,sz]."""
    sx, sy, sz = size
    X = np.array([sx, 0, 0.0]); Y = np.array([0, sy, 0.0]); Z = np.array([0, 0, sz])
    o = np.zeros(3)
    return [
        Chart(o, X, Y, tess, tess, part),  # floor, +z
        Chart(o + Z + Y, X, -Y, tess, tess, part),  # ceiling, -z
        Chart(o, Y, Z, tess, tess, part),  # x=0, +x
        Chart(o + X + Y, -Y, Z, tess, tess, part),  # x=sx, -x
        Chart(o + X, -X, Z, tess, tess, part),  # y=0, +y
        Chart(o + Y, X, Z, tess, tess, part),  # y=sy, -y
    ]


def _check_charts_normals(charts, want):
    for c, w in zip(charts, want):
        n = np.cross(c.du, c.dv)
        assert np.dot(n, w) > 0, (c, w)


# --------------------------------------------------------------------------
# lights
# --------------------------------------------------------------------------
def point_light(pos, color, rng_, power=1.0, radius=0.1, samples=4):
    return dict(type=api.LTR_LT_POINT, position=list(map(float, pos)), color_rgb=list(map(float, color)),
                range=float(rng_), power=float(power), light_radius=float(radius), shadow_sample_count=samples)


def spot_light(pos, direction, color, rng_, angle_out, angle_in, curve, power=1.0, radius=0.1, samples=4):
    return dict(type=api.LTR_LT_SPOT, position=list(map(float, pos)), direction=list(map(float, direction)),
                up_direction=[1.0, 0.0, 0.0], color_rgb=list(map(float, color)), range=float(rng_),
                power=float(power), light_radius=float(radius), shadow_sample_count=samples,
                spot_angle_out=float(angle_out), spot_angle_in=float(angle_in), spot_curve=float(curve))


def dir_light(direction, color, rng_, radius=0.1, samples=1):
    return dict(type=api.LTR_LT_DIRECT, direction=list(map(float, direction)), color_rgb=list(map(float, color)),
                range=float(rng_), power=1.0, light_radius=float(radius), shadow_sample_count=samples)


# --------------------------------------------------------------------------
# configurations
# --------------------------------------------------------------------------
def room_scene(seed=SEED_BASE + 0, lm=256, n_boxes=20, wall_tess=8, n_point=2, with_sun=True,
               bounce_count=1, ao_distance=0.0, ao_samples=17, blur_size=0.5, room=(10.0, 10.0, 4.0),
               name="C0"):
    """Room 10x10x4 m, 6 inward walls of wall_tess^2 quads, n_boxes axis-aligned
    boxes standing on the floor; one instance with one lm x lm atlas."""
    g = np.random.Generator(np.random.PCG64(seed))
    sx, sy, sz = room
    charts = room_charts(room, wall_tess, part=0)
    albedo = {0: (0.8, 0.8, 0.8)}
    for b in range(n_boxes):
        size = g.uniform(0.2, 1.5, 3)
        cx = g.uniform(size[0] / 2, sx - size[0] / 2)
        cy = g.uniform(size[1] / 2, sy - size[1] / 2)
        lo = np.array([cx - size[0] / 2, cy - size[1] / 2, 0.0])
        hi = lo + size
        charts += box_charts(lo, hi, part=1 + b)
        albedo[1 + b] = tuple(float(x) for x in g.uniform(0.2, 0.9, 3))
    parts = charts_to_part_arrays(charts, lm, lm)
    mesh = Mesh("room", parts, [Instance(IDENTITY.copy(), 1.0, 1, "room0")])
    lights = []
    for _ in range(n_point):
        p = (g.uniform(1, sx - 1), g.uniform(1, sy - 1), g.uniform(2.5, 3.5))
        lights.append(point_light(p, g.uniform(0.5, 1.0, 3), 8.0, radius=0.1, samples=4))
    if with_sun:
        d = np.array([g.uniform(-1, 1), g.uniform(-1, 1), g.uniform(0.5, 1.0)])
        diag = math.sqrt(sx * sx + sy * sy + sz * sz)
        lights.append(dir_light(d / np.linalg.norm(d), g.uniform(0.5, 1.0, 3), diag, radius=0.1, samples=4))
    cfg = dict(bounce_count=bounce_count, ao_distance=ao_distance, ao_num_samples=ao_samples,
               blur_size=blur_size, max_num_threads=1)
    return SceneSpec(name, [mesh], lights, cfg, lm_size=lm, albedo=albedo if bounce_count else None)


def value_noise(g, n, octaves=5, base_cells=4):
    """Fractal value noise on an (n+1)^2 vertex grid, range about [-1,1]."""
    out = np.zeros((n + 1, n + 1))
    amp, total = 1.0, 0.0
    cells = base_cells
    coords = np.linspace(0.0, 1.0, n + 1)
    for _ in range(octaves):
        lat = g.uniform(-1.0, 1.0, (cells + 1, cells + 1))
        f = coords * cells
        i0 = np.minimum(f.astype(int), cells - 1)
        t = f - i0
        t = t * t * (3 - 2 * t)
        a = lat[i0][:, i0]
        b = lat[i0][:, i0 + 1]
        c = lat[i0 + 1][:, i0]
        d = lat[i0 + 1][:, i0 + 1]
        ty = t[:, None]
        tx = t[None, :]
        out += amp * ((a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty)
        total += amp
        amp *= 0.5
        cells *= 2
    return out / total


def terrain_scene(seed=SEED_BASE + 1, grid=512, extent=200.0, amplitude=20.0, n_boxes=2000, tiles=16,
                  lm_total=2048, n_lights=32, sun=True, sun_range=300.0, name="C1",
                  box_size=(0.5, 4.0), light_range=(10.0, 40.0)):
    """Heightfield terrain (grid^2 quads over extent m) split into tiles x tiles
    meshes, plus n_boxes boxes merged into the tile under their centre; each
    tile has one instance with an (lm_total/tiles)^2 atlas."""
    g = np.random.Generator(np.random.PCG64(seed))
    H = value_noise(g, grid) * amplitude
    cell = extent / grid
    q = grid // tiles
    lm = lm_total // tiles
    tile_charts = [[[] for _ in range(tiles)] for _ in range(tiles)]
    for ty in range(tiles):
        for tx in range(tiles):
            o = np.array([tx * q * cell, ty * q * cell, 0.0])
            h = H[ty * q: ty * q + q + 1, tx * q: tx * q + q + 1]
            tile_charts[ty][tx].append(Chart(o, np.array([q * cell, 0, 0]), np.array([0, q * cell, 0]), q, q, 0,
                                             height=h.copy()))

    def height_at(x, y):
        fx, fy = x / cell, y / cell
        i, j = min(int(fx), grid - 1), min(int(fy), grid - 1)
        return H[j, i]

    tile_boxes = [[[] for _ in range(tiles)] for _ in range(tiles)]
    for b in range(n_boxes):
        size = g.uniform(box_size[0], box_size[1], 3)
        cx = g.uniform(size[0], extent - size[0])
        cy = g.uniform(size[1], extent - size[1])
        z0 = height_at(cx, cy) - 0.3
        lo = np.array([cx - size[0] / 2, cy - size[1] / 2, z0])
        tx = min(int(cx / (q * cell)), tiles - 1)
        ty = min(int(cy / (q * cell)), tiles - 1)
        tile_boxes[ty][tx].append((lo, lo + size))
    meshes = []
    for ty in range(tiles):
        for tx in range(tiles):
            bx = tile_boxes[ty][tx]
            quads = [box_quads([lo for lo, _ in bx], [hi for _, hi in bx], part=1)] if bx else []
            parts = charts_to_part_arrays(tile_charts[ty][tx], lm, lm, quads=quads)
            ident = f"tile_{tx}_{ty}"
            meshes.append(Mesh(ident, parts, [Instance(IDENTITY.copy(), 1.0, 1, ident)]))
    lights = []
    for i in range(n_lights):
        x, y = g.uniform(5, extent - 5, 2)
        z = height_at(x, y) + g.uniform(2.0, 8.0)
        col = g.uniform(0.3, 1.0, 3)
        rng_ = g.uniform(*light_range)
        if i % 2 == 0:
            lights.append(point_light((x, y, z), col, rng_, radius=0.5, samples=16))
        else:
            d = np.array([g.uniform(-0.5, 0.5), g.uniform(-0.5, 0.5), -1.0])
            lights.append(spot_light((x, y, z), d / np.linalg.norm(d), col, rng_, 60.0, 40.0, 1.0,
                                     radius=0.5, samples=16))
    if sun:
        d = np.array([0.35, -0.25, 0.9])
        lights.append(dir_light(d / np.linalg.norm(d), (1.0, 0.95, 0.9), sun_range, radius=0.5, samples=8))
    cfg = dict(bounce_count=0, ao_distance=0.0, max_num_threads=1)
    return SceneSpec(name, meshes, lights, cfg, lm_size=lm)


def city_scene(seed=SEED_BASE + 4, extent=1000.0, grid=256, amplitude=6.0, n_boxes=155_000, tiles=16,
               lm_total=4096, n_lights=64, name="C4", light_range=(20.0, 80.0)):
    """Seeded city block at scale: a gentle heightfield over extent m (grid^2
    quads) plus n_boxes boxes with lognormal(mu=1, sigma=0.6) m sides on a
    jittered grid, merged into tiles x tiles meshes (one instance and one
    (lm_total/tiles)^2 atlas each); n_lights point/spot lights above the
    roofs.  With the defaults: 131,072 + 1,860,000 = 1,991,072 triangles."""
    g = np.random.Generator(np.random.PCG64(seed))
    H = value_noise(g, grid) * amplitude
    cell = extent / grid
    q = grid // tiles
    lm = lm_total // tiles
    tile_charts = [[[] for _ in range(tiles)] for _ in range(tiles)]
    for ty in range(tiles):
        for tx in range(tiles):
            o = np.array([tx * q * cell, ty * q * cell, 0.0])
            h = H[ty * q: ty * q + q + 1, tx * q: tx * q + q + 1]
            tile_charts[ty][tx].append(Chart(o, np.array([q * cell, 0, 0]), np.array([0, q * cell, 0]), q, q, 0,
                                             height=h.copy()))

    def height_at(x, y):
        i, j = min(int(x / cell), grid - 1), min(int(y / cell), grid - 1)
        return H[j, i]

    side = int(np.ceil(np.sqrt(n_boxes)))
    pitch = extent / side
    sizes = np.minimum(g.lognormal(1.0, 0.6, (n_boxes, 3)), 0.9 * pitch * np.array([1, 1, 8]))
    jitter = g.uniform(-0.5, 0.5, (n_boxes, 2))
    tile_m = q * cell
    # every box at once: grid cell, jittered centre within the cell's slack,
    # floor at the terrain height under the centre minus 0.3 m
    b = np.arange(n_boxes)
    gx, gy = (b % side).astype(np.float64), (b // side).astype(np.float64)
    slack = np.maximum(pitch - sizes[:, :2], 0.0)
    cx = (gx + 0.5) * pitch + jitter[:, 0] * slack[:, 0]
    cy = (gy + 0.5) * pitch + jitter[:, 1] * slack[:, 1]
    hz = H[np.minimum((cy / cell).astype(np.int64), grid - 1), np.minimum((cx / cell).astype(np.int64), grid - 1)]
    lo = np.stack([cx - sizes[:, 0] / 2, cy - sizes[:, 1] / 2, hz - 0.3], 1)
    hi = lo + sizes
    tile = (np.minimum((cy / tile_m).astype(np.int64), tiles - 1) * tiles
            + np.minimum((cx / tile_m).astype(np.int64), tiles - 1))
    order = np.argsort(tile, kind="stable")
    bounds = np.searchsorted(tile[order], np.arange(tiles * tiles + 1))
    meshes = []
    for ty in range(tiles):
        for tx in range(tiles):
            t = ty * tiles + tx
            sel = order[bounds[t]:bounds[t + 1]]
            quads = [box_quads(lo[sel], hi[sel], part=1)] if len(sel) else []
            parts = charts_to_part_arrays(tile_charts[ty][tx], lm, lm, quads=quads)
            ident = f"block_{tx}_{ty}"
            meshes.append(Mesh(ident, parts, [Instance(IDENTITY.copy(), 1.0, 1, ident)]))
    lights = []
    for i in range(n_lights):
        x, y = g.uniform(10, extent - 10, 2)
        z = height_at(x, y) + g.uniform(6.0, 20.0)
        col = g.uniform(0.3, 1.0, 3)
        rng_ = g.uniform(*light_range)
        if i % 2 == 0:
            lights.append(point_light((x, y, z), col, rng_, radius=0.5, samples=16))
        else:
            d = np.array([g.uniform(-0.5, 0.5), g.uniform(-0.5, 0.5), -1.0])
            lights.append(spot_light((x, y, z), d / np.linalg.norm(d), col, rng_, 60.0, 40.0, 1.0,
                                     radius=0.5, samples=16))
    cfg = dict(bounce_count=0, ao_distance=0.0, max_num_threads=1)
    return SceneSpec(name, meshes, lights, cfg, lm_size=lm)


def basic_scene():
    """lighter_test.cpp:183-257: a quad with three instances and one point light."""
    pos = np.array([-1, -1, 0, 1, -1, 0, 1, 1, 0, -1, 1, 0], np.float32).reshape(4, 3)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
    tex = np.array([0.1, 0.1, 0.9, 0.1, 0.9, 0.9, 0.1, 0.9], np.float32).reshape(4, 2)
    idx = np.array([0, 1, 2, 2, 3, 0], np.uint32)
    part = Part(pos, nrm, tex, tex.copy(), idx, 1)
    # cos(0.5f) * 0.2f in float arithmetic with the C library's cosf/sinf
    libm = C.CDLL("libm.so.6")
    libm.cosf.restype = libm.sinf.restype = C.c_float
    libm.cosf.argtypes = libm.sinf.argtypes = [C.c_float]
    c = np.float32(libm.cosf(0.5)) * np.float32(0.2)
    s = np.float32(libm.sinf(0.5)) * np.float32(0.2)
    m1 = np.eye(4, dtype=np.float32)
    m2 = np.array([[c, s, 0, 0], [-s, c, 0, 0], [0, 0, 0.2, 0], [0, 0, 0.5, 1]], np.float32)
    m3 = np.array([[1.9, 0, 0, 0], [0, 1.9, 0, 0], [0, 0, 1.9, 0], [0, 0, -0.5, 1]], np.float32)
    mesh = Mesh("mesh1", [part], [Instance(m, 1.0, 1, "") for m in (m1, m2, m3)])
    lights = [point_light((-0.4, -0.4, 1.0), (0.9, 0.1, 0.05), 4.0, 1.0, 0.1, 9)]
    return SceneSpec("basic", [mesh], lights, dict(max_num_threads=1))


def sliver_scene(seed=SEED_BASE + 0x51, lm=64):
    """Edge cases of the triangle set (no counterpart among the reference's
    scenarios): a 4 x 4 m floor, 24 standing slivers whose cross products
    have a component within a few ulps of the 0.001 threshold of
    Triangle::CheckIsUseful (lighter_int.hpp:604-608; some are kept, some
    dropped), and a part of 7 indices whose last triangle takes its two
    missing indices from the next part (lighter.cpp:41-47,367-373; index
    counts other than 1 and 2 are accepted, lighter.cpp:1237).  One point
    light and a sun."""
    rng = np.random.default_rng(seed)
    f32 = np.float32
    # floor: 4 x 4 quads
    n = 4
    g = np.linspace(-2, 2, n + 1, dtype=f32)
    xx, yy = np.meshgrid(g, g)
    fpos = np.stack([xx.ravel(), yy.ravel(), np.zeros((n + 1) ** 2, f32)], 1).astype(f32)
    fnrm = np.tile(np.array([0, 0, 1], f32), ((n + 1) ** 2, 1))
    ft2 = (0.02 + 0.46 * (fpos[:, :2] + 2) / 4).astype(f32)
    fidx = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            fidx += [a, a + 1, a + n + 2, a + n + 2, a + n + 1, a]
    floor = Part(fpos, fnrm, ft2.copy(), ft2, np.array(fidx, np.uint32), 1)
    # slivers: P1, P1 + L e_u, P1 + h e_v with L h within ulps of 0.001; the
    # cross product then has the single component +-L h on the third axis
    spos, snrm, st2, sidx = [], [], [], []
    axes = [(0, 2, 1), (2, 1, 0), (1, 2, 0)]  # (u axis, v axis, normal axis)
    for k in range(24):
        u, v, w = axes[k % 3]
        L = f32(rng.uniform(0.2, 0.6))
        target = np.nextafter(f32(0.001), f32(1 if k % 2 else -1), dtype=f32)
        for _ in range(k % 4):
            target = np.nextafter(target, f32(1 if k % 2 else -1), dtype=f32)
        h = f32(target / L)
        P1 = np.array([rng.uniform(-1.6, 1.6), rng.uniform(-1.6, 1.6), rng.uniform(0.2, 1.0)], f32)
        if w == 2:
            P1[2] = f32(0.5 + 0.02 * k)
        P2 = P1.copy(); P2[u] = f32(P2[u] + L)
        P3 = P1.copy(); P3[v] = f32(P3[v] + h)
        nn = np.zeros(3, f32); nn[w] = 1
        base = len(spos)
        spos += [P1, P2, P3]
        snrm += [nn, nn, nn]
        cx, cy = 0.52 + 0.45 * (k % 6) / 6, 0.02 + 0.45 * (k // 6) / 4
        st2 += [(cx, cy), (cx + 0.06, cy), (cx, cy + 0.09)]
        sidx += [base, base + 1, base + 2]
    slivers = Part(np.array(spos, f32), np.array(snrm, f32), np.array(st2, f32), np.array(st2, f32),
                   np.array(sidx, np.uint32), 1)
    # a ragged part: a standing quad (vertices 0-3) plus vertices 4, 5 that
    # only the incomplete last triangle (4, next part's first two indices)
    # reaches
    rp = np.array([[-1.5, 1.2, 0.0], [-0.5, 1.2, 0.0], [-0.5, 1.2, 1.0], [-1.5, 1.2, 1.0],
                   [0.5, -1.5, 0.0], [1.5, -1.5, 0.0]], f32)
    rn = np.tile(np.array([0, -1, 0], f32), (6, 1))
    rt2 = np.array([[0.52, 0.55], [0.72, 0.55], [0.72, 0.75], [0.52, 0.75], [0.76, 0.55], [0.95, 0.55]], f32)
    ragged = Part(rp, rn, rt2.copy(), rt2, np.array([0, 1, 2, 2, 3, 0, 4], np.uint32), 1)
    # the part after it: a small standing triangle whose first two indices
    # (5, 2) are also valid in the ragged part's vertex range
    tp = np.array([[1.0, 0.5, 0.0], [1.4, 0.5, 0.0], [1.2, 0.5, 0.8], [1.0, 0.9, 0.0], [1.4, 0.9, 0.0],
                   [1.2, 0.9, 0.8]], f32)
    tn = np.tile(np.array([0, -1, 0], f32), (6, 1))
    tt2 = np.array([[0.52, 0.80], [0.60, 0.80], [0.56, 0.90], [0.64, 0.80], [0.72, 0.80], [0.68, 0.90]], f32)
    tail = Part(tp, tn, tt2.copy(), tt2, np.array([5, 2, 0, 3, 4, 5], np.uint32), 1)
    mesh = Mesh("slivers", [floor, slivers, ragged, tail], [Instance(IDENTITY.copy(), 1.0, 1, "")])
    lights = [point_light((0.3, 0.2, 2.5), (0.9, 0.8, 0.7), 8.0, 1.0, 0.1, 4),
              dir_light((0.3, 0.2, 1.0), (0.4, 0.4, 0.5), 10.0, radius=0.05)]
    return SceneSpec("slivers", [mesh], lights, dict(max_num_threads=1, blur_size=0.0), lm_size=lm)


# --------------------------------------------------------------------------
# building a scene through the ltr_* API
# --------------------------------------------------------------------------
class BuiltScene:
    """Owns the ctypes callbacks and keeps the scene pointer."""

    def __init__(self, lib, spec: SceneSpec, config_overrides: Optional[dict] = None):
        self.lib = lib
        self.spec = spec
        self.keep = []
        self.scene = lib.ltr_CreateScene()
        cfg = api.Config()
        lib.ltr_GetConfig(C.byref(cfg), self.scene)
        overrides = dict(spec.config)
        overrides.update(config_overrides or {})
        for k, v in overrides.items():
            if isinstance(getattr(cfg, k), C.Array):
                for i, x in enumerate(v):
                    getattr(cfg, k)[i] = x
            else:
                setattr(cfg, k, v)
        if spec.lm_size:
            # ltr_DefaultSizeFunc refuses any size above max_lightmap_size, and
            # the bake then uses default_width/height (lighter.cpp:19-27,325-330)
            cfg.default_width = spec.lm_size
            cfg.default_height = spec.lm_size
            cfg.max_lightmap_size = 0
        if spec.albedo is not None or spec.emission is not None or spec.sample_cb is not None:
            alb = spec.albedo or {}
            emi = spec.emission or {}
            extra = spec.sample_cb

            def sample_fn(cfgp, req):
                r = req.contents
                a = alb.get(r.part_id)
                if a is not None:
                    r.out_diffuse_color[0], r.out_diffuse_color[1], r.out_diffuse_color[2] = a
                e = emi.get(r.part_id)
                if e is not None:
                    r.out_emissive_color[0], r.out_emissive_color[1], r.out_emissive_color[2] = e
                if extra is not None:
                    extra(r)
                return 1

            cb = api.SAMPLE_FN(sample_fn)
            self.keep.append(cb)
            cfg.sample_fn = C.cast(cb, C.c_void_p).value
        self.config = cfg
        lib.ltr_SetConfig(self.scene, C.byref(cfg))
        for m in spec.meshes:
            mp = lib.ltr_CreateMesh(self.scene, m.ident.encode(), len(m.ident))
            for p in m.parts:
                info = api.MeshPartInfo(
                    p.positions.ctypes.data, p.normals.ctypes.data, p.tex1.ctypes.data, p.tex2.ctypes.data,
                    12, 12, 8, 8, p.indices.ctypes.data_as(C.POINTER(C.c_uint32)),
                    len(p.positions), len(p.indices), p.shadow)
                assert lib.ltr_MeshAddPart(mp, C.byref(info)) == 1
            for inst in m.instances:
                ii = api.MeshInstanceInfo()
                for r in range(4):
                    for c_ in range(4):
                        ii.matrix[r][c_] = float(inst.matrix[r, c_])
                ii.importance = inst.importance
                ii.shadow = inst.shadow
                ident = inst.ident.encode()
                self.keep.append(ident)
                ii.ident = ident
                ii.ident_size = len(ident)
                assert lib.ltr_MeshAddInstance(mp, C.byref(ii)) == 1
        for L in spec.lights:
            li = api.LightInfo()
            for k, v in L.items():
                if isinstance(getattr(li, k), C.Array):
                    for i, x in enumerate(v):
                        getattr(li, k)[i] = x
                else:
                    setattr(li, k, v)
            lib.ltr_LightAdd(self.scene, C.byref(li))
        for k, (p, n) in enumerate(spec.samples):
            si = api.SampleInfo()
            si.id = k
            for i in range(3):
                si.position[i] = p[i]
                si.normal[i] = n[i]
            lib.ltr_SampleAdd(self.scene, C.byref(si))

    def bake(self, poll_ms: int = 1, status_cb: Optional[Callable] = None):
        """ltr_Start + ltr_GetStatus polling; returns {stage: seconds}."""
        import time
        lib = self.lib
        st = api.WorkStatus()
        times = {}
        t_last = time.perf_counter()
        lib.ltr_Start(self.scene)
        cur = None
        while True:
            busy = lib.ltr_GetStatus(self.scene, C.byref(st))
            now = time.perf_counter()
            stage = st.stage.decode() if st.stage else None
            if cur is not None:
                times[cur] = times.get(cur, 0.0) + (now - t_last)
            t_last = now
            cur = stage
            if status_cb:
                status_cb(stage, st.completion)
            if not busy:
                break
            time.sleep(poll_ms / 1000.0)
        return times

    def outputs(self):
        """List of (uid, width, height, rgb (H,W,3) copy, normals or None)."""
        lib = self.lib
        info = api.WorkOutputInfo()
        lib.ltr_GetWorkOutputInfo(self.scene, C.byref(info))
        res = []
        for k in range(info.lightmap_count):
            wo = api.WorkOutput()
            assert lib.ltr_GetWorkOutput(self.scene, k, C.byref(wo)) == 1
            n = wo.width * wo.height
            rgb = np.ctypeslib.as_array(wo.lightmap_rgb, (n * 3,)).reshape(wo.height, wo.width, 3).copy()
            nm = None
            if wo.normals_xyzf:
                nm = np.ctypeslib.as_array(wo.normals_xyzf, (n * 4,)).reshape(wo.height, wo.width, 4).copy()
            res.append((wo.uid, wo.width, wo.height, rgb, nm))
        return res

    def sample_outputs(self):
        info = api.WorkOutputInfo()
        self.lib.ltr_GetWorkOutputInfo(self.scene, C.byref(info))
        out = np.zeros((info.sample_count, 3), np.float32)
        for i in range(info.sample_count):
            out[i] = info.samples[i].out_color[:]
        return out

    def destroy(self):
        if self.scene:
            self.lib.ltr_DestroyScene(self.scene)
            self.scene = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

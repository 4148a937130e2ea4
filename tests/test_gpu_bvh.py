"""The device scene's BVH built on the GPU (bvh_gpu.hip) against the host
build (bvh.cpp): the binned-SAH decisions depend only on sets of triangles
(bounds, bin counts and boxes, stable partitions, the split by position),
so the GPU's level-by-level build must give the host's tree exactly: the
4-wide node arrays identical byte for byte, the same leaf order and depth.
Soups: the room, the sliver scene (near-degenerate triangles), the terrain
workload (548,288 triangles), a city block, coincident centroids (every
split by position), tiny soups (1-3 triangles: the leaf-only root) and a
random soup.  That the tree answers every query exactly is checked by the
query, march and AO tests, which run on this build.

Local queries (distance queries of radius <= 2, segments of at most 2 m)
start at their cell of the entry grid, whose node lists at most four
subtrees holding every triangle within 2 m of the cell: every answer must
equal, bit for bit, the answer of the walk from the root."""
import ctypes as C

import numpy as np
import pytest

from helpers import Device, all_samples_and_tris, bits, prepared
from lighter_amd import _lib, scenes

pytestmark = pytest.mark.gpu


def _compare(lib, tris):
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    out = (C.c_double * 6)()
    ok = lib.lgx_BvhCompare(tris.ctypes.data_as(C.POINTER(C.c_float)), len(tris), out)
    assert ok == 1, _lib.last_error(lib)
    return list(out)


@pytest.mark.parametrize("name,make", [
    ("room", lambda: scenes.room_scene(lm=64, n_boxes=20)),
    ("sliver", scenes.sliver_scene),
    ("terrain", scenes.terrain_scene),
    ("city_block", lambda: scenes.city_scene(extent=200.0, grid=64, n_boxes=12000, tiles=4, lm_total=1024,
                                             n_lights=4)),
])
def test_gpu_bvh_equals_host(lib, name, make):
    b = prepared(lib, make())
    _, _, tris = all_samples_and_tris(lib, b)
    out = _compare(lib, tris)
    assert out[0] == 1 and out[1] == 1, (name, out)
    assert out[5] > 0
    print(f"{name}: {len(tris)} triangles, host {out[2]:.1f} ms, GPU {out[3]:.1f} ms ({out[4]:.1f} ms device)")
    b.destroy()


def test_gpu_bvh_special_soups(lib):
    rng = np.random.default_rng(77)
    one = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)
    soups = {
        "coincident": np.tile(one, 1000),                     # every split by position
        "one": one, "two": np.tile(one, 2), "three": np.tile(one, 3),
        "random": rng.uniform(-10, 10, 9 * 20000).astype(np.float32),
        "flat": np.concatenate([one + np.array([i, 0, 0] * 3, np.float32) for i in range(300)]),
    }
    for name, t in soups.items():
        out = _compare(lib, t)
        assert out[0] == 1 and out[1] == 1, (name, out)


def _entry_grid(lib, dev, on):
    cells = C.c_int64(0)
    _lib.check(lib.lgx_SetEntryGrid(dev.h, on, C.byref(cells)), lib)
    return cells.value


@pytest.mark.parametrize("name", ["room", "terrain", "city_block", "multiscale"])
def test_entry_grid_answers_as_the_root(lib, name):
    """Distance queries, closest-hit and occlusion segments (short ones from
    the grid, long ones from the root) and shadow marches with the entry
    grid on and off: identical bits.  Points lie near the surfaces, in the
    open, on cell faces and outside the grid."""
    rng = np.random.default_rng(80)
    if name == "multiscale":
        # small triangles scattered over 100 km beside metre-sized ones
        tiny = (rng.uniform(-5e4, 5e4, (4000, 1, 3)) + rng.uniform(0, 1e-3, (4000, 3, 3))).reshape(-1, 9)
        tris = np.ascontiguousarray(np.concatenate([tiny, rng.uniform(-5, 5, (4000, 9))]), np.float32)
        a = tris[rng.choice(len(tris), 20000), :3] + rng.normal(0, 1.0, (20000, 3))
    else:
        spec = {"room": lambda: scenes.room_scene(lm=48),
                "terrain": scenes.terrain_scene,
                "city_block": lambda: scenes.city_scene(extent=200.0, grid=64, n_boxes=12000, tiles=4,
                                                        lm_total=1024, n_lights=4)}[name]()
        b = prepared(lib, spec)
        pos, nrm, tris = all_samples_and_tris(lib, b)
        b.destroy()
        idx = rng.choice(len(pos), min(len(pos), 20000), replace=False)
        a = pos[idx] + nrm[idx] * rng.uniform(0.0, 1.5, (len(idx), 1))
    t = tris.reshape(-1, 3)
    lo, hi = t.min(0), t.max(0)
    # open space, and points around and beyond the grown scene box
    a = np.concatenate([a, rng.uniform(lo - 3, hi + 3, (5000, 3))])
    a = np.ascontiguousarray(a, np.float32)
    d = rng.normal(size=a.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    near = np.ascontiguousarray(a + d * rng.uniform(0.05, 2.0, (len(a), 1)), np.float32)
    far = np.ascontiguousarray(a + d * rng.uniform(2.0, 40.0, (len(a), 1)), np.float32)
    k = rng.choice([0.0, 0.5, 2.0], len(a)).astype(np.float32)
    dev = Device(lib, tris)
    try:
        res = []
        for on in (0, 1):
            cells = _entry_grid(lib, dev, on)
            res.append((dev.distance(a), dev.closest(a, near), dev.occluded(a, near), dev.closest(a, far),
                        dev.occluded(a, far), dev.shadow(a, far, k)))
        assert cells > 1000, cells
        for x, y in zip(res[0][:5], res[1][:5]):
            assert np.array_equal(bits(x) if x.dtype == np.float32 else x, bits(y) if y.dtype == np.float32 else y)
        assert np.array_equal(bits(res[0][5][0]), bits(res[1][5][0]))
        assert np.array_equal(res[0][5][1], res[1][5][1])
        dist = res[1][0]
        assert (dist < 2).mean() > 0.2 and (dist == 2).mean() > 0.05
        assert 0 < res[1][2].mean() < 1
        print(f"{name}: {len(tris)} triangles, {cells} cells")
    finally:
        dev.close()

"""The device scene's BVH built on the GPU (bvh_gpu.hip) against the host
build (bvh.cpp): the binned-SAH decisions depend only on sets of triangles
(bounds, bin counts and boxes, stable partitions, the split by position),
so the GPU's level-by-level build must give the host's tree exactly: the
4-wide node arrays identical byte for byte, the same leaf order and depth.
Soups: the room, the sliver scene (near-degenerate triangles), the terrain
workload (548,288 triangles), a city block, coincident centroids (every
split by position), tiny soups (1-3 triangles: the leaf-only root) and a
random soup.  That the tree answers every query exactly is checked by the
query, march and AO tests, which run on this build."""
import ctypes as C

import numpy as np
import pytest

from helpers import all_samples_and_tris, prepared
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

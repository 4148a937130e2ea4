"""Lightmap rasterisation and sample compaction on the GPU (raster.hip;
lighter.cpp:32-66, 392-420, lighter_math.cpp:245-302) against the host
rasteriser, which the golden fixtures and the sliver-scene reference test pin
to the reference bit for bit (test_golden.py, test_edge_geometry.py).  Both
evaluate the per-pixel functions of raster.h; the GPU resolves the drawing
order (two passes, later triangles overwriting earlier ones) by the largest
draw key per texel.  With the placement on the host in both runs, every
sample (position, normal, texel index, uv0 / part / area) must be identical,
on the room, the sliver scene (near-degenerate triangles and a ragged part),
the terrain workload (256 instances) and a dense city block."""
import ctypes as C

import numpy as np
import pytest

from helpers import HOST_PLACEMENT, PLACE_BIT, bits, prepared, prod_samples
from lighter_amd import scenes

pytestmark = pytest.mark.gpu


def samples_of(lib, b):
    return [prod_samples(lib, b.scene, k) for k in range(lib.lgx_InstanceCount(b.scene))]


@pytest.mark.parametrize("name,make", [
    ("room", lambda: scenes.room_scene(lm=96, n_boxes=20)),
    ("sliver", scenes.sliver_scene),
    ("terrain", scenes.terrain_scene),
    ("city_block", lambda: scenes.city_scene(extent=200.0, grid=64, n_boxes=12000, tiles=4, lm_total=1024,
                                             n_lights=4)),
])
def test_gpu_raster_equals_host(lib, name, make):
    spec = make()
    bh = prepared(lib, spec, HOST_PLACEMENT, max_num_threads=16)  # host rasterisation and placement
    bg = prepared(lib, spec, PLACE_BIT, max_num_threads=16)       # GPU rasterisation, host placement
    sh, sg = samples_of(lib, bh), samples_of(lib, bg)
    assert len(sh) == len(sg)
    total = 0
    for k, (h, g) in enumerate(zip(sh, sg)):
        for a, c in zip(h, g):  # pos, nrm, loc, info
            assert a.shape == c.shape, (name, k)
            assert np.array_equal(bits(a) if a.dtype == np.float32 else a,
                                  bits(c) if c.dtype == np.float32 else c), (name, k)
        total += len(h[0])
    t = (C.c_double * 3)()
    lib.lgx_RasterTimes(bg.scene, t)
    assert int(t[2]) == total and total > 0
    assert t[1] >= total
    lib.lgx_RasterTimes(bh.scene, t)
    assert list(t) == [0.0] * 3
    bh.destroy()
    bg.destroy()


def _overlap_scene():
    """Three parts drawn over one texel square: quads at z = 0 and z = 1 with
    the same chart, then a triangle at z = 2 over half of it."""
    f32 = np.float32
    quad_t = np.array([[0.1, 0.1], [0.9, 0.1], [0.9, 0.9], [0.1, 0.9]], f32)
    up = np.tile(np.array([0, 0, 1], f32), (4, 1))
    parts = []
    for z in (0.0, 1.0):
        pos = np.array([[0, 0, z], [2, 0, z], [2, 2, z], [0, 2, z]], f32)
        parts.append(scenes.Part(pos, up, quad_t.copy(), quad_t.copy(), np.array([0, 1, 2, 2, 3, 0], np.uint32), 1))
    tri_p = np.array([[0, 0, 2], [2, 0, 2], [2, 2, 2]], f32)
    parts.append(scenes.Part(tri_p, up[:3].copy(), quad_t[:3].copy(), quad_t[:3].copy(),
                             np.array([0, 1, 2], np.uint32), 1))
    mesh = scenes.Mesh("overlap", parts, [scenes.Instance(scenes.IDENTITY.copy(), 1.0, 1, "")])
    return scenes.SceneSpec("overlap", [mesh], [], dict(max_num_threads=1, blur_size=0.0), lm_size=32)


def test_gpu_raster_last_draw_wins(lib):
    """Charts drawn over each other in one atlas: the GPU's samples equal the
    host's, and no sample keeps the first quad's values (part 0), which the
    second quad (part 1) covers texel for texel in both passes; the
    triangle's part (2) appears too."""
    spec = _overlap_scene()
    bh = prepared(lib, spec, HOST_PLACEMENT, max_num_threads=1)
    bg = prepared(lib, spec, PLACE_BIT, max_num_threads=1)
    h, g = samples_of(lib, bh)[1], samples_of(lib, bg)[1]  # instance 0 holds the external samples
    for a, c in zip(h, g):
        assert np.array_equal(bits(a) if a.dtype == np.float32 else a, bits(c) if c.dtype == np.float32 else c)
    part = g[3][:, 2]
    assert set(np.unique(part).tolist()) == {1.0, 2.0}, np.unique(part)
    assert len(part) > 300
    bh.destroy()
    bg.destroy()

"""Lightmap export on the GPU (extrapolation, Gaussian blur, 2x downsample;
lighter.cpp:849-958, lighter_math.cpp:357-424) against the product's host
restatement, whose full bakes the golden fixtures and the live reference
build pin (test_golden.py, test_gpu_stages.py): complete bakes through
ltr_Start, identical but for the export stage, must give bit-identical
images, for blur radii with one to three taps per side, with and without
the downsample and the normal-map data, on one atlas, on the three
instances of the basic scenario and on 16 terrain tiles."""
import ctypes as C

import numpy as np
import pytest

from helpers import HOST_FINALIZE, bits
from lighter_amd import api, scenes

pytestmark = pytest.mark.gpu


def bake_outputs(lib, libc, spec, host_mask, **cfg):
    libc.srand(1)
    b = scenes.BuiltScene(lib, spec, dict(max_num_threads=16, **cfg))
    assert lib.lgx_SetHostStages(b.scene, host_mask) == 1
    b.bake()
    st = api.WorkStatus()
    lib.ltr_GetStatus(b.scene, C.byref(st))
    assert st.stage == b"finished", st.stage
    t = (C.c_double * 5)()
    lib.lgx_FinalizeTimes(b.scene, t)
    outs = b.outputs()
    b.destroy()
    return outs, list(t)


def assert_same(host, gpu):
    assert len(host) == len(gpu) > 0
    for (u1, w1, h1, rgb1, nm1), (u2, w2, h2, rgb2, nm2) in zip(host, gpu):
        assert (u1, w1, h1) == (u2, w2, h2)
        assert np.array_equal(bits(rgb1), bits(rgb2)), (u1, np.abs(rgb1 - rgb2).max())
        assert (nm1 is None) == (nm2 is None)
        if nm1 is not None:
            assert np.array_equal(bits(nm1), bits(nm2)), u1


@pytest.mark.parametrize("blur,ds2x,nm", [(0.0, 0, 0), (0.5, 0, 0), (1.7, 0, 1), (0.5, 1, 0), (2.5, 1, 1)])
def test_gpu_export_equals_host_room(lib, libc, blur, ds2x, nm):
    spec = scenes.room_scene(lm=64, n_boxes=6, bounce_count=0)
    cfg = dict(blur_size=blur, ds2x=ds2x, generate_normalmap_data=nm)
    host, th = bake_outputs(lib, libc, spec, HOST_FINALIZE, **cfg)
    gpu, tg = bake_outputs(lib, libc, spec, 0, **cfg)
    assert_same(host, gpu)
    assert th[4] == 0 and tg[4] == 64 * 64
    # the export changed the image: extrapolation filled texels beyond the samples
    assert np.count_nonzero(gpu[0][3].sum(axis=2)) > 0


@pytest.mark.parametrize("name,make", [
    ("basic", scenes.basic_scene),
    ("terrain_tiles", lambda: scenes.terrain_scene(grid=128, extent=50.0, n_boxes=100, tiles=4, lm_total=512,
                                                   n_lights=8, name="C1-small")),
])
def test_gpu_export_equals_host_instances(lib, libc, name, make):
    spec = make()
    for cfg in (dict(), dict(blur_size=1.2, ds2x=1)):
        host, _ = bake_outputs(lib, libc, spec, HOST_FINALIZE, **cfg)
        gpu, t = bake_outputs(lib, libc, spec, 0, **cfg)
        assert_same(host, gpu)
        print(f"{name} {cfg}: {len(gpu)} images, {int(t[4])} texels, kernels {t[2]:.3f} ms")


def test_gpu_normal_map_data_equals_host(lib, libc):
    """Normal-map data (lighter.cpp:505-513, 959-1030) from the GPU pass
    over the kept shadow factors against the host restatement's
    contribution lists, bit for bit, with point, spot and directional
    lights (every light's pow and acos evaluated alike on both sides), with
    and without the downsample."""
    spec = scenes.room_scene(lm=64, n_boxes=8, n_point=2, bounce_count=0)
    spec.lights.append(scenes.spot_light((5, 5, 3.5), (0.1, 0.0, -1.0), (0.9, 0.8, 0.7), 9.0, 50.0, 30.0, 0.7,
                                         radius=0.2))
    for ds2x in (0, 1):
        cfg = dict(generate_normalmap_data=1, ds2x=ds2x, blur_size=0.5)
        host, _ = bake_outputs(lib, libc, spec, HOST_FINALIZE, **cfg)
        gpu, _ = bake_outputs(lib, libc, spec, 0, **cfg)
        assert_same(host, gpu)
        nm = gpu[0][4]
        assert nm is not None
        md = nm.reshape(-1, 4)[:, 3]
        assert (md < 1).mean() > 0.05  # the min-dot term does something


"""The C-ABI library: exports, layouts and the host-only API behaviour
(no GPU needed).  Reference behaviour cited per test."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from lighter_amd import _lib, api, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("lighter.h", "lighter_gpu.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"LTRAPI[^;]*?\b((?:ltr|lgx)_\w+)\s*\(", txt, re.S))
    return names


def test_exports_every_declared_symbol():
    names = declared_symbols()
    assert len([n for n in names if n.startswith("ltr_")]) == 17
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = names - exported
    assert not missing, missing
    # nothing but the C entry points leaks out
    assert {e for e in exported if not e.startswith(("ltr_", "lgx_"))} <= {"_init", "_fini"}


def test_python_mirror_layout():
    api.check_layout()
    assert C.sizeof(_lib.LgxLight) == 64


def test_config_defaults(lib):
    """ltr_GetConfig(cfg, NULL) (lighter.cpp:1171-1208)."""
    cfg = api.Config()
    lib.ltr_GetConfig(C.byref(cfg), None)
    assert cfg.max_num_threads == 0x7FFF and cfg.max_lightmap_size == 1024
    assert (cfg.default_width, cfg.default_height) == (64, 64)
    assert cfg.global_size_factor == 4 and cfg.ao_num_samples == 17
    assert abs(cfg.max_correct_dist - 0.1) < 1e-7 and cfg.max_correct_angle == 60
    assert abs(cfg.ao_multiplier - 1.2) < 1e-7 and cfg.ao_falloff == 1 and cfg.blur_size == 0.5
    assert cfg.bounce_count == 0 and cfg.ao_distance == 0 and not cfg.sample_fn
    assert cfg.size_fn == C.cast(lib.ltr_DefaultSizeFunc, C.c_void_p).value


@pytest.mark.ref
def test_config_defaults_equal_reference(lib, ref):
    a, b = api.Config(), api.Config()
    lib.ltr_GetConfig(C.byref(a), None)
    ref.lib.ltr_GetConfig(C.byref(b), None)
    for name, _ in api.Config._fields_:
        if name in ("size_fn",):
            continue
        va, vb = getattr(a, name), getattr(b, name)
        if isinstance(va, C.Array):
            va, vb = list(va), list(vb)
        assert va == vb, name


def test_next_power_of_two(lib):
    for x, e in [(1, 1), (2, 2), (3, 4), (5, 8), (1000, 1024), (1024, 1024), (1025, 2048), (0, 0)]:
        assert lib.ltr_NextPowerOfTwo(x) == e


def test_default_size_func(lib):
    """lighter.cpp:7-28: 4*sqrt(area)*importance rounded up to a power of two."""
    cfg = api.Config()
    lib.ltr_GetConfig(C.byref(cfg), None)
    out = (C.c_uint32 * 2)()
    assert lib.ltr_DefaultSizeFunc(C.byref(cfg), b"", 0, b"", 0, 100.0, 1.0, out) == 1
    assert list(out) == [64, 64]
    assert lib.ltr_DefaultSizeFunc(C.byref(cfg), b"", 0, b"", 0, 1e6, 1.0, out) == 0  # > max size
    assert lib.ltr_DefaultSizeFunc(C.byref(cfg), b"", 0, b"", 0, 0.0, 1.0, out) == 1
    assert list(out) == [1, 1]


def test_mesh_add_part_rejects_short_index_lists(lib):
    """lighter.cpp:1237: index_count 1 or 2 is refused."""
    s = lib.ltr_CreateScene()
    m = lib.ltr_CreateMesh(s, b"m", 1)
    pos = np.zeros((3, 3), np.float32)
    uv = np.zeros((3, 2), np.float32)
    idx = np.array([0, 1, 2], np.uint32)
    for count, ok in [(1, 0), (2, 0), (3, 1), (0, 1)]:
        info = api.MeshPartInfo(pos.ctypes.data, pos.ctypes.data, uv.ctypes.data, uv.ctypes.data, 12, 12, 8, 8,
                                idx.ctypes.data_as(C.POINTER(C.c_uint32)), 3, count, 1)
        assert lib.ltr_MeshAddPart(m, C.byref(info)) == ok
    lib.ltr_DestroyScene(s)


def test_status_and_outputs_before_start(lib):
    s = lib.ltr_CreateScene()
    st = api.WorkStatus()
    assert lib.ltr_GetStatus(s, C.byref(st)) == 1
    assert st.stage == b"not started"
    info = api.WorkOutputInfo()
    lib.ltr_GetWorkOutputInfo(s, C.byref(info))
    assert info.lightmap_count == 0 and info.sample_count == 0
    wo = api.WorkOutput()
    assert lib.ltr_GetWorkOutput(s, 0, C.byref(wo)) == 0
    lib.ltr_DestroyScene(s)


def test_empty_scene_bake_finishes_without_gpu(lib):
    """An empty scene has no GPU stage (lighter.cpp:1050): it finishes."""
    s = lib.ltr_CreateScene()
    lib.ltr_Start(s)
    st = api.WorkStatus()
    for _ in range(2000):
        if not lib.ltr_GetStatus(s, C.byref(st)):
            break
        lib.ltr_Sleep(1)
    assert st.stage == b"finished" and st.completion == 1.0
    lib.ltr_DestroyScene(s)


def test_samples_only_scene_without_lights_finishes(lib):
    """No lights, no radiosity, no AO: with the host sample placement no
    GPU stage runs; external samples come back with the ambient colour
    (lighter.cpp:394-399, 854-862)."""
    spec = scenes.basic_scene()
    spec.lights = []
    spec.samples = [((0, 0, 1), (0, 0, 1)), ((1, 1, 1), (0, 1, 0))]
    spec.config = dict(spec.config, ambient_color=[0.25, 0.5, 0.75])
    b = scenes.BuiltScene(lib, spec)
    lib.lgx_SetHostStages(b.scene, 7)  # rasterisation, placement and export on the host
    times = b.bake()
    assert "rendering lightmaps" not in times
    outs = b.outputs()
    assert [o[0] for o in outs] == [1, 2, 3]
    np.testing.assert_array_equal(b.sample_outputs(), np.array([[0.25, 0.5, 0.75]] * 2, np.float32))
    b.destroy()


def test_bake_without_device_reports_failed(lib):
    """A GPU stage without a HIP device fails loudly: the status ends with
    "failed" instead of silently running on the CPU."""
    if lib.lgx_HasDevice():
        pytest.skip("a HIP device is present")
    b = scenes.BuiltScene(lib, scenes.basic_scene())
    st = api.WorkStatus()
    lib.ltr_Start(b.scene)
    while lib.ltr_GetStatus(b.scene, C.byref(st)):
        lib.ltr_Sleep(1)
    assert st.stage == b"failed"
    assert lib.lgx_StageDirect(b.scene) == 0
    assert "HIP" in _lib.last_error(lib)
    b.destroy()


def test_abort_stops_bake(lib):
    """ltr_Abort is a no-op in the reference (lighter.cpp:1154-1157); here the
    bake stops at the next stage boundary and reports finished with no
    outputs.  The abort is issued from the size callback, which the first
    stage calls (lighter.cpp:325-330), so it lands before any later stage
    whatever the thread timing."""
    spec = scenes.room_scene(lm=64)
    spec.lights = []
    b = scenes.BuiltScene(lib, spec)
    lib.lgx_SetHostStages(b.scene, 7)  # rasterisation, placement and export on the host: no GPU needed here
    calls = []

    def size_fn(cfgp, mesh, mesh_size, inst, inst_size, area, importance, out):
        calls.append(1)
        lib.ltr_Abort(b.scene)
        out[0] = out[1] = 64
        return 1

    cb = api.SIZE_FN(size_fn)
    cfg = api.Config()
    lib.ltr_GetConfig(C.byref(cfg), b.scene)
    cfg.size_fn = C.cast(cb, C.c_void_p).value
    lib.ltr_SetConfig(b.scene, C.byref(cfg))
    lib.ltr_Start(b.scene)
    st = api.WorkStatus()
    while lib.ltr_GetStatus(b.scene, C.byref(st)):
        lib.ltr_Sleep(1)
    assert calls, "the size callback did not run"
    assert st.stage == b"finished"
    info = api.WorkOutputInfo()
    lib.ltr_GetWorkOutputInfo(b.scene, C.byref(info))
    assert info.lightmap_count == 0
    b.destroy()


def test_second_bake_gives_the_same_images(lib):
    """A second ltr_Start bakes the scene again from its inputs: the same
    images at the same (downsampled) size, and no more outputs than the
    first bake made.  ds2x halves the exported image while the samples keep
    their full-resolution texel indices; a re-export must not scatter those
    into the halved image.  Rasterisation, placement and export on the
    host, no lights: no GPU stage runs."""
    spec = scenes.room_scene(lm=64, bounce_count=0)
    spec.lights = []
    b = scenes.BuiltScene(lib, spec, dict(ds2x=1, blur_size=1.0))
    assert lib.lgx_SetHostStages(b.scene, 7) == 1
    b.bake()
    first = b.outputs()
    assert len(first) == 1 and first[0][1] == first[0][2] == 32
    b.bake()
    second = b.outputs()
    assert len(second) == len(first)
    for f, s in zip(first, second):
        assert f[:3] == s[:3]
        np.testing.assert_array_equal(f[3].view(np.uint32), s[3].view(np.uint32))
    # a staged export after the bake uses the full-resolution atlas again
    assert lib.lgx_StageFinalize(b.scene) == 1, _lib.last_error(lib)
    third = b.outputs()
    assert [o[:3] for o in third] == [o[:3] for o in first]
    np.testing.assert_array_equal(third[0][3].view(np.uint32), first[0][3].view(np.uint32))
    b.destroy()


def test_host_stage_switch(lib):
    """lgx_SetHostStages takes the three documented bits and refuses others;
    a scene baked with every host form and no lights needs no GPU stage, and
    reports no GPU rasterisation, placement or export times."""
    spec = scenes.room_scene(lm=64, bounce_count=0)
    spec.lights = []
    b = scenes.BuiltScene(lib, spec)
    assert lib.lgx_SetHostStages(b.scene, 8) == 0
    assert "unknown bits" in _lib.last_error(lib)
    assert lib.lgx_SetHostStages(b.scene, 7) == 1
    st = api.WorkStatus()
    lib.ltr_Start(b.scene)
    while lib.ltr_GetStatus(b.scene, C.byref(st)):
        lib.ltr_Sleep(1)
    assert st.stage == b"finished"
    t = (C.c_double * 5)()
    lib.lgx_PlacementTimes(b.scene, t)
    assert list(t) == [0.0] * 5
    lib.lgx_FinalizeTimes(b.scene, t)
    assert list(t) == [0.0] * 5
    lib.lgx_RasterTimes(b.scene, t)
    assert list(t)[:3] == [0.0] * 3
    outs = b.outputs()
    assert len(outs) == 1 and outs[0][1] == outs[0][2] == 64
    b.destroy()

"""lgx_DirectLightNoAlloc: the direct-light pass with no device allocation,
on caller-owned device buffers (hipMalloc'd by the test), against
lgx_DirectLight on the same device and work: lightmap, weights and shadow
factors bit for bit, on two calls in a row (the second with fewer work
items, so the scratch array holds the first call's contents when it
starts), and the documented refusals (too many items, an item naming a
sample beyond the device's samples, a host array where a device buffer is
expected) as errors that leave the entry point usable."""
import ctypes as C

import numpy as np
import pytest

from hipmem import DeviceBuffer
from lighter_amd import _lib, scenes, workloads

pytestmark = pytest.mark.gpu

fp = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
MAX_ITEMS = 1 << 22


def P(a):
    return a.ctypes.data_as(fp)


@pytest.fixture(scope="module")
def room(lib):
    if not lib.lgx_HasDevice():
        pytest.skip("no HIP device")
    spec = scenes.room_scene(lm=96, n_point=3)
    b, pos, nrm, tris, items, off = workloads.direct_workload(lib, spec, 4)
    dev = lib.lgx_DeviceCreate(P(tris), len(tris))
    assert dev, _lib.last_error(lib)
    yield dict(lib=lib, spec=spec, pos=pos, nrm=nrm, items=items, off=off, dev=dev,
               lights=workloads.make_lights(spec))
    lib.lgx_DeviceDestroy(dev)
    b.destroy()


def _reset(r):
    S = len(r["pos"])
    _lib.check(r["lib"].lgx_SetSamples(r["dev"], P(r["pos"]), P(r["nrm"]), S, P(np.zeros((S, 3), np.float32))),
               r["lib"])


def _lightmap(r):
    lm = np.zeros((len(r["pos"]), 3), np.float32)
    _lib.check(r["lib"].lgx_ReadLightmap(r["dev"], P(lm)), r["lib"])
    return lm


def _existing(r, nl, items, off):
    lib = r["lib"]
    _reset(r)
    _lib.check(lib.lgx_DirectLight(r["dev"], r["lights"], nl, items.ctypes.data_as(u32p), len(items),
                                   off.ctypes.data_as(u32p), 1), lib)
    w = np.zeros(len(items), np.float32)
    v = np.zeros(len(items), np.float32)
    _lib.check(lib.lgx_ReadWeights(r["dev"], P(w), P(v), len(items)), lib)
    return _lightmap(r), w, v


def _no_alloc(r, nl, items, off):
    lib = r["lib"]
    _reset(r)
    d_items = DeviceBuffer.of(items)
    d_w = DeviceBuffer.of(np.full(len(items), 7.0, np.float32))
    d_v = DeviceBuffer.of(np.full(len(items), 7.0, np.float32))
    ok = lib.lgx_DirectLightNoAlloc(r["dev"], r["lights"], nl, d_items.addr, len(items),
                                    off.ctypes.data_as(u32p), d_w.addr, d_v.addr)
    assert ok == 1, _lib.last_error(lib)
    out = _lightmap(r), d_w.download(np.float32, len(items)), d_v.download(np.float32, len(items))
    for b in (d_items, d_w, d_v):
        b.free()
    return out


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_no_alloc_matches_direct_light_twice(room):
    items, off = room["items"], room["off"]
    nl = len(room["lights"])
    assert 0 < len(items) <= MAX_ITEMS
    # first call: every light's items
    exp = _existing(room, nl, items, off)
    got = _no_alloc(room, nl, items, off)
    for g, e in zip(got, exp):
        assert np.array_equal(_bits(g), _bits(e))
    assert (exp[0] > 0).mean() > 0.2 and (exp[2] < 1).any()
    # second call, smaller: the first two lights' items only
    items2 = np.ascontiguousarray(items[: off[2]])
    off2 = np.ascontiguousarray(off[:3])
    assert len(items2) < len(items)
    exp2 = _existing(room, 2, items2, off2)
    got2 = _no_alloc(room, 2, items2, off2)
    for g, e in zip(got2, exp2):
        assert np.array_equal(_bits(g), _bits(e))


def test_no_alloc_refusals(room):
    lib = room["lib"]
    small = DeviceBuffer.of(np.zeros(16, np.uint32))
    out = DeviceBuffer.of(np.zeros(16, np.float32))
    # more items than the scratch holds: refused before any device access
    off = np.array([0, MAX_ITEMS + 1], np.uint32)
    assert lib.lgx_DirectLightNoAlloc(room["dev"], room["lights"], 1, small.addr, MAX_ITEMS + 1,
                                      off.ctypes.data_as(u32p), out.addr, None) == 0
    assert "exceed" in _lib.last_error(lib)
    # an item naming a sample beyond the device's samples
    bad = np.full(16, len(room["pos"]) + 5, np.uint32)
    d_bad = DeviceBuffer.of(bad)
    off = np.array([0, 16], np.uint32)
    assert lib.lgx_DirectLightNoAlloc(room["dev"], room["lights"], 1, d_bad.addr, 16,
                                      off.ctypes.data_as(u32p), out.addr, None) == 0
    assert "sample beyond" in _lib.last_error(lib)
    # a host array where a device buffer is expected
    host = np.zeros(16, np.uint32)
    assert lib.lgx_DirectLightNoAlloc(room["dev"], room["lights"], 1, host.ctypes.data, 16,
                                      off.ctypes.data_as(u32p), out.addr, None) == 0
    assert "device buffer" in _lib.last_error(lib)
    # the entry point still works after the refusals
    items = np.ascontiguousarray(room["items"][: room["off"][1]])
    off = np.array([0, len(items)], np.uint32)
    got = _no_alloc(room, 1, items, off)
    exp = _existing(room, 1, items, off)
    for g, e in zip(got, exp):
        assert np.array_equal(_bits(g), _bits(e))

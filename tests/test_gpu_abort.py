"""Abort of the direct-light march from inside the persistent kernel.

The reference's ltr_Abort is a no-op (lighter.cpp:1154-1157); here a bake's
worker raises a flag in pinned host memory, one claim of work in 64 relays
it to a device flag, and every wave reads the device flag at each claim, so
a long march stops within about one claim per wave instead of at the next
stage boundary.  The device-level switch lgx_DeviceAbort raises the device
flag before the launch, which makes the outcome deterministic: every wave leaves at its first claim (no march finishes and
no distance query runs, by the kernel's own counters), the run returns 0
with "direct light aborted", the lightmap keeps its colours, and once the
switch is cleared the same work gives the same weights bit for bit as
before.
(The stage-boundary abort of a whole bake runs on the CPU in test_abi.py.)
"""
import ctypes as C

import numpy as np
import pytest

from helpers import Device, P, all_samples_and_tris, bits, prepared
from lighter_amd import _lib, scenes, workloads

pytestmark = pytest.mark.gpu
u32p = C.POINTER(C.c_uint32)


def test_device_abort_stops_the_march_and_leaves_the_lightmap(lib):
    spec = scenes.room_scene(lm=64, n_boxes=12, n_point=3)
    b = prepared(lib, spec)
    pos, nrm, tris = all_samples_and_tris(lib, b)
    N = len(pos)
    dev = Device(lib, tris)
    lm0 = np.full((N, 3), 0.125, np.float32)
    _lib.check(lib.lgx_SetSamples(dev.h, P(pos), P(nrm), N, P(lm0)), lib)
    lights = workloads.make_lights(spec)
    L = len(lights)
    items = np.concatenate([(np.uint32(l) << np.uint32(26)) | np.arange(N, dtype=np.uint32) for l in range(L)])
    off = np.arange(L + 1, dtype=np.uint32) * N

    def run():
        return lib.lgx_DirectLight(dev.h, lights, L, items.ctypes.data_as(u32p), len(items),
                                   off.ctypes.data_as(u32p), 1)

    def lightmap():
        out = np.zeros((N, 3), np.float32)
        _lib.check(lib.lgx_ReadLightmap(dev.h, P(out)), lib)
        return out

    def weights():
        w = np.zeros(len(items), np.float32)
        _lib.check(lib.lgx_ReadWeights(dev.h, P(w), None, len(items)), lib)
        return w

    assert run() == 1, _lib.last_error(lib)
    w1, lm1 = weights(), lightmap()
    assert (lm1 > 0.125).mean() > 0.2  # the lights reached the room
    _lib.check(lib.lgx_DeviceAbort(dev.h, 1), lib)
    lib.lgx_SetCounting(dev.h, 1)
    cnt = (C.c_uint64 * 6)()
    for _ in range(2):  # the switch stays on until cleared
        assert run() == 0
        assert "direct light aborted" in _lib.last_error(lib)
        assert np.array_equal(bits(lightmap()), bits(lm1))
        lib.lgx_ReadCounters(dev.h, cnt)
        assert cnt[4] == 0 and cnt[3] == 0, list(cnt)  # no march finished, no distance query ran
    lib.lgx_SetCounting(dev.h, 0)
    _lib.check(lib.lgx_DeviceAbort(dev.h, 0), lib)
    assert run() == 1, _lib.last_error(lib)
    assert np.array_equal(bits(weights()), bits(w1))
    lm2 = lightmap()
    assert np.all(lm2 >= lm1) and np.any(lm2 > lm1)  # the second run added its light
    dev.close()
    b.destroy()

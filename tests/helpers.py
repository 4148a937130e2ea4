"""Shared helpers for the tests: scenes prepared through the product's host
stages, and array access through the lgx_* C-ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

from lighter_amd import _lib, scenes

fp = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)


def P(a):
    return a.ctypes.data_as(fp)


def instance_info(lib, scene, i):
    info = (C.c_int32 * 4)()
    _lib.check(lib.lgx_InstanceInfo(scene, i, info), lib)
    return tuple(info)


def prod_samples(lib, scene, i):
    w, h, n, nt = instance_info(lib, scene, i)
    pos = np.zeros((n, 3), np.float32)
    nrm = np.zeros((n, 3), np.float32)
    loc = np.zeros(n, np.uint32)
    rad = np.zeros((n, 4), np.float32)
    lib.lgx_GetSamples(scene, i, P(pos), P(nrm), loc.ctypes.data_as(u32p), P(rad))
    return pos, nrm, loc, rad


def prod_tris(lib, scene, i):
    nt = instance_info(lib, scene, i)[3]
    out = np.zeros((nt, 9), np.float32)
    if nt:
        lib.lgx_GetTriangles(scene, i, P(out))
    return out


def prod_lightmap(lib, scene, i):
    n = instance_info(lib, scene, i)[2]
    out = np.zeros((n, 3), np.float32)
    _lib.check(lib.lgx_GetLightmap(scene, i, P(out)), lib)
    return out


# lgx_SetHostStages bits (include/lighter_gpu.h)
PLACE_BIT, FINALIZE_BIT, RASTER_BIT = 1, 2, 4
GPU_PLACEMENT = 0                        # rasterisation and placement on the GPU
HOST_PLACEMENT = PLACE_BIT | RASTER_BIT  # the host preparation: no GPU stage
HOST_FINALIZE = FINALIZE_BIT
HOST_STAGES = HOST_PLACEMENT | HOST_FINALIZE


def prepared(lib, spec, placement=GPU_PLACEMENT, **config):
    """Build spec through the product library and run the preparation
    stages: rasterisation and sample placement on the GPU (the default) or
    with the host restatements (placement = HOST_PLACEMENT), or any other
    lgx_SetHostStages mask."""
    b = scenes.BuiltScene(lib, spec, config or None)
    _lib.check(lib.lgx_SetHostStages(b.scene, placement), lib)
    _lib.check(lib.lgx_Prepare(b.scene), lib)
    return b


def all_samples_and_tris(lib, b):
    """Concatenated samples (instance order) and the union of shadow tris."""
    n = lib.lgx_InstanceCount(b.scene)
    pos, nrm, tris = [], [], []
    for i in range(n):
        p, q, _, _ = prod_samples(lib, b.scene, i)
        pos.append(p)
        nrm.append(q)
        tris.append(prod_tris(lib, b.scene, i))
    return np.concatenate(pos), np.concatenate(nrm), np.concatenate(tris)


class Device:
    """lgx_Device wrapper."""

    def __init__(self, lib, tris):
        self.lib = lib
        tris = np.ascontiguousarray(tris, np.float32)
        self.h = lib.lgx_DeviceCreate(P(tris), len(tris))
        if not self.h:
            raise RuntimeError(_lib.last_error(lib))

    def close(self):
        if self.h:
            self.lib.lgx_DeviceDestroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def distance(self, p):
        p = np.ascontiguousarray(p, np.float32)
        out = np.zeros(len(p), np.float32)
        _lib.check(self.lib.lgx_QueryDistance(self.h, P(p), len(p), P(out)), self.lib)
        return out

    def shadow(self, a, b, k):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        k = np.ascontiguousarray(k, np.float32)
        out = np.zeros(len(a), np.float32)
        st = np.zeros(len(a), np.uint32)
        _lib.check(self.lib.lgx_QueryShadow(self.h, P(a), P(b), P(k), len(a), P(out), st.ctypes.data_as(u32p)),
                   self.lib)
        return out, st

    def closest(self, a, b):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        out = np.zeros(len(a), np.float32)
        _lib.check(self.lib.lgx_QueryClosest(self.h, P(a), P(b), len(a), P(out)), self.lib)
        return out

    def occluded(self, a, b):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        out = np.zeros(len(a), np.uint8)
        _lib.check(self.lib.lgx_QueryOccluded(self.h, P(a), P(b), len(a), out.ctypes.data_as(C.POINTER(C.c_uint8))),
                   self.lib)
        return out


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)

"""Loader for liblighter_amd.so (built in-tree by ``lighter_amd.build``).

Raises if the library is missing: there is no CPU fallback for the baking
stages.
"""
from __future__ import annotations

import ctypes as C
import os

from . import api

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblighter_amd.so")

fp = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
vp = C.c_void_p


class LgxLight(C.Structure):
    """lgx_Light (include/lighter_gpu.h)."""
    _fields_ = [("type", C.c_int32), ("pos", C.c_float * 3), ("dir", C.c_float * 3), ("color", C.c_float * 3),
                ("range", C.c_float), ("power", C.c_float), ("radius", C.c_float),
                ("spot_angle_out", C.c_float), ("spot_angle_in", C.c_float), ("spot_curve", C.c_float)]


LGX_SIGNATURES = {
    "lgx_LastError": (C.c_char_p, []),
    "lgx_HasDevice": (C.c_int, []),
    "lgx_Prepare": (C.c_int, [vp]),
    "lgx_SetHostStages": (C.c_int, [vp, C.c_int]),
    "lgx_PlacementTimes": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "lgx_FinalizeTimes": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "lgx_RasterTimes": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "lgx_DeviceAbort": (C.c_int, [vp, C.c_int]),
    "lgx_BvhCompare": (C.c_int, [fp, C.c_size_t, C.POINTER(C.c_double)]),
    "lgx_StageDirect": (C.c_int, [vp]),
    "lgx_StageRadiosity": (C.c_int, [vp]),
    "lgx_StageAO": (C.c_int, [vp]),
    "lgx_StageFinalize": (C.c_int, [vp]),
    "lgx_InstanceCount": (C.c_int, [vp]),
    "lgx_InstanceInfo": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int32)]),
    "lgx_GetSamples": (C.c_int, [vp, C.c_int, fp, fp, u32p, fp]),
    "lgx_GetLightmap": (C.c_int, [vp, C.c_int, fp]),
    "lgx_SetLightmap": (C.c_int, [vp, C.c_int, fp]),
    "lgx_GetTriangles": (C.c_int, [vp, C.c_int, fp]),
    "lgx_LightInstances": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int32), C.c_int]),
    "lgx_RadLinkCount": (C.c_size_t, [vp]),
    "lgx_RadGetLinks": (C.c_int, [vp, u32p, u32p, fp]),
    "lgx_StageTimes": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "lgx_DeviceCreate": (vp, [fp, C.c_size_t]),
    "lgx_DeviceDestroy": (None, [vp]),
    "lgx_DeviceInfo": (C.c_int, [vp, C.POINTER(C.c_int64)]),
    "lgx_SetEntryGrid": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int64)]),
    "lgx_QueryDistance": (C.c_int, [vp, fp, C.c_size_t, fp]),
    "lgx_QueryShadow": (C.c_int, [vp, fp, fp, fp, C.c_size_t, fp, u32p]),
    "lgx_QueryClosest": (C.c_int, [vp, fp, fp, C.c_size_t, fp]),
    "lgx_QueryOccluded": (C.c_int, [vp, fp, fp, C.c_size_t, C.POINTER(C.c_uint8)]),
    "lgx_SetSamples": (C.c_int, [vp, fp, fp, C.c_size_t, fp]),
    "lgx_ReadLightmap": (C.c_int, [vp, fp]),
    "lgx_DirectLight": (C.c_int, [vp, C.POINTER(LgxLight), C.c_int, u32p, C.c_size_t, u32p, C.c_int]),
    "lgx_DirectLightSetup": (C.c_int, [vp, C.POINTER(LgxLight), C.c_int, u32p, C.c_size_t, u32p, C.c_int]),
    "lgx_DirectLightRun": (C.c_int, [vp]),
    "lgx_DirectLightNoAlloc": (C.c_int, [vp, C.POINTER(LgxLight), C.c_int, C.c_void_p, C.c_size_t, u32p,
                                        C.c_void_p, C.c_void_p]),
    "lgx_ReadWeights": (C.c_int, [vp, fp, fp, C.c_size_t]),
    "lgx_AmbientOcclusion": (C.c_int, [vp, fp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, fp]),
    "lgx_RadLinks": (C.c_int, [vp, fp, fp, C.c_size_t, C.POINTER(C.c_size_t)]),
    "lgx_RadReadLinks": (C.c_int, [vp, u32p, u32p, fp]),
    "lgx_RadBounce": (C.c_int, [vp, fp, fp, fp, C.c_int, fp]),
    "lgx_RadSolve": (C.c_int, [vp, C.POINTER(C.c_double), C.c_uint32, fp, fp, C.c_int, C.c_int, C.c_int,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "lgx_RadTiles": (C.c_int, [vp, C.POINTER(C.c_uint64)]),
    "lgx_RadSolveGrad": (C.c_int, [vp, C.POINTER(C.c_double), C.c_uint32, fp, fp, C.c_int, C.c_int,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "lgx_DirectLightGrad": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "lgx_SetCounting": (C.c_int, [vp, C.c_int]),
    "lgx_SetLinkSearch": (C.c_int, [vp, C.c_int]),
    "lgx_PrepareTimes": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "lgx_SplitTreeCompare": (C.c_int, [C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_double)]),
    "lgx_ReadCounters": (C.c_int, [vp, C.POINTER(C.c_uint64)]),
    "lgx_ReferenceCounts": (C.c_int, [vp, fp, C.c_size_t, C.POINTER(C.c_uint64)]),
    "lgx_ReferenceCountsAO": (C.c_int, [vp, fp, C.c_size_t, fp, C.c_int, C.c_float, C.c_float, C.c_float,
                                        C.POINTER(C.c_uint64)]),
    "lgx_ReferenceCountsLinks": (C.c_int, [vp, fp, C.c_size_t, C.POINTER(C.c_uint64)]),
    "lgx_DeviceTimes": (C.c_int, [vp, C.POINTER(C.c_double)]),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C lighter_amd/csrc`")
    lib = api.bind_ltr(C.CDLL(path, mode=os.RTLD_LOCAL))
    for name, (res, args) in LGX_SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def last_error(lib=None) -> str:
    lib = lib or load()
    msg = lib.lgx_LastError()
    return msg.decode() if msg else ""


def check(ok, lib=None):
    if not ok:
        raise RuntimeError("liblighter_amd: " + last_error(lib))
    return ok

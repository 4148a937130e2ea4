"""ctypes mirror of lighter's public C API (``include/lighter.h``).

The structures below have the exact x86-64 layout of the reference header
(lighter.h:64-233); ``check_layout`` asserts the sizes and offsets the C side
also ``static_assert``s.  ``bind_ltr`` declares the 17 ``ltr_*`` entry points
(lighter.h:191-258) on any shared library that exports them, so the same
Python code drives ``liblighter_amd.so`` and, in tests, the reference build.
"""
from __future__ import annotations

import ctypes as C

u32 = C.c_uint32
VEC3 = C.c_float * 3
VEC4 = C.c_float * 4
MAT4 = VEC4 * 4

LTR_LT_POINT = 1
LTR_LT_SPOT = 2
LTR_LT_DIRECT = 3


class MeshPartInfo(C.Structure):  # lighter.h:64-79
    _fields_ = [
        ("positions_f3", C.c_void_p),
        ("normals_f3", C.c_void_p),
        ("texcoords1_f2", C.c_void_p),
        ("texcoords2_f2", C.c_void_p),
        ("stride_positions", u32),
        ("stride_normals", u32),
        ("stride_texcoords1", u32),
        ("stride_texcoords2", u32),
        ("indices", C.POINTER(u32)),
        ("vertex_count", u32),
        ("index_count", u32),
        ("shadow", C.c_int),
    ]


class MeshInstanceInfo(C.Structure):  # lighter.h:81-89
    _fields_ = [
        ("matrix", MAT4),
        ("importance", C.c_float),
        ("shadow", C.c_int),
        ("ident", C.c_char_p),
        ("ident_size", C.c_size_t),
    ]


class LightInfo(C.Structure):  # lighter.h:91-106
    _fields_ = [
        ("type", u32),
        ("position", VEC3),
        ("direction", VEC3),
        ("up_direction", VEC3),
        ("color_rgb", VEC3),
        ("range", C.c_float),
        ("power", C.c_float),
        ("light_radius", C.c_float),
        ("shadow_sample_count", C.c_int),
        ("spot_angle_out", C.c_float),
        ("spot_angle_in", C.c_float),
        ("spot_curve", C.c_float),
    ]


class SampleInfo(C.Structure):  # lighter.h:108-115
    _fields_ = [
        ("id", u32),
        ("position", VEC3),
        ("normal", VEC3),
        ("out_color", VEC3),
    ]


class SampleRequest(C.Structure):  # lighter.h:117-133
    _fields_ = [
        ("position", VEC3),
        ("normal", VEC3),
        ("tex0u", C.c_float),
        ("tex0v", C.c_float),
        ("tex1u", C.c_float),
        ("tex1v", C.c_float),
        ("part_id", u32),
        ("mesh_ident", C.c_char_p),
        ("mesh_ident_size", C.c_size_t),
        ("inst_ident", C.c_char_p),
        ("inst_ident_size", C.c_size_t),
        ("out_diffuse_color", VEC3),
        ("out_emissive_color", VEC3),
    ]


class Config(C.Structure):
    pass


SIZE_FN = C.CFUNCTYPE(C.c_int, C.POINTER(Config), C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                      C.c_float, C.c_float, C.POINTER(u32))
SAMPLE_FN = C.CFUNCTYPE(C.c_int, C.POINTER(Config), C.POINTER(SampleRequest))

Config._fields_ = [  # lighter.h:135-189
    ("userdata", C.c_void_p),
    ("size_fn", C.c_void_p),
    ("max_num_threads", C.c_int),
    ("max_tree_memory", C.c_size_t),
    ("max_lightmap_size", u32),
    ("default_width", u32),
    ("default_height", u32),
    ("global_size_factor", C.c_float),
    ("max_correct_dist", C.c_float),
    ("max_correct_angle", C.c_float),
    ("clear_color", VEC3),
    ("ambient_color", VEC3),
    ("bounce_count", C.c_int),
    ("sample_fn", C.c_void_p),
    ("ao_distance", C.c_float),
    ("ao_multiplier", C.c_float),
    ("ao_falloff", C.c_float),
    ("ao_effect", C.c_float),
    ("ao_divergence", C.c_float),
    ("ao_color_rgb", VEC3),
    ("ao_num_samples", C.c_int),
    ("blur_size", C.c_float),
    ("ds2x", C.c_int),
    ("generate_normalmap_data", C.c_int),
]


class WorkOutputInfo(C.Structure):  # lighter.h:203-209
    _fields_ = [
        ("lightmap_count", u32),
        ("sample_count", u32),
        ("samples", C.POINTER(SampleInfo)),
    ]


class WorkOutput(C.Structure):  # lighter.h:211-223
    _fields_ = [
        ("uid", u32),
        ("mesh_ident", C.c_char_p),
        ("mesh_ident_size", C.c_size_t),
        ("inst_ident", C.c_char_p),
        ("inst_ident_size", C.c_size_t),
        ("lightmap_rgb", C.POINTER(C.c_float)),
        ("normals_xyzf", C.POINTER(C.c_float)),
        ("width", u32),
        ("height", u32),
    ]


class WorkStatus(C.Structure):  # lighter.h:228-233
    _fields_ = [("completion", C.c_float), ("stage", C.c_char_p)]


# x86-64 sizes / offsets of the reference ABI.
ABI_SIZES = {
    MeshPartInfo: 72, MeshInstanceInfo: 88, LightInfo: 80, SampleInfo: 40,
    SampleRequest: 104, Config: 144, WorkOutput: 64, WorkOutputInfo: 16, WorkStatus: 16,
}
ABI_OFFSETS = {
    (Config, "size_fn"): 8, (Config, "max_num_threads"): 16, (Config, "bounce_count"): 80,
    (Config, "sample_fn"): 88, (Config, "ao_distance"): 96, (Config, "ao_num_samples"): 128,
    (Config, "blur_size"): 132, (Config, "generate_normalmap_data"): 140,
    (WorkOutput, "lightmap_rgb"): 40, (WorkOutput, "width"): 56,
}


def check_layout() -> None:
    for cls, size in ABI_SIZES.items():
        assert C.sizeof(cls) == size, (cls.__name__, C.sizeof(cls), size)
    for (cls, field), off in ABI_OFFSETS.items():
        assert getattr(cls, field).offset == off, (cls.__name__, field)


LTR_SYMBOLS = (
    "ltr_DefaultSizeFunc", "ltr_CreateScene", "ltr_DestroyScene", "ltr_Start", "ltr_Abort",
    "ltr_GetStatus", "ltr_Sleep", "ltr_GetConfig", "ltr_SetConfig", "ltr_CreateMesh",
    "ltr_MeshAddPart", "ltr_MeshAddInstance", "ltr_LightAdd", "ltr_SampleAdd",
    "ltr_GetWorkOutputInfo", "ltr_GetWorkOutput", "ltr_NextPowerOfTwo",
)


def bind_ltr(lib: C.CDLL) -> C.CDLL:
    """Declare argument/return types of the ltr_* API on ``lib``."""
    vp = C.c_void_p
    sig = {
        "ltr_CreateScene": (vp, []),
        "ltr_DestroyScene": (None, [vp]),
        "ltr_Start": (None, [vp]),
        "ltr_Abort": (None, [vp]),
        "ltr_GetStatus": (C.c_int, [vp, C.POINTER(WorkStatus)]),
        "ltr_Sleep": (None, [C.c_int]),
        "ltr_GetConfig": (None, [C.POINTER(Config), vp]),
        "ltr_SetConfig": (C.c_int, [vp, C.POINTER(Config)]),
        "ltr_CreateMesh": (vp, [vp, C.c_char_p, C.c_size_t]),
        "ltr_MeshAddPart": (C.c_int, [vp, C.POINTER(MeshPartInfo)]),
        "ltr_MeshAddInstance": (C.c_int, [vp, C.POINTER(MeshInstanceInfo)]),
        "ltr_LightAdd": (None, [vp, C.POINTER(LightInfo)]),
        "ltr_SampleAdd": (None, [vp, C.POINTER(SampleInfo)]),
        "ltr_GetWorkOutputInfo": (None, [vp, C.POINTER(WorkOutputInfo)]),
        "ltr_GetWorkOutput": (C.c_int, [vp, u32, C.POINTER(WorkOutput)]),
        "ltr_NextPowerOfTwo": (u32, [u32]),
        "ltr_DefaultSizeFunc": (C.c_int, [C.POINTER(Config), C.c_char_p, C.c_size_t, C.c_char_p,
                                          C.c_size_t, C.c_float, C.c_float, C.POINTER(u32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib

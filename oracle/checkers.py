"""ctypes access to the two CPU checkers — TEST INFRASTRUCTURE ONLY.

* ``Oracle``: the C restatement of the hot path (oracle/liboracle.so).
* ``Reference``: the reference baker compiled from its own sources plus the
  stage-by-stage probe (oracle/_ref/liblighter_ref.so), loaded RTLD_LOCAL so its
  ``ltr_*`` symbols never meet the product's.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "liblighter_ref.so")

fp = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
vp = C.c_void_p


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def _p(a):
    return a.ctypes.data_as(fp)


class OrcLight(C.Structure):
    _fields_ = [("type", C.c_int32), ("pos", C.c_float * 3), ("dir", C.c_float * 3),
                ("color", C.c_float * 3), ("range", C.c_float), ("power", C.c_float),
                ("radius", C.c_float), ("spot_out_deg", C.c_float), ("spot_in_deg", C.c_float),
                ("spot_curve", C.c_float)]


def orc_light_from_info(li: dict, normalise=True) -> OrcLight:
    """Light as the reference stores it after ltr_LightAdd (lighter.cpp:1282-1298)."""
    L = OrcLight()
    L.type = li["type"]
    L.pos[:] = li.get("position", [0, 0, 0])
    d = np.asarray(li.get("direction", [0, 0, 0]), np.float32)
    if normalise:
        d = normalize_f32(d)
    L.dir[:] = [float(x) for x in d]
    L.color[:] = li.get("color_rgb", [0, 0, 0])
    L.range = li.get("range", 0.0)
    L.power = li.get("power", 0.0)
    L.radius = li.get("light_radius", 0.0)
    L.spot_out_deg = li.get("spot_angle_out", 0.0)
    L.spot_in_deg = li.get("spot_angle_in", 0.0)
    L.spot_curve = li.get("spot_curve", 0.0)
    return L


def normalize_f32(v):
    """Vec3::Normalized in float32 (lighter_int.hpp:404-415)."""
    v = np.asarray(v, np.float32)
    l2 = np.float32(v[0] * v[0]) + np.float32(v[1] * v[1])
    l2 = np.float32(l2 + np.float32(v[2] * v[2]))
    if l2 == 0:
        return np.zeros(3, np.float32)
    inv = np.float32(1.0) / np.sqrt(l2, dtype=np.float32)
    return (v * inv).astype(np.float32)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        L.orc_seg_tri.restype = C.c_float
        L.orc_seg_tri.argtypes = [fp] * 5
        L.orc_point_tri_dist.restype = C.c_float
        L.orc_point_tri_dist.argtypes = [fp] * 4
        L.orc_tri_useful.restype = C.c_int
        L.orc_tri_useful.argtypes = [fp]
        L.orc_spiral.restype = None
        L.orc_spiral.argtypes = [fp, C.c_float, C.c_int, C.c_int, fp]
        L.orc_scene_distance.restype = C.c_float
        L.orc_scene_distance.argtypes = [fp, C.c_size_t, fp]
        L.orc_shadow_factor.restype = C.c_float
        L.orc_shadow_factor.argtypes = [fp, C.c_size_t, fp, fp, C.c_float, u32p]
        L.orc_closest_hit.restype = C.c_float
        L.orc_closest_hit.argtypes = [fp, C.c_size_t, fp, fp]
        L.orc_occluded.restype = C.c_int
        L.orc_occluded.argtypes = [fp, C.c_size_t, fp, fp]
        L.orc_direct_light.restype = C.c_size_t
        L.orc_direct_light.argtypes = [fp, C.c_size_t, C.c_size_t, fp, fp, C.POINTER(OrcLight), fp, fp]
        L.orc_ao.restype = None
        L.orc_ao.argtypes = [fp, C.c_size_t, C.c_size_t, fp, fp, fp, C.c_float, C.c_float, C.c_float,
                             C.c_float, fp, C.c_int, fp, fp]
        L.orc_rad_links.restype = C.c_size_t
        L.orc_rad_links.argtypes = [fp, C.c_size_t, C.c_size_t, fp, fp, u32p, u32p, fp, C.c_size_t]
        L.orc_rad_bounce_f32.restype = None
        L.orc_rad_bounce_f32.argtypes = [C.c_size_t, u32p, u32p, fp, fp]
        dp = C.POINTER(C.c_double)
        L.orc_rad_series_f64.restype = None
        L.orc_rad_series_f64.argtypes = [C.c_size_t, u32p, u32p, fp, dp, dp, dp, C.c_int, C.c_int, dp]

    # -- primitives
    def seg_tri(self, l1, l2, p1, p2, p3):
        a = [_f(x) for x in (l1, l2, p1, p2, p3)]
        return self.lib.orc_seg_tri(*[_p(x) for x in a])

    def point_tri_dist(self, p, t0, t1, t2):
        a = [_f(x) for x in (p, t0, t1, t2)]
        return self.lib.orc_point_tri_dist(*[_p(x) for x in a])

    def spiral(self, d, randoff, i, n):
        d = _f(d)
        out = np.zeros(3, np.float32)
        self.lib.orc_spiral(_p(d), randoff, i, n, _p(out))
        return out

    # -- queries over a triangle soup (T,9)
    def scene_distance(self, tris, p):
        tris, p = _f(tris), _f(p)
        return self.lib.orc_scene_distance(_p(tris), len(tris), _p(p))

    def shadow_factor(self, tris, a, b, k):
        tris, a, b = _f(tris), _f(a), _f(b)
        steps = C.c_uint32(0)
        v = self.lib.orc_shadow_factor(_p(tris), len(tris), _p(a), _p(b), k, C.byref(steps))
        return v, steps.value

    def closest_hit(self, tris, a, b):
        tris, a, b = _f(tris), _f(a), _f(b)
        return self.lib.orc_closest_hit(_p(tris), len(tris), _p(a), _p(b))

    def occluded(self, tris, a, b):
        tris, a, b = _f(tris), _f(a), _f(b)
        return self.lib.orc_occluded(_p(tris), len(tris), _p(a), _p(b))

    # -- stages
    def direct_light(self, tris, pos, nrm, light: OrcLight, lightmap):
        tris, pos, nrm = _f(tris), _f(pos), _f(nrm)
        lm = _f(lightmap).copy()
        w = np.zeros(len(pos), np.float32)
        n = self.lib.orc_direct_light(_p(tris), len(tris), len(pos), _p(pos), _p(nrm), C.byref(light), _p(lm), _p(w))
        return lm, w, n

    def ao(self, tris, pos, nrm, randoff, lightmap, distance, multiplier=1.2, falloff=1.0, effect=0.0,
           color=(0, 0, 0), n=17):
        tris, pos, nrm, randoff = _f(tris), _f(pos), _f(nrm), _f(randoff)
        lm = _f(lightmap).copy()
        col = _f(color)
        fac = np.zeros(len(pos), np.float32)
        self.lib.orc_ao(_p(tris), len(tris), len(pos), _p(pos), _p(nrm), _p(randoff), distance, multiplier,
                        falloff, effect, _p(col), n, _p(lm), _p(fac))
        return lm, fac

    def rad_links(self, tris, pos, nrm, cap=None):
        tris, pos, nrm = _f(tris), _f(pos), _f(nrm)
        N = len(pos)
        rowmap = np.zeros(2 * N, np.uint32)
        if cap is None:
            cap = max(1024, N * 64)
        other = np.zeros(cap, np.uint32)
        fac = np.zeros(cap, np.float32)
        n = self.lib.orc_rad_links(_p(tris), len(tris), N, _p(pos), _p(nrm), rowmap.ctypes.data_as(u32p),
                                   other.ctypes.data_as(u32p), _p(fac), cap)
        if n > cap:
            return self.rad_links(tris, pos, nrm, cap=n)
        return rowmap, other[:n].copy(), fac[:n].copy()

    def rad_bounce_f32(self, rowmap, other, factor, colors13):
        c = _f(colors13).copy()
        rowmap = np.ascontiguousarray(rowmap, np.uint32)
        other = np.ascontiguousarray(other, np.uint32)
        factor = _f(factor)
        self.lib.orc_rad_bounce_f32(len(c), rowmap.ctypes.data_as(u32p), other.ctypes.data_as(u32p), _p(factor),
                                    _p(c))
        return c

    def rad_series_f64(self, rowmap, other, factor, area, diffuse3, E, bounces):
        dp = C.POINTER(C.c_double)
        rowmap = np.ascontiguousarray(rowmap, np.uint32)
        other = np.ascontiguousarray(other, np.uint32)
        factor = _f(factor)
        area = np.ascontiguousarray(area, np.float64)
        diffuse3 = np.ascontiguousarray(diffuse3, np.float64)
        E = np.ascontiguousarray(E, np.float64)
        N, R = E.shape
        out = np.zeros_like(E)
        self.lib.orc_rad_series_f64(N, rowmap.ctypes.data_as(u32p), other.ctypes.data_as(u32p), _p(factor),
                                    area.ctypes.data_as(dp), diffuse3.ctypes.data_as(dp), E.ctypes.data_as(dp),
                                    R, bounces, out.ctypes.data_as(dp))
        return out


def reference_available(path=REF_SO) -> bool:
    return os.path.exists(path)


class Reference:
    """The reference baker built from source, with the stage probe."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref` where the reference checkout exists")
        import sys
        sys.path.insert(0, os.path.dirname(HERE))
        from lighter_amd import api
        self.lib = api.bind_ltr(C.CDLL(path, mode=os.RTLD_LOCAL))
        L = self.lib
        for name, res, args in [
            ("rp_prepare", C.c_int, [vp, C.c_int]),
            ("rp_prepare_collision", C.c_int, [vp, C.c_int]),
            ("rp_shadow_factor_batch_mt", None, [vp, C.c_size_t, fp, fp, fp, fp, C.c_int]),
            ("rp_ao_factor_batch_mt", None, [vp, C.c_size_t, fp, fp, fp, C.c_int, C.c_float, C.c_float, fp,
                                             C.c_int]),
            ("rp_light_instances", C.c_int, [vp, C.c_int, C.POINTER(C.c_int32), C.c_int]),
            ("rp_direct", None, [vp, C.c_int, C.c_int]),
            ("rp_rad_links", None, [vp]),
            ("rp_rad_bounce", None, [vp, C.c_int]),
            ("rp_rad_commit", None, [vp]),
            ("rp_ao", None, [vp]),
            ("rp_finalize", None, [vp]),
            ("rp_num_instances", C.c_int, [vp]),
            ("rp_instance_info", None, [vp, C.c_int, C.POINTER(C.c_int32)]),
            ("rp_get_samples", None, [vp, C.c_int, fp, fp, u32p, fp]),
            ("rp_get_tris", None, [vp, C.c_int, fp]),
            ("rp_instance_shadow", C.c_int, [vp, C.c_int]),
            ("rp_get_lightmap", None, [vp, C.c_int, fp]),
            ("rp_set_lightmap", None, [vp, C.c_int, fp]),
            ("rp_num_links", C.c_size_t, [vp]),
            ("rp_num_patches", C.c_size_t, [vp]),
            ("rp_get_links", None, [vp, u32p, u32p, fp]),
            ("rp_get_rad_colors", None, [vp, fp]),
            ("rp_get_rad_geoms", None, [vp, fp]),
            ("rp_distance", C.c_float, [vp, fp]),
            ("rp_shadow_factor", C.c_float, [vp, fp, fp, C.c_float]),
            ("rp_closest_hit", C.c_float, [vp, fp, fp]),
            ("rp_visibility", C.c_int, [vp, fp, fp]),
            ("rp_shadow_factor_batch", None, [vp, C.c_size_t, fp, fp, C.c_float, fp]),
            ("rp_closest_hit_batch", None, [vp, C.c_size_t, fp, fp, fp]),
            ("rp_visibility_batch", None, [vp, C.c_size_t, fp, fp, C.POINTER(C.c_uint8)]),
            ("rp_seg_tri", C.c_float, [fp] * 5),
            ("rp_spiral", None, [fp, C.c_float, C.c_int, C.c_int, fp]),
            ("rp_point_tri_dist_capped", C.c_float, [fp] * 4),
            ("rp_tri_useful", C.c_int, [fp]),
        ]:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args

    # -- per-instance state
    def instances(self, scene):
        out = []
        for i in range(self.lib.rp_num_instances(scene)):
            info = (C.c_int32 * 4)()
            self.lib.rp_instance_info(scene, i, info)
            out.append(tuple(info))
        return out

    def samples(self, scene, i):
        w, h, n, _ = self.instances(scene)[i]
        pos = np.zeros((n, 3), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        loc = np.zeros(n, np.uint32)
        rad = np.zeros((n, 4), np.float32)
        self.lib.rp_get_samples(scene, i, _p(pos), _p(nrm), loc.ctypes.data_as(u32p) if i else None,
                                _p(rad) if i else None)
        return pos, nrm, loc, rad

    def tris(self, scene, i):
        nt = self.instances(scene)[i][3]
        out = np.zeros((nt, 9), np.float32)
        if nt:
            self.lib.rp_get_tris(scene, i, _p(out))
        return out

    def shadow_tris(self, scene):
        """Union of the triangles the scene queries test (lighter.cpp:159,123,265)."""
        parts = [self.tris(scene, i) for i in range(self.lib.rp_num_instances(scene))
                 if self.lib.rp_instance_shadow(scene, i)]
        parts = [p for p in parts if len(p)]
        return np.concatenate(parts) if parts else np.zeros((0, 9), np.float32)

    def lightmap(self, scene, i):
        n = self.instances(scene)[i][2]
        out = np.zeros((n, 3), np.float32)
        self.lib.rp_get_lightmap(scene, i, _p(out))
        return out

    def set_lightmap(self, scene, i, rgb):
        rgb = _f(rgb)
        self.lib.rp_set_lightmap(scene, i, _p(rgb))

    def links(self, scene):
        N = self.lib.rp_num_patches(scene)
        nl = self.lib.rp_num_links(scene)
        rowmap = np.zeros(2 * N, np.uint32)
        other = np.zeros(nl, np.uint32)
        fac = np.zeros(nl, np.float32)
        self.lib.rp_get_links(scene, rowmap.ctypes.data_as(u32p), other.ctypes.data_as(u32p), _p(fac))
        return rowmap, other, fac

    def rad_colors(self, scene):
        N = self.lib.rp_num_patches(scene)
        out = np.zeros((N, 13), np.float32)
        self.lib.rp_get_rad_colors(scene, _p(out))
        return out

    def rad_geoms(self, scene):
        N = self.lib.rp_num_patches(scene)
        out = np.zeros((N, 6), np.float32)
        self.lib.rp_get_rad_geoms(scene, _p(out))
        return out

    def light_instances(self, scene, light):
        buf = (C.c_int32 * 65536)()
        n = self.lib.rp_light_instances(scene, light, buf, 65536)
        return list(buf[:n])

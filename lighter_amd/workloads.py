"""Direct-light workloads over the lgx_* C-ABI: the host stages of a scene
through the product, its samples and shadow-casting triangles, and the work
items of every light (lighter.cpp:1075-1096: each light renders the sample
container and the instances the instance tree reports in its range, in
instance order).  Used by the benchmark and the GPU tests."""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import _lib, scenes

fp = C.POINTER(C.c_float)


def _P(a):
    return a.ctypes.data_as(fp)


def host_cores():
    """Host threads for the CPU legs and the reference baseline: every CPU
    this process may run on (sched_getaffinity), limited by the cgroup CPU
    quota and by OMP_NUM_THREADS when those are set (a GPU box allots each
    GPU a fixed CPU share and states it there).  Returns (threads, info)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(-(-int(q) // int(per))))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    omp = int(omp) if omp and omp.isdigit() and int(omp) > 0 else None
    n = min([aff] + [x for x in (quota, omp) if x])
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return max(1, n), dict(threads=max(1, n), affinity_cpus=aff, cgroup_quota_cpus=quota, omp_num_threads=omp,
                           os_cpu_count=os.cpu_count(), cpu_model=model)


def direct_workload(lib, spec, nthreads, radinfo=None):
    """Build spec through the product library and run its preparation;
    returns (built scene, positions, normals, shadow triangles, work items
    (light << 26) | sample grouped by light, per-light item offsets).  The
    built scene carries prepare_s.  radinfo (a list) receives the per-texel
    radiosity info (uv, part id, world area) of each instance when given."""
    b = scenes.BuiltScene(lib, spec, dict(max_num_threads=nthreads))
    t0 = time.perf_counter()
    _lib.check(lib.lgx_Prepare(b.scene), lib)
    b.prepare_s = time.perf_counter() - t0
    n_inst = lib.lgx_InstanceCount(b.scene)
    pos, nrm, tris, counts = [], [], [], []
    for i in range(n_inst):
        info = (C.c_int32 * 4)()
        lib.lgx_InstanceInfo(b.scene, i, info)
        n, nt = info[2], info[3]
        p = np.zeros((n, 3), np.float32)
        q = np.zeros((n, 3), np.float32)
        r = np.zeros((n, 4), np.float32) if radinfo is not None else None
        lib.lgx_GetSamples(b.scene, i, _P(p), _P(q), None, _P(r) if r is not None else None)
        if radinfo is not None:
            radinfo.append(r)
        t = np.zeros((nt, 9), np.float32)
        if nt:
            lib.lgx_GetTriangles(b.scene, i, _P(t))
        pos.append(p)
        nrm.append(q)
        tris.append(t)
        counts.append(n)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    items, light_off = [], [0]
    buf = (C.c_int32 * (n_inst + 1))()
    for l in range(len(spec.lights)):
        k = lib.lgx_LightInstances(b.scene, l, buf, n_inst + 1)
        ids = list(buf[:k])
        s = np.concatenate([np.arange(off[i], off[i + 1], dtype=np.uint32) for i in ids])
        items.append((np.uint32(l) << np.uint32(26)) | s)
        light_off.append(light_off[-1] + len(s))
    return (b, np.concatenate(pos), np.concatenate(nrm), np.concatenate(tris),
            np.concatenate(items).astype(np.uint32), np.array(light_off, np.uint32))


def make_lights(spec, radius_scale=None, colors=None):
    """lgx_Light array of spec's lights, directions normalised as
    ltr_LightAdd does (lighter.cpp:1288).  radius_scale multiplies every
    light's radius (the penumbra constant k); colors (L x 3) replaces the
    colours."""
    arr = (_lib.LgxLight * len(spec.lights))()
    for i, L in enumerate(spec.lights):
        d = np.asarray(L.get("direction", [0, 0, 0]), np.float32)
        l2 = np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])
        l2 = np.float32(l2 + np.float32(d[2] * d[2]))
        if l2 > 0:
            d = (d * (np.float32(1) / np.sqrt(l2, dtype=np.float32))).astype(np.float32)
        a = arr[i]
        a.type = L["type"]
        a.pos[:] = L.get("position", [0, 0, 0])
        a.dir[:] = [float(x) for x in d]
        a.color[:] = L.get("color_rgb", [0, 0, 0]) if colors is None else [float(x) for x in colors[i]]
        a.range = L.get("range", 0)
        a.power = L.get("power", 0)
        a.radius = L.get("light_radius", 0) * (radius_scale if radius_scale is not None else 1)
        a.spot_angle_out = L.get("spot_angle_out", 0)
        a.spot_angle_in = L.get("spot_angle_in", 0)
        a.spot_curve = L.get("spot_curve", 0)
    return arr



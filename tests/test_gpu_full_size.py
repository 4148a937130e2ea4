"""Shadow marches at the full size of the terrain direct-light workload
(512 x 512 heightfield + 2,000 boxes = 548,288 triangles, 2048^2 atlas).

CalcInvShadowFactor (lighter.cpp:190-207) at this size: the brute-force
oracle (every triangle at every step) still finishes a few dozen marches in
seconds, so a seeded sample is compared bit for bit; on 20,000 marches from
real texels to their lights the checks are ones that hold at any size:
  * the march's steps (t += min(h, 1)) do not depend on the penumbra factor
    k, so the step counts agree across k;
  * each step divides the same h by min(t k, 2), so the factor is
    non-increasing in k, bit for bit;
  * k = 0 gives a hard shadow: every factor is 0 or 1.
"""
import numpy as np
import pytest

from helpers import Device, all_samples_and_tris, bits, prepared
from lighter_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def terrain(lib):
    spec = scenes.terrain_scene()
    b = prepared(lib, spec)
    pos, nrm, tris = all_samples_and_tris(lib, b)
    assert len(tris) == 548288
    dev = Device(lib, tris)
    yield spec, pos, nrm, tris, dev
    dev.close()
    b.destroy()


def marches(spec, pos, nrm, n, seed):
    """n (from, to) pairs: random texels and, for each, a light that faces it
    (point/spot lights within range, or the sun along its direction)."""
    rng = np.random.default_rng(seed)
    frm, to = [], []
    lights = spec.lights
    while len(frm) < n:
        i = int(rng.integers(len(pos)))
        li = lights[int(rng.integers(len(lights)))]
        p = pos[i].astype(np.float64)
        if "direction" in li and "position" not in li:
            # a directional light's vector points toward the light
            # (lighter.cpp:586-589)
            d = np.asarray(li["direction"], np.float64)
            target = p + d / np.linalg.norm(d) * float(li["range"])
        else:
            target = np.asarray(li["position"], np.float64)
            if np.linalg.norm(target - p) > float(li["range"]):
                continue
        if np.dot(target - p, nrm[i]) <= 0:
            continue
        frm.append(pos[i] + nrm[i] * np.float32(0.005))
        to.append(target)
    return np.asarray(frm, np.float32), np.asarray(to, np.float32)


def test_full_size_marches_bitexact(orc, terrain):
    spec, pos, nrm, tris, dev = terrain
    frm, to = marches(spec, pos, nrm, 16, seed=11)
    k = np.full(len(frm), 0.5, np.float32)  # the workload's light radius
    got, steps = dev.shadow(frm, to, k)
    exp = np.array([orc.shadow_factor(tris, a, b, kk)[0] for a, b, kk in zip(frm, to, k)], np.float32)
    assert np.array_equal(bits(got), bits(exp)), (got, exp)


def test_full_size_march_properties(terrain):
    spec, pos, nrm, tris, dev = terrain
    frm, to = marches(spec, pos, nrm, 20000, seed=12)
    res, st = {}, {}
    for kk in (0.0, 0.25, 0.5, 2.0):
        res[kk], st[kk] = dev.shadow(frm, to, np.full(len(frm), kk, np.float32))
    for kk in (0.25, 0.5, 2.0):
        assert np.array_equal(st[kk], st[0.0]), kk
    assert np.all((res[0.0] == 0.0) | (res[0.0] == 1.0))
    assert np.all(res[0.25] <= res[0.0]) and np.all(res[0.5] <= res[0.25]) and np.all(res[2.0] <= res[0.5])
    assert np.all((res[2.0] >= 0.0) & (res[2.0] <= 1.0))
    # a real mix of lit, shadowed and penumbra marches
    r = res[0.5]
    assert (r == 0).mean() > 0.05 and (r == 1).mean() > 0.05 and ((r > 0) & (r < 1)).mean() > 0.05


def test_every_march_equals_the_reference_semantics(lib, terrain):
    """All 3.8 M marches of the workload's direct light, marched again on
    the GPU with the reference's query semantics (a radius-2 distance query
    per step on a plain binary SAH tree: no scene-box exit, no Lipschitz or
    need bound, a different tree): every shadow factor of the product's
    march is bit-identical, which pins the exactness of its shortcuts at
    full size.  The same pass gives the reference-shaped traversal counts
    the roofline figure is stated on."""
    import ctypes as C

    from helpers import P
    from lighter_amd import _lib
    spec, pos, nrm, tris, dev = terrain
    b = prepared(lib, spec)
    n_inst = lib.lgx_InstanceCount(b.scene)
    counts = []
    for i in range(n_inst):
        info = (C.c_int32 * 4)()
        lib.lgx_InstanceInfo(b.scene, i, info)
        counts.append(info[2])
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    items, light_off = [], [0]
    buf = (C.c_int32 * (n_inst + 1))()
    for li in range(len(spec.lights)):
        k = lib.lgx_LightInstances(b.scene, li, buf, n_inst + 1)
        s = np.concatenate([np.arange(off[i], off[i + 1], dtype=np.uint32) for i in buf[:k]])
        items.append((np.uint32(li) << np.uint32(26)) | s)
        light_off.append(light_off[-1] + len(s))
    b.destroy()
    items = np.concatenate(items).astype(np.uint32)
    light_off = np.array(light_off, np.uint32)
    arr = (_lib.LgxLight * len(spec.lights))()
    for i, L in enumerate(spec.lights):
        d = np.asarray(L.get("direction", [0, 0, 0]), np.float32)
        l2 = np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])) + np.float32(d[2] * d[2])
        if l2 > 0:
            d = (d * (np.float32(1) / np.sqrt(np.float32(l2)))).astype(np.float32)
        a = arr[i]
        a.type = L["type"]
        a.pos[:] = L.get("position", [0, 0, 0])
        a.dir[:] = [float(x) for x in d]
        a.color[:] = L.get("color_rgb", [0, 0, 0])
        a.range = L.get("range", 0)
        a.power = L.get("power", 0)
        a.radius = L.get("light_radius", 0)
        a.spot_angle_out = L.get("spot_angle_out", 0)
        a.spot_angle_in = L.get("spot_angle_in", 0)
        a.spot_curve = L.get("spot_curve", 0)
    u32p = C.POINTER(C.c_uint32)
    _lib.check(lib.lgx_SetSamples(dev.h, P(pos), P(nrm), len(pos), None), lib)
    _lib.check(lib.lgx_DirectLightSetup(dev.h, arr, len(arr), items.ctypes.data_as(u32p), len(items),
                                        light_off.ctypes.data_as(u32p), 1), lib)
    _lib.check(lib.lgx_DirectLightRun(dev.h), lib)
    out = (C.c_uint64 * 5)()
    _lib.check(lib.lgx_ReferenceCounts(dev.h, P(tris), len(tris), out), lib)
    visits, tests, queries, differ, marches = (int(x) for x in out)
    print(f"{marches} marches, {queries} queries, {visits} node visits, {tests} triangle tests, {differ} differ")
    assert marches > 3_000_000
    assert differ == 0
    assert queries > marches and visits > queries and tests > 0


def test_every_ao_factor_equals_the_canonical_traversal(lib, terrain):
    """AO at full size (64 rays of 2 m for each of the 2,989,242 texels):
    every texel's factor from the product's traversal (4-wide tree, nearest
    first) equals, bit for bit, the factor with the rays traced with the
    canonical closest-hit semantics on a plain binary SAH tree; the closest
    hit is order independent, so any difference would be a traversal bug."""
    import ctypes as C

    from helpers import P
    from lighter_amd import _lib
    spec, pos, nrm, tris, dev = terrain
    _lib.check(lib.lgx_SetSamples(dev.h, P(pos), P(nrm), len(pos), None), lib)
    ro = np.random.default_rng(13).uniform(0, 1, len(pos)).astype(np.float32)
    out = (C.c_uint64 * 5)()
    _lib.check(lib.lgx_ReferenceCountsAO(dev.h, P(tris), len(tris), P(ro), 64, 2.0, 1.2, 1.0, out), lib)
    visits, tests, rays, differ, texels = (int(x) for x in out)
    print(f"{texels} texels, {rays} rays, {visits} node visits, {tests} triangle tests, {differ} differ")
    assert texels == len(pos) and rays == 64 * len(pos)
    assert differ == 0
    assert visits > rays and tests > 0

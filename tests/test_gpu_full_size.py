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

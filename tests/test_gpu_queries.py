"""GPU scene queries against the C oracle (brute force over all triangles).

The queries are the building blocks of every hot-path stage:
distance (lighter.cpp:150-188), shadow march (190-207), closest segment hit
(254-289) and occlusion (112-147).  They use only + - * / sqrt, so the
gfx950 results must equal the oracle's bit for bit.
"""
import numpy as np
import pytest

from helpers import Device, all_samples_and_tris, bits, prepared
from lighter_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def room(lib):
    spec = scenes.room_scene(lm=48)
    b = prepared(lib, spec)
    pos, nrm, tris = all_samples_and_tris(lib, b)
    dev = Device(lib, tris)
    yield spec, pos, nrm, tris, dev
    dev.close()
    b.destroy()


def test_distance_bitexact(orc, room):
    _, pos, nrm, tris, dev = room
    rng = np.random.default_rng(1)
    sel = slice(None, None, 5)
    near = pos[sel] + nrm[sel] * rng.uniform(0.0, 0.5, (len(pos[sel]), 1)).astype(np.float32)
    far = rng.uniform([-3, -3, -3], [13, 13, 7], (1500, 3))
    pts = np.concatenate([near, far]).astype(np.float32)
    got = dev.distance(pts)
    exp = np.array([orc.scene_distance(tris, p) for p in pts], np.float32)
    assert np.array_equal(bits(got), bits(exp))
    assert (got < 2.0).mean() > 0.3  # the test exercises real triangle distances


def test_shadow_march_bitexact(orc, room):
    spec, pos, nrm, tris, dev = room
    rng = np.random.default_rng(2)
    idx = rng.choice(len(pos), 300, replace=False)
    frm = (pos[idx] + nrm[idx] * np.float32(0.005)).astype(np.float32)
    lp = np.asarray(spec.lights[0]["position"], np.float32)
    to = np.broadcast_to(lp, frm.shape).copy()
    k = rng.choice([0.0, 0.1, 0.5, 2.0], len(idx)).astype(np.float32)
    got, steps = dev.shadow(frm, to, k)
    exp = []
    exp_steps = []
    for a, b_, kk in zip(frm, to, k):
        v, s = orc.shadow_factor(tris, a, b_, kk)
        exp.append(v)
        exp_steps.append(s)
    exp = np.array(exp, np.float32)
    assert np.array_equal(bits(got), bits(exp))
    # steps may end earlier on the GPU once the march has left the scene box
    assert np.all(steps <= np.array(exp_steps))
    assert 0 < (got == 0).mean() < 1


def test_segments_bitexact(orc, room):
    _, pos, nrm, tris, dev = room
    rng = np.random.default_rng(3)
    idx = rng.choice(len(pos), 400, replace=False)
    a = (pos[idx] + nrm[idx] * np.float32(0.002)).astype(np.float32)
    d = rng.normal(size=(400, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    b_ = (a + d * np.float32(2.0)).astype(np.float32)
    # plus long segments between sample pairs (radiosity visibility)
    j = rng.choice(len(pos), 400)
    a2 = np.concatenate([a, pos[idx]]).astype(np.float32)
    b2 = np.concatenate([b_, pos[j]]).astype(np.float32)
    got = dev.closest(a2, b2)
    exp = np.array([orc.closest_hit(tris, x, y) for x, y in zip(a2, b2)], np.float32)
    assert np.array_equal(bits(got), bits(exp))
    occ = dev.occluded(a2, b2)
    exp_occ = np.array([orc.occluded(tris, x, y) for x, y in zip(a2, b2)], np.uint8)
    assert np.array_equal(occ, exp_occ)
    assert 0.05 < occ.mean() < 0.95


def test_empty_and_single_triangle(lib, orc):
    rng = np.random.default_rng(4)
    pts = rng.uniform(-2, 2, (64, 3)).astype(np.float32)
    empty = Device(lib, np.zeros((0, 9), np.float32))
    assert np.all(empty.distance(pts) == 2.0)
    assert np.all(empty.closest(pts, -pts) == 2.0)
    assert not empty.occluded(pts, -pts).any()
    v, st = empty.shadow(pts, -pts, np.full(64, 0.1, np.float32))
    assert np.all(v == 1.0)
    empty.close()
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    one = Device(lib, tri)
    exp = np.array([orc.scene_distance(tri, p) for p in pts], np.float32)
    assert np.array_equal(bits(one.distance(pts)), bits(exp))
    a = np.array([[0.2, 0.2, 1.0], [0.2, 0.2, -1.0], [2, 2, 1]], np.float32)
    b_ = np.array([[0.2, 0.2, -1.0], [0.3, 0.3, 1.0], [2, 2, -1]], np.float32)
    exp = np.array([orc.closest_hit(tri, x, y) for x, y in zip(a, b_)], np.float32)
    assert np.array_equal(bits(one.closest(a, b_)), bits(exp))
    assert list(one.occluded(a, b_)) == [1, 1, 0]
    one.close()

"""Radiosity link search: the uniform grid against the all-pairs enumeration.

Int_RDGenLinks (lighter.cpp:728-759) tests every pair i < j.  A pair can
only pass the factor threshold f = dA dB / (pi l^4) >= 0.001 when l^2 <=
|nA| |nB| / (0.001 pi), so the product searches a uniform grid of cells that
wide (every such pair lies in the same or a neighbouring cell) with the
per-pair arithmetic of the all-pairs loop, and assembles the symmetric CSR
by a per-row sort.  lgx_SetLinkSearch(dev, 1) switches to the all-pairs
enumeration, mode 2 forces the grid, and the default takes all pairs where
the grid would not prune.  Every mode must give the same links and factors
bit for bit, also above 262,144 patches (where a pair bitmask would take
N^2 / 8 = 8.6 GB); patches with non-finite positions or normals take the
all-pairs path, whose links must then match the oracle's loop.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import Device, P, all_samples_and_tris, bits, prepared, prod_samples
from lighter_amd import _lib, scenes

pytestmark = pytest.mark.gpu
u32p = C.POINTER(C.c_uint32)


def links(lib, dev, pos, nrm, mode):
    """mode 0: automatic, 1: all pairs, 2: the grid"""
    _lib.check(lib.lgx_SetLinkSearch(dev.h, mode), lib)
    nl = C.c_size_t(0)
    _lib.check(lib.lgx_RadLinks(dev.h, P(pos), P(nrm), len(pos), C.byref(nl)), lib)
    rowmap = np.zeros(2 * len(pos), np.uint32)
    other = np.zeros(nl.value, np.uint32)
    fac = np.zeros(nl.value, np.float32)
    _lib.check(lib.lgx_RadReadLinks(dev.h, rowmap.ctypes.data_as(u32p), other.ctypes.data_as(u32p), P(fac)), lib)
    return rowmap, other, fac


SCENES = {
    "room_96": lambda: scenes.room_scene(lm=96, n_boxes=20, bounce_count=1),
    "city_block": lambda: scenes.city_scene(extent=200.0, grid=32, amplitude=2.0, n_boxes=144, tiles=4,
                                            lm_total=160, n_lights=2),
    # 316,375 patches: above 262,144
    "city_block_316k": lambda: scenes.city_scene(extent=200.0, grid=32, amplitude=2.0, n_boxes=144, tiles=4,
                                                 lm_total=640, n_lights=2),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_grid_search_finds_the_all_pairs_links(lib, name):
    b = prepared(lib, SCENES[name]())
    pos, nrm, tris = all_samples_and_tris(lib, b)
    pn = scenes.normalize_rows_f32(nrm)
    if name.endswith("316k"):
        assert len(pos) > 262144
    dev = Device(lib, tris)
    try:
        rm_g, ot_g, fa_g = links(lib, dev, pos, pn, 2)
        rm_a, ot_a, fa_a = links(lib, dev, pos, pn, 1)
        rm_d, ot_d, fa_d = links(lib, dev, pos, pn, 0)
        assert np.array_equal(rm_d, rm_a) and np.array_equal(ot_d, ot_a)
        print(f"{name}: {len(pos)} patches, {len(ot_g)} links")
        assert len(ot_g) > 1000
        assert np.array_equal(rm_g, rm_a)
        assert np.array_equal(ot_g, ot_a)
        assert np.array_equal(bits(fa_g), bits(fa_a))
        # rows list their links by ascending patch index, every factor passes
        for i in np.flatnonzero(rm_g[1::2] > 1)[:2000]:
            o, c = rm_g[2 * i], rm_g[2 * i + 1]
            assert np.all(np.diff(ot_g[o:o + c].astype(np.int64)) > 0)
        assert np.all(fa_g >= np.float32(0.001))
    finally:
        dev.close()
        b.destroy()


def test_non_finite_normals_take_the_all_pairs_loop(lib, orc):
    """A NaN normal and an infinite normal component: no cell width bounds
    such pairs, so the search enumerates all pairs, and their tests compare
    as the reference's do (NaN fails neither `<=` nor `<`, so the pair is
    tested for occlusion and links with a NaN factor; +inf links with an
    infinite one).  The links must equal the oracle's loop."""
    spec = scenes.room_scene(lm=32, n_boxes=4, bounce_count=1)
    b = prepared(lib, spec)
    pos, nrm, tris = all_samples_and_tris(lib, b)
    pn = scenes.normalize_rows_f32(nrm)
    pn[5] = np.nan
    pn[17] = (np.inf, 0.0, 0.0)
    dev = Device(lib, tris)
    try:
        rm, ot, fa = links(lib, dev, pos, pn, 2)
        rm2, ot2, fa2 = orc.rad_links(tris, pos, pn)
        assert np.array_equal(rm, rm2)
        assert np.array_equal(ot, ot2)
        nan = np.isnan(fa)
        assert np.array_equal(nan, np.isnan(fa2)) and nan.any()
        assert np.array_equal(bits(fa[~nan]), bits(fa2[~nan]))
        assert len(ot) > 100
    finally:
        dev.close()
        b.destroy()


def test_bake_with_bounces_above_the_old_patch_limit(lib):
    """The staged bake with bounce_count = 1 on 316,375 patches (above
    262,144): it completes, the radiosity stage finds links, and the bounce
    only adds light to the direct-light lightmaps."""
    spec = scenes.city_scene(extent=200.0, grid=32, amplitude=2.0, n_boxes=144, tiles=4, lm_total=640,
                             n_lights=2)
    b = prepared(lib, spec, bounce_count=1)
    try:
        n = lib.lgx_InstanceCount(b.scene)
        _lib.check(lib.lgx_StageDirect(b.scene), lib)
        sizes = [len(prod_samples(lib, b.scene, i)[0]) for i in range(n)]
        assert sum(sizes) > 262144
        before = []
        for i in range(n):
            lm = np.zeros((sizes[i], 3), np.float32)
            _lib.check(lib.lgx_GetLightmap(b.scene, i, P(lm)), lib)
            before.append(lm)
        _lib.check(lib.lgx_StageRadiosity(b.scene), lib)
        assert lib.lgx_RadLinkCount(b.scene) > 1000
        gained = 0
        for i in range(n):
            lm = np.zeros((sizes[i], 3), np.float32)
            _lib.check(lib.lgx_GetLightmap(b.scene, i, P(lm)), lib)
            assert np.all(np.isfinite(lm))
            assert np.all(lm >= before[i])
            gained += int(np.sum(np.any(lm > before[i], axis=1)))
        assert gained > 100
    finally:
        b.destroy()


def test_c1_links_beyond_32_bit_offsets(lib):
    """C1's 2,989,242 texels have about 2.2 G links, 4.5 G CSR entries: the
    row offsets go 64-bit and the per-row sort runs in pieces of at most
    2^31 entries.  The search is repeatable, one bounce only adds light, and
    the entry points that hand out 32-bit offsets (the reference's link
    list, the solver forms) refuse with an error instead of wrapping."""
    from lighter_amd import workloads
    rad = []
    b, pos, nrm, tris, _, _ = workloads.direct_workload(lib, scenes.terrain_scene(), 16, rad)
    rad = np.concatenate(rad)
    pn = scenes.normalize_rows_f32(nrm)
    N = len(pos)
    dev = Device(lib, tris)
    try:
        counts = []
        for _ in range(2):
            nl = C.c_size_t(0)
            _lib.check(lib.lgx_RadLinks(dev.h, P(pos), P(pn), N, C.byref(nl)), lib)
            counts.append(nl.value)
        print(f"C1: {N} patches, {counts[0]} links")
        assert counts[0] == counts[1] and 2 * counts[0] >= 2 ** 32
        diff = np.full((N, 3), 0.6, np.float32)
        emit = np.ascontiguousarray(np.random.default_rng(3).uniform(0, 1, (N, 3)).astype(np.float32))
        area = np.ascontiguousarray(rad[:, 3])
        total = np.zeros((N, 3), np.float32)
        _lib.check(lib.lgx_RadBounce(dev.h, P(diff), P(emit), P(area), 1, P(total)), lib)
        assert np.all(np.isfinite(total)) and np.all(total >= emit) and np.mean(total > emit) > 0.5
        rowmap = np.zeros(2 * N, np.uint32)
        one = np.zeros(1, np.uint32)
        assert lib.lgx_RadReadLinks(dev.h, rowmap.ctypes.data_as(u32p), one.ctypes.data_as(u32p),
                                    P(np.zeros(1, np.float32))) == 0
        assert "2^32" in _lib.last_error(lib)
        E = np.ones((N, 3))
        out = np.zeros((N, 3))
        t = (C.c_double * 4)()
        dp = C.POINTER(C.c_double)
        assert lib.lgx_RadSolve(dev.h, E.ctypes.data_as(dp), 3, P(diff), P(area), 3, 1, 0,
                                out.ctypes.data_as(dp), t) == 0
        assert "2^32" in _lib.last_error(lib)
    finally:
        dev.close()
        b.destroy()

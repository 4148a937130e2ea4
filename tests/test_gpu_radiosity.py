"""Radiosity solver forms on the GPU against the FP64 series of the C oracle.

total = sum_{k=0..b} K^k E, K = F diag(area) diag(diffuse_{c%3}) (closed form
of lighter.cpp:767-797, 1104-1123) for many right-hand sides.  Tolerances
(relative L2 error against FP64), written per form:
  sparse FP64 1e-12, sparse FP32 1e-5, dense FP32 1e-5,
  BF16 tiles (MFMA) 2e-5 unrefined / 1e-6 after one refinement sweep,
  FP16 tiles (MFMA) 2e-6 unrefined / 1e-7 after one refinement sweep,
  FP32 tiles (FP32 matrix cores) 1e-5,
  FP16 x 3 (compensated product every bounce) 1e-5, BF16 x 3 1e-6.
The 16-bit tile forms apply the largest link factors (the top binades, a few
hundred near-coincident pairs where surfaces meet) in FP32 and the rest in
16 bits; rounded to 16 bits those few factors alone would leave errors near
the format's unit roundoff (about 1e-2 for BF16, 3e-4 for FP16).
"""
import ctypes as C

import numpy as np
import pytest

from helpers import Device, P, all_samples_and_tris, prepared
from lighter_amd import _lib, scenes

pytestmark = pytest.mark.gpu

TOL = {0: 1e-12, 1: 1e-5, 2: 1e-5, 3: 2e-5, 4: 2e-6, 5: 1e-5, 6: 1e-5, 7: 1e-6}
TOL_REFINED = {3: 1e-6, 4: 1e-7}


def scene_links(lib, spec):
    b = prepared(lib, spec)
    pos, nrm, tris = all_samples_and_tris(lib, b)
    pn = scenes.normalize_rows_f32(nrm)
    dev = Device(lib, tris)
    nl = C.c_size_t(0)
    _lib.check(lib.lgx_RadLinks(dev.h, P(pos), P(pn), len(pos), C.byref(nl)), lib)
    N = len(pos)
    rowmap = np.zeros(2 * N, np.uint32)
    other = np.zeros(nl.value, np.uint32)
    fac = np.zeros(nl.value, np.float32)
    u32p = C.POINTER(C.c_uint32)
    _lib.check(lib.lgx_RadReadLinks(dev.h, rowmap.ctypes.data_as(u32p), other.ctypes.data_as(u32p), P(fac)), lib)
    return b, dev, (dev, N, rowmap, other, fac)


@pytest.fixture(scope="module")
def links(lib):
    b, dev, out = scene_links(lib, scenes.room_scene(lm=40, n_boxes=6, bounce_count=1))
    yield out
    dev.close()
    b.destroy()


@pytest.fixture(scope="module")
def city_links(lib):
    """A 200 m city block in 4 x 4 tiles (16,988 patches): links reach a few
    tens of metres, so most 256-row x 32-column tiles of H hold none."""
    spec = scenes.city_scene(extent=200.0, grid=32, amplitude=2.0, n_boxes=144, tiles=4, lm_total=160,
                             n_lights=2)
    b, dev, out = scene_links(lib, spec)
    yield out
    dev.close()
    b.destroy()


def solve(lib, dev, E, diff, area, form, bounces, refine=0):
    N, R = E.shape
    E = np.ascontiguousarray(E, np.float64)
    out = np.zeros_like(E)
    t = (C.c_double * 4)()
    dp = C.POINTER(C.c_double)
    _lib.check(lib.lgx_RadSolve(dev.h, E.ctypes.data_as(dp), R, P(diff), P(area), form, bounces, refine,
                                out.ctypes.data_as(dp), t), lib)
    return out, list(t)


@pytest.mark.parametrize("form", [0, 1, 2, 3, 4, 5, 6, 7])
def test_solver_forms_against_fp64(lib, orc, links, form):
    dev, N, rowmap, other, fac = links
    rng = np.random.default_rng(30 + form)
    R = 12
    diff = rng.uniform(0.2, 0.9, (N, 3)).astype(np.float32)
    area = rng.uniform(0.005, 0.05, N).astype(np.float32)
    E = rng.uniform(0, 1, (N, R))
    E[rng.uniform(size=N) < 0.5] *= 1e-3  # dark patches: wide dynamic range
    bounces = 4
    exp = orc.rad_series_f64(rowmap, other, fac, area, diff, E, bounces)
    got, _ = solve(lib, dev, E, diff, area, form, bounces)
    rel = np.linalg.norm(got - exp) / np.linalg.norm(exp)
    assert rel < TOL[form], (form, rel)
    # the series must actually move light around
    assert np.linalg.norm(exp - E) / np.linalg.norm(E) > 1e-3
    if form in TOL_REFINED:
        got, _ = solve(lib, dev, E, diff, area, form, bounces, refine=1)
        rel1 = np.linalg.norm(got - exp) / np.linalg.norm(exp)
        assert rel1 < TOL_REFINED[form], (form, rel, rel1)
        assert rel1 < rel


def test_full_width_192_columns(lib, orc, links):
    """The production shape: 192 right-hand sides (64 setups x RGB)."""
    dev, N, rowmap, other, fac = links
    rng = np.random.default_rng(40)
    R = 192
    diff = rng.uniform(0.2, 0.9, (N, 3)).astype(np.float32)
    area = rng.uniform(0.005, 0.05, N).astype(np.float32)
    E = rng.uniform(0, 1, (N, R))
    exp = orc.rad_series_f64(rowmap, other, fac, area, diff, E, 3)
    for form, tol in [(1, 1e-5), (2, 1e-5), (3, 3e-2), (4, 3e-3), (5, 1e-5), (6, 1e-5), (7, 1e-4)]:
        got, _ = solve(lib, dev, E, diff, area, form, 3)
        assert np.linalg.norm(got - exp) / np.linalg.norm(exp) < tol, form


def stored_tiles(lib, dev):
    out = (C.c_uint64 * 2)()
    _lib.check(lib.lgx_RadTiles(dev.h, out), lib)
    return int(out[0]), int(out[1])


def tiles_holding_links(N, rowmap, other, order):
    """256 x 32 tiles of the symmetric link matrix that hold a link, with the
    patches at positions order (position -> patch)."""
    Np = max(512, -(-N // 256) * 256)
    pos = np.empty(N, np.int64)
    pos[order] = np.arange(N)
    i = pos[np.repeat(np.arange(N, dtype=np.int64), rowmap[1::2].astype(np.int64))]
    j = pos[other.astype(np.int64)]
    nks = Np // 32
    t = np.concatenate([(i >> 8) * nks + (j >> 5), (j >> 8) * nks + (i >> 5)])
    return len(np.unique(t)), (Np // 256) * nks


@pytest.mark.parametrize("form", [3, 4, 5, 6, 7])
def test_tile_skipping_on_a_city_block(lib, orc, city_links, form):
    """The 16-bit forms multiply only the tiles that hold links; on a city
    block most tiles are empty, and the series and its refinement still meet
    the dense tolerances (and the product equals the FP32 form's to the
    16-bit rounding)."""
    dev, N, rowmap, other, fac = city_links
    rng = np.random.default_rng(50 + form)
    R = 24
    diff = rng.uniform(0.2, 0.9, (N, 3)).astype(np.float32)
    area = rng.uniform(0.01, 0.06, N).astype(np.float32)
    E = rng.uniform(0, 1, (N, R))
    E[rng.uniform(size=N) < 0.5] *= 1e-3
    bounces = 4
    exp = orc.rad_series_f64(rowmap, other, fac, area, diff, E, bounces)
    assert np.linalg.norm(exp - E) / np.linalg.norm(E) > 1e-3
    got, _ = solve(lib, dev, E, diff, area, form, bounces)
    rel = np.linalg.norm(got - exp) / np.linalg.norm(exp)
    assert rel < TOL[form], (form, rel)
    stored, total = stored_tiles(lib, dev)
    # the Morton order stores fewer tiles than the patches' own order would
    unordered, total2 = tiles_holding_links(N, rowmap, other, np.arange(N))
    assert total == total2 and 0 < stored < unordered and stored < 0.5 * total, (stored, unordered, total)
    if form in TOL_REFINED:
        got1, _ = solve(lib, dev, E, diff, area, form, bounces, refine=1)
        rel1 = np.linalg.norm(got1 - exp) / np.linalg.norm(exp)
        assert rel1 < TOL_REFINED[form], (form, rel, rel1)
    # rows without links keep their emission exactly
    lonely = np.ones(N, bool)
    lonely[np.repeat(np.arange(N), rowmap[1::2])] = False
    lonely[other] = False
    if lonely.any():
        assert np.array_equal(got[lonely], E[lonely])


def test_patches_without_links_keep_their_emission(lib):
    """Coplanar patches facing the same way never link (a <= 0.001): every
    form returns E itself, and the 16-bit forms store no tile at all (their
    product has no work and must still write zeros)."""
    n = 300  # not a multiple of 256
    g = np.random.default_rng(7)
    pos = np.zeros((n, 3), np.float32)
    pos[:, :2] = g.uniform(-5, 5, (n, 2))
    nrm = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    tri = np.array([[-10, -10, -1, 10, -10, -1, 0, 10, -1]], np.float32)
    dev = Device(lib, tri)
    try:
        nl = C.c_size_t(0)
        _lib.check(lib.lgx_RadLinks(dev.h, P(pos), P(nrm), n, C.byref(nl)), lib)
        assert nl.value == 0
        E = g.uniform(0, 1, (n, 5))
        diff = np.full((n, 3), 0.5, np.float32)
        area = np.full(n, 0.1, np.float32)
        for form in range(8):
            got, _ = solve(lib, dev, E, diff, area, form, 3, refine=1 if form in (3, 4) else 0)
            assert np.array_equal(got, E), form
        assert stored_tiles(lib, dev)[0] == 0
    finally:
        dev.close()


@pytest.mark.parametrize("n", [0, 1])
def test_zero_and_one_patch(lib, n):
    """Degenerate patch sets: no patch at all (every form succeeds and writes
    nothing), and a single patch (no pairs: every form returns E)."""
    tri = np.array([[-10, -10, -1, 10, -10, -1, 0, 10, -1]], np.float32)
    dev = Device(lib, tri)
    try:
        pos = np.zeros((max(n, 1), 3), np.float32)
        nrm = np.tile(np.array([0, 0, 1], np.float32), (max(n, 1), 1))
        nl = C.c_size_t(7)
        _lib.check(lib.lgx_RadLinks(dev.h, P(pos), P(nrm), n, C.byref(nl)), lib)
        assert nl.value == 0
        E = np.full((n, 3), 0.25)
        diff = np.full((max(n, 1), 3), 0.5, np.float32)
        area = np.full(max(n, 1), 0.1, np.float32)
        for form in range(8):
            got, _ = solve(lib, dev, E, diff, area, form, 2, refine=1 if form in (3, 4) else 0)
            assert got.shape == (n, 3) and np.array_equal(got, E), form
    finally:
        dev.close()


@pytest.mark.parametrize("name", ["room_96", "city_block"])
def test_every_link_equals_the_canonical_any_hit(lib, name):
    """Link visibility at scale: every candidate pair that passes the
    geometric tests (Int_RDGenLinks, lighter.cpp:728-759) is tested again
    with the canonical any-hit walk (nearest child first, ending at the
    first hit with r < 1) on a plain binary SAH tree; the product's link
    set must agree pair for pair (occlusion is order independent)."""
    if name == "room_96":
        spec = scenes.room_scene(lm=96, n_boxes=20, bounce_count=1)
    else:
        spec = scenes.city_scene(extent=200.0, grid=32, amplitude=2.0, n_boxes=144, tiles=4, lm_total=160,
                                 n_lights=2)
    b = prepared(lib, spec)
    pos, nrm, tris = all_samples_and_tris(lib, b)
    pn = scenes.normalize_rows_f32(nrm)
    dev = Device(lib, tris)
    nl = C.c_size_t(0)
    lib.lgx_SetCounting(dev.h, 1)
    _lib.check(lib.lgx_RadLinks(dev.h, P(pos), P(pn), len(pos), C.byref(nl)), lib)
    lib.lgx_SetCounting(dev.h, 0)
    out = (C.c_uint64 * 5)()
    _lib.check(lib.lgx_ReferenceCountsLinks(dev.h, P(tris), len(tris), out), lib)
    visits, tests, pairs, differ, links = (int(x) for x in out)
    print(f"{name}: {len(pos)} patches, {pairs} candidate pairs, {links} links, {differ} differ")
    assert links == nl.value and pairs >= links > 1000
    assert differ == 0
    dev.close()
    b.destroy()

"""Radiosity solver forms on the GPU against the FP64 series of the C oracle.

total = sum_{k=0..b} K^k E, K = F diag(area) diag(diffuse_{c%3}) (closed form
of lighter.cpp:767-797, 1104-1123) for many right-hand sides.  Tolerances
(relative L2 error against FP64), written per form:
  sparse FP64 1e-12, sparse FP32 1e-5, dense FP32 1e-5,
  dense BF16 (MFMA) 3e-2 unrefined / 1e-4 after one refinement sweep,
  dense FP16 (MFMA) 3e-3 unrefined / 1e-5 after one refinement sweep.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import Device, P, all_samples_and_tris, prepared
from lighter_amd import _lib, scenes
from oracle.checkers import normalize_f32

pytestmark = pytest.mark.gpu

TOL = {0: 1e-12, 1: 1e-5, 2: 1e-5, 3: 3e-2, 4: 3e-3}
TOL_REFINED = {3: 1e-4, 4: 1e-5}


@pytest.fixture(scope="module")
def links(lib):
    spec = scenes.room_scene(lm=40, n_boxes=6, bounce_count=1)
    b = prepared(lib, spec)
    pos, nrm, tris = all_samples_and_tris(lib, b)
    pn = np.array([normalize_f32(v) for v in nrm], np.float32)
    dev = Device(lib, tris)
    nl = C.c_size_t(0)
    _lib.check(lib.lgx_RadLinks(dev.h, P(pos), P(pn), len(pos), C.byref(nl)), lib)
    N = len(pos)
    rowmap = np.zeros(2 * N, np.uint32)
    other = np.zeros(nl.value, np.uint32)
    fac = np.zeros(nl.value, np.float32)
    u32p = C.POINTER(C.c_uint32)
    _lib.check(lib.lgx_RadReadLinks(dev.h, rowmap.ctypes.data_as(u32p), other.ctypes.data_as(u32p), P(fac)), lib)
    yield dev, N, rowmap, other, fac
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


@pytest.mark.parametrize("form", [0, 1, 2, 3, 4])
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
    for form, tol in [(2, 1e-5), (3, 3e-2), (4, 3e-3)]:
        got, _ = solve(lib, dev, E, diff, area, form, 3)
        assert np.linalg.norm(got - exp) / np.linalg.norm(exp) < tol, form

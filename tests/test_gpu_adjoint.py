"""Reverse mode of the bake on the GPU (lgx_RadSolveGrad, lgx_DirectLightGrad).

The reference has no derivatives, so the checks rest on the forward pieces
already pinned elsewhere (the oracle's FP64 bounce series, the GPU direct
light) and on what the mathematics fixes:

* adjoint identity: T is linear in E, so <g, T(E')> = <lam, E'> for any E'
  (T from the oracle's FP64 series over the same links);
* albedo derivative: central difference of the oracle's FP64 series along a
  random direction (T is a polynomial in the albedo);
* colour derivative: the lightmap is linear in a light's colour, so two GPU
  direct-light runs differing in one colour give <lam, L2 - L1> exactly;
* the whole chain through torch.autograd against central differences
  (h = 2^-10, about 1e-3) on 20 random parameters: tolerance 2e-5 relative
  in the sparse FP64 form, 1e-3 in the FP16 tile form and in the FP16 x 3
  form (the error-compensated 16-bit product on every bounce, forward and
  reverse).

Tolerances (relative to the norm of the reference quantity), per form:
  sparse FP64 1e-10, sparse FP32 1e-5, dense FP32 1e-5, BF16 tiles 1e-3,
  FP16 tiles 3e-4, FP32 tiles 1e-5, FP16 x 3 1e-5, BF16 x 3 1e-5.
The plain 16-bit forms' widest error is the albedo derivative (16-bit link
factors in both sweeps); total, adjoint identity and lam stay near 1e-6.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from adjoint_fn import BakeContext, BakeFn
from helpers import Device, P, all_samples_and_tris, prepared
from lighter_amd import _lib, scenes

pytestmark = pytest.mark.gpu

TOL = {0: 1e-10, 1: 1e-5, 2: 1e-5, 3: 1e-3, 4: 3e-4, 5: 1e-5, 6: 1e-5, 7: 1e-5}
dp = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)


def _make_lights(spec):
    arr = (_lib.LgxLight * len(spec.lights))()
    for i, L in enumerate(spec.lights):
        d = np.asarray(L.get("direction", [0, 0, 0]), np.float64)
        n = np.linalg.norm(d)
        a = arr[i]
        a.type = L["type"]
        a.pos[:] = L.get("position", [0, 0, 0])
        a.dir[:] = [float(x) for x in (d / n if n > 0 else d)]
        a.color[:] = L.get("color_rgb", [0, 0, 0])
        a.range = L.get("range", 0)
        a.power = L.get("power", 0)
        a.radius = L.get("light_radius", 0)
        a.spot_angle_out = L.get("spot_angle_out", 0)
        a.spot_angle_in = L.get("spot_angle_in", 0)
        a.spot_curve = L.get("spot_curve", 0)
    return arr


@pytest.fixture(scope="module")
def room(lib):
    """Room samples as patches, its links, and direct-light work for 3 point
    lights + the sun over every sample."""
    spec = scenes.room_scene(lm=32, n_boxes=4, n_point=3, bounce_count=1)
    b = prepared(lib, spec)
    pos, nrm, tris = all_samples_and_tris(lib, b)
    pn = scenes.normalize_rows_f32(nrm)
    N = len(pos)
    dev = Device(lib, tris)
    lm0 = np.zeros((N, 3), np.float32)
    _lib.check(lib.lgx_SetSamples(dev.h, P(pos), P(nrm), N, P(lm0)), lib)
    lights = _make_lights(spec)
    L = len(spec.lights)
    items = np.concatenate([(np.uint32(l) << np.uint32(26)) | np.arange(N, dtype=np.uint32) for l in range(L)])
    light_off = np.arange(L + 1, dtype=np.uint32) * N
    _lib.check(lib.lgx_DirectLight(dev.h, lights, L, items.ctypes.data_as(u32p), len(items),
                                   light_off.ctypes.data_as(u32p), 1), lib)
    w = np.zeros(len(items), np.float32)
    vis = np.zeros(len(items), np.float32)
    _lib.check(lib.lgx_ReadWeights(dev.h, P(w), P(vis), len(items)), lib)
    nl = C.c_size_t(0)
    _lib.check(lib.lgx_RadLinks(dev.h, P(pos), P(pn), N, C.byref(nl)), lib)
    rowmap = np.zeros(2 * N, np.uint32)
    other = np.zeros(nl.value, np.uint32)
    fac = np.zeros(nl.value, np.float32)
    _lib.check(lib.lgx_RadReadLinks(dev.h, rowmap.ctypes.data_as(u32p), other.ctypes.data_as(u32p), P(fac)), lib)
    rng = np.random.default_rng(7)
    area = rng.uniform(0.005, 0.05, N).astype(np.float32)
    yield dict(spec=spec, b=b, dev=dev, N=N, pos=pos, nrm=nrm, lights=lights, items=items, light_off=light_off,
               w=w, rowmap=rowmap, other=other, fac=fac, area=area, L=L)
    dev.close()
    b.destroy()


def _grad(lib, room, E, diff, g, form, bounces):
    N, R = E.shape
    E = np.ascontiguousarray(E, np.float64)
    g = np.ascontiguousarray(g, np.float64)
    tot, lam, dalb = np.zeros_like(E), np.zeros_like(E), np.zeros_like(E)
    t = (C.c_double * 4)()
    _lib.check(lib.lgx_RadSolveGrad(room["dev"].h, E.ctypes.data_as(dp), R, P(diff), P(room["area"]), form,
                                    bounces, g.ctypes.data_as(dp), tot.ctypes.data_as(dp),
                                    lam.ctypes.data_as(dp), dalb.ctypes.data_as(dp), t), lib)
    return tot, lam, dalb


@pytest.mark.parametrize("form", [0, 1, 2, 3, 4, 5, 6, 7])
def test_radiosity_adjoint(lib, orc, room, form):
    N = room["N"]
    rng = np.random.default_rng(40 + form)
    R, B = 6, 4
    diff = rng.uniform(0.2, 0.9, (N, 3)).astype(np.float32)
    E = rng.uniform(0, 1, (N, R))
    g = rng.normal(size=(N, R))
    series = lambda e, d: orc.rad_series_f64(room["rowmap"], room["other"], room["fac"], room["area"], d, e, B)
    tot, lam, dalb = _grad(lib, room, E, diff, g, form, B)
    exp_tot = series(E, diff)
    assert np.linalg.norm(tot - exp_tot) / np.linalg.norm(exp_tot) < TOL[form]
    # adjoint identity on random directions: <g, T(E')> = <lam, E'> per column
    for k in range(3):
        Ep = rng.normal(size=(N, R))
        lhs = (g * series(Ep, diff)).sum(0)
        rhs = (lam * Ep).sum(0)
        assert np.abs(lhs - rhs).max() <= TOL[form] * np.abs(lhs).max(), (form, lhs, rhs)
    # lam against the adjoint series computed from the oracle columns
    # (T is linear: lam_j = <g, T(e_j)>) on a few patches
    for j in rng.choice(N, 4, replace=False):
        ej = np.zeros((N, R))
        ej[j] = 1.0
        exact = (g * series(ej, diff)).sum(0)
        assert np.abs(lam[j] - exact).max() <= TOL[form] * np.abs(lam).max(), (form, j)
    # albedo: central difference of the FP64 series along a random direction
    dirn = rng.normal(size=(N, 3))
    h = 1e-4
    dp_ = (diff + h * dirn)
    dm_ = (diff - h * dirn)
    fd = ((g * (series(E, dp_) - series(E, dm_))).sum()) / (2 * h)
    # dalb is per column; column c belongs to channel c % 3
    an = sum((dalb[:, c] * dirn[:, c % 3]).sum() for c in range(R))
    assert abs(an - fd) <= max(TOL[form], 1e-7) * abs(fd), (form, an, fd)


def test_direct_light_color_gradient(lib, room):
    """Linearity in a light's colour: one GPU run with colour + 1 in one channel."""
    dev, N, L = room["dev"], room["N"], room["L"]
    rng = np.random.default_rng(3)
    lam = rng.normal(size=(N, 3))
    # dcolor from the kernel, for the current work (run in the fixture)
    lights = room["lights"]
    items, off = room["items"], room["light_off"]
    lm0 = np.zeros((N, 3), np.float32)

    def run(lts):
        _lib.check(lib.lgx_SetSamples(dev.h, P(room["pos"]), P(room["nrm"]), N, P(lm0)), lib)
        _lib.check(lib.lgx_DirectLight(dev.h, lts, L, items.ctypes.data_as(u32p), len(items),
                                       off.ctypes.data_as(u32p), 1), lib)
        out = np.zeros((N, 3), np.float32)
        _lib.check(lib.lgx_ReadLightmap(dev.h, P(out)), lib)
        return out.astype(np.float64)

    base = run(lights)
    dcol = np.zeros((L, 3))
    _lib.check(lib.lgx_DirectLightGrad(dev.h, np.ascontiguousarray(lam).ctypes.data_as(dp),
                                       dcol.ctypes.data_as(dp)), lib)
    assert np.abs(dcol).max() > 0
    for l in range(L):
        for c in range(3):
            lp = (_lib.LgxLight * L)()
            C.memmove(lp, lights, C.sizeof(lights))
            lp[l].color[c] = lp[l].color[c] + 1.0
            diff = (lam * (run(lp) - base)).sum()
            assert abs(diff - dcol[l, c]) <= 1e-5 * np.abs(dcol).max() + 1e-6, (l, c, diff, dcol[l, c])


@pytest.mark.parametrize("form,tol", [(0, 2e-5), (4, 1e-3), (6, 1e-3)])
def test_autograd_chain_matches_finite_differences(lib, room, form, tol):
    """Colours, albedo groups and emission through torch.autograd; 20 random
    parameters against central differences with h = 2^-10 of the same
    form's forward bake (sparse FP64; FP16 tiles with the largest factors in
    FP32, and FP16 x 3, on the matrix cores)."""
    N, L = room["N"], room["L"]
    rng = np.random.default_rng(11)
    bake = BakeContext(lib, room["dev"].h, room["items"], room["light_off"], room["w"], room["area"], form=form,
                       bounces=4)
    # the weights must be those of the fixture's run (the colour test re-ran it)
    _lib.check(lib.lgx_DirectLight(room["dev"].h, room["lights"], L, room["items"].ctypes.data_as(u32p),
                                   len(room["items"]), room["light_off"].ctypes.data_as(u32p), 1), lib)
    groups = rng.integers(0, 8, N)
    # albedo goes to the device as FP32: values on a 2^-12 grid and h = 2^-10
    # keep every perturbed value exact there
    q = lambda a: np.round(a * 4096) / 4096
    colors = torch.tensor(q(rng.uniform(0.3, 1.0, (L, 3))), requires_grad=True)
    galb = torch.tensor(q(rng.uniform(0.2, 0.8, (8, 3))), requires_grad=True)
    emis = torch.tensor(rng.uniform(0, 0.1, (N, 3)), requires_grad=True)
    target = torch.from_numpy(rng.uniform(0, 1, (N, 3)))

    def loss_of(col, ga, em):
        T = BakeFn.apply(col, ga[torch.from_numpy(groups)], em, bake)
        return 0.5 * ((T - target) ** 2).sum()

    loss = loss_of(colors, galb, emis)
    loss.backward()
    params = [colors, galb, emis]
    h = 2.0 ** -10
    checked = 0
    for k in range(20):
        p = params[k % 3]
        idx = tuple(int(rng.integers(0, s)) for s in p.shape)
        with torch.no_grad():
            vals = [q.detach().clone() for q in params]
            vals[k % 3][idx] += h
            fp_ = loss_of(*vals).item()
            vals[k % 3][idx] -= 2 * h
            fm_ = loss_of(*vals).item()
        fd = (fp_ - fm_) / (2 * h)
        an = p.grad[idx].item()
        assert abs(an - fd) <= tol * max(1.0, abs(fd)), (k, idx, an, fd)
        checked += 1
    assert checked == 20

"""Full-size parity on the 2M-triangle city block (1 km heightfield + 155,000
lognormal boxes, 1,991,072 triangles in 16 x 16 tiles, 4096^2 texels, 64
point/spot lights) and a reference sample of the terrain workload.

Against the reference baker built from its sources (oracle/_ref):
* shadow factors of a seeded sample of marched (texel, light) pairs bit for
  bit against CalcInvShadowFactor (lighter.cpp:190-207), on the city
  (150,000 marches) and the terrain (100,000);
* AO factors of 60,000 sampled city texels against the reference's AO ray
  loop (lighter.cpp:812-831) within 1/255 on at least 99.5 % (the GPU's
  azimuth sine/cosine are correctly rounded, the C library's are not).
Properties of every city march (no reference needed): the march's steps do
not depend on the penumbra constant k (lighter.cpp:197-204), so the fully
shadowed pairs (a step closer than 0.001) are the same for k = 0, k and 2k,
k = 0 gives 0 or 1, and the factor does not grow with k.
Radiosity over the city's first 65,536 texels (one 62.5 m tile), one RGB
setup: every form against the oracle's FP64 series on the same links; the
plain BF16 and FP16 MFMA forms (largest factors in FP32, the rest in 16
bits) at or below 2e-6, refined or compensated (hi + lo) at or below 1e-6,
the FP32 forms at or below 1e-6.
Colour gradient of an L2 loss over every texel against central differences
of the GPU bake on 20 light intensities, to 1e-3.
"""
import ctypes as C

import numpy as np
import pytest

from lighter_amd import _lib, scenes, workloads

pytestmark = [pytest.mark.gpu]

fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
SKIP = 0xFFFFFFFF


def P(a):
    return a.ctypes.data_as(fp)


def D(a):
    return a.ctypes.data_as(dp)


class Workload:
    """A prepared scene on the device with its direct-light work items."""

    def __init__(self, lib, spec, radinfo=False):
        self.lib, self.spec = lib, spec
        self.nth = workloads.host_cores()[0]
        self.radinfo = [] if radinfo else None
        (self.built, self.pos, self.nrm, self.tris, self.items,
         self.off) = workloads.direct_workload(lib, spec, self.nth, self.radinfo)
        if radinfo:
            self.radinfo = np.concatenate(self.radinfo)
        self.S = len(self.pos)
        self.dev = lib.lgx_DeviceCreate(P(self.tris), len(self.tris))
        assert self.dev, _lib.last_error(lib)

    def direct(self, lights, lightmap=None, keep_vis=1):
        """One direct-light pass over every item from the given lightmap
        (zeros by default); returns (lightmap, weights, shadow factors)."""
        lib = self.lib
        lm0 = np.zeros((self.S, 3), np.float32) if lightmap is None else lightmap
        _lib.check(lib.lgx_SetSamples(self.dev, P(self.pos), P(self.nrm), self.S, P(lm0)), lib)
        _lib.check(lib.lgx_DirectLightSetup(self.dev, lights, len(lights), self.items.ctypes.data_as(u32p),
                                            len(self.items), self.off.ctypes.data_as(u32p), keep_vis), lib)
        _lib.check(lib.lgx_DirectLightRun(self.dev), lib)
        lm = np.zeros((self.S, 3), np.float32)
        _lib.check(lib.lgx_ReadLightmap(self.dev, P(lm)), lib)
        w = np.zeros(len(self.items), np.float32)
        vis = np.zeros(len(self.items), np.float32) if keep_vis else None
        _lib.check(lib.lgx_ReadWeights(self.dev, P(w), P(vis) if keep_vis else None, len(self.items)), lib)
        return lm, w, vis

    def reference_scene(self, ref):
        br = scenes.BuiltScene(ref.lib, self.spec, dict(max_num_threads=self.nth))
        assert ref.lib.rp_prepare_collision(br.scene, self.nth) == 1
        return br

    def march_inputs(self, lights, idx):
        """(from, to, k) of the reference's march for work items idx
        (lighter.cpp:503, 541: from the sample offset 0.005 along its
        normal to the light's position)."""
        lid = self.items[idx] >> 26
        smp = self.items[idx] & ((1 << 26) - 1)
        frm = (self.pos[smp] + self.nrm[smp] * np.float32(0.005)).astype(np.float32)
        to = np.zeros_like(frm)
        k = np.zeros(len(idx), np.float32)
        for li in np.unique(lid):
            m = lid == li
            L = lights[int(li)]
            assert L.type in (1, 2), "point and spot lights only"
            to[m] = np.array(L.pos[:], np.float32)
            k[m] = L.radius
        return frm, to, k

    def close(self):
        if self.dev:
            self.lib.lgx_DeviceDestroy(self.dev)
            self.dev = None
        self.built.destroy()


def reference_marches(ref, br, nth, frm, to, k):
    out = np.zeros(len(frm), np.float32)
    ref.lib.rp_shadow_factor_batch_mt(br.scene, len(frm), P(frm), P(to), P(k), P(out), nth)
    return out


@pytest.fixture(scope="module")
def city(lib):
    if not lib.lgx_HasDevice():
        pytest.skip("no HIP device")
    spec = scenes.city_scene()
    w = Workload(lib, spec, radinfo=True)
    w.lights = workloads.make_lights(spec)
    w.lm, w.w, w.vis = w.direct(w.lights)
    w.marched = np.nonzero(w.w.view(np.uint32) != SKIP)[0]
    assert spec.num_triangles == 1991072 and w.S > 11_000_000 and len(w.marched) > 2_000_000
    yield w
    w.close()


@pytest.fixture(scope="module")
def city_ref(city, ref):
    """The reference's collision structures of the city (PreXForm, ColInfo,
    instance tree; lighter.cpp:1050-1069)."""
    br = city.reference_scene(ref)
    yield br
    br.destroy()


@pytest.mark.ref
def test_c4_shadow_factors_match_reference(city, ref, city_ref):
    """150,000 seeded marched pairs of the city, bit for bit against the
    reference's CalcInvShadowFactor on its own collision structures."""
    rng = np.random.default_rng(0x4C540041)
    idx = np.sort(rng.choice(city.marched, 150_000, replace=False))
    frm, to, k = city.march_inputs(city.lights, idx)
    exp = reference_marches(ref, city_ref, city.nth, frm, to, k)
    got = city.vis[idx]
    same = got.view(np.uint32) == exp.view(np.uint32)
    assert same.all(), (int((~same).sum()), got[~same][:5], exp[~same][:5])
    # the sample holds shadowed, penumbra and lit pairs
    assert (exp == 0).mean() > 0.01 and ((exp > 0) & (exp < 1)).mean() > 0.01 and (exp == 1).mean() > 0.01


def test_c4_march_properties_on_every_march(city):
    """Every march of the city with the penumbra constant scaled by 0 and 2:
    the factor lies in [0, 1]; k = 0 gives 0 or 1; the fully shadowed
    pairs are the same for every k; the factor does not grow with k."""
    _, _, v0 = city.direct(workloads.make_lights(city.spec, radius_scale=0.0))
    _, _, v2 = city.direct(workloads.make_lights(city.spec, radius_scale=2.0))
    m = city.marched
    v, v0, v2 = city.vis[m], v0[m], v2[m]
    assert np.all((v >= 0) & (v <= 1))
    assert np.all((v0 == 0) | (v0 == 1))
    assert np.array_equal(v0 == 0, v == 0) and np.array_equal(v2 == 0, v == 0)
    assert np.all(v2 <= v) and np.all(v <= v0)
    assert (v2 < v).mean() > 0.01  # the penumbrae do depend on k


@pytest.mark.ref
def test_c4_ao_matches_reference(city, ref, city_ref):
    """60,000 seeded city texels: 64 cosine-weighted rays of 2 m each with a
    per-texel rotation; the device's AO over a lightmap of ones
    (multiplier 1, no falloff, effect 0, black) leaves 1 - factor."""
    lib = city.lib
    rng = np.random.default_rng(0x4C540042)
    ro = rng.uniform(0, 1, city.S).astype(np.float32)
    col = np.zeros(3, np.float32)
    _lib.check(lib.lgx_SetSamples(city.dev, P(city.pos), P(city.nrm), city.S,
                                  P(np.ones((city.S, 3), np.float32))), lib)
    _lib.check(lib.lgx_AmbientOcclusion(city.dev, P(ro), 64, 2.0, 1.0, 0.0, 0.0, P(col)), lib)
    lm = np.zeros((city.S, 3), np.float32)
    _lib.check(lib.lgx_ReadLightmap(city.dev, P(lm)), lib)
    idx = np.sort(rng.choice(city.S, 60_000, replace=False))
    p, q, r = (np.ascontiguousarray(a[idx]) for a in (city.pos, city.nrm, ro))
    fac = np.zeros(len(idx), np.float32)
    ref.lib.rp_ao_factor_batch_mt(city_ref.scene, len(idx), P(p), P(q), P(r), 64, 2.0, 1.0, P(fac), city.nth)
    got = lm[idx, 0]
    exp = np.float32(1) - fac
    close = np.abs(got - exp) <= 1.0 / 255
    assert close.mean() >= 0.995, close.mean()
    assert np.mean(got.view(np.uint32) == exp.view(np.uint32)) >= 0.99
    assert (fac > 0).mean() > 0.2  # the sample sees occlusion


@pytest.fixture(scope="module")
def city_radiosity(lib, city):
    """Links among the city's first 65,536 texels; E from the lights' direct
    weights (one colour setup) plus a small uniform emission."""
    N = 65536
    pos = np.ascontiguousarray(city.pos[:N])
    nrm = np.ascontiguousarray(city.nrm[:N])
    pn = scenes.normalize_rows_f32(nrm)
    dev = lib.lgx_DeviceCreate(P(city.tris), len(city.tris))
    _lib.check(lib.lgx_SetSamples(dev, P(pos), P(nrm), N, None), lib)
    nl = C.c_size_t(0)
    _lib.check(lib.lgx_RadLinks(dev, P(pos), P(pn), N, C.byref(nl)), lib)
    rowmap = np.zeros(2 * N, np.uint32)
    other = np.zeros(nl.value, np.uint32)
    fac = np.zeros(nl.value, np.float32)
    _lib.check(lib.lgx_RadReadLinks(dev, rowmap.ctypes.data_as(u32p), other.ctypes.data_as(u32p), P(fac)), lib)
    lid = (city.items >> 26).astype(np.int64)
    smp = (city.items & ((1 << 26) - 1)).astype(np.int64)
    wv = np.where(city.w.view(np.uint32) == SKIP, 0.0, city.w.astype(np.float64))
    sel = smp < N
    Wp = np.zeros((N, len(city.lights)))
    np.add.at(Wp, (smp[sel], lid[sel]), wv[sel])
    cols = np.random.default_rng(0x4C540043).uniform(0.5, 1.0, (len(city.lights), 3))
    E = np.ascontiguousarray(Wp @ cols + 0.05)
    diff = np.full((N, 3), 0.6, np.float32)
    area = np.ascontiguousarray(city.radinfo[:N, 3])
    yield dict(dev=dev, N=N, nlinks=nl.value, rowmap=rowmap, other=other, fac=fac, E=E, diff=diff, area=area)
    lib.lgx_DeviceDestroy(dev)


@pytest.mark.parametrize("form,refine,tol", [(0, 0, 1e-12), (1, 0, 1e-6), (2, 0, 1e-6), (3, 0, 2e-6),
                                             (3, 1, 1e-6), (4, 0, 2e-6), (4, 1, 1e-6), (5, 0, 1e-6),
                                             (6, 0, 1e-6), (7, 0, 1e-6)])
def test_c4_radiosity_tile_against_fp64_oracle(lib, orc, city_radiosity, form, refine, tol):
    """sum_{k<=4} K^k E over the 65,536-patch tile (12 M links) in each form
    against the oracle's FP64 series on the same link list."""
    r = city_radiosity
    if "exp" not in r:
        r["exp"] = orc.rad_series_f64(r["rowmap"], r["other"], r["fac"], r["area"], r["diff"], r["E"], 4)
    assert r["nlinks"] > 5_000_000
    out = np.zeros_like(r["E"])
    t = (C.c_double * 4)()
    _lib.check(lib.lgx_RadSolve(r["dev"], D(r["E"]), 3, P(r["diff"]), P(r["area"]), form, 4, refine, D(out), t),
               lib)
    err = np.linalg.norm(out - r["exp"]) / np.linalg.norm(r["exp"])
    assert err <= tol, err
    assert np.linalg.norm(r["exp"] - r["E"]) > 0.05 * np.linalg.norm(r["E"])  # the bounces add light


def test_c4_colour_gradient_against_central_differences(lib, city):
    """L = 1/2 sum |lightmap(colours) - target|^2 over every texel and
    channel, target from colours perturbed by +-30 %: lgx_DirectLightGrad's
    dL/dcolour against central differences (h = 1e-3) of the GPU bake on 20
    seeded (light, channel) intensities."""
    rng = np.random.default_rng(0x4C540044)
    col0 = np.array([[L.color[c] for c in range(3)] for L in city.lights], np.float64)
    target, _, _ = city.direct(workloads.make_lights(city.spec, colors=col0 * rng.uniform(0.7, 1.3, col0.shape)),
                               keep_vis=0)
    target = target.astype(np.float64)

    def loss(col):
        lm, _, _ = city.direct(workloads.make_lights(city.spec, colors=col), keep_vis=0)
        r = lm.astype(np.float64) - target
        return 0.5 * float(np.sum(r * r)), r

    _, r0 = loss(col0)
    lam = np.ascontiguousarray(r0)
    dcol = np.zeros((len(city.lights), 3))
    _lib.check(lib.lgx_DirectLightGrad(city.dev, D(lam), D(dcol)), lib)
    h = 1e-3
    picks = rng.choice(len(city.lights) * 3, 20, replace=False)
    fd, an = [], []
    for p in picks:
        l, c = divmod(int(p), 3)
        cp, cm = col0.copy(), col0.copy()
        cp[l, c] += h
        cm[l, c] -= h
        fd.append((loss(cp)[0] - loss(cm)[0]) / (2 * h))
        an.append(dcol[l, c])
    fd, an = np.array(fd), np.array(an)
    err = np.linalg.norm(fd - an) / np.linalg.norm(an)
    assert err <= 1e-3, (err, fd, an)
    assert np.all(np.abs(fd - an) <= 1e-3 * np.abs(an) + 1e-6 * np.abs(an).max())


@pytest.mark.ref
def test_c1_shadow_factors_match_reference(lib, ref):
    """100,000 seeded marched pairs of the terrain workload (548,288
    triangles, 32 point/spot lights and a 300 m sun) bit for bit against
    the reference's CalcInvShadowFactor; the sun's back-facing texels are
    not marches (zero shading factor) and are left out."""
    spec = scenes.terrain_scene()
    w = Workload(lib, spec)
    lights = workloads.make_lights(spec)
    _, wt, vis = w.direct(lights)
    marched = np.nonzero(wt.view(np.uint32) != SKIP)[0]
    lid = w.items[marched] >> 26
    smp = w.items[marched] & ((1 << 26) - 1)
    keep = np.ones(len(marched), bool)
    for li in np.unique(lid):
        L = lights[int(li)]
        if L.type == 3:
            m = np.nonzero(lid == li)[0]
            d = np.array(L.dir[:], np.float32)
            keep[m[(w.nrm[smp[m]] @ d) <= 0]] = False
    marched = marched[keep]
    idx = np.sort(np.random.default_rng(0x4C540045).choice(marched, 100_000, replace=False))
    lid = w.items[idx] >> 26
    smp = w.items[idx] & ((1 << 26) - 1)
    frm = (w.pos[smp] + w.nrm[smp] * np.float32(0.005)).astype(np.float32)
    to = np.zeros_like(frm)
    k = np.zeros(len(idx), np.float32)
    for li in np.unique(lid):
        m = lid == li
        L = lights[int(li)]
        k[m] = L.radius
        if L.type == 3:
            d = np.array(L.dir[:], np.float32)
            to[m] = (w.pos[smp[m]] + d * np.float32(L.range)).astype(np.float32)
        else:
            to[m] = np.array(L.pos[:], np.float32)
    br = w.reference_scene(ref)
    exp = reference_marches(ref, br, w.nth, frm, to, k)
    br.destroy()
    got = vis[idx]
    w.close()
    same = got.view(np.uint32) == exp.view(np.uint32)
    assert same.all(), int((~same).sum())
    assert (lid == len(lights) - 1).sum() > 1000  # the sun's marches are in the sample

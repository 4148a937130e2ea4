#!/usr/bin/env python3
"""Benchmark of the baking hot path on one MI355X.

Headline: shadow rays per second (one ray = one soft-shadow march of a
(texel, light) pair with a nonzero shading factor, lighter.cpp:485-600) for
the direct-lighting stage of the 2M-triangle city block: a 1 km value-noise
heightfield plus 155,000 lognormal boxes (1,991,072 triangles) in 16 x 16
tile instances with 256 x 256 lightmaps each (4096 x 4096 texels in total)
and 64 point/spot lights.

A step is one pass of the direct-light stage (shading gate, march kernel,
in-order accumulation) over all work items, with the scene, samples and
work items resident in HBM.  Timing: a host wall clock around the K timed
steps (each step returns after its work has finished on the device), and
HIP events around each march launch on the library's stream.

The JSON line also carries the march kernel's roofline figure (algorithmic
bytes per launch from reference-shaped traversal counts / average kernel
time) and the reference CPU baker's shadow-march rate on the host cores,
on the same marches (built from the reference sources).  Extra keys: the
rest of the city bake (AO, radiosity forms, colour gradient, export,
placement), the terrain workload's direct light, the room radiosity forms
and AO, and the gradient workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from lighter_amd import _lib, scenes, workloads  # noqa: E402
from lighter_amd.workloads import host_cores, make_lights  # noqa: E402

fp = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)

HBM_PEAK_GBS = 8000.0
NODE_BYTES = 112  # 4-wide node: six box planes + child references (seven dwordx4 loads)
RAY_TRI_BYTES = 64  # RayTri: the precomputed segment-test operands
DIST_TRI_BYTES = 160
TRI_BOX_BYTES = 48  # the vertex quads a distance query's box pre-test reads
METRIC = "shadow rays/s + s per radiosity bounce @ stated error; GPU plain vs MFMA/adjoint"


def P(a):
    return a.ctypes.data_as(fp)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def prepare_workload(lib, spec, nthreads, radinfo=None):
    """workloads.direct_workload, with the placement figures attached."""
    t0 = time.perf_counter()
    w = workloads.direct_workload(lib, spec, nthreads, radinfo)
    w[0].placement = placement_info(lib, w[0])
    return w


def placement_info(lib, b):
    """Times of the last GPU rasterisation and sample placement of scene b."""
    t = (C.c_double * 5)()
    lib.lgx_PlacementTimes(b.scene, t)
    r = (C.c_double * 3)()
    lib.lgx_RasterTimes(b.scene, r)
    p = (C.c_double * 7)()
    lib.lgx_PrepareTimes(b.scene, p)
    return dict(samples=int(t[4]), kernel_ms=round(t[2], 3), flatten_ms=round(t[0], 1), upload_ms=round(t[1], 1),
                download_ms=round(t[3], 1), raster_kernels_ms=round(r[0], 3), raster_texels=int(r[1]),
                prepare_steps_ms=dict(transforms=round(p[0], 1), collision_setup=round(p[1], 1),
                                      instance_tree=round(p[2], 1), samples=round(p[3], 1),
                                      samples_rasterisation=round(p[4], 1), samples_placement=round(p[5], 1),
                                      samples_commit=round(p[6], 1)))


def placement_vs_host(lib, spec, nthreads, b_gpu, pos_gpu):
    """The preparation again with the host restatements of the
    rasterisation and the sample placement: its wall time beside the GPU
    run's, and whether every sample position is the same."""
    b = scenes.BuiltScene(lib, spec, dict(max_num_threads=nthreads))
    _lib.check(lib.lgx_SetHostStages(b.scene, 5), lib)  # host rasterisation and placement
    t0 = time.perf_counter()
    _lib.check(lib.lgx_Prepare(b.scene), lib)
    host_s = time.perf_counter() - t0
    pos = []
    for i in range(lib.lgx_InstanceCount(b.scene)):
        info = (C.c_int32 * 4)()
        lib.lgx_InstanceInfo(b.scene, i, info)
        p = np.zeros((info[2], 3), np.float32)
        lib.lgx_GetSamples(b.scene, i, P(p), None, None, None)
        pos.append(p)
    b.destroy()
    pos = np.concatenate(pos)
    same = pos.shape == pos_gpu.shape and bool(np.array_equal(pos.view(np.uint32), pos_gpu.view(np.uint32)))
    out = dict(b_gpu.placement, prepare_s_gpu=round(b_gpu.prepare_s, 2), prepare_s_host=round(host_s, 2),
               host_threads=nthreads, identical_samples=same)
    log(f"[bench] rasterisation and sample placement: GPU kernels {out['raster_kernels_ms']} ms (rasterisation, "
        f"{out['raster_texels']} texels) + {out['kernel_ms']} ms (placement, {out['samples']} samples); preparation "
        f"{b_gpu.prepare_s:.2f} s (GPU) vs {host_s:.2f} s (host, {nthreads} threads); identical: {same}")
    return out


def export_vs_host(lib, b):
    """The lightmap export ("exporting lightmaps") of prepared scene b on the
    GPU, then again with the host restatement (its outputs replace the
    first call's): both wall times and whether every image is identical.
    The scene's lightmaps hold the ambient colour here (the device-level
    light runs of this bench do not write them back), so this times the
    stage; tests/test_gpu_finalize.py compares lit bakes."""
    t0 = time.perf_counter()
    _lib.check(lib.lgx_StageFinalize(b.scene), lib)
    gpu_s = time.perf_counter() - t0
    t = (C.c_double * 5)()
    lib.lgx_FinalizeTimes(b.scene, t)
    outs_gpu = b.outputs()
    _lib.check(lib.lgx_SetHostStages(b.scene, 2), lib)  # host export
    t0 = time.perf_counter()
    _lib.check(lib.lgx_StageFinalize(b.scene), lib)  # replaces the output list
    host_s = time.perf_counter() - t0
    outs_host = b.outputs()
    n = len(outs_gpu)
    same = len(outs_host) == n and all(
        g[:3] == h[:3] and np.array_equal(g[3].view(np.uint32), h[3].view(np.uint32))
        for g, h in zip(outs_gpu, outs_host))
    out = dict(images=n, texels=int(t[4]), kernels_ms=round(t[2], 3), pack_ms=round(t[0], 1),
               upload_ms=round(t[1], 1), download_ms=round(t[3], 1), stage_s_gpu=round(gpu_s, 3),
               stage_s_host=round(host_s, 3), identical_images=bool(same))
    log(f"[bench] lightmap export: {n} images, {out['texels']} texels; GPU kernels {out['kernels_ms']} ms, stage "
        f"{gpu_s:.3f} s vs host {host_s:.3f} s; identical: {same}")
    return out


def link_reference_counts(lib, dev, tris):
    """The last lgx_RadLinks' candidate pairs (kept: it ran with counting on)
    tested again with the canonical any-hit walk on a plain binary SAH tree:
    node visits, triangle tests, pairs, and links that differ."""
    out = (C.c_uint64 * 5)()
    _lib.check(lib.lgx_ReferenceCountsLinks(dev, P(tris), len(tris), out), lib)
    v, t, pairs, diff, links = (int(x) for x in out)
    log(f"[bench] link visibility, reference-shaped: {pairs} candidate pairs, {v} node visits, {t} triangle tests; "
        f"{diff} of them differ from the product's links ({links} links)")
    return dict(candidate_pairs=pairs, node_visits=v, triangle_tests=t, links=links, links_differing=diff,
                bytes=v * 32 + t * 36)


def cpu_baseline(spec, pos, nrm, items, w, lights, target_s, nthreads, ao=None):
    """Reference CalcInvShadowFactor (built from the reference sources) on a
    seeded random sample of the marched work items.  With ao = (randoff,
    rays per texel, distance, seconds, dict), the reference's AO ray loop
    (Job_AORender_Inner, lighter.cpp:812-828) is also timed on a seeded
    sample of texels of the same scene, and the dict receives the baseline
    and the sampled texels' AO factors."""
    from oracle.checkers import Reference, reference_available
    if not reference_available():
        return None, None
    ref = Reference()
    b = scenes.BuiltScene(ref.lib, spec, dict(max_num_threads=nthreads))
    t0 = time.perf_counter()
    ref.lib.rp_prepare_collision(b.scene, nthreads)
    prep_s = time.perf_counter() - t0
    marched = np.nonzero(w.view(np.uint32) != 0xFFFFFFFF)[0]
    # directional back-facing texels are not marches (zero shading factor)
    rng = np.random.default_rng(12345)
    order = rng.permutation(marched)
    lid = items >> 26
    smp = items & ((1 << 26) - 1)

    def build(idx):
        l = lid[idx]
        s = smp[idx]
        frm = (pos[s] + nrm[s] * np.float32(0.005)).astype(np.float32)
        to = np.zeros_like(frm)
        k = np.zeros(len(idx), np.float32)
        keep = np.ones(len(idx), bool)
        for li in np.unique(l):
            m = l == li
            L = lights[int(li)]
            k[m] = L.radius
            if L.type == 3:
                d = np.array(L.dir[:], np.float32)
                to[m] = (pos[s[m]] + d * np.float32(L.range)).astype(np.float32)
                fn = nrm[s[m]] @ d
                keep[np.nonzero(m)[0][fn <= 0]] = False
            else:
                to[m] = np.array(L.pos[:], np.float32)
        return frm[keep], to[keep], k[keep], idx[keep]

    # calibrate on a small sample, then size the run to about target_s
    n0 = min(len(order), 2000)
    frm, to, k, sel = build(order[:n0])
    out = np.zeros(len(frm), np.float32)
    t0 = time.perf_counter()
    ref.lib.rp_shadow_factor_batch_mt(b.scene, len(frm), P(frm), P(to), P(k), P(out), nthreads)
    dt = max(time.perf_counter() - t0, 1e-6)
    n = int(min(len(order), max(n0, n0 * target_s / dt)))
    frm, to, k, sel = build(order[:n])
    out = np.zeros(len(frm), np.float32)
    t0 = time.perf_counter()
    ref.lib.rp_shadow_factor_batch_mt(b.scene, len(frm), P(frm), P(to), P(k), P(out), nthreads)
    dt = time.perf_counter() - t0
    if ao is not None:
        ao_baseline(ref, b, pos, nrm, nthreads, *ao)
    b.destroy()
    return dict(value=len(frm) / dt, unit="shadow rays/s", cores=nthreads, kind="reference",
                sample=f"{len(frm)} of {len(marched)} marched (texel, light) pairs, seeded uniform sample; "
                       f"reference CalcInvShadowFactor on {nthreads} host threads, {dt:.1f} s "
                       f"(collision set-up {prep_s:.1f} s not counted)"), (sel, out)


def ao_baseline(ref, b, pos, nrm, nthreads, ro, nsamples, distance, target_s, out):
    """The reference's AO rays (multiplier 1, no falloff) on a seeded sample
    of texels, sized to about target_s on nthreads host threads."""
    order = np.random.default_rng(777).permutation(len(pos))

    def run(n):
        idx = np.sort(order[:n])
        p, q, r = (np.ascontiguousarray(a[idx]) for a in (pos, nrm, ro))
        fac = np.zeros(n, np.float32)
        t0 = time.perf_counter()
        ref.lib.rp_ao_factor_batch_mt(b.scene, n, P(p), P(q), P(r), nsamples, distance, 1.0, P(fac), nthreads)
        return idx, fac, max(time.perf_counter() - t0, 1e-6)

    n0 = min(len(pos), 1024)
    _, _, dt = run(n0)
    n = int(min(len(pos), max(n0, n0 * target_s / dt)))
    idx, fac, dt = run(n)
    out["baseline"] = dict(value=round(n * nsamples / dt, 1), unit="AO rays/s", cores=nthreads, kind="reference",
                           sample=f"{n} of {len(pos)} texels, seeded uniform sample, {nsamples} rays of "
                                  f"{distance:g} m each; the reference's AO ray loop (CreateSpiralDirVector, "
                                  f"DistanceTestBBT) on {nthreads} host threads, {dt:.1f} s")
    out["texels"], out["factor"] = idx, fac


FORMS = {0: "sparse_f64", 1: "sparse_f32", 2: "dense_f32", 3: "tiles_bf16_mfma", 4: "tiles_f16_mfma",
         5: "tiles_f32_mfma", 6: "tiles_f16x3_mfma", 7: "tiles_bf16x3_mfma"}
FORM_RUNS = [(0, 0), (1, 0), (2, 0), (5, 0), (3, 0), (3, 1), (4, 0), (4, 1), (6, 0), (7, 0)]
FP32_FORMS = ("sparse_f32", "dense_f32", "tiles_f32_mfma")


FP32_MATRIX_PEAK_TFLOPS = 157.3


def tile_counts(lib, dev):
    out = (C.c_uint64 * 2)()
    _lib.check(lib.lgx_RadTiles(dev, out), lib)
    return int(out[0]), int(out[1])


def form_rooflines(forms, stored, Np, R=192):
    """Rooflines of the tile forms per bounce: the FP32 tiles against the
    FP32 matrix-core peak (flops = stored tiles x 256 x 32 x R x 2); the
    16-bit tiles against HBM (bytes = the matrix planes' stored tiles x 16 KB
    + the operand planes R x Np x 2 B + the FP32 result R x Np x 4 B)."""
    for key, v in forms.items():
        sec = v["s_per_bounce"]
        if key == "tiles_f32_mfma":
            fl = stored * 256 * 32 * R * 2
            v["roofline"] = dict(bound="mfma", achieved=round(fl / sec / 1e12, 2), peak=FP32_MATRIX_PEAK_TFLOPS,
                                 unit="TFLOP/s", frac=round(fl / sec / 1e12 / FP32_MATRIX_PEAK_TFLOPS, 4))
        elif key in ("tiles_bf16_mfma", "tiles_f16_mfma", "tiles_f16x3_mfma", "tiles_bf16x3_mfma"):
            planes = 2 if "x3" in key else 1
            by = planes * (stored * 16384 + R * Np * 2) + R * Np * 4
            v["roofline"] = dict(bound="hbm", achieved=round(by / sec / 1e9, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                                 frac=round(by / sec / 1e9 / HBM_PEAK_GBS, 4))


def compare_at_1e4(forms):
    """The fastest 16-bit MFMA form at <= 1e-4 relative error against FP64,
    and its speed against every FP32 form and the fastest of them."""
    best = min((v["s_per_bounce"], k) for k, v in forms.items()
               if ("bf16" in k or "f16" in k) and v["rel_l2_err_vs_fp64"] <= 1e-4)
    fp32 = min((forms[k]["s_per_bounce"], k) for k in FP32_FORMS)
    return dict(form=best[1], s_per_bounce=best[0], rel_l2_err_vs_fp64=forms[best[1]]["rel_l2_err_vs_fp64"],
                fastest_fp32_form=fp32[1], speedup_vs_fastest_fp32=round(fp32[0] / best[0], 2),
                speedup_vs=dict((k, round(forms[k]["s_per_bounce"] / best[0], 2)) for k in FP32_FORMS))


def read_links(lib, dev, N, nlinks):
    """The link list of the last lgx_RadLinks: per-row (offset, count), the
    other patch and the factor of each link (rows i < j, as the reference
    stores them)."""
    rowmap = np.zeros(2 * N, np.uint32)
    other = np.zeros(nlinks, np.uint32)
    fac = np.zeros(nlinks, np.float32)
    _lib.check(lib.lgx_RadReadLinks(dev, rowmap.ctypes.data_as(u32p), other.ctypes.data_as(u32p), P(fac)), lib)
    return rowmap, other, fac


def tile_fraction(lib, dev):
    """Fraction of the 256-row x 32-column tiles of the link matrix that the
    16-bit forms store and multiply (those holding a link, Morton order)."""
    out = (C.c_uint64 * 2)()
    _lib.check(lib.lgx_RadTiles(dev, out), lib)
    return out[0] / out[1] if out[1] else 0.0


def room_patches(lib, nth, n_patches, lm0, **room_kw):
    """The room's texels as radiosity patches: grow the atlas until the room
    instance has n_patches samples, keep the first n_patches."""
    lm = lm0
    while True:
        spec = scenes.room_scene(lm=lm, **room_kw)
        b = scenes.BuiltScene(lib, spec, dict(max_num_threads=nth))
        _lib.check(lib.lgx_Prepare(b.scene), lib)
        info = (C.c_int32 * 4)()
        lib.lgx_InstanceInfo(b.scene, 1, info)
        if info[2] >= n_patches or lm >= 2048:
            break
        b.destroy()
        # covered texels grow with the atlas area
        lm = max(lm + 8, int(np.ceil(lm * np.sqrt(n_patches / max(info[2], 1)) * 1.01 / 8)) * 8)
    n_all, nt = info[2], info[3]
    pos = np.zeros((n_all, 3), np.float32)
    nrm = np.zeros((n_all, 3), np.float32)
    rad = np.zeros((n_all, 4), np.float32)
    lib.lgx_GetSamples(b.scene, 1, P(pos), P(nrm), None, P(rad))
    tris = np.zeros((nt, 9), np.float32)
    lib.lgx_GetTriangles(b.scene, 1, P(tris))
    N = min(n_patches, n_all)
    return spec, b, pos[:N].copy(), nrm[:N].copy(), rad[:N].copy(), tris, N, lm


def links_bench(lib, nth, lm=64):
    """Radiosity link generation (Int_RDGenLinks, lighter.cpp:667-765) on the
    room at a smaller atlas: the reference's own pass (serial, as it runs)
    timed beside lgx_RadLinks on the same patches and triangles (U = one
    patch pair examined)."""
    try:
        from oracle.checkers import Reference
        ref = Reference()
    except Exception as e:  # the reference build is optional on a bare install
        log(f"[bench] link baseline skipped: {e}")
        return None
    spec = scenes.room_scene(lm=lm, bounce_count=1, name="C2-links")
    br = scenes.BuiltScene(ref.lib, spec)
    ref.lib.rp_prepare(br.scene, nth)
    t0 = time.perf_counter()
    ref.lib.rp_rad_links(br.scene)
    cpu_s = time.perf_counter() - t0
    geo = ref.rad_geoms(br.scene)
    nl_ref = int(ref.lib.rp_num_links(br.scene))
    tris = np.ascontiguousarray(ref.shadow_tris(br.scene), np.float32)
    br.destroy()
    N = len(geo)
    pairs = N * (N - 1) // 2
    dev = device_create(lib, tris)
    pos = np.ascontiguousarray(geo[:, :3])
    nrm = np.ascontiguousarray(geo[:, 3:])
    nl = C.c_size_t(0)
    tm = (C.c_double * 6)()
    ms = []
    for _ in range(3):
        _lib.check(lib.lgx_RadLinks(dev, P(pos), P(nrm), N, C.byref(nl)), lib)
        lib.lgx_DeviceTimes(dev, tm)
        ms.append(tm[3])
    lib.lgx_DeviceDestroy(dev)
    gpu_ms = float(np.median(ms))
    out = dict(patches=N, pairs=pairs, links=int(nl.value), reference_links=nl_ref, gpu_ms=round(gpu_ms, 3),
               gpu_pairs_per_s=pairs / (gpu_ms / 1e3),
               cpu_baseline=dict(value=pairs / cpu_s, unit="patch pairs/s", cores=1, kind="reference",
                                 sample=f"the reference's Int_RDGenLinks over all {N} patches of the room at a "
                                        f"{lm}^2 atlas ({pairs} pairs), one thread as it runs, {cpu_s:.1f} s"))
    out["speedup_vs_cpu"] = round(out["gpu_pairs_per_s"] / out["cpu_baseline"]["value"], 1)
    log(f"[bench] links: {N} patches, {nl.value} links (reference {nl_ref}); GPU {gpu_ms:.2f} ms, "
        f"reference {cpu_s:.2f} s: {out['speedup_vs_cpu']}x")
    return out


def ao_roofline(lib, dev, ro, S, n, distance, ms, tris):
    """Algorithmic bytes of one AO launch over the timed launch's duration
    ms: reference-shaped (the same rays with the canonical closest-hit
    semantics on a plain binary SAH tree, node visits x 32 B + triangle
    tests x 36 B, lgx_ReferenceCountsAO, which also compares every texel's
    factor with the product's), and on this build's tree (wide-node visits x
    112 B + segment tests x 64 B); both plus 52 B per texel (position,
    normal, spiral offset, lightmap read and write)."""
    col = np.zeros(3, np.float32)
    lib.lgx_SetCounting(dev, 1)
    _lib.check(lib.lgx_AmbientOcclusion(dev, P(ro), n, distance, 1.0, 0.0, 0.0, P(col)), lib)
    cnt = (C.c_uint64 * 6)()
    lib.lgx_ReadCounters(dev, cnt)
    lib.lgx_SetCounting(dev, 0)
    nodes, tests = int(cnt[0]), int(cnt[1])
    own = nodes * NODE_BYTES + tests * RAY_TRI_BYTES + S * 52
    refc = (C.c_uint64 * 5)()
    _lib.check(lib.lgx_ReferenceCountsAO(dev, P(tris), len(tris), P(ro), n, distance, 1.0, 0.0, refc), lib)
    rvis, rtri, rrays, rdiff, rtex = (int(x) for x in refc)
    alg = rvis * 32 + rtri * 36 + S * 52
    achieved = alg / (ms / 1e3) / 1e9
    return dict(bound="cache", achieved=round(achieved, 1), unit="GB/s", hbm_peak=HBM_PEAK_GBS,
                x_hbm_peak=round(achieved / HBM_PEAK_GBS, 4), kernel="k_ao", alg_bytes_per_launch=int(alg),
                reference_counts=dict(node_visits=rvis, triangle_tests=rtri, rays=rrays, texels=rtex,
                                      factors_differing=rdiff),
                this_build=dict(node_visits=nodes, triangle_tests=tests, bytes_per_launch=int(own),
                                achieved_gbs=round(own / (ms / 1e3) / 1e9, 1)),
                note="reference-shaped bytes: node visits x 32 B + triangle tests x 36 B on a plain binary SAH "
                     "tree (16 bins, <= 4 triangles per leaf), closest hit shrinking the segment, + 52 B per "
                     "texel; 2 m rays stay near their origin, so the nodes and triangles they read are served "
                     "by L2 and the Infinity Cache and HBM is not the bound")


def ao_bench(lib, dev, spec, S, nth, with_cpu, tris):
    """C2 ambient occlusion: 64 cosine-weighted rays of 2 m per texel on the
    device's S texels (U2 = one ray), beside the reference's AO stage on all
    texels of the same room on the host threads."""
    rng = np.random.default_rng(0x4C540012)
    ro = rng.uniform(0, 1, S).astype(np.float32)
    col = np.zeros(3, np.float32)
    tm = (C.c_double * 6)()
    ms = []
    for _ in range(3):
        _lib.check(lib.lgx_AmbientOcclusion(dev, P(ro), 64, 2.0, 1.0, 0.0, 0.0, P(col)), lib)
        lib.lgx_DeviceTimes(dev, tm)
        ms.append(tm[2])
    best = float(np.median(ms))
    out = dict(texels=S, rays=S * 64, ms=round(best, 3), rays_per_s=round(S * 64 / (best / 1e3), 1),
               roofline=ao_roofline(lib, dev, ro, S, 64, 2.0, best, tris))
    if with_cpu:
        from oracle.checkers import Reference, reference_available
        if reference_available():
            ref = Reference()
            br = scenes.BuiltScene(ref.lib, spec, dict(max_num_threads=nth))
            ref.lib.rp_prepare(br.scene, nth)
            n_ref = sum(i[2] for i in ref.instances(br.scene))
            t0 = time.perf_counter()
            ref.lib.rp_ao(br.scene)
            dt = time.perf_counter() - t0
            br.destroy()
            out["cpu_baseline"] = dict(value=round(n_ref * 64 / dt, 1), unit="AO rays/s", cores=nth,
                                       kind="reference",
                                       sample=f"the reference's AO stage over all {n_ref} texels of the room, "
                                              f"64 rays of 2 m each, {dt:.2f} s")
            out["speedup_vs_cpu"] = round(out["rays_per_s"] / out["cpu_baseline"]["value"], 1)
    log(f"[bench] C2 AO: {S * 64} rays in {best:.2f} ms ({out['rays_per_s'] / 1e9:.2f} G rays/s)"
        + (f"; CPU {out['cpu_baseline']['value'] / 1e6:.2f} M rays/s" if "cpu_baseline" in out else ""))
    return out


def reference_bounce_seconds(rowmap, other, fac, diff, E, area, N):
    """CPU leg: the reference's Int_RDBounce loop (lighter.cpp:767-797,
    restated in C in the oracle) for one RGB setup on the same links, one
    thread as the reference runs it; seconds per bounce, or None."""
    try:
        from oracle.checkers import Oracle
        orc = Oracle()
        c13 = np.concatenate([diff, E[:, :3], E[:, :3], np.zeros((N, 3)), area[:, None]], 1).astype(np.float32)
        t0 = time.perf_counter()
        orc.rad_bounce_f32(rowmap, other, fac, c13)
        return time.perf_counter() - t0
    except Exception as e:  # the checker is optional on a bare install
        log(f"[bench] reference bounce loop skipped: {e}")
        return None


def radiosity_bench(lib, nth, n_patches=32768, R=192, bounces=8, setups=64, with_cpu=True):
    """C2: radiosity over the room's texels as patches, R right-hand sides
    (setups x RGB) built from the lights' direct contributions."""
    spec, b, pos, nrm, rad, tris, N, _ = room_patches(lib, nth, n_patches, 200, bounce_count=bounces,
                                                      ao_distance=2.0, ao_samples=64, name="C2")
    dev = device_create(lib, tris)
    # direct light of each light on the patches
    _lib.check(lib.lgx_SetSamples(dev, P(pos), P(nrm), N, None), lib)
    lights = make_lights(spec)
    L = len(lights)
    items = np.concatenate([(np.uint32(l) << np.uint32(26)) | np.arange(N, dtype=np.uint32) for l in range(L)])
    off = (np.arange(L + 1) * N).astype(np.uint32)
    _lib.check(lib.lgx_DirectLight(dev, lights, L, items.ctypes.data_as(u32p), len(items),
                                   off.ctypes.data_as(u32p), 0), lib)
    w = np.zeros(len(items), np.float32)
    lib.lgx_ReadWeights(dev, P(w), None, len(items))
    w[w.view(np.uint32) == 0xFFFFFFFF] = 0
    W = w.reshape(L, N).T.astype(np.float64)
    ao = ao_bench(lib, dev, spec, N, nth, with_cpu, tris)
    rng = np.random.default_rng(0x4C540002)
    cols = rng.uniform(0.5, 1.0, (setups, L, 3))
    E = np.zeros((N, 3 * setups))
    for s in range(setups):
        for c in range(3):
            E[:, 3 * s + c] = W @ cols[s, :, c]
    E = np.ascontiguousarray(E[:, :R])
    # links (patch normals normalised again, as Int_RDGenLinks does)
    pn = scenes.normalize_rows_f32(nrm)
    nl = C.c_size_t(0)
    t0 = time.perf_counter()
    lib.lgx_SetCounting(dev, 1)  # keeps the candidate pairs for the reference-shaped pass
    _lib.check(lib.lgx_RadLinks(dev, P(pos), P(pn), N, C.byref(nl)), lib)
    lib.lgx_SetCounting(dev, 0)
    tm = (C.c_double * 6)()
    lib.lgx_DeviceTimes(dev, tm)
    link_ms = tm[3]
    log(f"[bench] C2: {N} patches, {nl.value} links ({link_ms:.1f} ms GPU, {time.perf_counter() - t0:.1f} s wall)")
    link_refc = link_reference_counts(lib, dev, tris)
    links_cmp = links_bench(lib, nth) if with_cpu else None
    alb = spec.albedo
    diff = np.array([alb.get(int(p), (1.0, 1.0, 1.0)) for p in rad[:, 2]], np.float32)
    area = np.ascontiguousarray(rad[:, 3])
    dp = C.POINTER(C.c_double)

    def run(form, refine=0):
        out = np.zeros_like(E)
        t = (C.c_double * 4)()
        _lib.check(lib.lgx_RadSolve(dev, E.ctypes.data_as(dp), R, P(diff), P(area), form, bounces, refine,
                                    out.ctypes.data_as(dp), t), lib)
        return out, list(t)

    ref64, t64 = run(0)
    res = {}
    rowmap, other, fac = read_links(lib, dev, N, nl.value)
    cpu_bounce_s = reference_bounce_seconds(rowmap, other, fac, diff, E, area, N) if with_cpu else None
    nrm64 = np.linalg.norm(ref64)
    for form, refine in FORM_RUNS:
        out, t = (ref64, t64) if form == 0 else run(form, refine)
        key = FORMS[form] + ("_refined" if refine else "")
        per_bounce = t[2] if not refine else (t[1] + t[3]) / bounces
        res[key] = dict(s_per_bounce=per_bounce / 1000.0, rel_l2_err_vs_fp64=float(np.linalg.norm(out - ref64) / nrm64),
                        series_ms=round(t[1], 3), refine_ms=round(t[3], 3), build_ms=round(t[0], 3))
        log(f"[bench] {key}: {per_bounce:.3f} ms/bounce, err {res[key]['rel_l2_err_vs_fp64']:.2e}")
    tiles = tile_fraction(lib, dev)
    Np = max(512, ((N + 255) // 256) * 256)
    form_rooflines(res, tile_counts(lib, dev)[0], Np)
    lib.lgx_DeviceDestroy(dev)
    b.destroy()
    mf = res["tiles_bf16_mfma"]["s_per_bounce"]
    gemm_bytes = tiles * Np * Np * 2 + 2 * 192 * Np * 2
    out = dict(workload=f"C2 room radiosity: {N} patches (texels of the 10x10x4 m room), {R} right-hand sides "
                        f"({setups} seeded lighting setups x RGB), {bounces} bounces", patches=N, rhs=R,
               bounces=bounces, links=int(nl.value), link_density=round(2 * nl.value / (N * (N - 1)), 5),
               link_reference_counts=link_refc,
               stored_tile_fraction=round(tiles, 4), link_generation_s=link_ms / 1000.0, forms=res,
               mfma_at_1e4=compare_at_1e4(res),
               mfma_roofline=dict(bound="hbm", achieved=round(gemm_bytes / mf / 1e9, 1), peak=HBM_PEAK_GBS,
                                  unit="GB/s", frac=round(gemm_bytes / mf / 1e9 / HBM_PEAK_GBS, 4),
                                  note="BF16 matrix (its stored tiles) once per bounce + right-hand sides in and out"),
               ao=ao, links_vs_reference=links_cmp, cpu_port_s_per_bounce_rgb=cpu_bounce_s,
               cpu_note="the reference's Int_RDBounce loop restated in C (oracle), one RGB setup, one host thread, "
                        "on the same link list; the reference itself runs it single-threaded")
    return out


def adjoint_bench(lib, nth, n_patches=65536, bounces=4, ckpt_lm=1024):
    """C3: gradient of an L2 bake loss with respect to 64 light colours (RGB)
    and 32 albedo groups (RGB), 288 parameters, through 4 bounces.

    The room (about 20k triangles, 64 point lights) with its texels as
    patches; the target is baked from parameters perturbed by +-30 %.  One
    gradient evaluation = forward bake (E = sum_l colour_l w_l from the
    checkpointed direct-light weights, then the bounces) + reverse sweep
    (lgx_RadSolveGrad, which recomputes the forward partial sums) + colour
    gradient (lgx_DirectLightGrad).  The dense FP16 MFMA form is checked
    against the sparse FP64 form's gradient.  Second part: the colour
    gradient over the checkpointed visibility of a ckpt_lm^2 atlas (every
    texel x 64 lights)."""
    kw = dict(seed=scenes.SEED_BASE + 3, n_boxes=31, wall_tess=40, n_point=64, with_sun=False, bounce_count=bounces,
              name="C3")
    spec, b, pos, nrm, rad, tris, N, lm = room_patches(lib, nth, n_patches, 256, **kw)
    dev = device_create(lib, tris)
    _lib.check(lib.lgx_SetSamples(dev, P(pos), P(nrm), N, None), lib)
    lights = make_lights(spec)
    L = len(lights)
    items = np.concatenate([(np.uint32(l) << np.uint32(26)) | np.arange(N, dtype=np.uint32) for l in range(L)])
    off = (np.arange(L + 1) * N).astype(np.uint32)
    _lib.check(lib.lgx_DirectLight(dev, lights, L, items.ctypes.data_as(u32p), len(items),
                                   off.ctypes.data_as(u32p), 0), lib)
    tm = (C.c_double * 6)()
    lib.lgx_DeviceTimes(dev, tm)
    direct_ms = tm[0]
    w = np.zeros(len(items), np.float32)
    lib.lgx_ReadWeights(dev, P(w), None, len(items))
    w[w.view(np.uint32) == 0xFFFFFFFF] = 0
    W = w.reshape(L, N).T.astype(np.float64)  # N x L
    pn = scenes.normalize_rows_f32(nrm)
    nl = C.c_size_t(0)
    _lib.check(lib.lgx_RadLinks(dev, P(pos), P(pn), N, C.byref(nl)), lib)
    lib.lgx_DeviceTimes(dev, tm)
    link_ms = tm[3]
    area = np.ascontiguousarray(rad[:, 3])
    groups = rad[:, 2].astype(np.int64) % 32
    col0 = np.array([[lights[l].color[c] for c in range(3)] for l in range(L)], np.float64)
    alb0 = np.array([spec.albedo.get(g, (0.8, 0.8, 0.8)) for g in range(32)], np.float64)
    rng = np.random.default_rng(0x4C540003)
    col_t = col0 * rng.uniform(0.7, 1.3, col0.shape)
    alb_t = np.clip(alb0 * rng.uniform(0.7, 1.3, alb0.shape), 0.0, 0.95)
    dp = C.POINTER(C.c_double)

    def forward(form, col, alb):
        E = np.ascontiguousarray(W @ col)
        diff = np.ascontiguousarray(alb[groups], np.float32)
        T = np.zeros_like(E)
        t = (C.c_double * 4)()
        _lib.check(lib.lgx_RadSolve(dev, E.ctypes.data_as(dp), 3, P(diff), P(area), form, bounces, 0,
                                    T.ctypes.data_as(dp), t), lib)
        return E, diff, T, list(t)

    _, _, target, _ = forward(4, col_t, alb_t)

    def gradient(form):
        t0 = time.perf_counter()
        E, diff, T, tf = forward(form, col0, alb0)
        g = np.ascontiguousarray(T - target)
        tot, lam, dalb = np.zeros_like(E), np.zeros_like(E), np.zeros_like(E)
        t = (C.c_double * 4)()
        _lib.check(lib.lgx_RadSolveGrad(dev, E.ctypes.data_as(dp), 3, P(diff), P(area), form, bounces,
                                        g.ctypes.data_as(dp), tot.ctypes.data_as(dp), lam.ctypes.data_as(dp),
                                        dalb.ctypes.data_as(dp), t), lib)
        dcol = np.zeros((L, 3))
        _lib.check(lib.lgx_DirectLightGrad(dev, np.ascontiguousarray(lam).ctypes.data_as(dp),
                                           dcol.ctypes.data_as(dp)), lib)
        lib.lgx_DeviceTimes(dev, tm)
        dgrp = np.zeros((32, 3))
        np.add.at(dgrp, groups, dalb)
        wall = time.perf_counter() - t0
        loss = 0.5 * float((g * g).sum())
        return dict(loss=loss, dcol=dcol, dalb=dgrp, wall_s=wall, fwd_series_ms=tf[1], rev_fwd_ms=t[1],
                    rev_sweep_ms=t[3], color_grad_ms=tm[5], build_ms=tf[0])

    g4 = gradient(4)   # builds the FP16 matrix
    g4 = gradient(4)   # timed with the matrix resident
    g6 = gradient(6)   # the compensated FP16 form (same matrix store)
    g0 = gradient(0)
    rel = lambda a, b_: float(np.linalg.norm(a - b_) / max(np.linalg.norm(b_), 1e-300))
    gpu_ms = g4["fwd_series_ms"] + g4["rev_fwd_ms"] + g4["rev_sweep_ms"] + g4["color_grad_ms"]
    log(f"[bench] C3: {N} patches, {len(tris)} tris, {L} lights, {nl.value} links; gradient (FP16 MFMA) "
        f"{gpu_ms:.2f} ms GPU / {g4['wall_s'] * 1e3:.1f} ms wall; vs FP64: colours {rel(g4['dcol'], g0['dcol']):.2e}, "
        f"albedo {rel(g4['dalb'], g0['dalb']):.2e}")
    lib.lgx_DeviceDestroy(dev)
    b.destroy()
    out = dict(workload=f"C3 room gradient: {len(tris)} tris, {N} patches (texels, {lm}^2 atlas), {L} point "
                        f"lights, 288 parameters (64 light RGB + 32 albedo groups RGB), L2 loss against a bake "
                        f"from +-30 % perturbed parameters, {bounces} bounces",
               patches=N, triangles=len(tris), links=int(nl.value), direct_light_ms=round(direct_ms, 3),
               link_generation_ms=round(link_ms, 3),
               gradient_ms_gpu=round(gpu_ms, 3), gradient_ms_wall=round(g4["wall_s"] * 1e3, 2),
               breakdown_ms=dict(forward_series=round(g4["fwd_series_ms"], 3),
                                 reverse_recompute=round(g4["rev_fwd_ms"], 3),
                                 reverse_sweep=round(g4["rev_sweep_ms"], 3),
                                 colour_gradient=round(g4["color_grad_ms"], 3)),
               form="tiles FP16 MFMA (FP32 accumulation)",
               rel_err_vs_fp64=dict(colours=rel(g4["dcol"], g0["dcol"]), albedo=rel(g4["dalb"], g0["dalb"])),
               f16x3=dict(form="tiles FP16 x 3 (error-compensated product every bounce, forward and reverse)",
                          gradient_ms_gpu=round(g6["fwd_series_ms"] + g6["rev_fwd_ms"] + g6["rev_sweep_ms"]
                                                + g6["color_grad_ms"], 3),
                          rel_err_vs_fp64=dict(colours=rel(g6["dcol"], g0["dcol"]),
                                               albedo=rel(g6["dalb"], g0["dalb"]))),
               fp64_gradient_ms_gpu=round(g0["fwd_series_ms"] + g0["rev_fwd_ms"] + g0["rev_sweep_ms"]
                                          + g0["color_grad_ms"], 3))
    # colour gradient over the checkpointed visibility of a ckpt_lm^2 atlas
    if ckpt_lm:
        spec2 = scenes.room_scene(lm=ckpt_lm, **kw)
        b2, pos2, nrm2, tris2, items2, off2 = prepare_workload(lib, spec2, nth)
        S = len(pos2)
        dev2 = device_create(lib, tris2)
        _lib.check(lib.lgx_SetSamples(dev2, P(pos2), P(nrm2), S, None), lib)
        lights2 = make_lights(spec2)
        _lib.check(lib.lgx_DirectLight(dev2, lights2, len(lights2), items2.ctypes.data_as(u32p), len(items2),
                                       off2.ctypes.data_as(u32p), 0), lib)
        lib.lgx_DeviceTimes(dev2, tm)
        d2_ms = tm[0]
        lam2 = np.ascontiguousarray(rng.normal(size=(S, 3)))
        dcol2 = np.zeros((len(lights2), 3))
        ms = []
        for _ in range(3):
            _lib.check(lib.lgx_DirectLightGrad(dev2, lam2.ctypes.data_as(dp), dcol2.ctypes.data_as(dp)), lib)
            lib.lgx_DeviceTimes(dev2, tm)
            ms.append(tm[5])
        gms = float(min(ms))
        ck_bytes = len(items2) * 4
        stream = len(items2) * 8 + S * 24  # weights + items streamed, lam gathered once per item in cache
        out["checkpoint"] = dict(atlas=f"{ckpt_lm}^2", samples=S, lights=len(lights2), items=int(len(items2)),
                                 visibility_bytes=ck_bytes, direct_light_ms=round(d2_ms, 3),
                                 colour_gradient_ms=round(gms, 4),
                                 achieved_gbs=round(stream / (gms / 1e3) / 1e9, 1))
        log(f"[bench] C3 checkpoint {ckpt_lm}^2: {S} samples x {len(lights2)} lights, {ck_bytes / 1e6:.0f} MB "
            f"weights; direct {d2_ms:.1f} ms, colour gradient {gms:.3f} ms")
        lib.lgx_DeviceDestroy(dev2)
        b2.destroy()
    return out


class DirectLeg:
    """One direct-light workload on the device: the host stages through the
    product, the device scene, the work items, W untimed and K timed passes
    of the direct-light stage, the roofline of the march kernel and the
    reference's shadow march on a sample of the same marches."""

    def __init__(self, lib, spec, nth, radinfo=False):
        self.lib, self.spec, self.nth = lib, spec, nth
        t0 = time.perf_counter()
        self.radinfo = [] if radinfo else None
        (self.built, self.pos, self.nrm, self.tris, self.items,
         self.light_off) = prepare_workload(lib, spec, nth, self.radinfo)
        if self.radinfo is not None:
            self.radinfo = np.concatenate(self.radinfo)
        self.host_s = time.perf_counter() - t0
        self.S = len(self.pos)
        log(f"[bench] {spec.name} host stages: {self.S} samples, {len(self.items)} work items ({self.host_s:.1f} s)")
        t0 = time.perf_counter()
        self.dev = device_create(lib, self.tris)
        if not self.dev:
            raise RuntimeError(_lib.last_error(lib))
        self.device_create_s = time.perf_counter() - t0
        info = (C.c_int64 * 3)()
        lib.lgx_DeviceInfo(self.dev, info)
        self.ntri, self.nodes_bvh, self.depth = int(info[0]), int(info[1]), int(info[2])
        cells = C.c_int64(0)
        _lib.check(lib.lgx_SetEntryGrid(self.dev, 1 if ENTRY_GRID else 0, C.byref(cells)), lib)
        self.grid_cells = int(cells.value) if ENTRY_GRID else 0
        log(f"[bench] {spec.name} device scene (BVH build, triangle forms, entry grid, upload): "
            f"{self.device_create_s:.2f} s; {self.ntri} tris, {self.nodes_bvh} nodes, depth {self.depth}, "
            f"{self.grid_cells} entry-grid cells")
        _lib.check(lib.lgx_SetSamples(self.dev, P(self.pos), P(self.nrm), self.S, None), lib)
        self.lights = make_lights(spec)
        self.setup(keep_vis=1)

    def setup(self, keep_vis):
        _lib.check(self.lib.lgx_DirectLightSetup(self.dev, self.lights, len(self.lights),
                                                 self.items.ctypes.data_as(u32p), len(self.items),
                                                 self.light_off.ctypes.data_as(u32p), keep_vis), self.lib)

    def run(self, steps, warmup):
        """Counters from one instrumented pass, then warmup + timed passes:
        wall time around the K passes, HIP-event time of each march launch."""
        lib = self.lib
        lib.lgx_SetCounting(self.dev, 1)
        _lib.check(lib.lgx_DirectLightRun(self.dev), lib)
        cnt = (C.c_uint64 * 6)()
        lib.lgx_ReadCounters(self.dev, cnt)
        lib.lgx_SetCounting(self.dev, 0)
        self.nodes, self.ntests, self.qsteps, self.queries, self.marches, self.npre = (int(x) for x in cnt)
        tm = (C.c_double * 6)()
        for _ in range(warmup):
            _lib.check(lib.lgx_DirectLightRun(self.dev), lib)
        march_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            _lib.check(lib.lgx_DirectLightRun(self.dev), lib)
            lib.lgx_DeviceTimes(self.dev, tm)
            march_ms.append(tm[1])
        wall = time.perf_counter() - t0
        self.steps, self.warmup = steps, warmup
        self.ms_per_step = wall * 1000.0 / steps
        self.march_avg = float(np.mean(march_ms))
        self.value = self.marches / (self.ms_per_step / 1000.0)
        log(f"[bench] {self.spec.name} direct light: {steps} steps, {self.ms_per_step:.2f} ms/step, march kernel "
            f"{self.march_avg:.2f} ms; {self.marches} marches, {self.qsteps} distance queries: "
            f"{self.value / 1e6:.1f} M shadow rays/s")
        self.w = np.zeros(len(self.items), np.float32)
        self.vis = np.zeros(len(self.items), np.float32)
        _lib.check(lib.lgx_ReadWeights(self.dev, P(self.w), P(self.vis), len(self.items)), lib)

    def roofline(self, workload_key):
        """Algorithmic bytes of one march launch: the reference-shaped count
        (the same marches on a plain binary SAH tree, 16 bins, <= 4
        triangles per leaf, a radius-2 query per step: node visits x 32 B +
        triangle tests x 36 B) plus the per-item I/O, over the average
        HIP-event duration of the timed launches."""
        lib = self.lib
        io_bytes = len(self.items) * (4 + 24 + 4 + 4)  # item, sample pos+nrm, weight, shadow factor
        refc = (C.c_uint64 * 5)()
        _lib.check(lib.lgx_ReferenceCounts(self.dev, P(self.tris), len(self.tris), refc), lib)
        rvis, rtri, rq, rbad, rm = (int(x) for x in refc)
        alg = rvis * 32 + rtri * 36 + io_bytes
        achieved = alg / (self.march_avg / 1000.0) / 1e9
        own = self.nodes * NODE_BYTES + self.ntests * DIST_TRI_BYTES + self.npre * TRI_BOX_BYTES + io_bytes
        log(f"[bench] {self.spec.name} reference-shaped counts: {rvis} node visits, {rtri} triangle tests, {rq} "
            f"queries over {rm} marches ({rbad} shadow factors differ from the product's): "
            f"{achieved:.0f} GB/s, {achieved / HBM_PEAK_GBS:.3f} of HBM peak")
        self.factors_differing = rbad
        return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=measured_traffic("k_march", dict(workload=workload_key, triangles=self.ntri,
                                                             work_items=int(len(self.items)),
                                                             marches=int(self.marches))),
                    kernel="k_march", alg_bytes_per_launch=int(alg), kernel_ms=round(self.march_avg, 3),
                    note="reference-shaped count: the same marches on a plain binary SAH tree (16 bins, <= 4 "
                         "triangles per leaf) with one radius-2 distance query per step, node visits x 32 B + "
                         "triangle tests x 36 B, plus 36 B of I/O per item, over the march kernel's average "
                         "HIP-event duration; traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE",
                    reference_counts=dict(node_visits=rvis, triangle_tests=rtri, distance_queries=rq, marches=rm,
                                          factors_differing=rbad),
                    this_build=dict(node_visits=self.nodes, triangle_tests=self.ntests, box_pretests=self.npre,
                                    distance_queries=self.qsteps, bytes_per_launch=int(own),
                                    achieved_gbs=round(own / (self.march_avg / 1000.0) / 1e9, 1),
                                    note="4-wide nodes x 112 B + full triangle tests x 160 B + box pre-tests x "
                                         "48 B + I/O on this build's own tree and query bounds"))

    def cpu(self, seconds, cores_info, ao=None):
        """CPU leg: the reference's CalcInvShadowFactor on a seeded sample of
        the marched pairs (all of them when they fit in `seconds`), and the
        bit-level agreement of the sampled shadow factors."""
        cpu, samp = cpu_baseline(self.spec, self.pos, self.nrm, self.items, self.w, self.lights, seconds,
                                 self.nth, ao=ao)
        parity = None
        if cpu is not None:
            cpu.update(cpu_model=cores_info.get("cpu_model"), affinity_cpus=cores_info.get("affinity_cpus"))
            sel, refvis = samp
            parity = dict(sampled=int(len(sel)),
                          bitexact_fraction=round(float(np.mean(self.vis[sel].view(np.uint32)
                                                                == refvis.view(np.uint32))), 6),
                          shadowed_agreement=round(float(np.mean((self.vis[sel] == 0) == (refvis == 0))), 6))
            log(f"[bench] {self.spec.name} reference march {cpu['value']:.0f} rays/s on {cpu['cores']} threads; "
                f"GPU/CPU {self.value / cpu['value']:.0f}x; sampled parity {parity}")
        return cpu, parity

    def close(self):
        if self.dev:
            self.lib.lgx_DeviceDestroy(self.dev)
            self.dev = None


def c4_extras(lib, leg, nth, cpu_seconds, with_cpu, n_patches=65536, bounces=4):
    """The rest of C4 on its scene: 64 AO rays of 2 m per texel, the colour
    gradient over every work item, 65,536-patch radiosity in every form,
    the lightmap export and the sample placement beside their host forms."""
    S, dev, tris, items = leg.S, leg.dev, leg.tris, leg.items
    out = {}
    # the device scene's BVH: the GPU build against the host build
    bv = (C.c_double * 6)()
    _lib.check(lib.lgx_BvhCompare(P(tris), len(tris), bv), lib)
    out["bvh_build"] = dict(triangles=int(len(tris)), wide_nodes=int(bv[5]), host_ms=round(bv[2], 1),
                            gpu_ms=round(bv[3], 1), gpu_device_ms=round(bv[4], 1),
                            identical_tree=bool(bv[0] == 1 and bv[1] == 1))
    log(f"[bench] C4 BVH build: GPU {bv[3]:.1f} ms ({bv[4]:.1f} ms device, collapse on the host) vs host "
        f"{bv[2]:.1f} ms; identical tree: {out['bvh_build']['identical_tree']}")
    rng = np.random.default_rng(0x4C540004)
    ro = rng.uniform(0, 1, S).astype(np.float32)
    col = np.zeros(3, np.float32)
    tm = (C.c_double * 6)()
    _lib.check(lib.lgx_AmbientOcclusion(dev, P(ro), 64, 2.0, 1.0, 0.0, 0.0, P(col)), lib)
    _lib.check(lib.lgx_AmbientOcclusion(dev, P(ro), 64, 2.0, 1.0, 0.0, 0.0, P(col)), lib)
    lib.lgx_DeviceTimes(dev, tm)
    out["ao"] = dict(rays=S * 64, ms=round(tm[2], 3), rays_per_s=round(S * 64 / (tm[2] / 1e3), 1),
                     roofline=ao_roofline(lib, dev, ro, S, 64, 2.0, tm[2], tris))
    rc = out["ao"]["roofline"]["reference_counts"]
    log(f"[bench] C4 AO: {S * 64} rays in {tm[2]:.1f} ms ({out['ao']['rays_per_s'] / 1e9:.2f} G rays/s); "
        f"{rc['factors_differing']} of {rc['texels']} factors differ from the canonical traversal's")
    # colour gradient of L = 1/2 |direct - target|^2 over every texel (64 lights)
    lm = np.zeros((S, 3), np.float32)
    lib.lgx_ReadLightmap(dev, P(lm))
    lam = np.ascontiguousarray(lm.astype(np.float64) * rng.uniform(-0.3, 0.3, (S, 1)))
    dcol = np.zeros((len(leg.lights), 3))
    dp = C.POINTER(C.c_double)
    leg.setup(keep_vis=0)
    _lib.check(lib.lgx_DirectLightRun(dev), lib)
    _lib.check(lib.lgx_DirectLightGrad(dev, lam.ctypes.data_as(dp), dcol.ctypes.data_as(dp)), lib)
    lib.lgx_DeviceTimes(dev, tm)
    out["colour_gradient"] = dict(ms=round(tm[5], 4), items=int(len(items)), weights_bytes=int(len(items) * 4))
    log(f"[bench] C4 colour gradient over {len(items)} items: {tm[5]:.3f} ms")
    # radiosity on the aligned window of n_patches consecutive texels that
    # receives the most direct light, 192 right-hand sides from the lights'
    # weights
    N = min(n_patches, S)
    lid = (items >> 26).astype(np.int64)
    smp = (items & ((1 << 26) - 1)).astype(np.int64)
    wv = np.where(leg.w.view(np.uint32) == 0xFFFFFFFF, 0.0, leg.w.astype(np.float64))
    lit = np.bincount(smp, weights=wv, minlength=S)
    starts = np.arange(0, S - N + 1, N)
    s0 = int(starts[np.argmax([lit[a:a + N].sum() for a in starts])])
    pn = scenes.normalize_rows_f32(leg.nrm[s0:s0 + N])
    ppos = np.ascontiguousarray(leg.pos[s0:s0 + N])
    rad_dev = device_create(lib, tris)
    _lib.check(lib.lgx_SetSamples(rad_dev, P(ppos), P(np.ascontiguousarray(leg.nrm[s0:s0 + N])), N, None), lib)
    nl = C.c_size_t(0)
    lib.lgx_SetCounting(rad_dev, 1)  # keeps the candidate pairs for the reference-shaped pass
    _lib.check(lib.lgx_RadLinks(rad_dev, P(ppos), P(pn), N, C.byref(nl)), lib)
    lib.lgx_SetCounting(rad_dev, 0)
    lib.lgx_DeviceTimes(rad_dev, tm)
    link_ms = tm[3]
    link_refc = link_reference_counts(lib, rad_dev, tris)
    sel = (smp >= s0) & (smp < s0 + N)
    Wp = np.zeros((N, len(leg.lights)))
    np.add.at(Wp, (smp[sel] - s0, lid[sel]), wv[sel])
    cols = rng.uniform(0.5, 1.0, (64, len(leg.lights), 3))
    # the lights' direct contributions plus a small uniform emission, so that
    # no column is empty where the lights do not reach a patch
    E = np.ascontiguousarray(np.concatenate([Wp @ cols[k] for k in range(64)], 1) + 0.05)
    diff = np.full((N, 3), 0.6, np.float32)
    area = np.ascontiguousarray(leg.radinfo[s0:s0 + N, 3])  # world area per texel, as the reference's patches carry
    R = E.shape[1]

    def run(form, refine=0):
        o = np.zeros_like(E)
        t = (C.c_double * 4)()
        _lib.check(lib.lgx_RadSolve(rad_dev, E.ctypes.data_as(dp), R, P(diff), P(area), form, bounces, refine,
                                    o.ctypes.data_as(dp), t), lib)
        return o, list(t)

    ref64, t64 = run(0)
    forms = {}
    for form, refine in FORM_RUNS:
        o, t = (ref64, t64) if form == 0 else run(form, refine)
        key = FORMS[form] + ("_refined" if refine else "")
        per = t[2] if not refine else (t[1] + t[3]) / bounces
        forms[key] = dict(s_per_bounce=per / 1e3,
                          rel_l2_err_vs_fp64=float(np.linalg.norm(o - ref64) / np.linalg.norm(ref64)))
        log(f"[bench] C4 radiosity {key}: {per:.3f} ms/bounce, err {forms[key]['rel_l2_err_vs_fp64']:.2e}")
    tiles = tile_fraction(lib, rad_dev)
    form_rooflines(forms, tile_counts(lib, rad_dev)[0], max(512, ((N + 255) // 256) * 256))
    lit_share = float((Wp @ cols[0]).sum() / E[:, :3].sum())  # direct light's share of E
    rad = dict(patches=N, first_patch=s0, direct_light_share_of_emission=round(lit_share, 4), rhs=R, bounces=bounces, links=int(nl.value), link_generation_ms=round(link_ms, 1),
               link_reference_counts=link_refc, stored_tile_fraction=round(tiles, 4), forms=forms)
    cmp = compare_at_1e4(forms)
    rad["mfma_at_1e4"] = cmp
    best = (cmp["s_per_bounce"], cmp["form"])
    log(f"[bench] C4 radiosity on texels {s0} .. {s0 + N - 1} (direct light {lit_share:.1%} of the emission): "
        f"{tiles:.4f} of the 256 x 32 tiles stored; fastest 16-bit form at <= 1e-4: "
        f"{cmp['form']} {cmp['s_per_bounce'] * 1e3:.3f} ms/bounce, {cmp['speedup_vs_fastest_fp32']}x the fastest "
        f"FP32 form ({cmp['fastest_fp32_form']})")
    if with_cpu:
        rowmap, other, fac = read_links(lib, rad_dev, N, nl.value)
        cpu_b = reference_bounce_seconds(rowmap, other, fac, diff, E, area, N)
        if cpu_b:
            sets = R / 3.0
            rad["cpu_port_s_per_bounce_rgb"] = cpu_b
            rad["mfma_vs_cpu_bounce"] = dict(
                note=f"GPU seconds per bounce for all {R} right-hand sides ({sets:.0f} RGB setups) against the "
                     "reference loop's seconds per bounce for one RGB setup times the number of setups",
                speedup=round(cpu_b * sets / best[0], 1))
    out["radiosity"] = rad
    out["adjoint"] = c4_adjoint(lib, leg, rad_dev, s0, N, Wp, diff, area, bounces, rng)
    lib.lgx_DeviceDestroy(rad_dev)
    out["radiosity_whole_scene"] = whole_scene_radiosity(lib, leg, tris, rng)
    out["bake_end_to_end"] = whole_bake(lib, leg.spec, nth)
    if with_cpu:
        # the reference's AO ray loop on a sample of texels, and the device's
        # AO over a lightmap of ones (multiplier 1, no falloff, effect 0,
        # black), which leaves 1 - factor in every channel
        ao_ref = {}
        from oracle.checkers import Reference, reference_available
        if reference_available():
            ref = Reference()
            br = scenes.BuiltScene(ref.lib, leg.spec, dict(max_num_threads=nth))
            ref.lib.rp_prepare_collision(br.scene, nth)
            ao_baseline(ref, br, leg.pos, leg.nrm, nth, ro, 64, 2.0, cpu_seconds, ao_ref)
            br.destroy()
        if ao_ref:
            _lib.check(lib.lgx_SetSamples(dev, P(leg.pos), P(leg.nrm), S, P(np.ones((S, 3), np.float32))), lib)
            _lib.check(lib.lgx_AmbientOcclusion(dev, P(ro), 64, 2.0, 1.0, 0.0, 0.0, P(col)), lib)
            lm1 = np.zeros((S, 3), np.float32)
            lib.lgx_ReadLightmap(dev, P(lm1))
            got = lm1[ao_ref["texels"], 0]
            exp = np.float32(1) - ao_ref["factor"]
            out["ao"]["cpu_baseline"] = ao_ref["baseline"]
            out["ao"]["speedup_vs_cpu"] = round(out["ao"]["rays_per_s"] / ao_ref["baseline"]["value"], 1)
            out["ao"]["parity_sample"] = dict(texels=int(len(got)),
                                              bitexact_fraction=float(np.mean(got.view(np.uint32)
                                                                              == exp.view(np.uint32))),
                                              within_1_255=float(np.mean(np.abs(got - exp) <= 1.0 / 255)),
                                              max_abs_diff=float(np.max(np.abs(got - exp))))
            log(f"[bench] C4 AO reference {ao_ref['baseline']['value'] / 1e6:.2f} M rays/s; GPU/CPU "
                f"{out['ao']['speedup_vs_cpu']}x; parity {out['ao']['parity_sample']}")
    out["lightmap_export"] = export_vs_host(lib, leg.built)
    out["sample_placement"] = placement_vs_host(lib, leg.spec, nth, leg.built, leg.pos)
    return out


def whole_scene_radiosity(lib, leg, tris, rng):
    """The bake's radiosity stage at C4's full size: links among all
    11,917,038 texels (Int_RDGenLinks' pairs, found on the uniform grid, the
    CSR with 64-bit row offsets) and one bounce over them for one RGB setup
    (diffuse 0.6, world area per texel, seeded emission)."""
    S = len(leg.pos)
    d = device_create(lib, tris)
    try:
        pn = scenes.normalize_rows_f32(leg.nrm)
        nl = C.c_size_t(0)
        _lib.check(lib.lgx_RadLinks(d, P(leg.pos), P(pn), S, C.byref(nl)), lib)
        tm = (C.c_double * 6)()
        lib.lgx_DeviceTimes(d, tm)
        link_ms = tm[3]
        diff = np.full((S, 3), 0.6, np.float32)
        emit = np.ascontiguousarray(rng.uniform(0, 1, (S, 3)).astype(np.float32))
        area = np.ascontiguousarray(leg.radinfo[:, 3])
        total = np.zeros((S, 3), np.float32)
        _lib.check(lib.lgx_RadBounce(d, P(diff), P(emit), P(area), 1, P(total)), lib)
        lib.lgx_DeviceTimes(d, tm)
        entries = 2 * int(nl.value)
        res = dict(patches=int(S), links=int(nl.value), csr_entries=entries, link_search_ms=round(link_ms, 1),
                   bounce_ms=round(tm[4], 2), csr_gb=round(entries * 8 / 1e9, 1),
                   bounce_csr_tb_per_s=round(entries * 8 / (tm[4] / 1e3) / 1e12, 2),
                   total_finite=bool(np.isfinite(total).all()))
        log(f"[bench] C4 whole-scene radiosity: {nl.value} links among {S} texels in {link_ms:.0f} ms, "
            f"one bounce {tm[4]:.1f} ms ({res['bounce_csr_tb_per_s']} TB/s of CSR)")
        return res
    finally:
        lib.lgx_DeviceDestroy(d)


def whole_bake(lib, spec, nth):
    """The C4 scene baked through the drop-in API (ltr_Start, polled with
    ltr_GetStatus): preparation, direct light, one radiosity bounce over
    all texels, 64 AO rays of 2 m per texel and the export; wall seconds per
    stage as the status reports them."""
    b = scenes.BuiltScene(lib, spec, dict(max_num_threads=nth, bounce_count=1, ao_num_samples=64,
                                          ao_distance=2.0))
    try:
        t0 = time.perf_counter()
        stages = b.bake(poll_ms=2)
        wall = time.perf_counter() - t0
        outs = b.outputs()
        res = dict(config="bounce_count 1, ao_num_samples 64, ao_distance 2 m", wall_s=round(wall, 2),
                   stages_s={k: round(v, 3) for k, v in stages.items()}, outputs=len(outs),
                   outputs_finite=bool(all(np.isfinite(o[3]).all() for o in outs)))
        log(f"[bench] C4 bake through ltr_Start (one bounce, AO 64 rays): {wall:.2f} s; {res['stages_s']}")
        return res
    finally:
        b.destroy()


def c4_adjoint(lib, leg, rad_dev, s0, N, Wp, diff, area, bounces, rng):
    """Gradient of L = 1/2 |T - target|^2 over the radiosity tile (texels
    s0 .. s0 + N - 1, 4 bounces) with respect to all 64 light colours (RGB): E =
    Wp col from the checkpointed direct-light weights, T = sum_k K^k E;
    lgx_RadSolveGrad gives lam = dL/dE, lgx_DirectLightGrad the colours'
    gradient from lam over every work item (lam is zero beyond the tile).
    The FP16 tile form and the FP16 x 3 form against the sparse FP64 form's
    gradient."""
    dp = C.POINTER(C.c_double)
    L = len(leg.lights)
    col0 = np.array([[leg.lights[l].color[c] for c in range(3)] for l in range(L)], np.float64)
    # the target is the bake at 40-60 % of the colours: the residual keeps one
    # sign, so the gradient's sums over texels do not cancel.  E is the direct
    # light alone: with a uniform floor in E the residual would be the small
    # difference of two large totals, and every form's gradient would measure
    # the rounding of that floor's bounces instead of the light's.
    target_col = col0 * rng.uniform(0.4, 0.6, col0.shape)
    tm = (C.c_double * 6)()

    def solve(form, col):
        E = np.ascontiguousarray(Wp @ col)
        T = np.zeros_like(E)
        t = (C.c_double * 4)()
        _lib.check(lib.lgx_RadSolve(rad_dev, E.ctypes.data_as(dp), 3, P(diff), P(area), form, bounces, 0,
                                    T.ctypes.data_as(dp), t), lib)
        return E, T

    _, target = solve(0, target_col)
    out = {}
    for form in (4, 4, 6, 0):  # form 4 twice: the first call builds its matrix store
        t0 = time.perf_counter()
        E, T = solve(form, col0)
        g = np.ascontiguousarray(T - target)
        tot, lam, dalb = np.zeros_like(E), np.zeros_like(E), np.zeros_like(E)
        t = (C.c_double * 4)()
        _lib.check(lib.lgx_RadSolveGrad(rad_dev, E.ctypes.data_as(dp), 3, P(diff), P(area), form, bounces,
                                        g.ctypes.data_as(dp), tot.ctypes.data_as(dp), lam.ctypes.data_as(dp),
                                        dalb.ctypes.data_as(dp), t), lib)
        lam_all = np.zeros((leg.S, 3))
        lam_all[s0:s0 + N] = lam
        dcol = np.zeros((L, 3))
        _lib.check(lib.lgx_DirectLightGrad(leg.dev, lam_all.ctypes.data_as(dp), dcol.ctypes.data_as(dp)), lib)
        lib.lgx_DeviceTimes(leg.dev, tm)
        out[form] = dict(dcol=dcol, wall_ms=(time.perf_counter() - t0) * 1e3, forward_ms=t[1],
                         reverse_ms=t[3], colour_grad_ms=tm[5])
    g4, g6, g0 = out[4], out[6], out[0]
    rel = lambda g: float(np.linalg.norm(g["dcol"] - g0["dcol"]) / np.linalg.norm(g0["dcol"]))
    ms = lambda g: round(g["forward_ms"] + g["reverse_ms"] + g["colour_grad_ms"], 3)
    res = dict(parameters=f"{L} light colours x RGB", patches=N, bounces=bounces, form="tiles_f16_mfma",
               gradient_ms_gpu=ms(g4),
               breakdown_ms=dict(forward_series=round(g4["forward_ms"], 3),
                                 reverse_sweep=round(g4["reverse_ms"], 3),
                                 colour_gradient=round(g4["colour_grad_ms"], 3)),
               gradient_ms_wall=round(g4["wall_ms"], 1), rel_err_vs_fp64=rel(g4),
               f16x3=dict(form="tiles_f16x3_mfma", gradient_ms_gpu=ms(g6), rel_err_vs_fp64=rel(g6)),
               fp64_gradient_ms_gpu=ms(g0))
    log(f"[bench] C4 gradient w.r.t. {L} light colours through the radiosity tile: {res['gradient_ms_gpu']} ms "
        f"(FP16 tiles, largest factors in FP32), rel err vs FP64 {rel(g4):.2e}; FP16 x 3 {ms(g6)} ms, "
        f"{rel(g6):.2e}")
    return res


C4_WORKLOAD = ("C4 city block direct light: 1 km heightfield (131,072 tris) + 155,000 lognormal(1, 0.6) m "
               "boxes (1,860,000 tris) = 1,991,072 tris in 16 x 16 tile instances, 4096^2 texels in total "
               "(256^2 per tile), 64 point/spot lights with 16 soft-shadow samples each")
C1_WORKLOAD = ("C1 terrain direct light: 512x512 heightfield + 2000 boxes (548,288 tris) in 16x16 tiles, "
               "2048x2048 texels, 32 point/spot + sun (range 300 m)")


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE x fetch factor + WRITE_SIZE x write factor, KiB per
    dispatch), when it was measured on this same workload; else None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        for k in t[kernel]:
            if k["workload"] == workload:
                return int(round((k["fetch_kib"] * t["fetch_factor"] + k["write_kib"] * t["write_factor"]) * 1024))
    except (OSError, KeyError, ValueError, TypeError):
        pass
    return None


ENTRY_GRID = True  # --no-entry-grid: every traversal starts at the root (for comparison)


def device_create(lib, tris):
    """lgx_DeviceCreate on a float32 (n, 9) triangle array."""
    dev = lib.lgx_DeviceCreate(P(tris), len(tris))
    if dev and not ENTRY_GRID:
        _lib.check(lib.lgx_SetEntryGrid(dev, 0, None), lib)
    return dev


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--headline-only", action="store_true", help="only the C4 direct-light line (no extra legs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-entry-grid", action="store_true",
                    help="start every traversal at the root instead of its entry-grid cell (same results)")
    ap.add_argument("--no-c4-extras", action="store_true", help="skip C4's AO, radiosity, gradient, export legs")
    ap.add_argument("--no-c1", action="store_true", help="skip the C1 terrain leg")
    ap.add_argument("--no-radiosity", action="store_true", help="skip the radiosity (C2) leg")
    ap.add_argument("--radiosity-patches", type=int, default=32768)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the gradient (C3) leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (scales the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_bytes=4 * n)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU sharding of this path "
              "is a design only and is not run.", file=sys.stderr)
        sys.exit(2)
    global ENTRY_GRID
    ENTRY_GRID = not args.no_entry_grid
    if args.headline_only:
        args.no_c4_extras = args.no_c1 = args.no_radiosity = args.no_adjoint = True
    with_cpu = not args.no_cpu_baseline

    lib = _lib.load()
    _lib.check(lib.lgx_HasDevice(), lib)
    nth, cores = host_cores()
    log(f"[bench] host: {cores}")

    # headline: C4 direct light
    t0 = time.perf_counter()
    spec = scenes.city_scene()
    log(f"[bench] C4 scene: {spec.num_triangles} tris, {len(spec.lights)} lights ({time.perf_counter() - t0:.1f} s)")
    leg = DirectLeg(lib, spec, nth, radinfo=not args.no_c4_extras)
    leg.run(args.steps, args.warmup)
    roofline = leg.roofline("C4")
    cpu, parity = leg.cpu(args.cpu_seconds, cores) if with_cpu else (None, None)
    extras = {}
    if not args.no_c4_extras:
        extras["c4"] = c4_extras(lib, leg, nth, args.cpu_seconds, with_cpu)
    head = dict(value=leg.value, ms=leg.ms_per_step, march=leg.march_avg, marches=leg.marches, qsteps=leg.qsteps,
                S=leg.S, items=len(leg.items), ntri=leg.ntri, host_s=leg.host_s, dev_s=leg.device_create_s,
                cells=leg.grid_cells)
    leg.close()
    leg.built.destroy()

    if not args.no_c1:
        s1 = scenes.terrain_scene()
        c1 = DirectLeg(lib, s1, nth)
        c1.run(args.steps, args.warmup)
        c1_roof = c1.roofline("C1")
        c1_cpu, c1_par = c1.cpu(min(args.cpu_seconds, 5.0), cores) if with_cpu else (None, None)
        extras["c1"] = dict(workload=C1_WORKLOAD, shadow_rays_per_s=round(c1.value, 1),
                            ms_per_step=round(c1.ms_per_step, 3), marches=int(c1.marches),
                            march_kernel_ms=round(c1.march_avg, 3), roofline=c1_roof, cpu_baseline=c1_cpu,
                            parity_sample=c1_par,
                            speedup_vs_cpu=round(c1.value / c1_cpu["value"], 1) if c1_cpu else None)
        c1.close()
        c1.built.destroy()
    if not args.no_radiosity:
        extras["radiosity"] = radiosity_bench(lib, nth, n_patches=args.radiosity_patches, with_cpu=with_cpu)
    if not args.no_adjoint:
        extras["adjoint"] = adjoint_bench(lib, nth)

    out = dict(metric=METRIC, value=round(head["value"], 1), unit="shadow rays/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(head["ms"], 3), higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f32", data="synthetic",
               config=dict(workload=C4_WORKLOAD, triangles=head["ntri"], samples=head["S"], lights=len(spec.lights),
                           work_items=head["items"], shadow_rays_per_step=head["marches"],
                           march_steps_per_step=head["qsteps"], march_kernel_ms=round(head["march"], 3),
                           host_stages_s=round(head["host_s"], 2), device_create_s=round(head["dev_s"], 2),
                           entry_grid_cells=head["cells"]),
               roofline=roofline, cpu_baseline=cpu, parity_sample=parity,
               speedup_vs_cpu=round(head["value"] / cpu["value"], 1) if cpu else None, host=cores)
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

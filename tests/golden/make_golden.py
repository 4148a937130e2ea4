"""Write the golden fixtures tests/golden/ref_<scenario>.npz from the
reference baker built from its own sources (oracle/_ref, `make -C oracle`).

For each scenario of scenarios.py: the mesh arrays (mesh1, rad1: read from
the reference's bin/*.data with its loader's rules, lighter_test.cpp:48-104),
the reference's samples per instance after its preprocessing, the world-space
shadow-casting triangles, the lightmap
per instance after the direct-light stage, and the final outputs of a full
bake (ltr_Start ... ltr_GetWorkOutput) with srand(1) and one thread.

Run here, where the reference checkout exists:
    python tests/golden/make_golden.py [/root/reference]
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from lighter_amd import api, scenes  # noqa: E402
from oracle.checkers import Reference  # noqa: E402
import scenarios  # noqa: E402

libc = C.CDLL("libc.so.6")
libc.strtof.restype = C.c_float
libc.strtof.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]

MESH_FILES = {"mesh1": "test-mesh.data", "rad1": "test-set2-mesh1.data"}


def _f(tok):
    return np.float32(libc.strtof(tok.encode(), None))


def load_data_mesh(path):
    """The loader's rules: v/vt/vn/f records, vt.y -> 1 - vt.y, vertices
    deduplicated on exact (P, N, T) equality in first-seen order."""
    P, T, Nn = [], [], []
    pos, nrm, tex, idx = [], [], [], []
    seen = {}
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            if t[0] == "v":
                P.append(tuple(_f(x) for x in t[1:4]))
            elif t[0] == "vt":
                u, v = _f(t[1]), _f(t[2])
                T.append((u, np.float32(np.float32(1) - v)))
            elif t[0] == "vn":
                Nn.append(tuple(_f(x) for x in t[1:4]))
            elif t[0] == "f":
                for c in t[1:4]:
                    a, b, n = (int(x) - 1 for x in c.split("/"))
                    key = (P[a], Nn[n], T[b])
                    k = tuple(float(x) for x in key[0] + key[1] + key[2])
                    if k not in seen:
                        seen[k] = len(pos)
                        pos.append(key[0])
                        nrm.append(key[1])
                        tex.append(key[2])
                    idx.append(seen[k])
    return (np.array(pos, np.float32), np.array(nrm, np.float32), np.array(tex, np.float32),
            np.array(idx, np.uint32))


def outputs(lib, scene):
    info = api.WorkOutputInfo()
    lib.ltr_GetWorkOutputInfo(scene, C.byref(info))
    out = {}
    for i in range(info.lightmap_count):
        wo = api.WorkOutput()
        assert lib.ltr_GetWorkOutput(scene, i, C.byref(wo))
        n = wo.width * wo.height * 3
        rgb = np.ctypeslib.as_array(wo.lightmap_rgb, shape=(n,)).copy().reshape(wo.height, wo.width, 3)
        out[f"out{i}_rgb"] = rgb
        out[f"out{i}_uid"] = np.array([wo.uid], np.int64)
    out["n_outputs"] = np.array([info.lightmap_count], np.int64)
    return out


def main(ref_root):
    ref = Reference()
    for name in scenarios.NAMES:
        fx = {}
        if name in MESH_FILES:
            p, n, t, i = load_data_mesh(os.path.join(ref_root, "bin", MESH_FILES[name]))
            fx.update(mesh_pos=p, mesh_nrm=n, mesh_tex=t, mesh_idx=i)
        spec = scenarios.spec_for(name, fx)
        # preprocessing and direct light, staged
        libc.srand(1)
        b = scenes.BuiltScene(ref.lib, spec)
        ref.lib.rp_prepare(b.scene, 1)
        ni = ref.lib.rp_num_instances(b.scene)
        for k in range(ni):
            pos, nrm, loc = ref.samples(b.scene, k)[:3]
            fx[f"inst{k}_pos"], fx[f"inst{k}_nrm"], fx[f"inst{k}_loc"] = pos, nrm, loc
        fx["shadow_tris"] = ref.shadow_tris(b.scene)
        ref.lib.rp_direct(b.scene, 0, len(spec.lights))
        for k in range(ni):
            fx[f"inst{k}_direct"] = ref.lightmap(b.scene, k)
        fx["n_instances"] = np.array([ni], np.int64)
        b.destroy()
        # full bake
        libc.srand(1)
        b = scenes.BuiltScene(ref.lib, spec)
        b.bake()
        fx.update(outputs(ref.lib, b.scene))
        b.destroy()
        np.savez_compressed(scenarios.fixture_path(name), **fx)
        print(name, {k: v.shape for k, v in fx.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

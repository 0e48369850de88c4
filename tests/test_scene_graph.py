"""scene.json I/O and scene-graph flattening (host code of libatx, no GPU).

* the reference's sample scene (sample-scene-data/scene.json, committed as
  tests/golden/sample_scene.json) loads with the known flattened spheres and
  is written back byte for byte;
* flattening equals the oracle's restatement of traverseSceneGraph bit for
  bit on random node trees with rotations and non-uniform scales;
* a missing file gives the default scene; malformed JSON is an error.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from ataraxia_amd import AtxError, SceneGraph, default_scene, write_ppm
from ataraxia_amd import _lib

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_sample_scene_known_answers_and_round_trip(tmp_path):
    path = os.path.join(GOLDEN, "sample_scene.json")
    g = SceneGraph.load_json(path)
    assert g.counts() == {"nodes": 3, "spheres": 3, "materials": 3, "lights": 1}
    sph = g.flatten()
    # root sphere at the origin; ChildNode1 at (2,0,0) holds the r=100 ground at
    # (0,-101,0); its GrandChildNode at (0,0,0) holds (1,0,0) -> (3,0,0)
    assert sph["center"].tolist() == [[0, 0, 0], [2, -101, 0], [3, 0, 0]]
    assert sph["radius"].tolist() == [1, 100, 1]
    assert sph["material"].tolist() == [0, 1, 2]
    st = g.settings
    assert st.accumulation and not st.sky_light and st.max_bounces == 25
    pos, d, fov = g.get_camera()
    assert fov == 45.0 and pos[2] == np.float32(8.049220085144043)
    m = g.materials()
    assert m["metallic"].tolist() == [1.0, 0.0, 1.0]
    assert m["albedo"][0][0] == np.float32(1.0219999551773071)
    li = g.lights()
    assert li["position"][0].tolist() == [20, 20, 0] and li["intensity"][0] == 1.0
    text = open(path).read()
    assert g.to_json() == text
    out = tmp_path / "out.json"
    g.save_json(str(out))
    assert out.read_text() == text
    g2 = SceneGraph.parse_json(text)
    assert np.array_equal(g2.flatten(), sph)


def oracle_flatten(nodes):
    """nodes: list of (parent, pos, quat_xyzw, scale, spheres); pre-order ids."""
    out = []

    def visit(i, parent_m):
        _, pos, q, s, sph = nodes[i]
        local = np.zeros(16, np.float32)
        O.lib.ora_node_local_transform(np.float32(pos).ctypes.data, np.float32(q).ctypes.data,
                                       np.float32(s).ctypes.data, local.ctypes.data)
        G = np.zeros(16, np.float32)
        O.lib.ora_mat4_mul(parent_m.ctypes.data, local.ctypes.data, G.ctypes.data)
        for c, r, m in sph:
            a = np.zeros(1, _lib.SPHERE_DTYPE)
            a["center"][0] = c
            a["radius"][0] = r
            a["material"][0] = m
            o = np.zeros(1, _lib.SPHERE_DTYPE)
            O.lib.ora_flatten_sphere(G.ctypes.data, a.ctypes.data, o.ctypes.data)
            out.append(o[0])
        for j, nd in enumerate(nodes):
            if nd[0] == i and j != i:
                visit(j, G)

    visit(0, np.eye(4, dtype=np.float32).ravel())
    return np.array(out, dtype=_lib.SPHERE_DTYPE)


def test_flatten_matches_oracle_on_random_trees():
    rng = np.random.default_rng(21)
    for trial in range(12):
        g = SceneGraph()
        nodes = []
        n_nodes = int(rng.integers(1, 7))
        for i in range(n_nodes):
            parent = -1 if i == 0 else int(rng.integers(0, i))
            pos = rng.uniform(-5, 5, 3).astype(np.float32)
            q = rng.normal(size=4)
            q = (q / np.linalg.norm(q)).astype(np.float32)
            if trial % 3 == 0:
                q = np.zeros(4, np.float32)         # value-initialised quaternion -> identity
            scl = rng.uniform(0.5, 2.0, 3).astype(np.float32)
            sph = [(rng.uniform(-3, 3, 3).astype(np.float32), np.float32(rng.uniform(0.2, 2)), int(rng.integers(0, 4)))
                   for _ in range(int(rng.integers(0, 4)))]
            if i == 0:
                g_id = 0
                lib = _lib.lib
                lib.atx_scene_set_node_transform(g._h, 0, (_lib.C.c_float * 3)(*pos), (_lib.C.c_float * 4)(*q),
                                                 (_lib.C.c_float * 3)(*scl))
            else:
                g_id = g.add_node(parent, f"n{i}", pos, q, scl)
            assert g_id == i
            for c, r, m in sph:
                g.add_sphere(i, c, float(r), m)
            nodes.append((parent, pos, q, scl, sph))
        got = g.flatten()
        exp = oracle_flatten(nodes)
        assert np.array_equal(got, exp), (trial, got, exp)


def test_default_editor_scene_as_graph():
    """The editor's start-up graph (Engine/src/main.cpp:235-265) flattens to
    the three spheres of default_scene()."""
    g = SceneGraph()
    g.add_sphere(0, (0, 0, 0), 1.0, 0)
    c1 = g.add_node(0, "ChildNode1", (2, 0, 0))
    g.add_sphere(c1, (0, 0, 0), 1.0, 1)
    c2 = g.add_node(c1, "GrandChildNode", (0, 2, 0))
    g.add_sphere(c2, (0, 0, 0), 1.0, 2)
    assert np.array_equal(g.flatten()["center"], default_scene().spheres["center"])


def test_missing_file_and_malformed_json(tmp_path):
    g = SceneGraph.load_json(str(tmp_path / "does-not-exist.json"))
    assert g.counts() == {"nodes": 1, "spheres": 0, "materials": 0, "lights": 0}
    assert g.settings.max_bounces == 15 and g.settings.accumulation
    for bad in ['{"camera": ', '[1, 2]', '{"camera": {"position": [0,0], "direction": [0,0,1], "fov": 45}}',
                open(os.path.join(GOLDEN, "sample_scene.json")).read() + " x"]:
        with pytest.raises(AtxError) as e:
            SceneGraph.parse_json(bad)
        assert e.value.code == _lib.ATX_ERR_IO


def test_write_ppm(tmp_path):
    img = np.array([[0xFF0000FF, 0xFF00FF00], [0xFFFF0000, 0xFF123456]], np.uint32)
    p = tmp_path / "a.ppm"
    write_ppm(str(p), img, flip_vertical=True)
    data = p.read_bytes()
    assert data.startswith(b"P6\n2 2\n255\n")
    px = data[len(b"P6\n2 2\n255\n"):]
    # flipped: the last row first; bytes are r, g, b
    assert list(px) == [0, 0, 255, 0x56, 0x34, 0x12, 255, 0, 0, 0, 255, 0]

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
        pos, q, s = (np.ascontiguousarray(v, np.float32) for v in (pos, q, s))
        local = np.zeros(16, np.float32)
        O.lib.ora_node_local_transform(pos.ctypes.data, q.ctypes.data, s.ctypes.data, local.ctypes.data)
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


# ---- closed-form pins of flattening (SceneNode.cpp:46, Renderer.cu:78-86,
# Utils.cpp:145).  Each case names the answer a plausible mistake would give.

S2 = math.sqrt(0.5)


def _scene_json(root, children=()):
    """A scene.json text with one root node and nested children; each node is
    (position, rotation as the file stores it [x, y, z, w], scale, spheres)."""
    import json

    def node(name, spec, kids):
        pos, rot, scl, sph = spec
        d = {"name": name, "transformation": {"position": list(pos), "rotation": list(rot), "scale": list(scl)}}
        if sph:
            d["spheres"] = [{"center": list(c), "radius": r, "materialIndex": 0} for c, r in sph]
        if kids:
            d["children"] = [node(f"{name}.{i}", k[0], k[1]) for i, k in enumerate(kids)]
        return d

    doc = {"camera": {"direction": [0, 0, -1], "fov": 45.0, "position": [0, 0, 5]},
           "lights": [], "materials": [{"F0": [0.04] * 3, "albedo": [1, 1, 1], "emissionColor": [0, 0, 0],
                                        "emissionIntensity": 0.0, "metallic": 0.0, "roughness": 0.5}],
           "sceneGraph": node("Root", root, list(children)),
           "settings": {"accumulation": True, "maxBounces": 4, "skyLight": False}}
    return json.dumps(doc)


def _oracle_node_flatten(parent_m, pos, rot_xyzw, scl, sph):
    # the arrays are bound to names so that they outlive the calls reading them
    pos, rot, scl, par = (np.ascontiguousarray(v, np.float32) for v in (pos, rot_xyzw, scl, parent_m))
    local = np.zeros(16, np.float32)
    O.lib.ora_node_local_transform(pos.ctypes.data, rot.ctypes.data, scl.ctypes.data, local.ctypes.data)
    G = np.zeros(16, np.float32)
    O.lib.ora_mat4_mul(par.ctypes.data, local.ctypes.data, G.ctypes.data)
    out = []
    for c, r in sph:
        a = np.zeros(1, _lib.SPHERE_DTYPE)
        a["center"][0], a["radius"][0] = c, r
        o = np.zeros(1, _lib.SPHERE_DTYPE)
        O.lib.ora_flatten_sphere(G.ctypes.data, a.ctypes.data, o.ctypes.data)
        out.append(o[0])
    return G, out


def test_flatten_quaternion_order_and_rotation():
    """scene.json stores [x, y, z, w] and glm::quat takes (w, x, y, z): 90
    degrees about z maps (1,0,0) to (0,1,0) (reading the file as w,x,y,z
    would rotate about y instead, mapping (1,0,0) to (0,0,-1)); 90 degrees
    about x maps (0,1,0) to (0,0,1)."""
    I = np.eye(4, dtype=np.float32).ravel()
    for rot, c, exp in (([0, 0, S2, S2], (1, 0, 0), (0, 1, 0)), ([S2, 0, 0, S2], (0, 1, 0), (0, 0, 1)),
                        ([0, S2, 0, S2], (0, 0, 1), (1, 0, 0)), ([0, 0, 1, 0], (1, 2, 0), (-1, -2, 0))):
        g = SceneGraph.parse_json(_scene_json(((0, 0, 0), rot, (1, 1, 1), [(c, 1.0)])))
        got = g.flatten()[0]
        np.testing.assert_allclose(got["center"], exp, atol=1e-6)
        assert abs(float(got["radius"]) - 1.0) < 1e-6
        _, (o,) = _oracle_node_flatten(I, (0, 0, 0), rot, (1, 1, 1), [(c, 1.0)])
        np.testing.assert_allclose(o["center"], exp, atol=1e-6)
        assert np.array_equal(o["center"], got["center"])


def test_flatten_scale_and_trs_order():
    """Radius scales by the mean of the transform's column lengths: scale
    (1, 2, 3) doubles it (with or without a rotation); the node transform is
    T * R * S, so position (1,0,0), 90 degrees about z and scale 2 take
    (1,0,0) to (1,2,0) (S * R * T would give (0,4,0))."""
    I = np.eye(4, dtype=np.float32).ravel()
    cases = [(((0, 0, 0), [0, 0, 0, 1], (1, 2, 3)), (1, 1, 1), 0.5, (1, 2, 3), 1.0),
             (((0, 0, 0), [0, 0, S2, S2], (1, 2, 3)), (0, 1, 0), 0.5, (-2, 0, 0), 1.0),
             (((1, 0, 0), [0, 0, S2, S2], (2, 2, 2)), (1, 0, 0), 1.5, (1, 2, 0), 3.0)]
    for (pos, rot, scl), c, r, exp_c, exp_r in cases:
        g = SceneGraph.parse_json(_scene_json((pos, rot, scl, [(c, r)])))
        got = g.flatten()[0]
        np.testing.assert_allclose(got["center"], exp_c, atol=2e-6)
        assert abs(float(got["radius"]) - exp_r) < 2e-6, (got["radius"], exp_r)
        _, (o,) = _oracle_node_flatten(I, pos, rot, scl, [(c, r)])
        assert np.array_equal(o["center"], got["center"]) and o["radius"] == got["radius"]
    # exactly representable case: identity rotation, scale (1,2,3): radius 0.5 -> 1.0 exactly
    g = SceneGraph.parse_json(_scene_json(((0, 0, 0), [0, 0, 0, 1], (1, 2, 3), [((1, 1, 1), 0.5)])))
    s = g.flatten()[0]
    assert s["radius"] == np.float32(1.0) and s["center"].tolist() == [1, 2, 3]


def test_flatten_parent_times_child():
    """global = parent * local: a parent at (2,0,0) rotated 90 degrees about z
    with a child at (1,0,0) puts the child's origin at (2,1,0) (local *
    parent would give (3,0,0)); the child's scale 3 triples its radius and
    the parent's rotation turns the child's offset sphere centre."""
    root = ((2, 0, 0), [0, 0, S2, S2], (1, 1, 1), [])
    child = (((1, 0, 0), [0, 0, 0, 1], (3, 3, 3), [((0, 0, 0), 1.0), ((1, 0, 0), 0.5)]), [])
    g = SceneGraph.parse_json(_scene_json(root, [child]))
    s = g.flatten()
    np.testing.assert_allclose(s["center"][0], (2, 1, 0), atol=2e-6)
    np.testing.assert_allclose(s["center"][1], (2, 4, 0), atol=4e-6)
    np.testing.assert_allclose(s["radius"], (3.0, 1.5), atol=4e-6)
    I = np.eye(4, dtype=np.float32).ravel()
    G, _ = _oracle_node_flatten(I, (2, 0, 0), [0, 0, S2, S2], (1, 1, 1), [])
    _, outs = _oracle_node_flatten(G, (1, 0, 0), [0, 0, 0, 1], (3, 3, 3), [((0, 0, 0), 1.0), ((1, 0, 0), 0.5)])
    for o, got in zip(outs, s):
        assert np.array_equal(o["center"], got["center"]) and o["radius"] == got["radius"]


def test_numbers_outside_the_fields_range_are_rejected():
    """A float field that would overflow (it would be written back as null,
    which no reader accepts) or an int field outside int32 is a parse error,
    not undefined behaviour; whatever parses writes back and re-parses."""
    text = open(os.path.join(GOLDEN, "sample_scene.json")).read()
    for old, new, field in (("0.03999999910593033", "1e39", "material.F0"),
                            ('"radius": 100.0', '"radius": -1e300', "sphere.radius"),
                            ('"materialIndex": 1', '"materialIndex": 3e9', "sphere.materialIndex"),
                            ('"maxBounces": 25', '"maxBounces": 1e10', "settings.maxBounces")):
        assert old in text, old
        with pytest.raises(AtxError) as e:
            SceneGraph.parse_json(text.replace(old, new, 1))
        assert e.value.code == _lib.ATX_ERR_IO and field in str(e.value) and "range" in str(e.value)
    g = SceneGraph.parse_json(text.replace('"radius": 100.0', '"radius": 3.4e38', 1))
    assert SceneGraph.parse_json(g.to_json()).flatten()["radius"][1] == np.float32(3.4e38)
    import json
    doc = json.loads(text)
    doc["materials"] = {"a": doc["materials"]}
    with pytest.raises(AtxError) as e:
        SceneGraph.parse_json(json.dumps(doc))
    assert "materials is not an array" in str(e.value)

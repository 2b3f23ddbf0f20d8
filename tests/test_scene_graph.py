"""Host ingestion: scene files into a scene graph and the graph into a frame's
draw list (include/softlight_amd/SL_SceneGraph.hpp, SL_SceneFileLoader.hpp;
the reference's SL_SceneGraph.cpp:223-299,925-1045,1180-1260 and the render
loop of tests/sl_instancing_test.cpp:233-280).

The C++ example softlight_amd/examples/scene_graph_draw.cpp drives the
headers: its host-only build (scene_graph_dump) writes the loaded buffers,
the graph and the draw list; the full build renders the file on the GPU.

What pins each part, and the mistake each check catches:

  loader buffers     a hand-written file whose vertices, indices, meshes,
                     materials, image texels and nodes are listed in the test
                     (fan triangulation, 1-based and negative indices, corners
                     without uv or normal, usemtl runs, de-duplication)
  loader errors      an index past the end, an unknown material, a missing
                     file: non-zero exit naming the line
  rotations          half turns about x, y, z give exact sign matrices; a
                     quarter turn about z maps x to y (a transposed rotation
                     maps it to -y)
  hierarchy          model matrices of a three-level tree against float64
                     products of scipy rotations (an independent quaternion
                     convention) and T*S*R per node, parent first
  update             a parent edited after the first update moves its whole
                     subtree and nothing else
  child queries      total / immediate children and ancestry against a
                     brute-force walk of the parent array
  draw list          folding consecutive sub-meshes of a node with one
                     material into one item draws the same image as one draw
                     per mesh (the reference's loop), bit for bit; a node
                     translation equals pre-translated vertices, bit for bit
GPU (marked gpu): the C++ program's frame against the oracle drawing the
dumped draw list (depth and order bit-exact, RGBA8 within 1 LSB).
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import softlight_amd
from softlight_amd import api as A
from softlight_amd import scenes
from softlight_amd.scene_files import draw_dump, run_dump

BUILD = os.path.dirname(softlight_amd.LIB_PATH)
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="session")
def dump_exe(tmp_path_factory):
    exe = os.path.join(BUILD, "scene_graph_dump")
    if os.path.exists(exe):
        return exe
    exe = str(tmp_path_factory.mktemp("sg") / "scene_graph_dump")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-DSL_SCENE_DUMP_ONLY", "-I",
                    os.path.join(ROOT, "include"), "-o", exe,
                    os.path.join(ROOT, "softlight_amd", "examples", "scene_graph_draw.cpp")], check=True)
    return exe


def name_ids(d):
    return {g["name"]: g["id"] for g in d["graph"]}


# ---------------------------------------------------------------------------
# loader
# ---------------------------------------------------------------------------
TINY_OBJ = """mtllib s.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vt 1 0
vt 1 1
vn 0 0 1
o quad
usemtl red
f 1/1/1 2/2/1 3/3/1 4//1
o tri
f -4 -3 -2
usemtl tex
f 1 3 4
f 1/1/1 2/2/1 3/3/1
"""
TINY_MTL = """newmtl red
Ka 0.1 0.2 0.3
Kd 0.9 0.1 0.1
newmtl tex
Kd 1 1 1
d 0.5
Ns 7
map_Kd t.ppm
"""


def tiny_scene(tmp_path):
    (tmp_path / "s.obj").write_text(TINY_OBJ)
    (tmp_path / "s.mtl").write_text(TINY_MTL)
    (tmp_path / "t.ppm").write_bytes(b"P6\n# comment\n2 1\n255\n\x01\x02\x03\x04\x05\x06")
    return tmp_path / "s.obj"


def test_loader_buffers(dump_exe, tmp_path):
    d = run_dump(dump_exe, tiny_scene(tmp_path), "", tmp_path)
    z = [0.0, 0.0, 1.0]
    expect_v = np.array([
        [0, 0, 0, 0, 0, *z], [1, 0, 0, 1, 0, *z], [1, 1, 0, 1, 1, *z], [0, 1, 0, 0, 0, *z],   # quad: 4//1 has no uv
        [0, 0, 0, 0, 0, *z], [1, 0, 0, 0, 0, *z], [1, 1, 0, 0, 0, *z],                        # tri: face normal of a ccw face
        [0, 1, 0, 0, 0, *z],
    ], dtype=np.float32)
    np.testing.assert_array_equal(d["vertices"], expect_v)
    # quad fan (0,1,2),(0,2,3); tri (4,5,6); tex face (4,6,7); the last face reuses the quad's corners
    np.testing.assert_array_equal(d["indices"], [0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7, 0, 1, 2])
    assert d["meshes"] == [(0, 6, 0), (6, 9, 0), (9, 15, 1)]
    red, tex = d["materials"]
    np.testing.assert_array_equal(red[:12].astype(np.float32),
                                  np.float32([0.1, 0.2, 0.3, 1, 0.9, 0.1, 0.1, 1, 0, 0, 0, 0]))
    assert red[13] == -1 and tex[13] == 0 and tex[12] == 7 and tex[7] == 0.5
    np.testing.assert_array_equal(d["images"][0], [[[1, 2, 3, 255], [4, 5, 6, 255]]])
    assert [(g["name"], g["parent"], g["type"]) for g in d["graph"]] == [("s.obj", -1, 0), ("quad", 0, 1), ("tri", 0, 1)]
    # nodes' sub-mesh runs become draw items: quad 1 item, tri 2 (material changes)
    assert d["draws"] == [(1, 0, 1, 0), (2, 1, 1, 0), (2, 2, 1, 1)]


@pytest.mark.parametrize("obj,mtl,needle", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n", None, "line 4"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -4\n", None, "line 4"),
    ("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl nope\nf 1 2 3\n", "newmtl a\nKd 1 1 1\n", "unknown material"),
    ("v 0 0 0\nv 1 0 0\nf 1 2\n", None, "face"),
    ("v 0 0 0\n", None, "no faces"),
])
def test_loader_errors(dump_exe, tmp_path, obj, mtl, needle):
    (tmp_path / "bad.obj").write_text(obj)
    if mtl:
        (tmp_path / "m.mtl").write_text(mtl)
    r = run_dump(dump_exe, tmp_path / "bad.obj", "", tmp_path, check=False)
    assert r.returncode == 3 and needle in r.stderr, r.stderr


def test_loader_missing_file(dump_exe, tmp_path):
    r = run_dump(dump_exe, tmp_path / "none.obj", "", tmp_path, check=False)
    assert r.returncode == 3 and "cannot open" in r.stderr


def test_loader_reused_for_a_second_file(tmp_path):
    """One loader object loads a file whose faces without usemtl take a
    default material appended after three named ones, then a file with no
    mtllib at all: every mesh of the second load must name a material of the
    second load's table (a default-material index kept from the first file
    would point past it)."""
    (tmp_path / "m.mtl").write_text("newmtl a\nKd 1 0 0\nnewmtl b\nKd 0 1 0\nnewmtl c\nKd 0 0 1\n")
    (tmp_path / "one.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\n"
                                      "f 2 4 3\nusemtl b\nf 1 2 3\n")
    (tmp_path / "two.obj").write_text("v 0 0 0\nv 2 0 0\nv 0 2 0\nf 1 2 3\n")
    exe = tmp_path / "loader_reuse"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    os.path.join(HERE, "native", "loader_reuse.cpp")], check=True)
    r = subprocess.run([str(exe), str(tmp_path / "one.obj"), str(tmp_path / "two.obj")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    first, second = map(int, r.stdout.split()[1:3])
    assert first == 4 and second == 1, r.stdout


# ---------------------------------------------------------------------------
# transforms and the hierarchy
# ---------------------------------------------------------------------------
def xf(p=(0, 0, 0), s=(1, 1, 1), q=(0, 0, 0, 1)):
    return " ".join(scenes._fmt(v) for v in (*p, *s, *q))


def test_rotation_closed_forms(dump_exe, tmp_path):
    obj = tiny_scene(tmp_path)
    h = np.float32(np.sqrt(0.5))
    cmds = "\n".join([
        f"empty rx - {xf(q=(1, 0, 0, 0))}",
        f"empty ry - {xf(q=(0, 1, 0, 0))}",
        f"empty rz - {xf(p=(4, -2, 0.5), s=(2, 3, 4), q=(0, 0, 1, 0))}",
        f"empty qz - {xf(q=(0, 0, h, h))}",
    ])
    d = run_dump(dump_exe, obj, cmds, tmp_path)
    ids = name_ids(d)
    M = {n: d["model"][ids[n]].T for n in ("rx", "ry", "rz", "qz")}     # to [row][col]
    np.testing.assert_array_equal(M["rx"], np.diag([1, -1, -1, 1]).astype(np.float32))
    np.testing.assert_array_equal(M["ry"], np.diag([-1, 1, -1, 1]).astype(np.float32))
    # (T*S)*R: the rotation's columns scaled per row, then the translation
    np.testing.assert_array_equal(M["rz"], np.float32([[-2, 0, 0, 4], [0, -3, 0, -2], [0, 0, 4, 0.5], [0, 0, 0, 1]]))
    np.testing.assert_allclose(M["qz"] @ np.float32([1, 0, 0, 1]), [0, 1, 0, 1], atol=2e-7)
    np.testing.assert_allclose(M["qz"] @ np.float32([0, 1, 0, 1]), [-1, 0, 0, 1], atol=2e-7)


def trs64(p, s, q):
    m = np.eye(4)
    m[:3, :3] = np.diag(s) @ Rotation.from_quat(q).as_matrix()
    m[:3, 3] = p
    return m


def random_tree(seed):
    """(name, parent, p, s, q) in depth-first order: two roots' subtrees,
    three levels, unit quaternions."""
    rng = np.random.default_rng(seed)

    def tr():
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        return (tuple(np.float32(rng.uniform(-3, 3, 3))), tuple(np.float32(rng.uniform(0.5, 2, 3))),
                tuple(np.float32(q)))
    nodes = []
    for a in range(2):
        nodes.append((f"a{a}", "-", *tr()))
        for b in range(2):
            nodes.append((f"a{a}b{b}", f"a{a}", *tr()))
            for c in range(2):
                nodes.append((f"a{a}b{b}c{c}", f"a{a}b{b}", *tr()))
    return nodes


def expected_models(nodes):
    world = {}
    for name, parent, p, s, q in nodes:
        local = trs64(p, s, q)
        world[name] = local if parent == "-" else world[parent] @ local
    return world


def test_hierarchy_against_float64(dump_exe, tmp_path):
    obj = tiny_scene(tmp_path)
    nodes = random_tree(3)
    cmds = "\n".join(f"inst {n} {p} quad {xf(*t)}" for n, p, *t in nodes)
    d = run_dump(dump_exe, obj, cmds, tmp_path)
    ids = name_ids(d)
    for name, M in expected_models(nodes).items():
        np.testing.assert_allclose(d["model"][ids[name]].T, M, rtol=0, atol=2e-5 * max(1.0, np.abs(M).max()),
                                   err_msg=name)


def test_update_moves_only_the_edited_subtree(dump_exe, tmp_path):
    obj = tiny_scene(tmp_path)
    nodes = random_tree(4)
    first = "\n".join(f"inst {n} {p} quad {xf(*t)}" for n, p, *t in nodes)
    d0 = run_dump(dump_exe, obj, first, tmp_path)
    moved = list(nodes)
    k = [n[0] for n in nodes].index("a0b1")
    n, p, _, s, q = moved[k]
    moved[k] = (n, p, (np.float32(1.5), np.float32(-0.25), np.float32(2.0)), s, q)
    edit = first + f"\nupdate\nxform a0b1 {xf(*moved[k][2:])}\n"
    d1 = run_dump(dump_exe, obj, edit, tmp_path)
    ids = name_ids(d1)
    exp = expected_models(moved)
    for name, M in exp.items():
        np.testing.assert_allclose(d1["model"][ids[name]].T, M, rtol=0, atol=2e-5 * max(1.0, np.abs(M).max()),
                                   err_msg=name)
        same = not name.startswith("a0b1")
        assert np.array_equal(d0["model"][ids[name]], d1["model"][ids[name]]) == same, name


def test_rotate_normalises(dump_exe, tmp_path):
    obj = tiny_scene(tmp_path)
    d = run_dump(dump_exe, obj, f"empty e - {xf()}\nrotate e 0 0 3 3\n", tmp_path)
    M = d["model"][name_ids(d)["e"]].T
    np.testing.assert_allclose(M[:3, :3], Rotation.from_quat([0, 0, 1, 1]).as_matrix(), atol=2e-7)


def test_child_queries(dump_exe, tmp_path):
    obj = tiny_scene(tmp_path)
    nodes = random_tree(5)
    cmds = "\n".join(f"empty {n} {p} {xf(*t)}" for n, p, *t in nodes)
    d = run_dump(dump_exe, obj, cmds, tmp_path)
    parent = [g["parent"] for g in d["graph"]]
    n = len(parent)

    def ancestors(i):
        out = []
        while parent[i] != -1:
            i = parent[i]
            out.append(i)
        return out
    for g in d["graph"]:
        i = g["id"]
        assert g["total"] == sum(1 for j in range(n) if i in ancestors(j)), g
        assert g["immediate"] == sum(1 for j in range(n) if parent[j] == i), g
        row = d["is_child"][i]
        assert row[:n] == "".join("1" if j in ancestors(i) else "0" for j in range(n)), g
        assert row[n] == "0"            # the reference rejects parentId >= nodeIndex first, the root id included


# ---------------------------------------------------------------------------
# draw lists on the oracle (and the GPU)
# ---------------------------------------------------------------------------
def sphere_file(tmp_path, **kw):
    return scenes.write_icosphere_scene(str(tmp_path), **kw)[0]


def sphere_commands(seed=6):
    """The file's root turned and moved, one sphere re-posed, two instances
    of spheres under other spheres, and a node drawing meshes 0 and 4 (one
    material: one folded item) and 5."""
    h = np.float32(np.sqrt(0.5))
    return "\n".join([
        f"xform spheres.obj {xf(p=(0.5, -0.25, -1), q=(0, np.float32(0.2588190451), 0, np.float32(0.9659258263)))}",
        f"xform sphere1 {xf(p=(1, 1, 0), s=(1.25, 1, 1))}",
        f"update",
        f"inst copyA sphere4 sphere2 {xf(p=(2, 0, 1), s=(0.5, 0.5, 0.5), q=(0, 0, h, h))}",
        f"inst copyB copyA sphere3 {xf(p=(-1, 1, 0), s=(0.7, 0.7, 0.7))}",
        f"mnode multi - 3 0 4 5 {xf(p=(-2, -1, 1))}",
    ])


def test_folded_draw_list_draws_like_per_mesh_draws(dump_exe, oracle_lib, tmp_path):
    d = run_dump(dump_exe, sphere_file(tmp_path, n_objects=5, level=2), sphere_commands(), tmp_path)
    assert any(c > 1 for _, _, c, _ in d["draws"])                 # the 'multi' node folds
    W, H = 200, 150
    a = draw_dump(oracle_lib, d, W, H)
    b = draw_dump(oracle_lib, d, W, H, per_mesh=True)
    assert np.array_equal(a["depth"].view(np.uint32), b["depth"].view(np.uint32))
    assert np.array_equal(a["color"], b["color"])
    assert (a["depth"] < 1.0).mean() > 0.05          # the spheres are in view


def test_folding_merges_same_material_runs(dump_exe, tmp_path):
    # a node's consecutive sub-meshes with one material (and mode) become one
    # item; a material change starts the next
    obj = tmp_path / "f.obj"
    (tmp_path / "f.mtl").write_text("newmtl a\nKd 1 1 1\nnewmtl b\nKd 1 0 0\n")
    obj.write_text("mtllib f.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\no x\nusemtl a\nf 1 2 3\nusemtl b\nf 1 2 3\n"
                   "usemtl a\nf 1 2 3\nf 1 2 3\n")
    d = run_dump(dump_exe, obj, f"mnode y - 4 0 2 1 0 {xf()}\nmnode z y 2 1 1 {xf()}", tmp_path)
    assert d["meshes"] == [(0, 3, 0), (3, 6, 1), (6, 12, 0)]
    ids = name_ids(d)
    assert d["draws"] == [(ids["x"], 0, 1, 0), (ids["x"], 1, 1, 1), (ids["x"], 2, 1, 0),
                          (ids["y"], 3, 2, 0), (ids["y"], 5, 1, 1), (ids["y"], 6, 1, 0),
                          (ids["z"], 7, 2, 1)]
    assert [m[:3] for m in d["draw_meshes"]] == [(0, 3, 0), (3, 6, 1), (6, 12, 0), (0, 3, 0), (6, 12, 0), (3, 6, 1),
                                                 (0, 3, 0), (3, 6, 1), (3, 6, 1)]


def test_node_translation_equals_pretranslated_vertices(dump_exe, oracle_lib, tmp_path):
    obj = sphere_file(tmp_path, n_objects=4, level=2)
    t = (np.float32(0.75), np.float32(-0.5), np.float32(-1.25))
    moved = run_dump(dump_exe, obj, f"xform spheres.obj {xf(p=t)}", tmp_path)
    plain = run_dump(dump_exe, obj, "", tmp_path)
    v = plain["vertices"].copy()
    v[:, 0:3] = v[:, 0:3] + np.float32(t)
    W, H = 160, 120
    a = draw_dump(oracle_lib, moved, W, H)
    b = draw_dump(oracle_lib, plain, W, H, vertices=v)
    assert np.array_equal(a["depth"].view(np.uint32), b["depth"].view(np.uint32))
    assert np.array_equal(a["color"], b["color"])


# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("W,H,n_objects,level,cover", [
    (160, 120, 1, 0, 0.0),          # 60 indices: 8-bit index buffer
    (320, 240, 6, 2, 0.05),         # 5,760 indices: 16-bit
    (640, 360, 12, 3, 0.05),
    (480, 270, 24, 3, 0.05),        # 92,160 indices: 32-bit
])
def test_gpu_cpp_program_matches_oracle(dump_exe, oracle_lib, tmp_path, W, H, n_objects, level, cover):
    obj = sphere_file(tmp_path, n_objects=n_objects, level=level)
    cmds = sphere_commands() if n_objects >= 5 else ""      # the commands name spheres 0..4
    d = run_dump(dump_exe, obj, cmds, tmp_path)
    exe = os.path.join(BUILD, "scene_graph_draw")
    cout, dout = tmp_path / "color.bin", tmp_path / "depth.bin"
    r = subprocess.run([exe, "render", str(obj), str(tmp_path / "nodes.txt"), str(W), str(H), str(cout), str(dout)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    color = np.fromfile(cout, dtype=np.uint8).reshape(H, W, 4)
    depth = np.fromfile(dout, dtype=np.float32).reshape(H, W)
    base = A.sl_std_uniforms.from_buffer_copy((tmp_path / "color.bin.uniforms").read_bytes())
    o = draw_dump(oracle_lib, d, W, H, base=base)
    assert np.array_equal(depth.view(np.uint32), o["depth"].view(np.uint32))
    assert int(np.abs(color.astype(int) - o["color"].astype(int)).max()) <= 1
    assert (depth < 1.0).mean() > cover


@pytest.mark.gpu
def test_gpu_draw_list_parity(dump_exe, gpu_lib, oracle_lib, tmp_path):
    d = run_dump(dump_exe, sphere_file(tmp_path, n_objects=8, level=3), sphere_commands(), tmp_path)
    g = draw_dump(gpu_lib, d, 480, 270)
    o = draw_dump(oracle_lib, d, 480, 270)
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert int(np.abs(g["color"].astype(int) - o["color"].astype(int)).max()) <= 1

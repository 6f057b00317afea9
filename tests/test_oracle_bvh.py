"""Oracle pins for the LBVH and the traversal: radix-tree definition of the
Karras hierarchy, box containment, key order, traversal == brute force over all
triangles, analytic hits, watertightness on shared edges, refit."""
import numpy as np
import pytest

from anerend_amd import scenegen
from anerend_amd.types import INSTANCE_DTYPE, MATERIAL_DTYPE
from oracle import oracle as O


def soup_scene(n=300, seed=5):
    """Random triangles as one mesh (instance identity)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, (n, 1, 3))
    p = (c + rng.normal(scale=0.6, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    mb = scenegen._MeshBuilder()
    nrm = np.tile([[0, 1, 0]], (len(p), 1)).astype(np.float32)
    mesh = mb.add(p, nrm, np.arange(len(p), dtype=np.uint32))
    v, i, m = mb.arrays()
    mats = np.zeros(1, MATERIAL_DTYPE)
    mats["bindless_indices"] = -1
    mats["base_col_fac"] = (0.5, 0.5, 0.5, 1)
    inst = np.array([scenegen.make_instance(mesh, 0, np.eye(3), (1, 1, 1), (0, 0, 0))], INSTANCE_DTYPE)
    return scenegen.Scene(v, i, m, mats, inst)


SCENES = {
    "room": lambda: scenegen.room_scene(1),
    "soup": lambda: soup_scene(),
    "terrain": lambda: scenegen.terrain_scene(9, cells=24, num_instances=40, spread=10.0),
}


@pytest.fixture(scope="module", params=list(SCENES))
def osc(request):
    return O.OracleScene(SCENES[request.param]())


def _delta(keys, i, j):
    n = len(keys)
    if j < 0 or j >= n:
        return -1
    i, j = int(i), int(j)
    a, b = int(keys[i]), int(keys[j])
    if a == b:
        return 64 + (32 - (i ^ j).bit_length())
    return 64 - (a ^ b).bit_length()


def test_lbvh_is_the_radix_tree_of_the_sorted_keys(osc):
    b = osc.bvh()
    keys, ids, L, R, P = b["keys"], b["ids"], b["left"], b["right"], b["parent"]
    n = len(keys)
    assert np.array_equal(np.sort(ids), np.arange(n))
    assert np.all(np.diff(keys.astype(np.float64)) >= 0)
    # each leaf and each non-root internal node is referenced exactly once
    kids = np.concatenate([L, R])
    leaves = np.sort(~kids[kids < 0])
    inner = np.sort(kids[kids >= 0])
    assert np.array_equal(leaves, np.arange(n))
    assert np.array_equal(inner, np.arange(1, n - 1))
    # every node covers a contiguous key range split where the common prefix of
    # the augmented keys is shortest (definition of the binary radix tree)
    def rng_of(c):
        if c < 0:
            return ~c, ~c
        lo_l, _ = rng_of(L[c])
        _, hi_r = rng_of(R[c])
        return lo_l, hi_r
    import sys
    sys.setrecursionlimit(10000)
    for i in range(n - 1):
        lo, hi = rng_of(i)
        llo, lhi = rng_of(L[i])
        rlo, rhi = rng_of(R[i])
        assert llo == lo and rhi == hi and lhi + 1 == rlo
        split = lhi
        dmin = min(_delta(keys, k, k + 1) for k in range(lo, hi))
        assert _delta(keys, split, split + 1) == dmin
        if i > 0:
            assert L[P[i]] == i or R[P[i]] == i


def test_boxes_enclose_children(osc):
    b = osc.bvh()
    tris = osc.world_tris()
    boxes, L, R, ids = b["boxes"], b["left"], b["right"], b["ids"]

    def box(c):
        if c < 0:
            t = tris[ids[~c]].reshape(3, 3)
            return np.concatenate([t.min(0), t.max(0)])
        return boxes[c]
    for i in range(len(L)):
        for c in (L[i], R[i]):
            cb = box(c)
            assert np.all(boxes[i][:3] <= cb[:3]) and np.all(boxes[i][3:] >= cb[3:])
        u = np.concatenate([np.minimum(box(L[i])[:3], box(R[i])[:3]), np.maximum(box(L[i])[3:], box(R[i])[3:])])
        assert np.array_equal(u, boxes[i])
    assert osc.max_depth <= 96


def _random_rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o, d


def test_traversal_equals_brute_force(osc):
    o, d = _random_rays(6000, 11, [-5, -1, -5], [5, 4, 5])
    for any_hit in (False, True):
        h, _ = osc.trace(o, d, any_hit=any_hit)
        hb = osc.trace(o, d, any_hit=any_hit, brute=True)
        if any_hit:
            assert np.array_equal(h["tri"] >= 0, hb["tri"] >= 0)
        else:
            assert np.array_equal(h, hb)


def test_analytic_room_hits():
    osc = O.OracleScene(scenegen.room_scene(1, num_boxes=0))
    o = np.array([[0, 1, 0], [0, 1, 0], [1, 1, 1], [-2, 0, 3]], np.float32)
    d = np.array([[0, -1, 0], [1, 0, 0], [0, 1, 0], [0, 0, -1]], np.float32)
    h, _ = osc.trace(o, d)
    assert np.all(h["tri"] >= 0)
    assert np.allclose(h["t"], [1.5, 5.0, 2.5, 8.0], rtol=1e-6)
    # t range: a wall closer than t_min or farther than t_max is not reported
    h, _ = osc.trace(o[:1], d[:1], tmin=1.6)
    assert h["tri"][0] != -1 and h["t"][0] > 1.6 or h["tri"][0] == -1
    h, _ = osc.trace(o[:1], d[:1], tmax=1.4)
    assert h["tri"][0] == -1
    # barycentrics reproduce the hit point
    tris = osc.world_tris()
    h, _ = osc.trace(o, d)
    for k in range(len(o)):
        t = tris[h["tri"][k]].reshape(3, 3).astype(np.float64)
        u, v = float(h["u"][k]), float(h["v"][k])
        p = (1 - u - v) * t[0] + u * t[1] + v * t[2]
        assert np.allclose(p, o[k] + h["t"][k] * d[k], atol=1e-5)


def test_watertight_on_shared_edges_and_vertices():
    """Rays aimed exactly at shared vertices and edge midpoints of a triangle
    grid must all hit (Woop et al. 2013 watertightness)."""
    scene = scenegen.terrain_scene(4, cells=16, num_instances=0, amplitude=0.0)
    osc = O.OracleScene(scene)
    xs = np.arange(-7, 8, 0.5, dtype=np.float32)
    X, Z = np.meshgrid(xs, xs)
    o = np.stack([X.ravel(), np.full(X.size, 3.0, np.float32), Z.ravel()], 1)
    d = np.tile(np.array([[0, -1, 0]], np.float32), (len(o), 1))
    h, _ = osc.trace(o, d)
    assert np.all(h["tri"] >= 0)
    assert np.allclose(h["t"], 3.0)
    # oblique rays through diagonal edges
    d2 = np.tile(np.array([[0.3, -1, 0.2]], np.float32), (len(o), 1))
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    inside = (np.abs(o[:, 0] + 0.9) < 7.5) & (np.abs(o[:, 2] + 0.6) < 7.5)
    h2, _ = osc.trace(o[inside], d2[inside])
    assert np.all(h2["tri"] >= 0)


def test_refit_moves_geometry_and_keeps_topology():
    scene = scenegen.room_scene(1, num_boxes=16)
    osc = O.OracleScene(scene)
    before = osc.bvh()
    inst = scene.instances.copy()
    inst["transform"][1:, 3] += 0.25        # shift boxes along x
    osc.refit(inst)
    after = osc.bvh()
    for k in ("keys", "ids", "left", "right", "parent"):
        assert np.array_equal(before[k], after[k])
    ref = O.OracleScene(scenegen.Scene(scene.vertices, scene.indices, scene.meshes, scene.materials, inst))
    assert np.array_equal(osc.world_tris(), ref.world_tris())
    o, d = _random_rays(3000, 2, [-4, 0, -4], [4, 3, 4])
    h, _ = osc.trace(o, d)
    assert np.array_equal(h, osc.trace(o, d, brute=True))
    inst["custom_index_and_mask"][3] ^= (1 << 10)
    with pytest.raises(ValueError):
        osc.refit(inst)


def test_degenerate_directions_are_no_rays(osc):
    """A zero or non-finite direction is no ray: a miss with no node visits.  A NaN
    direction would otherwise pass every slab test (fmaxf/fminf drop NaN) and
    visit the whole hierarchy."""
    d = np.array([[0, 0, 0], [np.nan, 0, 1], [np.inf, 0, 0], [0, -np.inf, np.nan]], np.float32)
    o = np.zeros((4, 3), np.float32)
    for any_hit in (False, True):
        h, ctr = osc.trace(o, d, any_hit=any_hit)
        assert np.all(h["tri"] == -1) and np.all(ctr == 0)
    h, ctr = osc.trace(o[:1], np.array([[0, -1, 0]], np.float32))
    assert ctr[0] > 0


def _exact_child_boxes(osc):
    """Exact boxes [n-1][2][2][3] of every internal node's two children."""
    b = osc.bvh()
    tris = osc.world_tris().reshape(-1, 3, 3)
    leaf = np.stack([tris.min(1), tris.max(1)], 1)[b["ids"]]      # sorted order
    node = b["boxes"].reshape(-1, 2, 3)
    out = []
    for side in ("left", "right"):
        c = b[side]
        out.append(np.where((c < 0)[:, None, None], leaf[np.where(c < 0, ~c, 0)], node[np.where(c >= 0, c, 0)]))
    return np.stack(out, 1)


@pytest.mark.parametrize("name", list(SCENES))
def test_quantized_nodes_are_conservative_and_tight(name):
    osc = O.OracleScene(SCENES[name]())
    osc.set_node32(True)
    q = osc.qboxes().astype(np.float64)
    ex = _exact_child_boxes(osc).astype(np.float64)
    assert np.all(q[:, :, 0] <= ex[:, :, 0]) and np.all(q[:, :, 1] >= ex[:, :, 1])
    # at most one step (extent / 127 of the node box, or the smallest step) plus one rounding
    lo = np.minimum(ex[:, 0, 0], ex[:, 1, 0])
    hi = np.maximum(ex[:, 0, 1], ex[:, 1, 1])
    step = np.maximum((hi - lo) / 127.0, 2.0 ** -126)[:, None, :]
    slack = step + np.abs(lo)[:, None, :] * 2.0 ** -23
    assert np.all(ex[:, :, 0] - q[:, :, 0] <= slack) and np.all(q[:, :, 1] - ex[:, :, 1] <= slack)


@pytest.mark.parametrize("name", list(SCENES))
def test_quantized_traversal_equals_brute_force(name):
    osc = O.OracleScene(SCENES[name]())
    o, d = _random_rays(6000, 11, [-5, -1, -5], [5, 4, 5])
    _, c_exact = osc.trace(o, d)
    osc.set_node32(True)
    for any_hit in (False, True):
        h, c = osc.trace(o, d, any_hit=any_hit)
        hb = osc.trace(o, d, any_hit=any_hit, brute=True)
        if any_hit:
            assert np.array_equal(h["tri"] >= 0, hb["tri"] >= 0)
        else:
            assert np.array_equal(h, hb)
            assert c[0] >= c_exact[0]      # looser boxes visit at least as many nodes


def test_quantized_nodes_follow_a_refit():
    scene = scenegen.room_scene(1, num_boxes=16)
    osc = O.OracleScene(scene)
    osc.set_node32(True)
    inst = scene.instances.copy()
    inst["transform"][1:, 3] += 0.25
    osc.refit(inst)
    q = osc.qboxes().astype(np.float64)
    ex = _exact_child_boxes(osc).astype(np.float64)
    assert np.all(q[:, :, 0] <= ex[:, :, 0]) and np.all(q[:, :, 1] >= ex[:, :, 1])
    o, d = _random_rays(3000, 2, [-4, 0, -4], [4, 3, 4])
    h, _ = osc.trace(o, d)
    assert np.array_equal(h, osc.trace(o, d, brute=True))

"""Structure of the BVHs our builder emits against the reference's layout contract.

The split decisions are ours (Embree 3.13.5 is not available), so topology is not compared;
what is checked is the contract the traversal relies on: fat pairs in depth-first pre-order
with the root dropped (AlignNodes2, utils/raytracing/cbvh.cpp:104-169; CreateFatTreeArray,
cbvh_fat.cpp:100-142), leaf words (cbvh.cpp:502-511, 645-655), escape indices, boxes that
contain their subtrees, and every primitive referenced exactly once.
"""
import numpy as np
import pytest

from conftest import host_scene

LEAF_BIT = 0x80000000


def walk_fat(nodes, visit_leaf):
    """Depth-first walk from pair 0; checks pre-order numbering and escape indices."""
    order = []
    stack = [(0, 0xFFFFFFFE)]
    while stack:
        p, parent_escape = stack.pop()
        order.append(p)
        pair = nodes[p]
        assert pair["left"]["escapeIndex"] == 2 * p + 1
        assert pair["right"]["escapeIndex"] == parent_escape
        kids = []
        for k, side in enumerate(("left", "right")):
            lo = int(pair[side]["leftOffset"])
            if lo & LEAF_BIT:
                visit_leaf(p, side, lo)
            else:
                kids.append((lo, 2 * p + 1 if k == 0 else parent_escape))
        for c in reversed(kids):
            stack.append(c)
    assert order == list(range(len(nodes))), "pairs are not in depth-first pre-order"
    return order


def subtree_boxes_contain(nodes, leaf_box):
    """Returns (min,max) of every subtree and asserts each stored child box contains it."""
    def rec(p):
        pair = nodes[p]
        out = []
        for side in ("left", "right"):
            lo = int(pair[side]["leftOffset"])
            if lo == 0xFFFFFFFF:
                continue
            if lo & LEAF_BIT:
                mn, mx = leaf_box(lo)
            else:
                mn, mx = rec(lo)
            if mn is None:
                continue
            assert np.all(pair[side]["boxMin"] <= mn + 0.0) and np.all(pair[side]["boxMax"] >= mx - 0.0), \
                f"child box of pair {p} does not contain its subtree"
            out.append((mn, mx))
        if not out:
            return None, None
        return np.min([o[0] for o in out], axis=0), np.max([o[1] for o in out], axis=0)
    return rec(0)


@pytest.mark.parametrize("subdiv,amp", [(3, 0.0), (4, 0.05)])
def test_triangle_blas_contract(lrt, subdiv, amp):
    v, f = lrt.gen_icosphere(subdiv, amp, 4.0, 1)
    s = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_Triangles3f(v, f)))
    nodes = s.get_array("nodes")
    pic = s.get_array("prim_id_count")
    prim = s.get_array("prim_indices")
    idx = s.get_array("indices").reshape(-1, 3)
    assert np.array_equal(idx, f[prim]), "reordered index buffer must follow m_primIndices"
    seen_leaf = []
    covered = np.zeros(f.shape[0], np.int32)

    def on_leaf(p, side, lo):
        leaf_id = lo & 0x0FFFFFFF
        assert (lo >> 28) & 7 == 1
        seen_leaf.append(leaf_id)
        w = int(pic[leaf_id])
        start, count = w & 0x0FFFFFFF, (w >> 28) & 7
        assert 1 <= count <= 2, "cbvh_embree2 packs at most 2 triangles per leaf"
        covered[start:start + count] += 1

    walk_fat(nodes, on_leaf)
    assert sorted(seen_leaf) == list(range(len(pic))), "every leaf word is referenced once"
    assert np.all(covered == 1), "every triangle is in exactly one leaf"
    verts = s.get_array("vert_pos").reshape(-1, 4)[:, :3]

    def leaf_box(lo):
        w = int(pic[lo & 0x0FFFFFFF])
        start, count = w & 0x0FFFFFFF, (w >> 28) & 7
        pts = verts[idx[start:start + count].reshape(-1)]
        return pts.min(axis=0), pts.max(axis=0)

    subtree_boxes_contain(nodes, leaf_box)


def test_box_blas_and_tlas_contract(lrt):
    rng = np.random.default_rng(5)
    v, f = lrt.gen_icosphere(1)
    mats = []
    for _ in range(37):
        m = np.eye(4, dtype=np.float32)
        m[:3, 3] = rng.uniform(-20, 20, 3)
        mats.append(m)
    def build(s):
        gid = s.AddGeom_Triangles3f(v, f)
        for m in mats:
            s.AddInstance(gid, m)
    s = host_scene(lrt, build)
    tlas = s.get_array("tlas")
    inst = s.get_array("instances")
    # stackless walk over escape indices that enters every node visits every instance once
    seen = []
    node = 0
    while node < 0xFFFFFFFE:
        n = tlas[node]
        if n["leftOffset"] & LEAF_BIT:
            seen.append(int(n["leftOffset"]) & 0x0FFFFFFF)
            node = int(n["escapeIndex"])
        else:
            kid = int(n["leftOffset"])
            # children boxes are inside their parent's
            for c in (kid, kid + 1):
                assert np.all(tlas[c]["boxMin"] >= n["boxMin"]) and np.all(tlas[c]["boxMax"] <= n["boxMax"])
            node = kid
    assert sorted(seen) == list(range(len(mats)))
    for i, m in enumerate(mats):
        assert np.allclose(inst[i]["transform"].reshape(4, 4).T, m)
        assert np.allclose(inst[i]["transformInv"].reshape(4, 4).T @ m, np.eye(4), atol=1e-5)


def test_frame_octree_leaf_boxes(lrt):
    sph = np.array([[0.1, -0.2, 0.05, 0.5], [-0.4, 0.3, 0.2, 0.3]], np.float32)
    nodes = lrt.sdf_spheres_frame_octree(sph, 5)
    s = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfFrameOctree(nodes)))
    orig = s.get_array("orig_nodes")
    fat = s.get_array("nodes")
    # one box per leaf whose corner values change sign (BVH2Common_host.cpp:1348-1383)
    leaves = [i for i in range(nodes.size) if nodes[i]["offset"] == 0
              and (nodes[i]["values"] <= 0).any() and (nodes[i]["values"] > 0).any()]
    assert sorted(int(x) for x in orig["leftOffset"]) == sorted(leaves)
    seen = []
    walk_fat(fat, lambda p, side, lo: seen.append(lo & 0x0FFFFFFF))
    assert sorted(seen) == list(range(orig.size))

    def leaf_box(lo):
        b = orig[lo & 0x0FFFFFFF]
        return b["boxMin"], b["boxMax"]
    subtree_boxes_contain(fat, leaf_box)


def test_oracle_tlas_refit_pinned(lrt, orc):
    """The oracle's TLAS refit (checker of the GPU refit): on a freshly built TLAS every stored box
    already is the union of its subtree's instance boxes, so a refit changes nothing; after one
    instance box grows, exactly that leaf and its ancestors change, each to a box containing the
    new one."""
    rng = np.random.default_rng(12)
    v, f = lrt.gen_icosphere(1)
    mats = []
    for _ in range(53):
        m = np.eye(4, dtype=np.float32)
        m[:3, 3] = rng.uniform(-30, 30, 3)
        mats.append(m)

    def build(s):
        gid = s.AddGeom_Triangles3f(v, f)
        for m in mats:
            s.AddInstance(gid, m)
    s = host_scene(lrt, build)
    tlas, inst = s.get_array("tlas"), s.get_array("instances")
    assert orc.tlas_refit(tlas, inst).tobytes() == tlas.tobytes()
    parent = s.get_array("tlas_parent")
    inst2 = inst.copy()
    k = 17
    inst2["boxMax"][k][:3] += 5.0
    r = orc.tlas_refit(tlas, inst2)
    changed = set(np.nonzero(np.any(r["boxMax"] != tlas["boxMax"], axis=1))[0].tolist())
    leaf = [i for i in range(tlas.size) if tlas["leftOffset"][i] & LEAF_BIT and (tlas["leftOffset"][i] & 0x0FFFFFFF) == k][0]
    path, i = [], leaf
    while i != 0xFFFFFFFF:
        path.append(i)
        i = int(parent[i])
    assert changed <= set(path) and leaf in changed
    assert np.all(r["boxMax"][path] >= inst2["boxMax"][k][:3])
    assert np.array_equal(r["boxMin"], tlas["boxMin"])
    # the parent array names every inner node's two children
    for i in range(tlas.size):
        lo = int(tlas["leftOffset"][i])
        if not lo & LEAF_BIT:
            assert parent[lo] == i and parent[lo + 1] == i


def test_gpu_builder_without_device_fails_loudly(lrt):
    """build names starting with "gpu" build on the HIP device; without one the add call fails
    with a HIP error instead of falling back to the host builder."""
    v, f = lrt.gen_icosphere(2)
    s = lrt.SceneRT(build_name="gpu_lbvh2")
    with pytest.raises(lrt.LrtError) as e:
        s.AddGeom_Triangles3f(v, f)
    assert "GPU BVH build" in str(e.value)

"""The global-octree builder (sdf_to_global_octree, sdf/build/sparse_octree_builder.cpp:37-264) on
the sphere-union field, pinned by what the field fixes rather than by a restatement of the
builder: every node's samples are the field at the node's brick lattice (corner + (i - pad) *
size / brick_size, x-major), the nodes form one tree of 8-child blocks below the root, and every
point of the zero set lies in a leaf at full depth (the exact distance field's surface is never
left in a coarser leaf). A stored sample at the wrong lattice point, a child block in another
order or a surface cell left unsplit fails one of these."""
import numpy as np
import pytest

from test_coctree_cpu import SPHERES


def field(p):
    c, r = SPHERES[:, :3].astype(np.float64), SPHERES[:, 3].astype(np.float64)
    return np.min(np.linalg.norm(p[..., None, :] - c, axis=-1) - r, axis=-1)


def walk(nodes):
    """(corner, size, depth) of every node reached from the root; each node reached once."""
    n = nodes.shape[0]
    corner, size, depth = np.zeros((n, 3)), np.zeros(n), np.full(n, -1)
    corner[0], size[0], depth[0] = -1.0, 2.0, 0
    stack = [0]
    while stack:
        i = stack.pop()
        off = int(nodes[i, 2])
        if off:
            assert off + 8 <= n
            for c in range(8):
                k = off + c
                assert depth[k] == -1, "a node is reached twice"
                corner[k] = corner[i] + 0.5 * size[i] * np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
                size[k], depth[k] = 0.5 * size[i], depth[i] + 1
                stack.append(k)
    assert np.all(depth >= 0), "a node is not reached from the root"
    return corner, size, depth


@pytest.mark.parametrize("depth,bs,pad", [(5, 1, 0), (4, 2, 1), (3, 4, 0)])
def test_samples_structure_and_surface_cover(lrt, depth, bs, pad):
    nodes, values = lrt.sdf_spheres_global_octree(SPHERES, depth, bs, pad)
    v = bs + 2 * pad + 1
    corner, size, dep = walk(nodes)
    assert dep.max() == depth
    # samples: the field at the brick lattice, float32-rounded
    a = (np.arange(v) - pad) / bs
    grid = np.stack(np.meshgrid(a, a, a, indexing="ij"), -1).reshape(-1, 3)
    for i in range(0, nodes.shape[0], max(1, nodes.shape[0] // 400)):
        want = field(corner[i] + size[i] * grid)
        got = values[int(nodes[i, 1]): int(nodes[i, 1]) + v ** 3].astype(np.float64)
        assert np.max(np.abs(got - want)) < 2e-6, f"node {i}"
    # cover: points on each sphere that are on the union's boundary lie in full-depth leaves
    rng = np.random.default_rng(1)
    u = rng.normal(size=(3000, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    pts = []
    for c in SPHERES:
        p = c[:3] + c[3] * u
        pts.append(p[np.abs(field(p)) < 1e-9])
    pts = np.concatenate(pts)
    pts = pts[np.all(np.abs(pts) < 0.999, axis=1)]
    assert len(pts) > 1000
    for p in pts:
        i = 0
        while int(nodes[i, 2]):
            rel = (p - corner[i]) / size[i]
            c = (int(rel[0] >= 0.5) << 2) | (int(rel[1] >= 0.5) << 1) | int(rel[2] >= 0.5)
            i = int(nodes[i, 2]) + c
        assert dep[i] == depth, f"surface point {p} in a depth-{dep[i]} leaf"

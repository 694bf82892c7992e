"""LBVH oracle (oracle/lbvh_oracle.c) checked against the defining invariants.

Morton order, the Karras binary radix tree over sorted keys, exact box
nesting, outward-rounded 8-bit child boxes and exact leaf coverage.  A build
that drops or duplicates a primitive, inverts a split or rounds a box inward
fails here; the GPU build is then compared with this oracle byte for byte.
"""

import struct

import numpy as np
import pytest

from ordinarylight_amd import workloads
from oracle import native as O

LEAF = 0x80000000


def spread10(v):
    out = 0
    for b in range(10):
        out |= ((v >> b) & 1) << (3 * b)
    return out


def decode_nodes(nodes):
    """Decode uint8 [N,64] nodes: origin, scales, (lo, hi) per child, children."""
    out = []
    for raw in nodes:
        words = struct.unpack("<3f13I", raw.tobytes())
        o = np.array(words[:3], np.float32)
        exps = words[3]
        scale = np.array([struct.unpack("<f", struct.pack("<I", ((exps >> (8 * a)) & 255) << 23))[0]
                          for a in range(3)], np.float32)
        qlo = [words[4 + a] for a in range(3)]
        qhi = [words[7 + a] for a in range(3)]
        children = words[10:14]
        boxes = []
        for c in range(4):
            lo = np.array([o[a] + np.float32((qlo[a] >> (8 * c)) & 255) * scale[a] for a in range(3)], np.float32)
            hi = np.array([o[a] + np.float32((qhi[a] >> (8 * c)) & 255) * scale[a] for a in range(3)], np.float32)
            boxes.append((lo, hi))
        out.append((children, boxes))
    return out


def mesh_cases():
    scene, _ = workloads.box_room(0)
    yield "room", scene
    scene, _ = workloads.terrain(5, nx=61, nz=41, chunk=20, emissive_quads=3)
    yield "terrain", scene
    rng = np.random.default_rng(3)
    from ordinarylight_amd.scene import Scene
    s = Scene()
    v = rng.normal(0, 5, (3000, 3)).astype(np.float32)
    # Duplicate centroids stress the index-augmented Karras keys.
    v[1500:] = v[:1500]
    s.add_mesh(v, np.arange(3000, dtype=np.uint32).reshape(-1, 3))
    yield "duplicates", s


@pytest.mark.parametrize("leaf", [1, 4, 8])
@pytest.mark.parametrize("case", ["room", "terrain", "duplicates"])
def test_lbvh_invariants(case, leaf):
    from ordinarylight_amd import packing
    scene = dict(mesh_cases())[case]
    p = packing.pack_scene(scene)
    r = O.lbvh_build(p.positions, p.indices, leaf)
    T = len(p.indices)
    tri = p.positions[p.indices]
    mn, mx = tri.min(1), tri.max(1)
    cen = (mn + mx) * np.float32(0.5)
    lo, hi = cen.min(0), cen.max(0)
    # Morton codes: 10-bit quantised centroids, interleaved x:y:z.
    q = np.clip(np.floor(((cen - lo) / np.where(hi > lo, hi - lo, 1)) * np.float32(1024)), 0, 1023).astype(np.uint32)
    expect = np.array([(spread10(a) << 2) | (spread10(b) << 1) | spread10(c) for a, b, c in q], np.uint32)
    assert np.array_equal(r["morton"], expect)
    # Stable sort by code.
    order = r["sorted"]
    assert np.array_equal(order, np.argsort(expect, kind="stable"))
    codes = expect[order]
    # Karras tree: every leaf once, ranges nest, boxes contain children.
    left = [int(x) for x in r["left"]]
    right = [int(x) for x in r["right"]]
    box = r["box"]
    seen = np.zeros(T, int)

    def rng_of(ref):
        if ref & LEAF:
            k = ref & ~LEAF
            return k, k
        a, b = rng_of(left[ref])
        c, d = rng_of(right[ref])
        assert b + 1 == c, "children ranges must be adjacent"
        return a, d

    if T > 1:
        a, b = rng_of(0)
        assert (a, b) == (0, T - 1)
        for i in range(T - 1):
            for child in (left[i], right[i]):
                if child & LEAF:
                    k = child & ~LEAF
                    seen[k] += 1
                    cmn, cmx = mn[order[k]], mx[order[k]]
                else:
                    cmn, cmx = box[child, :3], box[child, 4:7]
                assert np.all(box[i, :3] <= cmn) and np.all(box[i, 4:7] >= cmx)
        assert np.all(seen == 1)
    # Wide nodes: decoded child boxes contain the exact boxes; leaves tile [0, T).
    nodes = decode_nodes(r["nodes"])

    def exact_bounds(node_idx):
        los, his = [], []
        for k in nodes[node_idx][0]:
            k = int(k)
            if k == 0xFFFFFFFF:
                continue
            if k & LEAF:
                f, n = k & ((1 << 27) - 1), ((k >> 27) & 15) + 1
                los.append(mn[order[f:f + n]].min(0))
                his.append(mx[order[f:f + n]].max(0))
            else:
                a, b = exact_bounds(k)
                los.append(a)
                his.append(b)
        return np.min(los, axis=0), np.max(his, axis=0)

    covered = np.zeros(T, int)
    stack = [(0, None)]
    depth_seen = 0
    while stack:
        idx, _ = stack.pop()
        children, boxes = nodes[idx]
        for c, ch in enumerate(children):
            ch = int(ch)
            if ch == 0xFFFFFFFF:
                continue
            lo_c, hi_c = boxes[c]
            if ch & LEAF:
                first, count = ch & ((1 << 27) - 1), ((ch >> 27) & 15) + 1
                assert count <= max(leaf, 1)
                covered[first:first + count] += 1
                exact_lo = mn[order[first:first + count]].min(0)
                exact_hi = mx[order[first:first + count]].max(0)
            else:
                assert ch > idx
                stack.append((ch, None))
                exact_lo, exact_hi = exact_bounds(ch)
            assert np.all(lo_c <= exact_lo) and np.all(hi_c >= exact_hi)
        depth_seen += 1
    assert np.all(covered == 1)
    # Leaf-order triangles: v0 (+ original id bits), e1, e2.
    t = r["tri"]
    assert np.array_equal(t[:, 0, 3].view(np.uint32), order)
    assert np.array_equal(t[:, 0, :3], tri[order, 0])
    assert np.array_equal(t[:, 1, :3], tri[order, 1] - tri[order, 0])
    assert np.array_equal(t[:, 2, :3], tri[order, 2] - tri[order, 0])
    assert r["depth"] >= 1 and codes.size == T

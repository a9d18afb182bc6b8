"""The reference's BBH tests (test/src/test_bbh.cpp:57-170) restated against
the oracle's restatement of generateSimpleBBH (src/models/BBH.cpp:38-115),
which bench.py times as the CPU baseline. The mesh: `count` unit triangles
spread 10 apart along x, as the reference's spreadTriangleMesh."""
import numpy as np
import pytest

import oracle


def spread(count):
    v, f = [], []
    for i in range(count):
        x = 10.0 * i
        b = len(v)
        v += [(x, 0, 0), (x + 1, 0, 0), (x, 1, 0)]
        f.append((b, b + 1, b + 2))
    return np.float32(v), np.int32(f)


def structural_depth(links, i=0):
    l, r = links[i, 0], links[i, 1]
    d = 1
    if l != -1:
        d = max(d, 1 + structural_depth(links, l))
    if r != -1:
        d = max(d, 1 + structural_depth(links, r))
    return d


def test_root_covers_all_triangles():
    v, f = spread(8)
    _, links, _ = oracle.RefBBH(v, f).nodes()
    assert links[0, 3] == 0 and links[0, 4] == 8 and links[0, 2] == -1


def test_leaf_ranges_partition_all_triangles():
    v, f = spread(13)
    _, links, order = oracle.RefBBH(v, f).nodes()
    leaves = sorted((tb, te) for l, r, _, tb, te in links if l == -1 and r == -1)
    assert all(tb < te for tb, te in leaves)
    pos = 0
    for tb, te in leaves:
        assert tb == pos
        pos = te
    assert pos == 13 and sorted(order) == list(range(13))


@pytest.mark.parametrize("count", [1, 2, 4, 7, 16, 31])
def test_depth_field_matches_structural_depth(count):
    v, f = spread(count)
    b = oracle.RefBBH(v, f)
    assert b.depth() == structural_depth(b.nodes()[1])


def test_root_splits_along_x_by_median():
    v, f = spread(8)
    _, links, order = oracle.RefBBH(v, f).nodes()
    l, r = links[0, 0], links[0, 1]
    assert l != -1 and r != -1
    maxx = lambda i: v[f[order[i]]][:, 0].max()   # noqa: E731
    left = max(maxx(i) for i in range(links[l, 3], links[l, 4]))
    right = min(maxx(i) for i in range(links[r, 3], links[r, 4]))
    assert left <= right and links[l, 4] - links[l, 3] == 4


def test_leaves_have_no_children_and_parents_link_back():
    v, f = spread(7)
    _, links, _ = oracle.RefBBH(v, f).nodes()
    for i, (l, r, p, tb, te) in enumerate(links):
        assert (l == -1 and r == -1) == (te - tb == 1)
        for c in (l, r):
            if c != -1:
                assert links[c, 2] == i


def test_boxes_on_each_depth():
    v, f = spread(4)
    boxes, links, _ = oracle.RefBBH(v, f).nodes()
    depth = np.zeros(len(links), int)
    for i, (l, r, *_rest) in enumerate(links):   # depth-first creation: parents before children
        for c in (l, r):
            if c != -1:
                depth[c] = depth[i] + 1
    assert [int(np.sum(depth == d)) for d in range(3)] == [1, 2, 4]
    assert tuple(links[0, 3:5]) == (0, 4)
    assert np.allclose(boxes[0], [0, 0, 0, 31, 1, 0])  # the root box covers every triangle

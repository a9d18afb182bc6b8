"""Pin the CPU oracle's geometry against the reference's own known-answer
tests, published vectors and brute force.

Sources of the known answers (Frontier789/DevilRay):
  test/src/test_intersection.cpp  -- triangle KATs (:33-200)
  test/src/test_tracing.cpp       -- box interval / box hit KATs (:75-131),
                                     counters (:302-357), BBH == brute force (:242-272)
Philox-4x32-10 vectors are the published known-answer vectors of Random123.
"""
import math

import numpy as np
import pytest

import devilray_amd as dr
import oracle

UNIT = ([0, 0, 0], [1, 0, 0], [0, 1, 0])


# ------------------------------------------------------------------ Philox
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert list(oracle.philox(ctr, key)) == expect


def test_u01_range():
    assert oracle.lib().or_u01(0) == 0.0
    assert oracle.lib().or_u01(0xFFFFFFFF) == pytest.approx(1.0 - 2 ** -24)
    assert oracle.lib().or_u01(0xFFFFFFFF) < 1.0


# ------------------------------------------------- triangle tests: both modes
def ref_tri(p, d, tri=UNIT):
    return oracle.tri_reference(p, d, *tri)


def wt_tri(p, d, tri=UNIT, tmin=0.0, tmax=math.inf):
    return oracle.tri_watertight([*p, tmin, *d, tmax], *tri)


@pytest.mark.parametrize("mode", ["reference", "watertight"])
class TestTriangleKATs:
    """test/src/test_intersection.cpp, run against both triangle tests."""

    def hit(self, mode, p, d, tri=UNIT):
        if mode == "reference":
            r = ref_tri(p, d, tri)
            return None if r is None else (r[0], r[1])
        r = wt_tri(p, d, tri)
        if r is None or r[0] < 0:  # the reference accepts t >= 0 only
            return None
        t, u, v = r
        return t, np.array([1 - u - v, u, v], np.float32)

    def test_center_hit(self, mode):  # :33-49
        t, b = self.hit(mode, [0.2, 0.2, 5], [0, 0, -1])
        assert t == pytest.approx(5.0, abs=1e-5)
        assert np.all(b >= 0) and b.sum() == pytest.approx(1.0, abs=1e-5)

    def test_vertex_hit(self, mode):  # :51-59
        t, _ = self.hit(mode, [0, 0, 10], [0, 0, -1])
        assert t == pytest.approx(10.0, abs=1e-5)

    @pytest.mark.parametrize("p,d", [
        ([-0.01, 0.5, 1], [0, 0, -1]),   # just outside an edge (:62-67)
        ([0.2, 0.2, 1], [0, 0, 1]),      # behind the origin (:69-74)
        ([-1, 0.2, 1], [1, 0, 0]),       # parallel (:77-82)
        ([0.2, 0.2, 1], [0, 0, 0]),      # zero direction (:123-129)
        ([0, 0, 0], [math.nan] * 3),     # NaN direction (:131-137)
    ])
    def test_misses(self, mode, p, d):
        assert self.hit(mode, p, d) is None

    def test_degenerate_point_and_collinear(self, mode):  # :84-90, :114-121
        assert self.hit(mode, [0, 0, 1], [0, 0, -1], ([0, 0, 0], [0, 0, 0], [0, 0, 0])) is None
        assert self.hit(mode, [0.5, 0, 1], [0, 0, -1], ([0, 0, 0], [1, 0, 0], [2, 0, 0])) is None

    def test_no_backface_culling(self, mode):  # :92-101
        assert self.hit(mode, [0.2, 0.2, -1], [0, 0, 1]) is not None

    def test_distant(self, mode):  # :103-110
        t, _ = self.hit(mode, [0.2, 0.2, 10000], [0, 0, -1])
        assert t == pytest.approx(10000.0, abs=1e-2)

    def test_sliver(self, mode):  # :139-148
        assert self.hit(mode, [50, 1e-9, 1], [0, 0, -1], ([0, 0, 0], [100, 0, 0], [100, 2e-8, 0])) is not None

    @pytest.mark.parametrize("origin", [[0.2, 0.2, 10], [10, 0.2, 10], [-5, -5, 2], [0, 0, 1]])
    def test_barycentrics_from_four_origins(self, mode, origin):  # :150-200
        o = np.array(origin, np.float32)
        r = self.hit(mode, o, np.array([0.2, 0.2, 0], np.float32) - o)
        assert r is not None
        t, b = r
        assert t == pytest.approx(1.0, abs=1e-4)
        assert b.sum() == pytest.approx(1.0, abs=1e-5)
        assert b[0] == pytest.approx(0.6, abs=1e-4)
        assert b[1] == pytest.approx(0.2, abs=1e-4)
        assert b[2] == pytest.approx(0.2, abs=1e-4)
        assert self.hit(mode, o, np.array([5, 5, 0], np.float32) - o) is None


def test_watertight_closed_form_plane_hits():
    """Axis-aligned rays onto a slanted triangle: t and barycentrics from the
    plane equation computed in float64."""
    rng = np.random.default_rng(5)
    tri = np.array([[0.3, -0.2, 1.0], [2.0, 0.1, 1.5], [0.1, 1.7, 0.6]], np.float32)
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0]).astype(np.float64)
    for _ in range(200):
        b = rng.dirichlet([1, 1, 1])
        target = b @ tri.astype(np.float64)
        o = np.array([target[0], target[1], 5.0], np.float32)
        r = wt_tri(o, [0, 0, -1], tuple(tri))
        assert r is not None
        t_exact = (n @ (tri[0] - o.astype(np.float64))) / (n @ np.array([0, 0, -1.0]))
        assert r[0] == pytest.approx(t_exact, rel=1e-5)
        assert r[1] == pytest.approx(b[1], abs=1e-4) and r[2] == pytest.approx(b[2], abs=1e-4)


def test_watertight_never_misses_shared_edges_and_vertices():
    """Rays from the centre of a closed icosphere aimed exactly at vertices
    and edge midpoints must hit (no cracks, Woop et al. 2013); the
    reference's plane test is not required to."""
    v, f = dr.gen_icosphere(3, 0.0, 0)
    targets = [v[i] for i in range(0, len(v), 7)]
    for t in f[::11]:
        targets.append(0.5 * (v[t[0]] + v[t[1]]))
    rays = np.zeros(len(targets), oracle.RAY_DTYPE)
    for i, tg in enumerate(targets):
        rays[i] = (0, 0, 0, 0, tg[0], tg[1], tg[2], np.inf)
    hits = oracle.brute_closest(v, f, rays)
    assert np.all(hits["id"] >= 0)


# ------------------------------------------------------------- box KATs
def test_box_interval_kats():  # test_tracing.cpp:75-87
    assert oracle.box_interval_ref(-1, 1, -5, 1) == (4.0, 6.0)
    assert oracle.box_interval_ref(-1, 1, 5, -1) == (4.0, 6.0)


@pytest.mark.parametrize("p,d,expect", [
    ([0, 0, -5], [0, 0, 1], 4.0),       # hits from outside (:91-96)
    ([0, 0, 0], [0, 0, 1], 0.0),        # origin inside clamps to 0 (:98-103)
    ([0, 0, 5], [0, 0, 1], None),       # box behind (:105-109)
    ([0.5, -0.5, -5], [0, 0, 1], 4.0),  # axis-parallel inside slabs (:111-117)
    ([5, 0, -5], [0, 0, 1], None),      # parallel outside slab (:119-124)
    ([-5, -5, 0], [1, 1, 0], 4.0),      # corner graze (:126-131)
])
def test_box_kats(p, d, expect):
    got = oracle.box_ref([-1, -1, -1], [1, 1, 1], p, d)
    assert got == expect


# ------------------------------------------------- BVH / traversal vs brute
def scattered(rng, n):
    """Random small triangles in a 10x10x4 slab (cf. scatteredTriangleMesh)."""
    c = rng.uniform([-5, -5, -2], [5, 5, 2], size=(n, 3))
    v = (c[:, None, :] + rng.uniform(-0.6, 0.6, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    f = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return v, f


def random_rays(rng, n, lo=-8, hi=8, tmax=np.inf):
    r = np.zeros(n, oracle.RAY_DTYPE)
    o = rng.uniform(lo, hi, size=(n, 3))
    tgt = rng.uniform(-4, 4, size=(n, 3))
    d = tgt - o
    r["ox"], r["oy"], r["oz"] = o.T
    r["dx"], r["dy"], r["dz"] = d.T
    r["tmax"] = tmax
    return r


@pytest.mark.parametrize("max_leaf", [1, 4, 8])
def test_sah_bvh_matches_brute_force(max_leaf):
    rng = np.random.default_rng(11)
    v, f = scattered(rng, 400)
    bvh = oracle.BVH(v, f, max_leaf=max_leaf)
    assert bvh.validate() == 0
    rays = random_rays(rng, 3000)
    a, b = bvh.closest(rays), oracle.brute_closest(v, f, rays)
    assert np.array_equal(a["id"], b["id"])
    hit = a["id"] >= 0
    assert np.array_equal(a["t"][hit].view(np.uint32), b["t"][hit].view(np.uint32))
    assert np.array_equal(bvh.any(rays), oracle.brute_any(v, f, rays))
    short = random_rays(rng, 3000, tmax=1.0)
    assert np.array_equal(bvh.any(short), oracle.brute_any(v, f, short))


def test_sah_bvh_on_sphere_and_counts():
    v, f = dr.gen_icosphere(4, 0.05, 0)
    bvh = oracle.BVH(v, f)
    assert bvh.validate() == 0
    rays = dr.gen_rays_surface(v, f, 4000, seed=9, inward=True)
    h, nt, tt = bvh.closest(rays, counts=True)
    assert np.array_equal(h["id"], oracle.brute_closest(v, f, rays)["id"])
    assert np.all(tt < f.shape[0]) and np.all(nt >= 1)


def test_reference_bbh_counters():  # test_tracing.cpp:302-357
    v = np.array(UNIT, np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    bbh = oracle.RefBBH(v, f)
    r = np.zeros(2, oracle.RAY_DTYPE)
    r[0] = (0.2, 0.2, 5, 0, 0, 0, -1, np.inf)
    r[1] = (5, 5, 5, 0, 0, 0, -1, np.inf)
    h, bt, tt = bbh.closest(r, counts=True)
    assert h["id"][0] == 0 and h["t"][0] == pytest.approx(5.0)
    assert (bt[0], tt[0]) == (1, 1)
    assert h["id"][1] == -1 and (bt[1], tt[1]) == (1, 0)
    rng = np.random.default_rng(3)
    v2, f2 = scattered(rng, 64)
    b2 = oracle.RefBBH(v2, f2)
    one = np.zeros(1, oracle.RAY_DTYPE)
    one[0] = (0, 0, 5, 0, 0, 0, -1, np.inf)
    _, bt2, tt2 = b2.closest(one, counts=True)
    assert bt2[0] > 0 and tt2[0] < 64


def test_reference_bbh_matches_brute_force_of_reference_test():
    """BBH traversal == linear scan with the same reference triangle test
    (test_tracing.cpp:242-272)."""
    rng = np.random.default_rng(7)
    v, f = scattered(rng, 40)
    bbh = oracle.RefBBH(v, f)
    rays = random_rays(rng, 500)
    h = bbh.closest(rays)
    for i in range(len(rays)):
        p = [rays["ox"][i], rays["oy"][i], rays["oz"][i]]
        d = [rays["dx"][i], rays["dy"][i], rays["dz"][i]]
        best = None
        for k in range(len(f)):
            r = oracle.tri_reference(p, d, v[f[k, 0]], v[f[k, 1]], v[f[k, 2]])
            if r is not None and (best is None or r[0] < best[0]):
                best = (r[0], k)
        if best is None:
            assert h["id"][i] == -1
        else:
            assert h["t"][i] == pytest.approx(best[0], rel=1e-6)


def test_reference_and_watertight_agree_off_edges():
    """The two triangle tests agree on ids away from edges and ties."""
    v, f = dr.gen_icosphere(4, 0.05, 0)
    rays = dr.gen_rays_surface(v, f, 3000, seed=21, inward=True)
    a = oracle.BVH(v, f).closest(rays)
    b = oracle.RefBBH(v, f).closest(rays)
    assert np.mean(a["id"] == b["id"]) > 0.995

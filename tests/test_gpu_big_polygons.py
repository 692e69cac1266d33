"""Candidate polygons of more than sixteen vertices on the GPU: their plane
and vertex triplet (the element the reference's std::sort ranking puts first,
from the left-most chain of its partitions with its heap-sort fallback, over a
global triplet scratch) and whole dispatches on meshes with large polygons,
against the oracle, whose triplet choice the reference pins for polygons of
17-40 and 256-330 vertices (tests/test_tripsort.py)."""
import os

import numpy as np
import pytest

import meshes
from mcut_amd import meshgen as mg

pytestmark = pytest.mark.gpu

G = np.load(os.path.join(meshes.GOLDEN, "ref_big_polygons.npz"))
H = np.load(os.path.join(meshes.GOLDEN, "ref_huge_polygons.npz"))


def polygon_pair(i, G=G):
    """Golden polygon i as a one-face source mesh whose polygon-soup vertex
    list is the polygon's own order (the soup lists a source face from its
    second vertex, so the face starts with the last one), cut by a triangle
    through its centroid across its plane."""
    n = int(G["n"][i])
    P = np.ascontiguousarray(G["poly"][i][:n])
    src = (P, np.array([n - 1] + list(range(n - 1)), np.uint32), np.array([n], np.uint32))
    nrm = G["normal"][i]
    c = P.mean(axis=0)
    u = np.cross(nrm, [0.3, 0.5, 0.8])
    u /= np.linalg.norm(u)
    T = np.array([c + 0.7 * nrm, c - 0.7 * nrm + 3.0 * u, c - 0.7 * nrm - 3.0 * u])
    return src, (T, np.array([0, 1, 2], np.uint32), None)


def prism(n, r=1.0, h=1.0, centre=(0.0, 0.0, 0.0), twist=0.1):
    """Closed prism over a regular n-gon: two n-gon caps and n quads."""
    a = 2 * np.pi * np.arange(n) / n + twist
    ring = np.stack([r * np.cos(a), r * np.sin(a), np.zeros(n)], 1)
    V = np.vstack([ring, ring + [0, 0, h]]) + np.asarray(centre) - [0, 0, h / 2]
    faces = [list(range(n - 1, -1, -1)), list(range(n, 2 * n))]
    for i in range(n):
        j = (i + 1) % n
        faces.append([i, j, n + j, n + i])
    return V, np.array([x for f in faces for x in f], np.uint32), np.array([len(f) for f in faces], np.uint32)


@pytest.fixture(scope="module")
def ctx():
    from mcut_amd import frontend
    c = frontend.Context(0)
    yield c
    c.close()


def test_big_polygon_triplets_match_oracle(ctx):
    """Every golden polygon through a dispatch: the GPU's plane and triplet
    equal the oracle's, whose triplet choice equals the reference's on these
    polygons (tests/test_tripsort.py; the dispatch shifts the coordinates, so
    the dispatch's triplets are compared with the oracle on the same shifted
    polygon)."""
    from oracle import oracle as om
    for i in range(len(G["n"])):
        src, cut = polygon_pair(i)
        R = ctx.dispatch(src, cut)
        O = om.frontend(src, cut)
        pl = R.planes()
        assert R.status == O["status"], i
        assert np.array_equal(pl["face"], O["cand_face"]), i
        assert np.array_equal(pl["normal"].view(np.uint64), O["plane_n"].view(np.uint64)), i
        assert np.array_equal(pl["triple"], O["plane_triple"]), (i, int(G["n"][i]))
        e, g, p = R.registry()
        assert np.array_equal(e, O["reg_edge"]) and np.array_equal(p.view(np.uint64), O["reg_point"].view(np.uint64)), i


def test_big_polygon_dispatches_against_oracle(ctx):
    from oracle import oracle as om
    cases = [polygon_pair(i) for i in range(0, len(G["n"]), 7)]
    for n in (17, 24, 33, 40, 64, 120):
        cases.append((prism(n, centre=(0.05, 0.02, 0.0)), mg.tessellated_box((0.6, 0.3, 0.2), 1.1, 3)))
    for k, (src, cut) in enumerate(cases):
        print(f"big polygon case {k}", flush=True)
        R = ctx.dispatch(src, cut)
        O = om.frontend(src, cut)
        assert R.status == O["status"], (k, R.status, O["status"])
        pl = R.planes()
        assert np.array_equal(pl["face"], O["cand_face"]) and np.array_equal(pl["triple"], O["plane_triple"]), k
        e, g, p = R.registry()
        assert np.array_equal(e, O["reg_edge"]) and np.array_equal(g, O["reg_face"]), k
        assert np.array_equal(p.view(np.uint64), O["reg_point"].view(np.uint64)), k


def test_big_polygons_in_a_batch(ctx):
    from mcut_amd import frontend
    pairs = [polygon_pair(i) for i in range(0, len(G["n"]), 3)]
    pairs.append((prism(48), mg.tessellated_box((0.6, 0.3, 0.2), 1.1, 3)))
    B = ctx.dispatch_batch(pairs)
    plain = frontend.Context(0)
    try:
        for i, (s, c) in enumerate(pairs):
            R = plain.dispatch(s, c)
            assert B.status[i] == R.status, i
            e, g, p = B.pair_registry(i)
            e0, g0, p0 = R.registry()
            assert np.array_equal(e, e0) and np.array_equal(g, g0), i
            assert np.array_equal(p.view(np.uint64), p0.view(np.uint64)), i
    finally:
        plain.close()


def test_polygons_past_255_vertices_match_oracle(ctx):
    """Polygons of 256-330 vertices (vertex indices past 8 bits, up to 6 M
    triplets each, whose triplet the reference pins for the oracle) and
    closed prisms with 300-gon and 600-gon caps, alone and (up to the
    300-gon prism) in a batch."""
    from mcut_amd import frontend
    from oracle import oracle as om
    cases = [polygon_pair(i, H) for i in range(len(H["n"]))]
    cases.append((prism(300, centre=(0.05, 0.02, 0.0)), mg.tessellated_box((0.6, 0.3, 0.2), 1.1, 3)))
    big = 0
    # and caps of 600 vertices: 35.8 M triplets a cap, many more partitions deep
    cases = cases + [(prism(600, centre=(0.05, 0.02, 0.0)), mg.tessellated_box((0.6, 0.3, 0.2), 1.1, 3))]
    for k, (src, cut) in enumerate(cases):
        print(f"polygon case {k}: {int(np.max(src[2]))} vertices", flush=True)
        R = ctx.dispatch(src, cut)
        O = om.frontend(src, cut)
        assert R.status == O["status"], (k, R.status, O["status"])
        pl = R.planes()
        assert np.array_equal(pl["face"], O["cand_face"]), k
        assert np.array_equal(pl["normal"].view(np.uint64), O["plane_n"].view(np.uint64)), k
        assert np.array_equal(pl["triple"], O["plane_triple"]), k
        big += int((pl["triple"] > 255).any())
        e, g, p = R.registry()
        assert np.array_equal(e, O["reg_edge"]) and np.array_equal(g, O["reg_face"]), k
        assert np.array_equal(p.view(np.uint64), O["reg_point"].view(np.uint64)), k
    assert big >= 4  # triplets that need more than 8 bits per index
    batch = cases[:3] + cases[-2:-1]
    B = ctx.dispatch_batch(batch)
    plain = frontend.Context(0)
    try:
        for i, (s, c) in enumerate(batch):
            R = plain.dispatch(s, c)
            assert B.status[i] == R.status, i
            e, g, p = B.pair_registry(i)
            e0, g0, p0 = R.registry()
            assert np.array_equal(e, e0) and np.array_equal(g, g0), i
            assert np.array_equal(p.view(np.uint64), p0.view(np.uint64)), i
    finally:
        plain.close()

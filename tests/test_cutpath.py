"""Cut-path edges (kernel.cpp:3427-3615): the edges between intersection points
along which each (source face, cut face) pair is cut, in the order MCUT makes
them.

* The oracle's restatement (oracle/frontend_oracle.c: om_cutpath) against
  MCUT itself: mcut_amd/seam/_build/libmcut_probe.so is MCUT compiled
  unchanged with a recorder called right after it makes these edges
  (cutpath_probe.cpp), which also records its registry in its own point
  order; the oracle, given that order, must make the same edges, source and
  target, in the same order.
* The GPU (mcgpu_cutpath_edges) against the oracle, in registry order and in
  MCUT's order, and at full size (-m gpu).
* MCUT with its front end and cut-path edges on the GPU (libmcut_gpu.so)
  against the recording MCUT: same registry order, same edges, same per-face
  edge lists (-m gpu)."""
import os

import numpy as np
import pytest

import meshes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def mcut_order(rec, reg_edge, reg_face):
    """order[k] = index in (reg_edge, reg_face) of MCUT's point k; None when
    the two registries are not the same set (MCUT partitioned a floating
    polygon of its soup, which renumbers edges)."""
    idx = {(int(e), int(f)): i for i, (e, f) in enumerate(zip(reg_edge, reg_face))}
    try:
        order = [idx[(int(e), int(f))] for e, f in zip(rec["reg_edge"], rec["reg_face"])]
    except KeyError:
        return None
    if len(order) != len(idx):
        return None
    return np.array(order, np.uint32)


def cases():
    yield "config1_rotated", meshes.config1_rotated()
    for i in range(64):
        yield f"bench{i:03d}", meshes.bench_pair(i)
    yield "icosphere3", meshes.icosphere_pair(3, amp=0.02)
    yield "icosphere4", meshes.icosphere_pair(4, amp=0.005)


@pytest.fixture(scope="module")
def probe():
    from mcut_amd.seam import mcut_api
    if not os.path.exists(mcut_api.PROBE_LIB):
        pytest.skip("the recording MCUT build is missing (make -C mcut_amd/seam)")
    return mcut_api.Mcut(mcut_api.PROBE_LIB)


def record(probe, pair):
    from mcut_amd.seam import mcut_api
    before = probe.cutpath_record()["dispatches"]
    try:
        probe.dispatch(*pair)
    except mcut_api.McutError:
        pass  # MCUT may fail after the cut-path edges (its cutting stage)
    rec = probe.cutpath_record()
    return rec if rec["dispatches"] > before else None


def test_oracle_makes_mcuts_cutpath_edges(probe):
    from oracle import oracle as om
    compared = longer = 0
    for name, pair in cases():
        rec = record(probe, pair)
        if rec is None:
            continue  # MCUT stopped before its cut-path edges
        O = om.frontend(*pair, detail=False)
        order = mcut_order(rec, O["reg_edge"], O["reg_face"])
        if order is None:
            continue
        c = om.frontend(*pair, detail=False, cutpath=order)["cutpath"]
        assert c["short_group"] == 0, name
        assert np.array_equal(c["src"], rec["src"]) and np.array_equal(c["tgt"], rec["tgt"]), name
        compared += 1
        longer += int(np.count_nonzero(c["kinds"]))
    # 31 of the 67 cases (the others: MCUT stops before its cut-path edges, or
    # it partitioned a floating polygon of its soup, which renumbers edges
    # and leaves no common registry to compare in)
    assert compared >= 30 and longer > 0  # edges of longer groups (the midpoint tests) are among them


def test_order_matters(probe):
    """The same registry in another order makes other edges in another order
    (a group's first two points set its sort direction and a two-point
    group's orientation): the comparison above is sensitive to the order."""
    from oracle import oracle as om
    pair = meshes.icosphere_pair(4, amp=0.005)
    a = om.frontend(*pair, detail=False, cutpath=True)["cutpath"]
    n = a["src"].size
    rev = np.arange(n, dtype=np.uint32)[::-1].copy()
    O = om.frontend(*pair, detail=False)
    b = om.frontend(*pair, detail=False, cutpath=rev[:O["reg_edge"].size])["cutpath"]
    assert b["src"].size == n
    assert not np.array_equal(a["src"], rev[b["src"]])  # not just a renaming


@pytest.mark.gpu
def test_gpu_matches_oracle_in_registry_order():
    from mcut_amd import frontend
    from oracle import oracle as om
    ctx = frontend.Context(0)
    try:
        for name, pair in list(cases())[:40] + [("icosphere5", meshes.icosphere_pair(5, amp=0.005))]:
            print(f"cutpath case {name}", flush=True)
            R = ctx.dispatch(*pair)
            O = om.frontend(*pair, detail=False, cutpath=True)
            if R.status != 0 or O["status"] != 0 or R.counts["registry"] == 0:
                continue
            G, c = R.cutpath(), O["cutpath"]
            assert G["short_group"] == c["short_group"] and G["unsupported"] == 0, name
            for k in ("src", "tgt", "faces", "groups", "kinds"):
                assert np.array_equal(G[k], c[k]), (name, k)
    finally:
        ctx.close()


@pytest.mark.gpu
def test_gpu_matches_mcut_in_mcuts_order(probe):
    from mcut_amd import frontend
    ctx = frontend.Context(0)
    try:
        n = 0
        for name, pair in cases():
            rec = record(probe, pair)
            if rec is None:
                continue
            R = ctx.dispatch(*pair)
            e, g, _ = R.registry()
            order = mcut_order(rec, e, g)
            if order is None:
                continue
            print(f"cutpath in MCUT's order: {name}", flush=True)
            G = R.cutpath(order)
            assert np.array_equal(G["src"], rec["src"]) and np.array_equal(G["tgt"], rec["tgt"]), name
            n += 1
        assert n >= 30
    finally:
        ctx.close()


@pytest.mark.gpu
def test_gpu_cutpath_full_size():
    """config 2 (two icosphere-7 meshes, 19,838 points): GPU == oracle."""
    from mcut_amd import frontend, meshgen as mg
    from oracle import oracle as om
    src, cut = mg.config2()
    ctx = frontend.Context(0)
    try:
        R = ctx.dispatch(src, cut)
        G = R.cutpath()
        O = om.frontend(src, cut, detail=False, cutpath=True)
        c = O["cutpath"]
        print(f"config 2: {G['src'].size} cut-path edges in {G['device_ms']:.3f} ms", flush=True)
        assert G["src"].size == R.counts["registry"]  # closed intersection curves: one edge per point
        for k in ("src", "tgt", "faces", "groups", "kinds"):
            assert np.array_equal(G[k], c[k]), k
    finally:
        ctx.close()


@pytest.mark.gpu
def test_seam_cutpath_edges_match_mcut(probe):
    """MCUT with front end and cut-path edges on the GPU against MCUT
    unchanged: registry order, edges and every face's edge list identical."""
    from mcut_amd.seam import mcut_api
    if not os.path.exists(mcut_api.GPU_LIB):
        pytest.skip("the GPU front-end build is missing")
    gpu = mcut_api.Mcut(mcut_api.GPU_LIB)
    n = 0
    for name, pair in list(cases()) + [("icosphere5", meshes.icosphere_pair(5, amp=0.005))]:
        a = record(probe, pair)
        b = record(gpu, pair)
        assert (a is None) == (b is None), name
        if a is None:
            continue
        print(f"seam cutpath case {name}", flush=True)
        for k in ("reg_edge", "reg_face", "src", "tgt", "face_edges"):
            assert np.array_equal(a[k], b[k]), (name, k)
        n += 1
    assert n >= 55

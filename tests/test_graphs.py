"""Graph replay of dispatches (mcgpu_set_graphs): every replay must give
exactly what an eager dispatch gives -- status, counts, candidate pairs,
registry edges, faces and point bits -- with eager work and synchronisations
between the replays; a new shape records anew, and new data behind the same
shape replays."""
import numpy as np
import pytest

import meshes
from mcut_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def snapshot(R):
    e, g, p = R.registry()
    return dict(status=R.status, counts=R.counts, pairs=R.pairs(), e=e, g=g, p=p.view(np.uint64))


def same(a, b):
    assert a["status"] == b["status"] and a["counts"] == b["counts"]
    for k in ("pairs", "e", "g", "p"):
        assert np.array_equal(a[k], b[k]), k


@pytest.fixture(scope="module")
def ctxs():
    from mcut_amd import frontend
    eager, graph = frontend.Context(0), frontend.Context(0)
    graph.set_graphs(True)
    yield eager, graph
    eager.close()
    graph.close()


CASES = {"config1_rotated": meshes.config1_rotated, "bench005": lambda: meshes.bench_pair(5),
         "bench020": lambda: meshes.bench_pair(20), "icosphere5": lambda: meshes.icosphere_pair(5, amp=0.005)}


@pytest.mark.parametrize("name", list(CASES))
def test_replays_match_eager(ctxs, name):
    eager, graph = ctxs
    pair = CASES[name]()
    print(f"graphs: replays of {name}", flush=True)
    ref = snapshot(eager.dispatch(*pair))
    before = graph.graph_stats()
    for i in range(4):
        got = snapshot(graph.dispatch(*pair))
        same(ref, got)
        # eager work of the program's own between the replays
        same(ref, snapshot(eager.dispatch(*pair)))
    st = graph.graph_stats()
    assert st["recordings"] == before["recordings"] + 1  # recorded once for this shape
    assert st["replays"] == before["replays"] + 4
    assert st["graph_launches"] > 0


def test_new_data_same_shape_replays(ctxs):
    """Host input is uploaded into the context's own buffers, so moved
    coordinates keep the shape: the program replays on the new data."""
    eager, graph = ctxs
    src, cut = meshes.icosphere_pair(4, amp=0.005)
    graph.dispatch(src, cut)
    rec = graph.graph_stats()["recordings"]
    rng = np.random.default_rng(3)
    for k in range(3):
        V = np.array(cut[0], np.float64) + rng.uniform(-0.05, 0.05, size=3)
        moved = (V, cut[1], cut[2])
        ref = snapshot(eager.dispatch(src, moved))
        same(ref, snapshot(graph.dispatch(src, moved)))
    assert graph.graph_stats()["recordings"] == rec


def test_shape_change_records_anew(ctxs):
    eager, graph = ctxs
    a, b = meshes.bench_pair(1), meshes.bench_pair(2)
    rec0 = graph.graph_stats()["recordings"]
    for pair in (a, b, a):
        same(snapshot(eager.dispatch(*pair)), snapshot(graph.dispatch(*pair)))
    assert graph.graph_stats()["recordings"] >= rec0 + 2


def test_batch_replays_match_eager():
    from mcut_amd import frontend
    pairs = mg.config4_pairs(16, seed=23, tri_range=(200, 20_000), box_range=(12, 3_000))
    pairs += [meshes.bench_pair(i) for i in (0, 14, 33)]
    print("graphs: batch of 19 pairs", flush=True)
    eager, graph = frontend.Context(0), frontend.Context(0)
    try:
        graph.set_graphs(True)
        ref = eager.dispatch_batch(pairs)
        ref_regs = [tuple(np.ascontiguousarray(x).view(np.uint8) for x in ref.pair_registry(i))
                    for i in range(len(pairs))]
        ref_status = list(ref.status)
        for _ in range(3):
            R = graph.dispatch_batch(pairs)
            assert list(R.status) == ref_status
            for i in range(len(pairs)):
                got = tuple(np.ascontiguousarray(x).view(np.uint8) for x in R.pair_registry(i))
                assert all(np.array_equal(x, y) for x, y in zip(ref_regs[i], got)), i
        st = graph.graph_stats()
        assert st["recordings"] == 1 and st["replays"] >= 3
    finally:
        eager.close()
        graph.close()


def test_stage_timing_runs_eagerly():
    from mcut_amd import frontend
    c = frontend.Context(0)
    try:
        c.set_graphs(True)
        c.set_stage_timing(True)
        pair = meshes.bench_pair(5)
        R = c.dispatch(*pair)
        assert R.stage_ms()["traversal"] > 0
        assert c.graph_stats()["replays"] == 0
    finally:
        c.close()


def test_far_from_origin_pair_replays(ctxs):
    """A pair far from the origin (a sphere and a box of about 130 units near
    2^52) runs its broad phase a second time in shifted coordinates: every
    dispatch is two programs, the raw one and the shifted one, and every
    replay equals the eager dispatch.  A context keeps one program, so the
    two evict each other and each dispatch records both."""
    eager, graph = ctxs
    f = lambda m: (np.asarray(m[0], np.float64) * 60.0 + 2.0 ** 52, m[1], m[2])
    pair = (f(mg.uv_sphere(16, 8)), f(mg.tessellated_box((0.2, 0.1, 0.05), 1.3, 3, rotation=mg.random_rotation(4))))
    print("graphs: far-from-origin pair", flush=True)
    ref = snapshot(eager.dispatch(*pair))
    assert ref["status"] == 0 and ref["counts"]["registry"] > 0
    before = graph.graph_stats()
    for i in range(3):
        same(ref, snapshot(graph.dispatch(*pair)))
        same(ref, snapshot(eager.dispatch(*pair)))
    st = graph.graph_stats()
    assert st["recordings"] == before["recordings"] + 6
    assert st["replays"] == before["replays"] + 6

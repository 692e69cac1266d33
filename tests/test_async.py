"""The enqueued dispatch (mcgpu_dispatch_frontend_async / _query / _wait,
after mcEnqueueDispatch, frontend.cpp:690-706): the call returns with the work
queued, the host is free until the wait, and the wait gives exactly what the
blocking dispatch gives, including a dispatch whose lists overflow and are
grown and re-run inside the wait."""
import time

import numpy as np
import pytest

import meshes
from mcut_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def fan(n=300):
    """An open fan of n long triangles from the origin to a unit circle at
    x = 1: their boxes overlap one another wholesale."""
    t = np.linspace(0.0, 2.0 * np.pi, n, endpoint=False)
    V = np.concatenate([[[0.0, 0.0, 0.0]], np.stack([np.ones(n), np.cos(t), np.sin(t)], 1)])
    F = np.array([[0, 1 + i, 1 + (i + 1) % n] for i in range(n)], np.uint32).ravel()
    return V, F, None


def crowded_pair():
    """A fan against a sphere crossing its cone: about 68k candidate
    pairs for 1,580 faces, past the 16,384 a fresh context's pair list holds."""
    V, F, S = mg.icosphere(3)
    V = np.asarray(V) * 0.55 + np.array([0.5, 0.05, 0.03])
    return fan(), (V, np.asarray(F, np.uint32), S)


def far_pair():
    """A sphere and a box of about 130 units near 2^52: the raw-space
    tolerance is loose there, so the dispatch runs its broad phase a second
    time in shifted coordinates (inside the wait, for the enqueued form)."""
    f = lambda m: (np.asarray(m[0], np.float64) * 60.0 + 2.0 ** 52, m[1], m[2])
    src = mg.uv_sphere(16, 8)
    cut = mg.tessellated_box((0.2, 0.1, 0.05), 1.3, 3, rotation=mg.random_rotation(4))
    return f(src), f(cut)


def snapshot(R):
    e, g, p = R.registry()
    return dict(status=R.status, counts=R.counts, pairs=R.pairs(), e=e, g=g, p=p.view(np.uint64))


def same(a, b):
    assert a["status"] == b["status"] and a["counts"] == b["counts"]
    for k in ("pairs", "e", "g", "p"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", ["config1_rotated", "icosphere5", "crowded", "far"])
def test_async_equals_blocking(name):
    from mcut_amd import frontend
    pair = {"config1_rotated": meshes.config1_rotated, "icosphere5": lambda: meshes.icosphere_pair(5, amp=0.005),
            "crowded": crowded_pair, "far": far_pair}[name]()
    print(f"async case {name}", flush=True)
    a, b = frontend.Context(0), frontend.Context(0)
    try:
        ref = snapshot(a.dispatch(*pair))
        b.dispatch_async(*pair)
        host = sum(range(10000))  # host work while the GPU runs
        t0 = time.time()
        while not b.query() and time.time() - t0 < 30.0:
            time.sleep(0.0005)
        R = b.wait()
        assert host == 49995000
        same(ref, snapshot(R))
    finally:
        a.close()
        b.close()


def test_crowded_pair_overflows_and_matches_the_oracle():
    """The crowded pair does overflow a fresh context's lists (more raw pairs
    than the first capacity), and the grown re-run equals the oracle."""
    from mcut_amd import frontend
    from oracle import oracle as om
    src, cut = crowded_pair()
    ctx = frontend.Context(0)
    try:
        R = ctx.dispatch(src, cut)
        assert R.counts["raw_pairs"] > 16384
        O = om.frontend(src, cut, detail=False)
        assert R.status == O["status"]
        assert np.array_equal(R.pairs(), O["pairs"])
        e, g, p = R.registry()
        assert np.array_equal(e, O["reg_edge"]) and np.array_equal(g, O["reg_face"])
        assert np.array_equal(p.view(np.uint64), O["reg_point"].view(np.uint64))
    finally:
        ctx.close()


def test_pending_context_refuses_other_calls():
    from mcut_amd import frontend
    ctx = frontend.Context(0)
    try:
        pair = meshes.config1_rotated()
        ctx.dispatch_async(*pair)
        with pytest.raises(frontend.McgpuError) as ei:
            ctx.dispatch(*pair)
        assert ei.value.code == -1
        with pytest.raises(frontend.McgpuError) as ei:
            ctx.dispatch_batch([pair])
        assert ei.value.code == -1
        R = ctx.wait()
        assert R.status == 0 and R.counts["registry"] == 134
        with pytest.raises(frontend.McgpuError):
            ctx.wait()  # nothing pending
    finally:
        ctx.close()

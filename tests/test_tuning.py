"""Traversal launch shapes and the autotuner (mcgpu_set_tuning,
mcgpu_autotune): every shape -- 8, 16 or 32 source leaves per block, 2, 3 or
4 cut-tree levels per step, one wave per block (in Morton order or with each
XCD's blocks one contiguous Morton range) or a persistent grid over the eight
per-XCD work queues -- must leave every output of a dispatch unchanged:
status, counts, candidate pairs, registry edges, faces and point bits.  The
default shape is held to the oracle by test_gpu_parity.py."""
import numpy as np
import pytest

import meshes
from mcut_amd import meshgen as mg

pytestmark = pytest.mark.gpu

# the autotuner's candidates, and the shapes held to the default's outputs
# (those and runs of one block per XCD)
TUNED = [(ll, st, q, ch) for ll in (8, 16, 32) for st in (2, 3, 4) for q, ch in ((0, 1), (1, 16), (8, 1), (8, 4))]
SHAPES = TUNED + [(ll, st, 1, 1) for ll in (8, 16, 32) for st in (2, 3, 4)]


@pytest.fixture(scope="module")
def ctx():
    from mcut_amd import frontend
    c = frontend.Context(0)
    yield c
    c.close()


def snapshot(R):
    e, g, p = R.registry()
    return dict(status=R.status, counts=R.counts, pairs=R.pairs(), e=e, g=g, p=p.view(np.uint64))


def same(a, b):
    assert a["status"] == b["status"] and a["counts"] == b["counts"]
    for k in ("pairs", "e", "g", "p"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", ["config1_rotated", "bench005", "icosphere5"])
def test_every_shape_gives_the_same_dispatch(ctx, name):
    pair = {"config1_rotated": meshes.config1_rotated, "bench005": lambda: meshes.bench_pair(5),
            "icosphere5": lambda: meshes.icosphere_pair(5, amp=0.005)}[name]()
    print(f"tuning: shapes on {name}", flush=True)
    ctx.set_tuning(None)
    ref = snapshot(ctx.dispatch(*pair))
    t = ctx.tuning()
    assert (t["wave_leaves"], t["tree_stride"], t["queues"], t["chunk"]) == (32, 4, 1, 16)
    try:
        for ll, st, q, ch in SHAPES:
            ctx.set_tuning(wave_leaves=ll, tree_stride=st, queues=q, chunk=ch)
            got = snapshot(ctx.dispatch(*pair))
            t = ctx.tuning()
            assert (t["wave_leaves"], t["tree_stride"], t["queues"], t["chunk"]) == (ll, st, q, ch)
            same(ref, got)
    finally:
        ctx.set_tuning(None)


@pytest.mark.skip(reason="changing a context's CU split (new CU-masked stream pair, old pair destroyed) hung a "
                         "dispatch on the GPU once, cause not found; the split is no longer settable")
def test_split_override_gives_the_same_dispatch(ctx):
    from mcut_amd.frontend import lib, Tuning
    import ctypes as C
    pair = meshes.icosphere_pair(5, amp=0.005)
    ctx.set_tuning(None)
    ref = snapshot(ctx.dispatch(*pair))
    try:
        for pct in (40, 85):
            t = Tuning(wave_leaves=16, tree_stride=3, queues=8, chunk=1, broad_share_pct=pct)
            assert lib().mcgpu_set_tuning(ctx.h, C.byref(t)) == 0
            same(ref, snapshot(ctx.dispatch(*pair)))
    finally:
        ctx.set_tuning(None)


def test_bad_shapes_are_refused(ctx):
    from mcut_amd import frontend
    for kw in (dict(wave_leaves=12), dict(wave_leaves=64), dict(wave_leaves=16, tree_stride=5),
               dict(wave_leaves=16, queues=4), dict(wave_leaves=16, chunk=0)):
        with pytest.raises(frontend.McgpuError) as ei:
            ctx.set_tuning(**kw)
        assert ei.value.code == -1
    from mcut_amd.frontend import lib, Tuning
    import ctypes as C
    t = Tuning(wave_leaves=16, tree_stride=3, broad_share_pct=40)  # the split is not settable
    assert lib().mcgpu_set_tuning(ctx.h, C.byref(t)) == -1
    ctx.set_tuning(None)


def test_autotune_single_pair(ctx):
    pair = meshes.icosphere_pair(5, amp=0.005)
    print("tuning: autotune icosphere-5", flush=True)
    ref = snapshot(ctx.dispatch(*pair))
    rep = ctx.autotune()
    assert rep["candidates"] == len(TUNED)
    assert rep["wave_leaves"] in (8, 16, 32) and rep["tree_stride"] in (2, 3, 4) and rep["queues"] in (0, 1, 8)
    assert 0 < rep["traversal_ms"] <= rep["default_traversal_ms"] and rep["dispatch_ms"] > 0
    # the class's later dispatches run with the choice, unchanged in result
    R = ctx.dispatch(*pair)
    t = ctx.tuning()
    assert t["candidates"] == len(TUNED) and t["wave_leaves"] == rep["wave_leaves"]
    assert t["broad_share_pct"] == rep["broad_share_pct"]
    same(ref, snapshot(R))


def test_autotune_batch():
    """A batch tuned by its first dispatch (mcgpu_set_autotune): per-pair
    outputs unchanged, and the second dispatch runs with the choice."""
    from mcut_amd import frontend
    pairs = mg.config4_pairs(24, seed=17, tri_range=(200, 20_000), box_range=(12, 3_000))
    pairs += [meshes.bench_pair(i) for i in (0, 5, 14)]
    print("tuning: autotune a 27-pair batch", flush=True)
    c = frontend.Context(0)
    try:
        ref = c.dispatch_batch(pairs)
        ref_regs = [ref.pair_registry(i) for i in range(len(pairs))]
        ref_status = list(ref.status)
        c.set_autotune(True)

        def check(R):  # a result's arrays are read from the context: before its next dispatch
            assert list(R.status) == ref_status
            for i in range(len(pairs)):
                for x, y in zip(ref_regs[i], R.pair_registry(i)):
                    assert np.array_equal(np.ascontiguousarray(x).view(np.uint8),
                                          np.ascontiguousarray(y).view(np.uint8)), i

        check(c.dispatch_batch(pairs))
        t = c.tuning()
        assert t["candidates"] == len(TUNED)
        check(c.dispatch_batch(pairs))
        assert c.tuning() == t
    finally:
        c.close()

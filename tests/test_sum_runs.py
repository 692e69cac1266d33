"""Coordinate sums from run items (k_sum_blocks .. k_sum_verify).

The reference sums every coordinate of a mesh left to right (preproc.cpp:
2188-2202), and every shifted bit depends on those sums.  The GPU adds a
same-binade run of steps as one exact item when a verified interval proves it
exact, and falls back to the serial sum for a coordinate that fails.

CPU: a restatement of the decomposition (runs of steps whose predicted result
binade equals the previous one, the run's exact sum of RN_u(x), the start
interval that keeps every partial sum in [2^e + u, 2^(e+1)]) must give,
whenever its intervals hold, the left-to-right float sum bit for bit -- also
when the predictions are deliberately wrong, where the intervals must reject.

GPU: the dispatch's centre (sum / n of each mesh) bit for bit against the
oracle on vertex arrays built to stress the runs: ties, sums hovering about
zero, exactly representable grids, tiny and huge scales, long runs in one
binade."""
import math

import numpy as np
import pytest

import meshes


def _binade(a):
    if a == 0.0 or abs(a) < 2.0 ** -900 or abs(a) >= 2.0 ** 1001:
        return None
    return (math.frexp(abs(a))[1] - 1, a < 0)


def _items(x, approx, span):
    """(items, intervals) for values x; approx[i] is the predicted partial sum
    after x[i]; runs never cross a multiple of span (a thread range)."""
    items, iv = [], []
    run = None  # [key, q, qmin, qmax]

    def close():
        nonlocal run
        if run is None:
            return
        (e, neg), q, qmin, qmax = run
        u = math.ldexp(1.0, e - 52)
        L = -(1 << 53) if neg else (1 << 52) + 1
        H = -((1 << 52) + 1) if neg else (1 << 53)
        items.append(q * u)
        iv.append(((L - qmin) * u, (H - qmax) * u))  # exact: Python ints, then scaled
        run = None

    prev = None
    for i, xi in enumerate(x):
        if i % span == 0:
            close()
            prev = _binade(approx[i - 1] if i else 0.0)
        key = _binade(approx[i])
        ok = key is not None and key == prev
        if ok:
            u = math.ldexp(1.0, key[0] - 52)
            q = xi / u
            if abs(xi) >= math.ldexp(1.0, key[0] + 3) or q - math.floor(q) == 0.5:
                ok = False
        if ok:
            c = int(round(q))  # round half even; ties excluded above
            if run is None or run[0] != key:
                close()
                run = [key, 0, 0, 0]
            run[1] += c
            run[2] = min(run[2], run[1])
            run[3] = max(run[3], run[1])
        else:
            close()
            items.append(float(xi))
            iv.append((-math.inf, math.inf))
        prev = key
    close()
    return items, iv


def _chain(items, iv):
    """Left-to-right sum of the items; False when a run's start leaves its interval."""
    s = 0.0
    for it, (lo, hi) in zip(items, iv):
        if not (lo <= s <= hi):
            return s, False
        s = s + it
    return s, True


def _seq(x):
    s = 0.0
    for v in x:
        s = s + v
    return s


def _cases():
    rng = np.random.default_rng(3)
    out = {
        "sphere_x": meshes.icosphere_pair(5)[0][0][:, 0],
        "hover": rng.normal(size=6000),
        "positive": rng.uniform(0.5, 1.5, 6000),
        "grid": np.round(rng.uniform(-3, 3, 6000) * 4096) / 4096,
        "ties": (2.0 ** 52) + rng.integers(0, 4, 6000) + 0.5,
        "tiny": rng.normal(size=4000) * 1e-290,
        "huge": rng.normal(size=4000) * 1e290,
        "mixed": rng.normal(size=4000) * 10.0 ** rng.integers(-8, 8, 4000),
    }
    return out


@pytest.mark.parametrize("name", list(_cases()))
@pytest.mark.parametrize("span", [32, 7])
def test_runs_reproduce_the_sequential_sum(name, span):
    x = np.asarray(_cases()[name], dtype=np.float64)
    approx = np.cumsum(x)  # a prediction; any order would do
    items, iv = _items(x.tolist(), approx.tolist(), span)
    s, ok = _chain(items, iv)
    ref = _seq(x.tolist())
    if name == "ties":  # sums landing a unit off 2^e: a run may be rejected (the serial sum then runs)
        assert not ok or s == ref
        return
    assert ok, name
    assert s == ref, (name, s, ref)
    assert len(items) <= x.size


@pytest.mark.parametrize("name", ["hover", "positive", "mixed"])
def test_wrong_predictions_are_rejected_or_exact(name):
    """Predictions pushed off by large amounts: the intervals either reject the
    run (a failure the GPU resolves with the serial sum) or the result is exact."""
    x = np.asarray(_cases()[name], dtype=np.float64)
    rng = np.random.default_rng(5)
    for k in range(20):
        approx = np.cumsum(x) * (1.0 + rng.normal() * 10.0 ** rng.integers(-12, 0)) + rng.normal() * 50
        items, iv = _items(x.tolist(), approx.tolist(), 32)
        s, ok = _chain(items, iv)
        if ok:
            assert s == _seq(x.tolist()), (name, k)


def _remap(mesh, f):
    V, F, S = mesh
    return (f(np.array(V, dtype=np.float64)), F, S)


# Maps that keep the mesh's shape (the broad phase sees the same overlaps) and
# stress the runs: 30 fractional bits on a 1024 offset make a tie in about one
# step in eight once the partial sums pass 2^23.
GPU_CASES = {
    "ties_fractional_bits": lambda V: np.round(V * 2.0 ** 30) / 2.0 ** 30 + 1024.0,
    "hover_about_zero": lambda V: V - np.array([0.25, 0.0, 0.0]),
    "grid": lambda V: np.round(V * 4096) / 4096,
    "positive_runs": lambda V: V + 3.0,
    "scaled_offset": lambda V: V * 1000.0 + 5000.0,
    "small_scale": lambda V: V * 1e-3,
}


@pytest.fixture(scope="module")
def ctx():
    from mcut_amd import frontend
    c = frontend.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(GPU_CASES))
def test_centre_bits_on_adversarial_coordinates(ctx, name):
    from oracle import oracle as om
    f = GPU_CASES[name]
    src, cut = meshes.icosphere_pair(6, amp=0.01)
    src, cut = _remap(src, f), _remap(cut, f)
    print(f"sums: {name}", flush=True)
    print("sums: collapsed icosphere-6 pair", flush=True)
    R = ctx.dispatch(src, cut)
    O = om.frontend(src, cut, detail=False)
    assert np.array_equal(R.com.view(np.uint64), np.asarray(O["com"]).view(np.uint64))
    assert R.status == O["status"] and R.counts["pairs"] == O["counts"]["pairs"], (R.com, O["com"])
    assert np.array_equal(R.t.view(np.uint64), np.asarray(O["t"]).view(np.uint64))
    assert R.status == O["status"]


def test_collapsed_mesh_raw_boxes_stay_exact():
    """The icosphere-6 pair mapped by V * 4 + 2^52 + 0.5 collapses onto a few
    lattice points.  Rounded to float32 in the raw frame (neighbours 2^28
    apart near 2^52) every face box of the pair was the same box, and the raw
    broad phase counted 81,920^2 pairs, sizing lists of hundreds of GB.  The
    raw boxes are now taken relative to the source mesh's raw lower corner:
    there every coordinate is a small multiple of 1/2, exact in float32, so
    the raw-space pairs are the exact pairs widened only by the tolerance."""
    f = lambda V: V * 4 + 2.0 ** 52 + 0.5
    src, cut = meshes.icosphere_pair(6, amp=0.01)
    S, C = f(src[0]), f(cut[0])
    corner = S.min(axis=0)
    for X in (S, C):
        rel = X - corner
        assert np.all(rel.astype(np.float32).astype(np.float64) == rel)
        assert rel.max() < 16.0
    assert np.unique(S.astype(np.float32)).size <= 2  # the raw frame's float32 boxes collapse


@pytest.mark.gpu
def test_collapsed_mesh_runs_its_broad_phase_shifted(ctx):
    """Near 2^52 the shift rounds to whole units, so the raw-space broad
    phase's tolerance covers the collapsed pair's whole extent (its superset
    would be every face pair, 81,920^2).  The dispatch sees that its raw
    tolerance is loose, leaves the raw leaves empty and runs again with the
    broad phase on the shifted face boxes: MCUT's 29.7 M candidate pairs and
    its INVALID_CUT_MESH, with the centre's bits, as the oracle has them.  A
    normal dispatch on the same context then runs in raw space again."""
    from oracle import oracle as om
    f = lambda V: V * 4 + 2.0 ** 52 + 0.5
    src, cut = meshes.icosphere_pair(6, amp=0.01)
    src, cut = _remap(src, f), _remap(cut, f)
    print("sums: collapsed icosphere-6 pair", flush=True)
    R = ctx.dispatch(src, cut)
    O = om.frontend(src, cut, detail=False)
    assert np.array_equal(R.com.view(np.uint64), np.asarray(O["com"]).view(np.uint64))
    assert np.array_equal(R.t.view(np.uint64), np.asarray(O["t"]).view(np.uint64))
    assert R.status == O["status"] == om.INVALID_CUT_MESH
    assert R.counts["pairs"] == O["counts"]["pairs"] == 29_690_882
    assert R.counts["raw_pairs"] < 2 * R.counts["pairs"]  # shifted leaves, not the raw superset
    a = R.pairs()
    assert np.array_equal(a, O["pairs"])
    R2 = ctx.dispatch(*meshes.config1_rotated())
    assert R2.status == 0 and R2.counts["registry"] > 0

"""mcgpu_dispatch_narrow -- dispatch()'s front-end block on the inputs preproc
hands it (soup coordinates, face AABBs, candidate pairs) -- against
mcgpu_dispatch_frontend on the same pair: fed that dispatch's own shifted
coordinates, face boxes and candidate pairs, it must give the same status,
planes and registry bit for bit (and the full front end is held to the oracle
by test_gpu_parity.py)."""
import numpy as np
import pytest

import meshes
from mcut_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def soup_inputs(R, src, cut):
    V = R.ps_vertices()
    Vs = src[0].shape[0]

    def boxes(m, off, eps):
        Vv, F, S = m
        sizes = np.full(meshes.nfaces(m), 3, np.int64) if S is None else np.asarray(S, np.int64)
        starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        lo = np.array([V[off + np.asarray(F[s:s + k], np.int64)].min(0) for s, k in zip(starts.tolist(), sizes.tolist())])
        hi = np.array([V[off + np.asarray(F[s:s + k], np.int64)].max(0) for s, k in zip(starts.tolist(), sizes.tolist())])
        return np.hstack([lo - eps, hi + eps]) if eps else np.hstack([lo, hi])

    s2 = (V[:Vs].copy(), src[1], src[2])
    c2 = (V[Vs:].copy(), cut[1], cut[2])
    return s2, c2, boxes(src, 0, 0.0), boxes(cut, Vs, R.eps), R.pairs()


def cases():
    yield "config1_rotated", meshes.config1_rotated()
    yield "config1", mg.config1()
    for i in (0, 2, 5, 13, 14, 20, 33, 40, 52, 61):
        yield f"bench{i:03d}", meshes.bench_pair(i)
    yield "icosphere4", meshes.icosphere_pair(4)


@pytest.mark.parametrize("name,pair", list(cases()), ids=[c[0] for c in cases()])
def test_narrow_equals_full_front_end(name, pair):
    from mcut_amd import frontend
    src, cut = pair
    a, b = frontend.Context(0), frontend.Context(0)
    try:
        R = a.dispatch(src, cut)
        s2, c2, sb, cb, pairs = soup_inputs(R, src, cut)
        N = b.dispatch_narrow(s2, c2, sb, cb, pairs)
        assert N.status == R.status, (name, N.status_name, R.status_name)
        assert N.counts["pairs"] == R.counts["pairs"] and N.counts["tests"] == R.counts["tests"]
        assert N.partial_cut == R.partial_cut
        for k in ("face", "k", "triple"):
            assert np.array_equal(N.planes()[k], R.planes()[k]), (name, k)
        assert np.array_equal(N.planes()["normal"].view(np.uint64), R.planes()["normal"].view(np.uint64))
        e, g, p = N.registry()
        e0, g0, p0 = R.registry()
        assert np.array_equal(e, e0) and np.array_equal(g, g0) and np.array_equal(p.view(np.uint64), p0.view(np.uint64))
        assert np.array_equal(np.stack(N.edges()), np.stack(R.edges()))
    finally:
        a.close()
        b.close()


def corner_edges(R, src, cut):
    """Each face corner's soup edge (input vertex i to i + 1), looked up in the
    dispatch's own edge table."""
    v0, v1, _, _ = R.edges()
    by_key = {(min(a, b), max(a, b)): e for e, (a, b) in enumerate(zip(v0.tolist(), v1.tolist()))}
    out = []
    for m, off in ((src, 0), (cut, src[0].shape[0])):
        V, F, S = m
        sizes = np.full(meshes.nfaces(m), 3, np.int64) if S is None else np.asarray(S, np.int64)
        ce, at = [], 0
        for n in sizes.tolist():
            f = [int(x) + off for x in F[at:at + n]]
            ce += [by_key[(min(f[i], f[(i + 1) % n]), max(f[i], f[(i + 1) % n]))] for i in range(n)]
            at += n
        out.append(np.array(ce, np.uint32))
    return out


@pytest.mark.parametrize("name,pair", list(cases()), ids=[c[0] for c in cases()])
def test_narrow_with_given_topology(name, pair):
    """The same with the soup's edge numbering handed in (corner edges and
    edge table) instead of numbered on the device."""
    from mcut_amd import frontend
    src, cut = pair
    a, b = frontend.Context(0), frontend.Context(0)
    try:
        R = a.dispatch(src, cut)
        if R.status == -2 and name in ("bench061",):
            pytest.skip("rejected cut mesh: no edge table to hand in")
        s2, c2, sb, cb, pairs = soup_inputs(R, src, cut)
        topo = corner_edges(R, src, cut) + list(R.edges())
        N = b.dispatch_narrow(s2, c2, sb, cb, pairs, topology=topo)
        assert N.status == R.status, (name, N.status_name, R.status_name)
        e, g, p = N.registry()
        e0, g0, p0 = R.registry()
        assert np.array_equal(e, e0) and np.array_equal(g, g0) and np.array_equal(p.view(np.uint64), p0.view(np.uint64))
        assert np.array_equal(np.stack(N.edges()), np.stack(R.edges()))
    finally:
        a.close()
        b.close()

"""Batched front end (mcgpu_dispatch_batch): every pair of a batch must come
out exactly as its own mcgpu_dispatch_frontend call gives it -- status,
candidate pairs, registry edges, faces and point bits -- and those plain
dispatches are held to the oracle by test_gpu_parity.py; a few pairs are
checked against the oracle directly as well."""
import numpy as np
import pytest

import meshes
from mcut_amd import meshgen as mg


def _mix():
    """Pairs of every kind the batch must keep apart: the seeded sphere/box
    mix at small sizes, the reference's benchmark pairs (polygons, quads),
    a general-position violation, a touching-free rotation of it, an empty
    intersection and a cut mesh far larger than its source."""
    pairs = mg.config4_pairs(40, seed=11, tri_range=(200, 20_000), box_range=(12, 3_000))
    pairs += [meshes.bench_pair(i) for i in (0, 1, 2, 5, 14, 20, 33, 40)]
    pairs.append(mg.config1())
    pairs.append(meshes.config1_rotated())
    pairs.append((mg.uv_sphere(16, 8), mg.tessellated_box((5.0, 0.0, 0.0), 1.0, 2)))
    pairs.append((mg.uv_sphere(8, 4), mg.tessellated_box((0.2, 0.1, 0.0), 1.3, 12)))
    return pairs


def test_pack_batch_layout():
    from mcut_amd import frontend
    pairs = [meshes.bench_pair(0), mg.config1(), meshes.bench_pair(1)]
    P = frontend.pack_batch(pairs)
    assert P["num_pairs"] == 3
    for side, k in (("src", 0), ("cut", 1)):
        V, F, S = P[side]
        vo, fo = P[side + "_vo"], P[side + "_fo"]
        assert vo[0] == 0 and fo[0] == 0 and vo[-1] == V.shape[0] and fo[-1] == (S.size if S is not None else F.size // 3)
        at = 0
        for i, pr in enumerate(pairs):
            Vi, Fi, Si = pr[k]
            assert np.array_equal(V[vo[i]:vo[i + 1]], Vi)
            n = Fi.size // 3 if Si is None else Si.size
            assert fo[i + 1] - fo[i] == n
            assert np.array_equal(F[at:at + Fi.size], Fi)  # indices stay local to their mesh
            at += Fi.size


def test_config4_generator():
    """Deterministic, within the stated size ranges, and every sphere passes
    the reference's face-normal check."""
    a = mg.config4_pairs(24, seed=5)
    b = mg.config4_pairs(24, seed=5)
    for (s0, c0), (s1, c1) in zip(a, b):
        assert np.array_equal(s0[0], s1[0]) and np.array_equal(c0[0], c1[0])
    for s, c in a:
        fs, fc = s[1].size // 3, c[1].size // 3
        assert 320 <= fs <= 200_000 and fs % 20 == 0
        assert 12 <= fc <= 12 * 32 * 32 and fc % 12 == 0
        assert mg.newell_min_sq(s[0], s[1]) >= 1e-9


@pytest.fixture(scope="module")
def ctxs():
    from mcut_amd import frontend
    a, b = frontend.Context(0), frontend.Context(0)
    yield a, b
    a.close()
    b.close()


def _check_against_plain(B, i, R):
    assert B.status[i] == R.status, (i, B.status[i], R.status)
    e, g, p = B.pair_registry(i)
    e0, g0, p0 = R.registry()
    assert np.array_equal(e, e0) and np.array_equal(g, g0), i
    assert np.array_equal(p.view(np.uint64), p0.view(np.uint64)), i


@pytest.mark.gpu
def test_batch_matches_plain_dispatch(ctxs):
    bctx, pctx = ctxs
    pairs = _mix()
    print(f"batch of {len(pairs)} pairs", flush=True)
    B = bctx.dispatch_batch(pairs)
    print(f"batch done: {B.counts}", flush=True)
    T = B.candidate_pairs()
    tot_reg = 0
    for i, (src, cut) in enumerate(pairs):
        R = pctx.dispatch(src, cut)
        if i % 10 == 0:
            print(f"plain dispatch {i}", flush=True)
        _check_against_plain(B, i, R)
        mine = T[T[:, 0] == i][:, 1:]
        assert np.array_equal(mine, R.pairs()), i
        tot_reg += R.counts["registry"]
    assert B.counts["registry"] == tot_reg
    assert len(set(B.status.tolist())) > 1  # the mix holds more than one outcome


@pytest.mark.gpu
def test_batch_against_oracle(ctxs):
    from oracle import oracle as om
    bctx, _ = ctxs
    pairs = [meshes.config1_rotated(), meshes.bench_pair(2)] + mg.config4_pairs(3, seed=12, tri_range=(200, 5_000))
    B = bctx.dispatch_batch(pairs)
    T = B.candidate_pairs()
    for i, (src, cut) in enumerate(pairs):
        O = om.frontend(src, cut, detail=False)
        assert B.status[i] == O["status"], i
        assert np.array_equal(T[T[:, 0] == i][:, 1:], O["pairs"]), i
        if O["status"] == om.SUCCESS or O["stage"] >= 2:
            e, g, p = B.pair_registry(i)
            assert np.array_equal(e, O["reg_edge"]) and np.array_equal(g, O["reg_face"]), i
            assert np.array_equal(p.view(np.uint64), O["reg_point"].view(np.uint64)), i


@pytest.mark.gpu
def test_batch_of_one_and_repeat(ctxs):
    """A batch of one pair equals the plain dispatch, and a second batch call
    on the same context (smaller) is not disturbed by the first."""
    bctx, pctx = ctxs
    src, cut = meshes.icosphere_pair(5, amp=0.005)
    B = bctx.dispatch_batch([(src, cut)])
    R = pctx.dispatch(src, cut)
    _check_against_plain(B, 0, R)
    small = [meshes.config1_rotated(), meshes.bench_pair(0)]
    B2 = bctx.dispatch_batch(small)
    for i, (s, c) in enumerate(small):
        _check_against_plain(B2, i, pctx.dispatch(s, c))


@pytest.mark.gpu
def test_batch_invalid_pair_rejects_the_batch(ctxs):
    from mcut_amd import frontend
    bctx, _ = ctxs
    V, F, S = mg.uv_sphere(8, 4)
    bad = F.copy()
    bad[4] = V.shape[0]  # one past this pair's own vertices (valid in the concatenation)
    pairs = [meshes.config1_rotated(), ((V, bad, None), mg.tessellated_box((0, 0, 0), 1.0, 2))]
    with pytest.raises(frontend.McgpuError) as ei:
        bctx.dispatch_batch(pairs)
    assert ei.value.code == -1
    B = bctx.dispatch_batch(pairs[:1])  # the context stays usable
    assert B.status[0] == 0


@pytest.mark.gpu
def test_plain_and_batch_alternate_on_one_context():
    """A context's first batch makes the batch's own broad-phase and
    edge-numbering streams; plain and batched dispatches then alternate on
    the same buffers and each equals a dispatch on a fresh context."""
    from mcut_amd import frontend
    a, ref = frontend.Context(0), frontend.Context(0)
    try:
        big = meshes.icosphere_pair(5, amp=0.005)
        small = [meshes.config1_rotated(), meshes.bench_pair(0)]
        for rnd in range(2):
            print(f"alternation {rnd}: plain", flush=True)
            R = a.dispatch(*big)
            R0 = ref.dispatch(*big)
            assert R.status == R0.status and np.array_equal(R.pairs(), R0.pairs())
            e, g, p = R.registry()
            e0, g0, p0 = R0.registry()
            assert np.array_equal(e, e0) and np.array_equal(g, g0)
            assert np.array_equal(p.view(np.uint64), p0.view(np.uint64))
            print(f"alternation {rnd}: batch", flush=True)
            B = a.dispatch_batch(small + [big])
            for i, (s, c) in enumerate(small + [big]):
                _check_against_plain(B, i, ref.dispatch(s, c))
    finally:
        a.close()
        ref.close()


@pytest.mark.gpu
def test_device_polygon_batch_matches_host_batch(ctxs):
    """Two edge numberings of a batch: host input (corner counts known on the
    host: one packed table region per pair) and device input with face sizes
    on the device (the two-table hash over the whole batch) give the same
    per-pair outputs."""
    from mcut_amd import frontend
    from test_zz_no_alloc import Hip
    bctx, pctx = ctxs
    pairs = [meshes.bench_pair(i) for i in (0, 1, 2, 5, 14, 20, 33, 40)] + [meshes.config1_rotated()]
    pairs += mg.config4_pairs(6, seed=29, tri_range=(200, 5_000), box_range=(12, 1_000))
    H = bctx.dispatch_batch(pairs)
    ref = [tuple(np.ascontiguousarray(x).copy() for x in H.pair_registry(i)) for i in range(len(pairs))]
    ref_status = H.status.copy()
    P = frontend.pack_batch(pairs)
    hip = Hip()  # device buffers from the HIP runtime libmcgpu links
    bufs = []

    def up(side):
        V, F, S = P[side]
        if S is None:
            S = np.full(F.size // 3, 3, np.uint32)
        ptrs = [hip.upload(np.ascontiguousarray(a)) for a in (np.asarray(V, np.float64), np.asarray(F, np.uint32),
                                                             np.asarray(S, np.uint32))]
        bufs.extend(ptrs)
        return frontend.device_mesh(ptrs[0], ptrs[1], ptrs[2], V.shape[0], S.size)

    try:
        ms, mc = up("src"), up("cut")
        D = bctx.dispatch_batch_device(P["num_pairs"], ms, mc, [P["src_vo"], P["src_fo"], P["cut_vo"], P["cut_fo"]])
        assert np.array_equal(D.status, ref_status)
        for i in range(len(pairs)):
            got = D.pair_registry(i)
            for x, y in zip(ref[i], got):
                assert np.array_equal(np.ascontiguousarray(x).view(np.uint8),
                                      np.ascontiguousarray(y).view(np.uint8)), i
    finally:
        for p in bufs:
            hip.free(p)


# ---------------------------------------------------------------------------
# Non-manifold soups in a batch.  The batch numbers each pair's edges in a
# table region of its own, with corners handled in blocks of consecutive
# corners; these soups put the extra corners of a shared edge thousands of
# corners away from the first two (another 2,048-corner block), once as a
# second closed sphere glued on by an edge (four corners, two each way) and
# once as a fin placed before or after a sphere (three corners, two the same
# way).  MCUT refuses such a soup (add_face returns null_face,
# hmesh.cpp:610-627, and mcDispatch throws std::invalid_argument,
# preproc.cpp:2338-2350): MC_INVALID_VALUE, here MCGPU_ERROR_INVALID_VALUE
# for the whole batch, as for a plain dispatch.
# ---------------------------------------------------------------------------
_BLOCK = 2048


def _corners(F):
    F = np.asarray(F, np.int64).reshape(-1, 3)
    return F.reshape(-1), np.roll(F, -1, axis=1).reshape(-1)


def _edge_in_one_block(F, corner_base, lo_block=0):
    """An edge of triangle list F whose two corners (soup corners
    corner_base + c) lie in one block at or after lo_block: (a, b, block)."""
    a, d = _corners(F)
    where = {}
    for c in range(a.size):
        where.setdefault((min(a[c], d[c]), max(a[c], d[c])), []).append(c)
    for (u, v), cs in where.items():
        if len(cs) == 2:
            t0, t1 = ((corner_base + x) // _BLOCK for x in cs)
            if t0 == t1 and t0 >= lo_block:
                return int(u), int(v), t0
    raise AssertionError("no edge with both corners in one block")


def _glued_spheres():
    """Two closed spheres whose faces follow each other, the second with two
    vertices renamed onto an edge of the first: that edge has four corners,
    two (one each way) in a block of the first sphere and two in a block of
    the second."""
    V, F, _ = mg.geodesic_sphere(12)  # 2,880 faces: 8,640 corners, over four blocks
    a, b, ta = _edge_in_one_block(F, 0)
    nv = V.shape[0]
    c, d, tb = _edge_in_one_block(F, F.size, lo_block=F.size // _BLOCK + 1)
    F2 = F.astype(np.int64) + nv
    F2[F2 == c + nv] = a
    F2[F2 == d + nv] = b
    V2 = V * 0.5 + np.array([3.0, 0.0, 0.0])
    return np.vstack([V, V2]), np.concatenate([F, F2.astype(np.uint32)]), (a, b, ta, tb)


def _sphere_and_fin(first):
    """A closed sphere and one more triangle on an edge of it with the
    sphere's own direction: three corners, two in one block of the sphere and
    the fin's in another block (the fin placed before or after the sphere)."""
    V, F, _ = mg.geodesic_sphere(12)
    lo = 1 if first else 0  # the fin's corners sit in block 0 when it comes first
    a, b, t = _edge_in_one_block(F, 3 if first else 0, lo_block=lo)
    A, B = _corners(F)
    k = int(np.nonzero((A == a) & (B == b) | (A == b) & (B == a))[0][0])
    u, v = int(A[k]), int(B[k])  # a corner of the edge, u -> v
    z = V.shape[0]
    fin = np.array([u, v, z], np.uint32)
    V2 = np.vstack([V, [[2.0, 2.0, 2.0]]])
    F2 = np.concatenate([fin, F]) if first else np.concatenate([F, fin])
    fin_block = 0 if first else F.size // _BLOCK
    assert fin_block != t
    return V2, F2, (u, v, t, fin_block)


def test_nonmanifold_soups_split_their_corners_over_blocks():
    """The soups below hold what they claim: every corner of the shared edge,
    the two blocks, and exactly one non-manifold edge."""
    V, F, (a, b, ta, tb) = _glued_spheres()
    A, D = _corners(F)
    on = np.nonzero(((A == a) & (D == b)) | ((A == b) & (D == a)))[0]
    assert on.size == 4 and ta != tb
    assert sorted(set((on // _BLOCK).tolist())) == sorted({ta, tb})
    keys = np.minimum(A, D) * (1 << 32) + np.maximum(A, D)
    _, cnt = np.unique(keys, return_counts=True)
    assert (cnt > 2).sum() == 1
    for first in (False, True):
        V, F, (u, v, t, ft) = _sphere_and_fin(first)
        A, D = _corners(F)
        on = np.nonzero(((A == u) & (D == v)) | ((A == v) & (D == u)))[0]
        assert on.size == 3 and ((A[on] == u) & (D[on] == v)).sum() == 2
        assert set((on // _BLOCK).tolist()) == {t, ft}


@pytest.mark.gpu
def test_batch_refuses_nonmanifold_soups_split_over_blocks(ctxs):
    from mcut_amd import frontend
    bctx, pctx = ctxs
    box = mg.tessellated_box((0.0, 0.0, 0.0), 1.3, 4)
    good = meshes.config1_rotated()
    V, F, _ = _glued_spheres()
    soups = [(V, F)] + [_sphere_and_fin(first)[:2] for first in (False, True)]
    for k, (V, F) in enumerate(soups):
        for side in ("src", "cut"):
            pair = ((V, F, None), box) if side == "src" else (box, (V, F, None))
            print(f"non-manifold soup {k} as {side}", flush=True)
            with pytest.raises(frontend.McgpuError) as ei:
                pctx.dispatch(*pair)
            assert ei.value.code == -1
            with pytest.raises(frontend.McgpuError) as ei:
                bctx.dispatch_batch([good, pair, good])
            assert ei.value.code == -1
    B = bctx.dispatch_batch([good])  # the context stays usable
    assert B.status[0] == 0

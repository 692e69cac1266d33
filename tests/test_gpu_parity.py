"""HIP front end (through libmcgpu's C ABI) against the CPU oracle, bit for bit.

Every intermediate the reference produces is compared: shift parameters,
shifted polygon-soup vertices, the halfedge edge table, candidate face pairs,
candidate-face planes and vertex triplets, every edge-face test with its two
orient3d signs, segment type and point-in-polygon result, the intersection
registry (edge, face, point bits) and the dispatch status.
"""
import numpy as np
import pytest

from oracle import oracle as om
import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from mcut_amd import frontend
    c = frontend.Context(0)
    yield c
    c.close()


def _sorted_pairs(a):
    a = np.asarray(a).reshape(-1, 2)
    return a[np.lexsort((a[:, 1], a[:, 0]))]


def compare(ctx, src, cut, detail=True, **kw):
    print(f"compare: Fs={meshes.nfaces(src)} Fc={meshes.nfaces(cut)} {kw}", flush=True)
    R = ctx.dispatch(src, cut, keep_tests=detail, **kw)
    okw = {}
    if kw.get("flags", 0) & (1 << 15):
        okw["enforce_gp"] = True
    O = om.frontend(src, cut, detail=detail, **okw)
    assert np.array_equal(R.com, O["com"]) and np.array_equal(R.t, O["t"]) and R.eps == O["eps"]
    assert R.status == O["status"], (R.status_name, O["status"], O["stage"])
    assert R.counts["pairs"] == O["counts"]["pairs"]
    assert np.array_equal(R.pairs(), O["pairs"])
    if detail:
        assert np.array_equal(R.ps_vertices(), O["ps_vertices"])
        v0, v1, f0, f1 = R.edges()
        assert np.array_equal(v0, O["edge_v0"]) and np.array_equal(v1, O["edge_v1"])
        assert np.array_equal(f0, O["edge_f0"]) and np.array_equal(f1, O["edge_f1"])
    if O["stage"] >= 1 and not detail:
        assert R.counts["tests"] == O["counts"]["tests"]
        if O["status"] == om.SUCCESS or O["stage"] >= 2:
            e, g, p = R.registry()
            assert np.array_equal(e, O["reg_edge"]) and np.array_equal(g, O["reg_face"])
            assert np.array_equal(p.view(np.uint64), O["reg_point"].view(np.uint64))
    if O["stage"] >= 1 and detail:  # the oracle got past the face-property stage
        P = R.planes()
        assert np.array_equal(P["face"], O["cand_face"])
        assert np.array_equal(P["normal"], O["plane_n"]) and np.array_equal(P["d"], O["plane_d"])
        assert np.array_equal(P["k"], O["plane_k"])
        ot = O["plane_triple"].copy()
        gt = P["triple"].copy()
        coll = ot[:, 0] < 0
        assert np.array_equal(gt[:, 0] < 0, coll)
        assert np.array_equal(gt[~coll], ot[~coll])
        assert R.counts["tests"] == O["counts"]["tests"]
        if detail:
            T = R.tests()
            assert np.array_equal(T["edge"], O["test_edge"]) and np.array_equal(T["face"], O["test_face"])
            assert np.array_equal(T["qsign"], O["qsign"]) and np.array_equal(T["rsign"], O["rsign"])
            assert np.array_equal(T["type"], O["type"])
            assert np.array_equal(T["pip"], O["pip"])
        if O["status"] == om.SUCCESS or O["stage"] >= 2:
            e, g, p = R.registry()
            assert np.array_equal(e, O["reg_edge"]) and np.array_equal(g, O["reg_face"])
            assert np.array_equal(p.view(np.uint64), O["reg_point"].view(np.uint64))
    return R, O


def test_config1_rotated(ctx):
    R, O = compare(ctx, *meshes.config1_rotated())
    assert R.status == 0 and R.counts["registry"] == 134


def test_config1_general_position_violation(ctx):
    from mcut_amd import meshgen as mg
    R, O = compare(ctx, *mg.config1())
    assert R.status_name == "GENERAL_POSITION_VIOLATION"


def test_config1_perturbed_retry(ctx):
    """Reference's enforce-general-position retry: same candidates, cut mesh
    moved by the reference's perturbation vector (golden fixture)."""
    import os
    from mcut_amd import meshgen as mg
    cases = np.load(os.path.join(meshes.GOLDEN, "ref_cases.npz"))
    pert = cases["config1_enforce/perturbation"]
    src, cut = mg.config1()
    R = ctx.dispatch(src, cut, flags=1 << 15)
    assert R.status_name == "GENERAL_POSITION_VIOLATION"
    nv = src[0].shape[0] + cut[0].shape[0]
    R2 = ctx.redispatch_perturbed(pert, 1, nv)
    O = om.frontend(src, cut, enforce_gp=True, enforcement_count=1, perturbation=pert)
    assert R2.status == O["status"] == 0
    e, g, p = R2.registry()
    assert np.array_equal(e, O["reg_edge"]) and np.array_equal(g, O["reg_face"])
    assert np.array_equal(p.view(np.uint64), O["reg_point"].view(np.uint64))
    want = {tuple(x) for x in cases["config1_enforce/points"].tolist()}
    assert {tuple(x) for x in p.tolist()} == want


def test_enforced_retry_draws_the_reference_perturbation():
    """mcgpu_dispatch_enforced on config 1 with ENFORCE_GENERAL_POSITION in a
    fresh context: one perturbation, bit-identical to the one the reference
    drew (golden fixture), and the oracle's registry for it."""
    import os
    from mcut_amd import frontend, meshgen as mg
    cases = np.load(os.path.join(meshes.GOLDEN, "ref_cases.npz"))
    want = cases["config1_enforce/perturbation"]
    src, cut = mg.config1()
    c = frontend.Context(0)
    try:
        R, p, k = c.dispatch_enforced(src, cut)
        assert k == 1 and np.array_equal(p.view(np.uint64), np.asarray(want, np.float64).view(np.uint64))
        O = om.frontend(src, cut, enforce_gp=True, enforcement_count=0, perturbation=p)
        assert R.status == O["status"] == 0
        e, g, pt = R.registry()
        assert np.array_equal(e, O["reg_edge"]) and np.array_equal(g, O["reg_face"])
        assert np.array_equal(pt.view(np.uint64), O["reg_point"].view(np.uint64))
        # the sequence continues in the context: a second enforced dispatch draws the next vector
        R2, p2, k2 = c.dispatch_enforced(src, cut)
        assert k2 >= 1 and not np.array_equal(p2, p)
    finally:
        c.close()


@pytest.mark.parametrize("i", range(64))
def test_benchmark_pairs(ctx, i):
    compare(ctx, *meshes.bench_pair(i))


@pytest.mark.parametrize("lvl,amp", [(3, 0.02), (4, 0.02), (5, 0.005), (6, 0.005), (6, 0.02)])
def test_icosphere_pairs(ctx, lvl, amp):
    compare(ctx, *meshes.icosphere_pair(lvl, amp=amp))


def test_icosphere7_benchmark_pair(ctx):
    """The benchmark workload at full size: 655,360 triangles, about 20k
    registry entries (several tiles of the registry sort)."""
    from mcut_amd import meshgen as mg
    R, O = compare(ctx, *mg.config2(level=7), detail=False)
    assert R.status == 0 and R.counts["registry"] > 3 * 4096


@pytest.mark.parametrize("seed", range(6))
def test_random_box_cuts(ctx, seed):
    """Tessellated boxes (triangles and quads) at random poses through an icosphere."""
    from mcut_amd import meshgen as mg
    rng = np.random.default_rng(100 + seed)
    V, F, _ = mg.icosphere(4)
    V = mg.radial_displace(V, 0.01, seed=seed)
    Rm = mg.random_rotation(seed=200 + seed)
    box = mg.tessellated_box(rng.uniform(-0.3, 0.3, 3), rng.uniform(0.6, 1.4, 3), int(rng.integers(2, 7)),
                             rotation=Rm, triangulate=bool(seed % 2))
    compare(ctx, (V, F, None), box)


@pytest.mark.parametrize("axis,sign", [(2, 1.0), (2, -1.0), (0, 1.0)])
def test_boxes_touching_at_eps(ctx, axis, sign):
    """A cut patch offset from a source patch by eps + k * 5e-17, so that the
    shifted boxes go from overlapping to disjoint inside the sweep.  The GPU's
    broad phase runs on raw boxes widened by a rounding bound and then applies
    the exact shifted test: its pairs must equal the oracle's at every k."""
    from mcut_amd import meshgen as mg
    V, F, _ = mg.grid_patch(6)
    src = (V.copy(), F, None)
    base = 1.0 if axis == 0 else 0.0  # lateral case: next patch over, touching one column
    counts = set()
    for k in range(-12, 13):
        Vc = V.copy()
        Vc[:, axis] += sign * (base + 0.0001414213562373095 + k * 5e-17)
        R, O = compare(ctx, src, (Vc, F, None))
        counts.add(R.counts["pairs"])
    assert 0 in counts and len(counts) > 1  # the sweep crosses the boundary


def test_empty_intersection(ctx):
    from mcut_amd import meshgen as mg
    src = mg.uv_sphere(16, 8)
    cut = mg.tessellated_box((5.0, 0.0, 0.0), 1.0, 2)
    R, O = compare(ctx, src, cut)
    assert R.counts["pairs"] == 0 and R.counts["registry"] == 0 and R.status == 0


@pytest.mark.parametrize("src_name,cut_name", [("bunny", "bunnyCuttingPlanePartial"),
                                               ("cube-flattened", "cube-flattened-with-holes")])
def test_reference_test_meshes(ctx, src_name, cut_name):
    """The reference's own test meshes: the bunny cut partly by a plane
    (tests/source/dispatchFilterFlags.cpp:155-204, which expects success) and
    a flattened cube against one with holes (tests/source/polygonWithHoles.cpp:
    95-140, which expects MC_INVALID_OPERATION, the front end's
    general-position violation)."""
    import os
    from mcut_amd import meshgen as mg
    d = os.path.join(meshes.GOLDEN, "meshes")
    src, cut = mg.read_off(os.path.join(d, f"{src_name}.off")), mg.read_off(os.path.join(d, f"{cut_name}.off"))
    R, O = compare(ctx, src, cut)
    assert R.status == (om.SUCCESS if src_name == "bunny" else om.GENERAL_POSITION_VIOLATION)


def test_invalid_input_rejected(ctx):
    from mcut_amd import frontend, meshgen as mg
    V, F, S = mg.uv_sphere(8, 4)
    bad = F.copy()
    bad[5] = V.shape[0] + 7  # vertex index out of range
    with pytest.raises(frontend.McgpuError) as ei:
        ctx.dispatch((V, bad, None), mg.tessellated_box((0, 0, 0), 1.0, 2))
    assert ei.value.code == -1
    # the context stays usable
    compare(ctx, *meshes.config1_rotated())


def test_coordinates_beyond_the_broad_phase_bound(ctx):
    """Coordinates whose squared extent overflows (1e200) leave the raw-space
    broad phase without a finite rounding bound: the dispatch reports
    MCGPU_ERROR_UNSUPPORTED and the context stays usable."""
    from mcut_amd import frontend, meshgen as mg
    V, F, _ = mg.uv_sphere(8, 4)
    Vc, Fc, Sc = mg.tessellated_box((0, 0, 0), 1.0, 2)
    with pytest.raises(frontend.McgpuError) as ei:
        ctx.dispatch((V * 1e200, F, None), (Vc * 1e200, Fc, Sc))
    assert ei.value.code == -4
    compare(ctx, *meshes.config1_rotated())


def _brute_pairs(a, b):
    """All (i, j) with inclusive overlap of boxes a[i] and b[j] (exact float64 compares)."""
    ov = np.ones((a.shape[0], b.shape[0]), bool)
    for j in range(3):
        ov &= (a[:, None, j] <= b[None, :, 3 + j]) & (a[:, None, 3 + j] >= b[None, :, j])
    i, k = np.nonzero(ov)
    return np.stack([i, k], 1).astype(np.uint32)


@pytest.mark.parametrize("case,scale", [("sphere-box", 60.0), ("sphere-box", 100.0), ("icosphere-2", 60.0)])
def test_coordinates_far_from_the_origin_run_the_broad_phase_shifted(ctx, case, scale):
    """Meshes of 100-250 units near 2^52, where the shift rounds to whole
    units: the raw-space tolerance (16 units of rounding) is more than a
    sixteenth of the pair's extent, so the dispatch reruns its broad phase on
    the shifted face boxes.  Every intermediate equals the oracle's, these
    pairs reach the narrow phase and succeed, and the raw-space superset is
    the shifted leaves' (close to the exact pairs), not the loose one."""
    from mcut_amd import meshgen as mg
    if case == "sphere-box":
        src, cut = mg.uv_sphere(16, 8), mg.tessellated_box((0.2, 0.1, 0.05), 1.3, 3, rotation=mg.random_rotation(4))
    else:
        src, cut = meshes.icosphere_pair(2, amp=0.01)
    f = lambda m: (np.asarray(m[0], np.float64) * scale + 2.0 ** 52, m[1], m[2])
    R, O = compare(ctx, f(src), f(cut))
    assert R.status == om.SUCCESS and R.counts["registry"] > 0
    assert R.counts["raw_pairs"] < 2 * R.counts["pairs"]


@pytest.mark.parametrize("seed", range(4))
def test_broad_phase_boxes(ctx, seed):
    """mcgpu_broad_phase against brute force on random boxes, with boxes that
    touch exactly (shared faces, zero-width boxes) and a few empty boxes."""
    rng = np.random.default_rng(300 + seed)
    n_s, n_c = int(rng.integers(1, 3000)), int(rng.integers(1, 3000))

    def boxes(n):
        lo = rng.integers(0, 64, (n, 3)).astype(np.float64) / 8.0  # coarse grid: exact touching is common
        ext = rng.integers(0, 6, (n, 3)).astype(np.float64) / 8.0
        b = np.concatenate([lo, lo + ext], 1)
        b[rng.random(n) < 0.01, :3] += 100.0  # inverted (empty) boxes
        return b
    a, b = boxes(n_s), boxes(n_c)
    got = ctx.broad_phase(a, b)
    want = _brute_pairs(a, b)
    want = want[np.lexsort((want[:, 1], want[:, 0]))]
    assert np.array_equal(got, want), (got.shape, want.shape)


def test_broad_phase_mesh_boxes(ctx):
    """Face boxes of a mesh pair, as build_oibvh makes them: the pairs equal
    the oracle's broad phase on the same pair."""
    from mcut_amd import meshgen as mg
    src, cut = meshes.icosphere_pair(5, amp=0.005)
    O = om.frontend(src, cut, detail=True)
    ps = O["ps_vertices"]
    ns = src[0].shape[0]

    def face_boxes(F, off, eps):
        T = F.reshape(-1, 3) + off
        P = ps[T]
        return np.concatenate([P.min(1) - eps, P.max(1) + eps], 1)
    a = face_boxes(src[1], 0, 0.0)
    b = face_boxes(cut[1], ns, O["eps"])
    got = ctx.broad_phase(a, b)
    assert np.array_equal(got, O["pairs"])


def test_stage_timing_switch(ctx):
    """Stage events only when asked for: none by default, all twelve stages and
    the traversal kernel's own time after mcgpu_set_stage_timing, none again
    once it is off."""
    src, cut = meshes.icosphere_pair(4)
    assert ctx.dispatch(src, cut).stage_ms() == {}
    ctx.set_stage_timing(True)
    try:
        st = ctx.dispatch(src, cut).stage_ms()
    finally:
        ctx.set_stage_timing(False)
    assert len(st) == 13 and all(v >= 0.0 for v in st.values()) and st["traversal"] > 0.0
    assert 0.0 < st["traversal_kernel"] <= st["traversal"] * 1.5
    assert ctx.dispatch(src, cut).stage_ms() == {}

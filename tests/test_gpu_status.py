"""Dispatch statuses and the partial-cut flag on the GPU, against the kernel
statuses the reference reports (tests/golden/ref_status.npz, written from MCUT
compiled from its sources) and against the oracle."""
import os

import numpy as np
import pytest

import meshes

pytestmark = pytest.mark.gpu

STATUS = np.load(os.path.join(meshes.GOLDEN, "ref_status.npz"))
NAMES = {0: "", -1: "INVALID_SRC_MESH", -2: "INVALID_CUT_MESH", -3: "INVALID_MESH_INTERSECTION",
         -4: "GENERAL_POSITION_VIOLATION"}


@pytest.fixture(scope="module")
def ctx():
    from mcut_amd import frontend
    c = frontend.Context(0)
    yield c
    c.close()


def quad_case(k):
    src = (STATUS[f"quad{k}/src"], np.array([0, 1, 2, 3], np.uint32), np.array([4], np.uint32))
    cut = (STATUS[f"quad{k}/cut"], np.array([0, 1, 2], np.uint32), None)
    return src, cut


def test_benchmark_statuses_match_reference_reports(ctx):
    for i in range(64):
        R = ctx.dispatch(*meshes.bench_pair(i))
        assert NAMES[R.status] == str(STATUS[f"bench{i:03d}/status"]), (i, R.status_name)


@pytest.mark.parametrize("k", range(5))
def test_invalid_mesh_intersection(ctx, k):
    from mcut_amd import frontend
    from oracle import oracle as om
    src, cut = quad_case(k)
    print(f"quad case {k}", flush=True)
    R = ctx.dispatch(src, cut)
    assert R.status_name == str(STATUS[f"quad{k}/status"]) == "INVALID_MESH_INTERSECTION"
    assert R.counts["registry"] == 1 and R.partial_cut == 0
    O = om.frontend(src, cut, detail=False)
    e, g, p = R.registry()
    assert np.array_equal(e, O["reg_edge"]) and np.array_equal(p.view(np.uint64), O["reg_point"].view(np.uint64))
    R = ctx.dispatch(src, cut, flags=frontend.ENFORCE_GENERAL_POSITION)
    assert R.status_name == str(STATUS[f"quad{k}/enforce_status"])


def test_partial_cut_flag_and_early_exit(ctx):
    """Pair 013 of the reference's benchmarks is a partial cut: the reference
    stops after its front end under MC_DISPATCH_REQUIRE_THROUGH_CUTS (two
    input components, computeSeams.cpp:356-398) and cuts without it."""
    from mcut_amd import frontend
    from oracle import oracle as om
    src, cut = meshes.bench_pair(13)
    assert int(STATUS["bench013_through/n_cc"]) == 2 and int(STATUS["bench013_plain/n_cc"]) > 2
    for flags, through in ((frontend.REQUIRE_THROUGH_CUTS, True), (0, False)):
        R = ctx.dispatch(src, cut, flags=flags)
        O = om.frontend(src, cut, require_through_cuts=through, detail=False)
        assert R.status == O["status"] == 0
        assert R.partial_cut == O["partial"] == 1
    R = ctx.dispatch(*meshes.config1_rotated(), flags=frontend.REQUIRE_THROUGH_CUTS)
    assert R.status == 0 and R.partial_cut == 0


def test_partial_cut_in_a_batch(ctx):
    """Per-pair statuses and partial-cut flags of a batch under
    MC_DISPATCH_REQUIRE_THROUGH_CUTS, against the oracle pair by pair: quad
    case 0's one point lies on the cut triangle's border edge (a partial cut,
    so the early exit comes before the per-face-pair check), case 1's on a
    source edge (an invalid mesh intersection)."""
    from mcut_amd import frontend
    from oracle import oracle as om
    pairs = [meshes.config1_rotated(), meshes.bench_pair(13), quad_case(0), quad_case(1), meshes.bench_pair(2)]
    B = ctx.dispatch_batch(pairs, flags=frontend.REQUIRE_THROUGH_CUTS)
    want_p, want_s = [], []
    for src, cut in pairs:
        O = om.frontend(src, cut, require_through_cuts=True, detail=False)
        want_p.append(O["partial"])
        want_s.append(O["status"])
    assert B.partial_cut.tolist() == want_p == [0, 1, 1, 0, 0]
    assert B.status.tolist() == want_s
    assert NAMES[want_s[3]] == "INVALID_MESH_INTERSECTION"

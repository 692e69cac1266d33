"""Pins of the oracle's parts that the reference's per-function vectors leave
open, against the reference's own outputs (golden fixtures written by
tests/golden/make_golden.py from MCUT compiled from its sources) and against
what the mathematics fixes:

* the edge-face test list (expansion of candidate pairs to edges, dedup and
  edge-box cull, kernel.cpp:1779-2177), rebuilt here by brute force from the
  pinned candidate pairs, ps edge table and face boxes;
* the kernel statuses the reference reports (INVALID_MESH_INTERSECTION,
  general-position violations, invalid meshes) and the partial-cut early exit
  of MC_DISPATCH_REQUIRE_THROUGH_CUTS (tests/source/computeSeams.cpp:356-398);
* the two subdivision-7 icospheres (327,680 triangles each) at full size:
  shift parameters and candidate pairs against the reference's broad phase,
  and every intersection point against exact rational arithmetic.
"""
import hashlib
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as om
from mcut_amd import exact, meshgen as mg
import meshes

GOLD = meshes.GOLDEN
STATUS = np.load(os.path.join(GOLD, "ref_status.npz"))
CONFIG2 = np.load(os.path.join(GOLD, "ref_config2.npz"))
NAMES = {0: "", -1: "INVALID_SRC_MESH", -2: "INVALID_CUT_MESH", -3: "INVALID_MESH_INTERSECTION",
         -4: "GENERAL_POSITION_VIOLATION"}


# ------------------------------------------------------- expansion and cull

def brute_force_tests(R, src, cut):
    """{(edge, face)}: for every candidate pair (f, g) and every edge e of f,
    the test (e, g), and of g, (e, f); kept when the edge's box (its two
    shifted endpoints) meets the other face's box (cut faces enlarged by eps),
    with the inclusive test of math.h:931-941."""
    V = R["ps_vertices"]
    Vs = src[0].shape[0]
    ns = meshes.nfaces(src)

    def face_lists(m, voff):
        Vv, F, S = m
        sizes = np.full(meshes.nfaces(m), 3, np.int64) if S is None else np.asarray(S, np.int64)
        starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        return [np.asarray(F[s:s + k], np.int64) + voff for s, k in zip(starts.tolist(), sizes.tolist())]

    faces = face_lists(src, 0) + face_lists(cut, Vs)
    lo = np.array([V[f].min(0) for f in faces])
    hi = np.array([V[f].max(0) for f in faces])
    lo[ns:] -= R["eps"]
    hi[ns:] += R["eps"]
    v0, v1, f0, f1 = R["edge_v0"], R["edge_v1"], R["edge_f0"], R["edge_f1"]
    edges_of = {}
    for e in range(v0.size):
        for f in (f0[e], f1[e]):
            if f >= 0:
                edges_of.setdefault(int(f), []).append(e)
    elo = np.minimum(V[v0], V[v1])
    ehi = np.maximum(V[v0], V[v1])
    out = set()
    for sf, cfl in R["pairs"]:
        cf = int(cfl) + ns
        for f, g in ((int(sf), cf), (cf, int(sf))):
            for e in edges_of[f]:
                if np.all(elo[e] <= hi[g]) and np.all(ehi[e] >= lo[g]):
                    out.add((e, g))
    return out


def expansion_cases():
    yield "config1_rot", meshes.config1_rotated()
    yield "ico3", meshes.icosphere_pair(3)
    yield "ico4", meshes.icosphere_pair(4)
    for i in range(64):
        yield f"bench{i:03d}", meshes.bench_pair(i)


def test_test_list_equals_brute_force_expansion_and_cull(oracle_lib):
    checked = 0
    for name, (src, cut) in expansion_cases():
        R = om.frontend(src, cut)
        if R["counts"]["pairs"] <= 0 or R["stage"] < 1:  # rejected before the narrow phase
            continue
        want = brute_force_tests(R, src, cut)
        got = set(zip(R["test_edge"].tolist(), R["test_face"].tolist()))
        assert got == want, (name, len(got), len(want))
        checked += 1
    assert checked >= 55


def test_brute_force_expansion_detects_a_dropped_cull(oracle_lib):
    """The check can fail: without the edge-box cull the brute-force list is
    larger than the reference-rule list on an icosphere pair."""
    src, cut = meshes.icosphere_pair(3)
    R = om.frontend(src, cut)
    R2 = dict(R)
    R2["eps"] = R["eps"] * 1e6  # wildly enlarged cut boxes stand for a cull that keeps everything
    assert len(brute_force_tests(R2, src, cut)) > len(brute_force_tests(R, src, cut))


# ------------------------------------------------------------ statuses

def test_front_end_statuses_against_reference_kernel_reports(oracle_lib):
    """Every benchmark pair the reference ends with a front-end status (its
    debug report names it) gets that status from the oracle, and the
    others succeed in the front end."""
    seen = set()
    for i in range(64):
        ref = str(STATUS[f"bench{i:03d}/status"])
        src, cut = meshes.bench_pair(i)
        R = om.frontend(src, cut, detail=False)
        assert NAMES[R["status"]] == ref, (i, R["status"], ref)
        seen.add(ref)
    assert {"GENERAL_POSITION_VIOLATION", "INVALID_CUT_MESH"} <= seen


def quad_case(k):
    src = (STATUS[f"quad{k}/src"], np.array([0, 1, 2, 3], np.uint32), np.array([4], np.uint32))
    cut = (STATUS[f"quad{k}/cut"], np.array([0, 1, 2], np.uint32), None)
    return src, cut


@pytest.mark.parametrize("k", range(5))
def test_single_point_face_pair_is_an_invalid_mesh_intersection(oracle_lib, k):
    """A non-planar quad cut by a triangle that crosses it once: one (source
    face, cut face) pair holds a single point (kernel.cpp:3439-3452).  The
    reference reports INVALID_MESH_INTERSECTION, and a general-position
    violation under ENFORCE_GENERAL_POSITION."""
    src, cut = quad_case(k)
    assert str(STATUS[f"quad{k}/status"]) == "INVALID_MESH_INTERSECTION"
    R = om.frontend(src, cut, detail=False)
    assert R["status"] == om.INVALID_MESH_INTERSECTION
    assert R["counts"]["registry"] == 1
    assert NAMES[om.frontend(src, cut, enforce_gp=True, detail=False)["status"]] == \
        str(STATUS[f"quad{k}/enforce_status"])


def test_partial_cut_with_through_cuts_required(oracle_lib):
    """Benchmark pair 013 cuts the source partially.  With
    MC_DISPATCH_REQUIRE_THROUGH_CUTS the reference stops after the front end
    with success and only the two input components (its own test asserts 2,
    computeSeams.cpp:397); without the flag it cuts (8 components)."""
    assert int(STATUS["bench013_through/result"]) == 0 and int(STATUS["bench013_through/n_cc"]) == 2
    assert int(STATUS["bench013_through/n_input_cc"]) == 2
    assert int(STATUS["bench013_plain/result"]) == 0 and int(STATUS["bench013_plain/n_cc"]) > 2
    src, cut = meshes.bench_pair(13)
    R = om.frontend(src, cut, require_through_cuts=True, detail=False)
    assert R["status"] == om.SUCCESS and R["partial"] == 1 and R["stage"] == 3  # the early exit
    R = om.frontend(src, cut, detail=False)
    assert R["status"] == om.SUCCESS and R["partial"] == 1 and R["stage"] == 4
    # a through cut is not partial
    R = om.frontend(*meshes.config1_rotated(), require_through_cuts=True, detail=False)
    assert R["status"] == om.SUCCESS and R["partial"] == 0


# ------------------------------------------------- config 2 at full size

@pytest.fixture(scope="module")
def config2(oracle_lib):
    src, cut = mg.config2(level=7)
    R = om.frontend(src, cut)
    return src, cut, R


def test_config2_shift_and_pairs_against_reference_broad_phase(config2):
    src, cut, R = config2
    assert np.array_equal(R["com"], CONFIG2["com"])
    assert np.array_equal(R["t"], CONFIG2["t"])
    assert R["eps"] == float(CONFIG2["eps"])
    p = np.ascontiguousarray(R["pairs"].astype(np.uint32))
    assert p.shape[0] == int(CONFIG2["n_pairs"])
    assert hashlib.sha256(p.tobytes()).hexdigest() == str(CONFIG2["pairs_sha256"])


def test_config2_points_against_exact_rationals(config2):
    """Every one of the registry points of the full-size pair lies within
    1e-12 (relative, per coordinate) of the exact segment/plane intersection,
    and every registry edge strictly straddles its face's plane."""
    src, cut, R = config2
    tri = exact.triangle_soup_faces(src[1], cut[1], src[0].shape[0])
    err, straddle = exact.registry_point_errors(R["ps_vertices"], tri, R["edge_v0"], R["edge_v1"], R["reg_edge"],
                                                R["reg_face"], R["reg_point"])
    assert err.size == R["counts"]["registry"] > 10000
    assert straddle.all()
    print(f"config 2: {err.size} points, max relative error {err.max():.3e}")
    assert err.max() <= 1e-12, err.max()


# ------------------------------------------------- the exact helper itself

def test_exact_point_errors_agree_with_fractions(oracle_lib):
    """The integer-scaled computation against a direct Fraction computation
    on config 1 (rotated), entry by entry, and a planted error of 1e-9 in one
    point is found."""
    src, cut = meshes.config1_rotated()
    R = om.frontend(src, cut)
    V = R["ps_vertices"]
    tri = exact.triangle_soup_faces(src[1], cut[1], src[0].shape[0])
    err, straddle = exact.registry_point_errors(V, tri, R["edge_v0"], R["edge_v1"], R["reg_edge"], R["reg_face"],
                                                R["reg_point"])
    assert straddle.all()

    def det(a, b, c, d):
        A = [x - y for x, y in zip(a, d)]
        B = [x - y for x, y in zip(b, d)]
        C = [x - y for x, y in zip(c, d)]
        return A[0] * (B[1] * C[2] - B[2] * C[1]) - A[1] * (B[0] * C[2] - B[2] * C[0]) + A[2] * (B[0] * C[1] - B[1] * C[0])

    for i, (e, g, p) in enumerate(zip(R["reg_edge"], R["reg_face"], R["reg_point"])):
        q = [Fraction(x) for x in V[R["edge_v0"][e]]]
        r = [Fraction(x) for x in V[R["edge_v1"][e]]]
        a, b, c = ([Fraction(x) for x in V[j]] for j in tri[g])
        oq, orr = det(a, b, c, q), det(a, b, c, r)
        t = oq / (oq - orr)
        want = max(float(abs(Fraction(pi) - (qi + t * (ri - qi))) / max(abs(qi + t * (ri - qi)), Fraction(1)))
                   for pi, qi, ri in zip(p, q, r))
        assert abs(err[i] - want) <= 1e-6 * max(want, 1e-300), (i, err[i], want)
    P = R["reg_point"].copy()
    P[3, 1] *= 1 + 1e-9
    err2, _ = exact.registry_point_errors(V, tri, R["edge_v0"], R["edge_v1"], R["reg_edge"], R["reg_face"], P)
    assert err2[3] > 1e-10 and err2.max() > 1e-10

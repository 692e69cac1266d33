"""Serving (mcgpu_server): every request's status and registry must equal
mcgpu_dispatch_frontend on (resident source, that cut) -- alone and in the
mix, whether requests overlap in slots or are admitted one at a time by a
small HBM budget.  The plain dispatch is held to the oracle by
test_gpu_parity.py; one request is checked against the oracle directly."""
import numpy as np
import pytest

from mcut_amd import meshgen as mg


def test_crack_sheet_generator():
    a = mg.crack_sheet(3)
    b = mg.crack_sheet(3)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for k in range(12):
        V, F, _ = mg.crack_sheet(k)
        nf = F.size // 3
        m = int(round(np.sqrt(nf / 2)))
        assert 2 * m * m == nf and V.shape[0] == (m + 1) ** 2
        assert 1_500 <= nf <= 51_000
        assert np.abs(np.linalg.norm(V, axis=1)).max() < 3.0


def _requests(n, small=True):
    rng = (500, 4_000) if small else (2_000, 50_000)
    return [mg.crack_sheet(k, seed=9, tri_range=rng) for k in range(n)]


@pytest.fixture(scope="module")
def served():
    from mcut_amd import frontend
    src = mg.config5_source(n=30)  # 18,000 triangles
    cuts = _requests(16)
    mv = max(c[0].shape[0] for c in cuts)
    mf = max(c[1].size // 3 for c in cuts)
    srv = frontend.Server(src, mv, mf, slots=4, hbm_budget=8 << 30)
    ctx = frontend.Context(0)
    want = []
    for c in cuts:  # each registry is read before the next dispatch replaces it
        r = ctx.dispatch(src, c)
        want.append((r.status, *r.registry()))
    yield src, cuts, srv, want, ctx
    srv.close()
    ctx.close()


def _check(srv, res, want):
    for i, (st, e, g, p) in enumerate(want):
        assert res["status"][i] == st, (i, res["status"][i], st)
        assert res["num_registry"][i] == e.size, i
        e2, g2, p2 = srv.registry(i, int(res["num_registry"][i]))
        assert np.array_equal(e2, e) and np.array_equal(g2, g), i
        assert np.array_equal(p2.view(np.uint64), p.view(np.uint64)), i
    assert np.all(res["start_s"] >= res["arrival_s"]) and np.all(res["done_s"] >= res["start_s"])


@pytest.mark.gpu
def test_served_all_at_once(served):
    """16 requests arriving together: four run at a time in overlapping slots."""
    src, cuts, srv, want, _ = served
    assert srv.slots == 4
    print(f"source edges {srv.source_edges}, slot bytes {srv.slot_bytes}", flush=True)
    res = srv.run(cuts, np.zeros(len(cuts)), keep_registries=True)
    _check(srv, res, want)
    assert sum(e.size for _, e, _, _ in want) > 0


@pytest.mark.gpu
def test_served_staggered_arrivals(served):
    """Poisson arrivals faster than service: the queue builds and drains; results unchanged."""
    src, cuts, srv, want, _ = served
    rng = np.random.default_rng(6)
    arr = np.cumsum(rng.exponential(1.0 / 4000.0, len(cuts)))
    res = srv.run(cuts, arr, keep_registries=True)
    _check(srv, res, want)


@pytest.mark.gpu
def test_served_budget_admits_one_at_a_time(served):
    """A budget below two requests' footprint: requests never overlap."""
    from mcut_amd import frontend
    src, cuts, _, want, _ = served
    mv = max(c[0].shape[0] for c in cuts)
    mf = max(c[1].size // 3 for c in cuts)
    srv = frontend.Server(src, mv, mf, slots=4, hbm_budget=1)
    try:
        res = srv.run(cuts, np.zeros(len(cuts)), keep_registries=True)
        _check(srv, res, want)
        order = np.argsort(res["start_s"])
        assert np.all(res["start_s"][order][1:] >= res["done_s"][order][:-1])
    finally:
        srv.close()


@pytest.mark.gpu
def test_served_request_against_oracle(served):
    from oracle import oracle as om
    src, cuts, srv, _, _ = served
    res = srv.run(cuts[:3], np.zeros(3), keep_registries=True)
    for i in range(3):
        O = om.frontend(src, cuts[i], detail=False)
        assert res["status"][i] == O["status"]
        e, g, p = srv.registry(i, int(res["num_registry"][i]))
        assert np.array_equal(e, O["reg_edge"]) and np.array_equal(g, O["reg_face"])
        assert np.array_equal(p.view(np.uint64), O["reg_point"].view(np.uint64))

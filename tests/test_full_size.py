"""The batched mix and the served stream at their full stated sizes, through
size-independent properties: every pair of the 4096-pair batch (130 M faces)
comes out exactly as its own plain dispatch gives it, and every request cut
from the resident 165,620-triangle source is bit-identical alone, in a mix of
overlapping slots and under Poisson arrivals.  (The plain dispatch is held to
the oracle by test_gpu_parity.py.)"""
import numpy as np
import pytest

from mcut_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def test_config4_batch_equals_plain_dispatch_at_full_size():
    from mcut_amd import frontend
    print("generating 4096 pairs (seed 5)", flush=True)
    pairs = mg.config4_pairs(4096, seed=5)
    P = frontend.pack_batch(pairs)
    nfaces = int(P["src_fo"][-1]) + int(P["cut_fo"][-1])
    assert nfaces > 100_000_000
    bctx = frontend.Context(0)
    try:
        B = bctx.dispatch_batch(packed=P)
        print(f"batch done: {B.counts}", flush=True)
        be, bg, bp = B.registry()
        be, bg, bp = be.copy(), bg.copy(), bp.copy()
        offs = B.reg_offsets.astype(np.int64)
        status_b = B.status.copy()
        partial_b = B.partial_cut.copy()
        T = B.candidate_pairs()
        npairs_b = np.bincount(T[:, 0].astype(np.int64), minlength=len(pairs))
        del T
    finally:
        bctx.close()
    pctx = frontend.Context(0)
    try:
        for i, (src, cut) in enumerate(pairs):
            if i % 512 == 0:
                print(f"plain dispatch {i}", flush=True)
            R = pctx.dispatch(src, cut)
            assert status_b[i] == R.status and partial_b[i] == R.partial_cut, (i, status_b[i], R.status)
            assert npairs_b[i] == R.counts["pairs"], i
            e, g, p = R.registry()
            a, b = offs[i], offs[i + 1]
            assert np.array_equal(be[a:b], e) and np.array_equal(bg[a:b], g), i
            assert np.array_equal(bp[a:b].view(np.uint64), p.view(np.uint64)), i
    finally:
        pctx.close()
    assert offs[-1] == be.size > 1_000_000


def test_config5_requests_identical_alone_and_in_the_mix_at_full_size():
    from mcut_amd import frontend
    src = mg.config5_source(91)
    assert src[1].size // 3 == 165_620
    cuts = [mg.crack_sheet(k, seed=6) for k in range(24)]
    sizes = [c[1].size // 3 for c in cuts]
    assert min(sizes) >= 1_500 and max(sizes) <= 51_000 and max(sizes) > 20_000
    ctx = frontend.Context(0)
    want = []
    try:
        for c in cuts:
            r = ctx.dispatch(src, c)
            want.append((r.status, *r.registry()))
    finally:
        ctx.close()
    print(f"plain dispatches done: {sum(w[1].size for w in want)} registry entries", flush=True)
    mv = max(c[0].shape[0] for c in cuts)
    mf = max(sizes)

    def check(srv, res):
        for i, (st, e, g, p) in enumerate(want):
            assert res["status"][i] == st, i
            assert res["num_registry"][i] == e.size, i
            e2, g2, p2 = srv.registry(i, int(res["num_registry"][i]))
            assert np.array_equal(e2, e) and np.array_equal(g2, g), i
            assert np.array_equal(p2.view(np.uint64), p.view(np.uint64)), i

    srv = frontend.Server(src, mv, mf, slots=4, hbm_budget=16 << 30)
    try:
        check(srv, srv.run(cuts, np.zeros(len(cuts)), keep_registries=True))  # all at once: overlapping slots
        arr = np.cumsum(np.random.default_rng(6).exponential(1.0 / 1000.0, len(cuts)))
        check(srv, srv.run(cuts, arr, keep_registries=True))  # Poisson arrivals at 1000 requests/s
    finally:
        srv.close()
    alone = frontend.Server(src, mv, mf, slots=4, hbm_budget=1)  # one request in flight at a time
    try:
        res = alone.run(cuts, np.zeros(len(cuts)), keep_registries=True)
        check(alone, res)
        order = np.argsort(res["start_s"])
        assert np.all(res["start_s"][order][1:] >= res["done_s"][order][:-1])
    finally:
        alone.close()

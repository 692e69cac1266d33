"""Precision stage: the FP32 orient3d pre-filter ahead of the FP64 filter,
the degenerate-heavy sectioning workload, and the predicate stage alone.

The FP32 bound is pinned on the CPU: a float32 restatement of the kernel's
arithmetic (same operation order, IEEE round to nearest, no contraction) on
adversarial near-coplanar inputs must never decide a sign that differs from
the exact sign (the oracle's expansion arithmetic, itself pinned to exact
rationals in test_oracle.py)."""
import numpy as np
import pytest

from mcut_amd import meshgen as mg

F32_BOUND = np.float32(9.5367431640625e-07)  # 16 * 2^-24
LO, HI = np.float32(2.0 ** -40), np.float32(2.0 ** 40)


def f32_filter(pa, pb, pc, pd):
    """Vectorised restatement of orient3d_filter_f32: +1 / -1 when decided, 0 when not."""
    d = [(P[:, j] - pd[:, j]).astype(np.float32) for P in (pa, pb, pc) for j in range(3)]
    adx, ady, adz, bdx, bdy, bdz, cdx, cdy, cdz = d
    ok = np.ones(pa.shape[0], bool)
    for x in d:
        a = np.abs(x)
        ok &= (a == 0) | ((a >= LO) & (a <= HI))
    with np.errstate(all="ignore"):
        bdxcdy, cdxbdy = bdx * cdy, cdx * bdy
        cdxady, adxcdy = cdx * ady, adx * cdy
        adxbdy, bdxady = adx * bdy, bdx * ady
        det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady)
        perm = (np.abs(bdxcdy) + np.abs(cdxbdy)) * np.abs(adz) + (np.abs(cdxady) + np.abs(adxcdy)) * np.abs(bdz) + \
               (np.abs(adxbdy) + np.abs(bdxady)) * np.abs(cdz)
        eb = F32_BOUND * perm
    assert det.dtype == np.float32 and eb.dtype == np.float32
    s = np.where(det > eb, 1, np.where(-det > eb, -1, 0))
    return np.where(ok, s, 0)


def _near_coplanar(rng, n, scale):
    """Three points of a random plane and a fourth on it (snapped to doubles),
    moved off it by a random amount down to the rounding level, at a given
    coordinate scale; a quarter exactly coplanar on an integer lattice."""
    base = rng.uniform(-1, 1, (n, 3)) * scale + rng.uniform(-4, 4, (n, 3)) * scale
    u = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    a = base
    b = base + u * scale * rng.uniform(1e-6, 1, (n, 1))
    c = base + v * scale * rng.uniform(1e-6, 1, (n, 1))
    s, t = rng.uniform(-2, 2, (n, 1)), rng.uniform(-2, 2, (n, 1))
    d = a + s * (b - a) + t * (c - a)
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    off = 10.0 ** rng.uniform(-17, -2, (n, 1)) * scale * rng.choice([-1, 1], (n, 1))
    d = d + nrm * off
    q = n // 4  # exact coplanarity: lattice points
    L = rng.integers(-50, 50, (q, 3, 3)).astype(np.float64)
    w = rng.integers(-3, 4, (q, 2)).astype(np.float64)
    a[:q], b[:q], c[:q] = L[:, 0], L[:, 1], L[:, 2]
    d[:q] = L[:, 0] + w[:, :1] * (L[:, 1] - L[:, 0]) + w[:, 1:] * (L[:, 2] - L[:, 0])
    return a, b, c, d


@pytest.mark.parametrize("scale", [1e-6, 1e-3, 1.0, 37.5, 1e4])
def test_f32_prefilter_never_wrong(scale):
    from oracle import oracle as om
    rng = np.random.default_rng(int(scale * 1000) % 9973)
    a, b, c, d = _near_coplanar(rng, 6000, scale)
    s = f32_filter(a, b, c, d)
    decided = np.nonzero(s)[0]
    assert decided.size > 1000  # the filter settles most well-separated cases
    for i in decided:
        assert om.orient3d_sign(a[i], b[i], c[i], d[i]) == s[i], (scale, i)
    # exactly coplanar lattice cases are never decided
    assert np.all(s[:1500] == 0)


def test_f32_prefilter_declines_out_of_range():
    z = np.zeros((1, 3))
    big = np.array([[2.0 ** 41, 1.0, 1.0]])
    tiny = np.array([[2.0 ** -41, 1.0, 1.0]])
    e = np.eye(3)
    assert f32_filter(z + big, z + e[0:1], z + e[1:2], z)[0] == 0
    assert f32_filter(z + tiny, z + e[0:1], z + e[1:2], z + e[2:3] * 0.5)[0] == 0
    assert f32_filter(e[0:1], e[1:2], e[2:3], z)[0] != 0


def test_config3_generator():
    (V, F, _), (Vc, Fc, _) = mg.config3(cells=64)
    assert V.shape == (65 * 65, 3) and F.size == 3 * 2 * 64 * 64
    assert np.all(V[:, 2] * 4096 == np.round(V[:, 2] * 4096)) and V[:, 2].min() == 0 and V[:, 2].max() == 0.2 * 4096 // 1 / 4096
    assert np.all(Vc[:, 2] == 410.0 / 4096.0) and Fc.size == 3 * 2 * 64 * 64
    assert np.any(V[:, 2] == 410.0 / 4096.0)  # vertices exactly on the cut plane


@pytest.fixture(scope="module")
def ctx():
    from mcut_amd import frontend
    c = frontend.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
def test_config3_status(ctx):
    """The full heightfield (2,097,152 faces): candidate faces' normals fall
    below the reference's 1e-9 threshold, so the dispatch reports an invalid
    source mesh, as the reference does."""
    from oracle import oracle as om
    src, cut = mg.config3()
    R = ctx.dispatch(src, cut)
    O = om.frontend(src, cut, detail=False)
    assert R.status == O["status"] == om.INVALID_SRC_MESH
    assert np.array_equal(R.pairs(), O["pairs"])


@pytest.mark.gpu
def test_config3_scaled_parity(ctx):
    """Faces 1 x 1 (scale = cells): planes valid, thousands of vertices on
    the cut plane exactly; every sign, type and point-in-polygon result must
    match the oracle."""
    import test_gpu_parity as tp
    src, cut = mg.config3(cells=256, scale=256.0)
    R, O = tp.compare(ctx, src, cut, detail=True)
    assert R.status == O["status"]
    assert R.counts["escalated"] > 0 and R.counts["decided_f32"] > R.counts["decided_f64"]


@pytest.mark.gpu
def test_predicate_stage_signs(ctx):
    """The predicate stage alone gives the narrow phase's signs on every test,
    the exact signs on a sample (all zeros included), and counts that add up."""
    from oracle import oracle as om
    src, cut = mg.config3(cells=256, scale=256.0)
    R = ctx.dispatch(src, cut, keep_tests=True)
    T = R.tests()
    P = R.predicate_stage(signs=True)
    n = R.counts["tests"]
    assert P["f32"] + P["f64"] + P["exact"] == 2 * n
    assert P["negative"] + P["zero"] + P["positive"] == 2 * n
    assert P["zero"] > 0 and P["exact"] >= P["zero"]
    narrow = T["type"] != ord("0")
    assert np.array_equal(P["qsign"][narrow], T["qsign"][narrow])
    assert np.array_equal(P["rsign"][narrow], T["rsign"][narrow])
    # exact signs on a sample: polygon-soup vertex lists and plane triplets
    ps = R.ps_vertices()
    pl = R.planes()
    tri_of = {int(f): t for f, t in zip(pl["face"], pl["triple"])}
    Fs, Vs = src[1].size // 3, src[0].shape[0]
    v0, v1, _, _ = R.edges()

    def face_vertex(g, i):
        if g < Fs:
            return int(src[1][3 * g + (i + 1) % 3])
        c = g - Fs
        return Vs + int(cut[1][3 * c + (i + 2) % 3])
    rng = np.random.default_rng(1)
    zeros = np.nonzero(P["qsign"] == 0)[0][:200]
    pick = np.concatenate([zeros, rng.choice(n, 300, replace=False)])
    for i in pick:
        g = int(T["face"][i])
        t = tri_of[g]
        t = (0, 1, 2) if t[0] < 0 else t
        a, b, c = (ps[face_vertex(g, int(k))] for k in t)
        e = int(T["edge"][i])
        assert om.orient3d_sign(a, b, c, ps[v0[e]]) == P["qsign"][i], i
        assert om.orient3d_sign(a, b, c, ps[v1[e]]) == P["rsign"][i], i

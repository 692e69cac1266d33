"""Deterministic synthetic mesh generators and an OFF reader.

Every generator returns ``(V, F, S)``: ``V`` float64 ``(nv, 3)``, ``F`` uint32
flat face-index array, ``S`` uint32 face sizes or ``None`` when every face is a
triangle (the convention of ``mcDispatch``'s ``pFaceSizes == NULL``,
mcut.h:1008-1020, preproc.cpp:206).

Randomness comes from a counter-based SplitMix64 stream (``u01``), so a value
depends only on ``(seed, stream, index)`` and never on call order or platform.
"""
from __future__ import annotations

import math
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def u64(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """n 64-bit words of the SplitMix64 stream (seed, stream), from index start."""
    base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
    idx = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = base + idx * _GOLD
    return _mix64(z)


def u01(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """n doubles uniform on [0,1) with 53 random bits each."""
    return (u64(seed, stream, n, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


# --------------------------------------------------------------------------
# primitive shapes
# --------------------------------------------------------------------------

def uv_sphere(slices: int = 32, stacks: int = 16, r: float = 1.0, phi0: float = 0.0):
    """UV sphere, poles on +-z, ring k at theta=k*pi/stacks, vertex j at
    phi=phi0+2*pi*j/slices; caps are fans, bands are quads split into two
    triangles; outward (counter-clockwise seen from outside) winding."""
    verts = [(0.0, 0.0, r)]
    for k in range(1, stacks):
        th = k * math.pi / stacks
        for j in range(slices):
            ph = phi0 + 2.0 * math.pi * j / slices
            verts.append((r * math.sin(th) * math.cos(ph), r * math.sin(th) * math.sin(ph), r * math.cos(th)))
    verts.append((0.0, 0.0, -r))
    south = len(verts) - 1

    def ring(k, j):
        return 1 + (k - 1) * slices + (j % slices)

    faces = []
    for j in range(slices):
        faces.append((0, ring(1, j), ring(1, j + 1)))
    for k in range(1, stacks - 1):
        for j in range(slices):
            a, b = ring(k, j), ring(k, j + 1)
            c, d = ring(k + 1, j), ring(k + 1, j + 1)
            faces.append((a, c, d))
            faces.append((a, d, b))
    for j in range(slices):
        faces.append((ring(stacks - 1, j), south, ring(stacks - 1, j + 1)))
    return np.asarray(verts, dtype=np.float64), np.asarray(faces, dtype=np.uint32).ravel(), None


def tessellated_box(center=(0.0, 0.0, 0.0), edge=1.0, m: int = 4, rotation=None, triangulate=True):
    """Axis-aligned cube of side ``edge`` (scalar or 3-vector) centred at
    ``center``; each face split into m x m quads (then into two triangles when
    ``triangulate``).  Grid vertices are shared between faces; winding is
    outward.  ``rotation`` (3x3) is applied about the centre."""
    edge = np.broadcast_to(np.asarray(edge, dtype=np.float64), (3,))
    half = edge / 2.0
    key_to_id = {}
    verts = []

    def vid(ix, iy, iz):
        key = (ix, iy, iz)
        if key not in key_to_id:
            key_to_id[key] = len(verts)
            verts.append((-half[0] + edge[0] * ix / m, -half[1] + edge[1] * iy / m, -half[2] + edge[2] * iz / m))
        return key_to_id[key]

    faces = []
    sizes = []
    # each face: (fixed axis, fixed value index, u axis, v axis) with u x v = outward normal
    for axis in range(3):
        for side in (0, m):
            u_ax, v_ax = [(1, 2), (2, 0), (0, 1)][axis]
            if side == 0:
                u_ax, v_ax = v_ax, u_ax
            for a in range(m):
                for b in range(m):
                    def p(du, dv):
                        c = [0, 0, 0]
                        c[axis] = side
                        c[u_ax] = a + du
                        c[v_ax] = b + dv
                        return vid(*c)
                    q = (p(0, 0), p(1, 0), p(1, 1), p(0, 1))
                    if triangulate:
                        faces.append((q[0], q[1], q[2]))
                        faces.append((q[0], q[2], q[3]))
                        sizes += [3, 3]
                    else:
                        faces.append(q)
                        sizes.append(4)
    V = np.asarray(verts, dtype=np.float64)
    if rotation is not None:
        V = V @ np.asarray(rotation, dtype=np.float64).T
    V = V + np.asarray(center, dtype=np.float64)
    F = np.asarray([i for f in faces for i in f], dtype=np.uint32)
    return V, F, (None if triangulate else np.asarray(sizes, dtype=np.uint32))


_ICO_V = None


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
                  (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
                  (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
                  (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
                  (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    return v, f


def icosphere(level: int):
    """Icosphere: icosahedron, ``level`` rounds of 1->4 midpoint subdivision
    with re-projection to the unit sphere.  20*4^level faces."""
    V, F = _icosahedron()
    for _ in range(level):
        nv = V.shape[0]
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
        lo = np.minimum(e[:, 0], e[:, 1])
        hi = np.maximum(e[:, 0], e[:, 1])
        key = lo * (1 << 32) + hi
        uk, inv = np.unique(key, return_inverse=True)
        ulo = uk >> 32
        uhi = uk & 0xFFFFFFFF
        mid = V[ulo] + V[uhi]
        mid /= np.linalg.norm(mid, axis=1)[:, None]
        V = np.concatenate([V, mid], axis=0)
        m = (inv + nv).reshape(3, -1)
        m01, m12, m20 = m[0], m[1], m[2]
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        F = np.concatenate([np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1),
                            np.stack([c, m20, m12], 1), np.stack([m01, m12, m20], 1)], axis=0)
    return V, F.astype(np.uint32).ravel(), None


_GEODESIC = {}


def geodesic_sphere(n: int):
    """Class-I geodesic sphere of frequency n: each icosahedron face is split
    into n^2 triangles on a barycentric grid, points projected to the unit
    sphere.  20*n^2 faces, 10*n^2+2 vertices (grid points shared between
    faces through their (icosahedron vertex, weight) sets, so a shared point
    gets one id and one coordinate)."""
    if n in _GEODESIC:
        V, F = _GEODESIC[n]
        return V.copy(), F.copy(), None
    V0, F0 = _icosahedron()
    I, J = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    ok = (I + J) <= n
    gi, gj = I[ok], J[ok]                          # grid points of one face, i-major
    G = gi.size
    lmap = -np.ones((n + 2, n + 2), np.int64)
    lmap[gi, gj] = np.arange(G)
    w = np.stack([n - gi - gj, gi, gj], 1)         # weights of (a, b, c)
    vid = F0[:, None, :].repeat(G, 1)              # (20, G, 3)
    ww = np.broadcast_to(w[None], (20, G, 3))
    code = np.where(ww > 0, vid * (n + 1) + ww, -1).reshape(-1, 3)
    code.sort(axis=1)
    uk, inv = np.unique(code, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    P = np.zeros((uk.shape[0], 3))
    for k in range(3):
        c = uk[:, k]
        has = c >= 0
        P[has] += V0[c[has] // (n + 1)] * (c[has] % (n + 1))[:, None].astype(np.float64)
    P /= n
    P /= np.linalg.norm(P, axis=1)[:, None]
    # cells (i, j), i + j <= n - 1, i-major: an up triangle each, a down one when i + j <= n - 2
    ci, cj = gi[gi + gj <= n - 1], gj[gi + gj <= n - 1]
    up = np.stack([lmap[ci, cj], lmap[ci + 1, cj], lmap[ci, cj + 1]], 1)
    dn = np.stack([lmap[ci + 1, cj], lmap[ci + 1, cj + 1], lmap[ci, cj + 1]], 1)
    both = np.stack([up, dn], 1)                   # (cells, 2, 3)
    keep = np.stack([np.ones(ci.size, bool), ci + cj <= n - 2], 1)
    local = both[keep]                             # per-face local triangles in order
    Fall = (local[None, :, :] + (np.arange(20) * G)[:, None, None]).reshape(-1)
    F = inv[Fall].astype(np.uint32)
    _GEODESIC[n] = (P, F)
    return P.copy(), F.copy(), None


_BOX = {}


def _box_grid(m: int):
    """Integer grid coordinates (0..m) and triangles of tessellated_box's
    m x m-per-face cube (cached per m)."""
    if m not in _BOX:
        V, F, _ = tessellated_box((0.0, 0.0, 0.0), float(m), m)
        _BOX[m] = (V + m / 2.0, F)
    return _BOX[m]


def random_rotation(seed: int, stream: int = 0) -> np.ndarray:
    """Uniform random rotation (Shoemake's subgroup algorithm for a unit quaternion)."""
    u1, u2, u3 = u01(seed, stream, 3)
    return rotation_from_uniforms(u1, u2, u3)


def rotation_from_uniforms(u1: float, u2: float, u3: float) -> np.ndarray:
    """Shoemake's map of three U[0,1) draws to a uniform random rotation matrix."""
    a = math.sqrt(1.0 - u1)
    b = math.sqrt(u1)
    w = a * math.sin(2 * math.pi * u2)
    x = a * math.cos(2 * math.pi * u2)
    y = b * math.sin(2 * math.pi * u3)
    z = b * math.cos(2 * math.pi * u3)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]], dtype=np.float64)


def radial_displace(V: np.ndarray, amp: float, seed: int, stream: int = 1) -> np.ndarray:
    """Scale each vertex radially by 1 + U(-amp, amp)."""
    s = 1.0 + amp * (2.0 * u01(seed, stream, V.shape[0]) - 1.0)
    return V * s[:, None]


def grid_patch(m: int, side: float = 1.0):
    """Open planar m x m quad grid on [-side/2, side/2]^2 at z=0, triangulated."""
    g = np.linspace(-side / 2.0, side / 2.0, m + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    V = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    a = (i * (m + 1) + j).ravel()
    b = a + (m + 1)
    c = b + 1
    d = a + 1
    F = np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], 1).reshape(-1, 3)
    return V, F.astype(np.uint32).ravel(), None


def value_noise_fbm(x: np.ndarray, y: np.ndarray, octaves: int, seed: int) -> np.ndarray:
    """2-D fBm of lattice value noise (smoothstep interpolation, lacunarity 2, gain 0.5)."""
    out = np.zeros_like(x)
    amp, freq = 1.0, 4.0
    for o in range(octaves):
        xs, ys = x * freq, y * freq
        xi, yi = np.floor(xs).astype(np.int64), np.floor(ys).astype(np.int64)
        fx, fy = xs - xi, ys - yi
        sx, sy = fx * fx * (3 - 2 * fx), fy * fy * (3 - 2 * fy)

        def lat(ix, iy):
            k = (ix.astype(np.uint64) * np.uint64(73856093)) ^ (iy.astype(np.uint64) * np.uint64(19349663))
            with np.errstate(over="ignore"):
                z = _mix64(k + np.uint64((seed * 1000003 + o) & 0xFFFFFFFF) * _GOLD)
            return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

        n00, n10, n01, n11 = lat(xi, yi), lat(xi + 1, yi), lat(xi, yi + 1), lat(xi + 1, yi + 1)
        nx0 = n00 + (n10 - n00) * sx
        nx1 = n01 + (n11 - n01) * sx
        out += amp * (nx0 + (nx1 - nx0) * sy)
        amp *= 0.5
        freq *= 2.0
    return out


# --------------------------------------------------------------------------
# benchmark workloads
# --------------------------------------------------------------------------

def config1(enforce_variant: bool = False, phi0: float = 0.0):
    """UV sphere r=1 (32 slices x 16 stacks, 960 tris) and an axis-aligned cube
    of edge 1.2 centred at (0.4, 0.3, 0.2), 4x4 quads per face triangulated
    (192 tris)."""
    src = uv_sphere(32, 16, 1.0, phi0)
    cut = tessellated_box((0.4, 0.3, 0.2), 1.2, 4)
    return src, cut


def config2(level: int = 7, amp: float = 0.02):
    """Two icospheres (20*4^level tris each), radial noise U(-amp, amp) from
    seeds 1 and 2; the second is rotated by a random rotation (seed 3) about
    its centre and translated by +0.5 in x."""
    V, F, _ = icosphere(level)
    A = radial_displace(V, amp, seed=1)
    B = radial_displace(V, amp, seed=2)
    R = random_rotation(seed=3)
    B = B @ R.T
    B[:, 0] += 0.5
    return (A, F, None), (B.copy(), F.copy(), None)


def config4_pairs(num_pairs: int = 4096, seed: int = 5, amp: float = 0.02, tri_range=(200, 200_000),
                  box_range=(12, 12_000)):
    """Batched mix: independent (sphere, box) pairs.  Pair k draws, from its
    own stream: sphere triangles T ~ logU(tri_range) -> Class-I geodesic
    frequency n = round(sqrt(T / 20)) clamped to [4, 100], radius 1, radial
    noise U(-amp, amp); box triangles T_b ~ logU(box_range) -> m = round(sqrt(
    T_b / 12)) quads per face edge, triangulated, uniform random rotation,
    edge s ~ U(0.5, 1.5), centre at distance 1 + s/2 - delta along a uniform
    random direction, delta ~ U(0.1, 0.6).  A draw whose sphere has a face of
    squared Newell length below 1e-9 (rejected by the reference,
    math.cpp:151-152) is redrawn."""
    pairs = []
    lt, lb = math.log(tri_range[0]), math.log(box_range[0])
    rt, rb = math.log(tri_range[1] / tri_range[0]), math.log(box_range[1] / box_range[0])
    for k in range(num_pairs):
        for a in range(16):
            st = 1000 + 16 * k + a
            u = u01(seed, st, 12)
            n = int(min(100, max(4, round(math.sqrt(math.exp(lt + u[0] * rt) / 20.0)))))
            m = int(max(1, round(math.sqrt(math.exp(lb + u[1] * rb) / 12.0))))
            R = rotation_from_uniforms(u[2], u[3], u[4])
            sc = 0.5 + u[5]
            z = 2.0 * u[6] - 1.0
            ph = 2.0 * math.pi * u[7]
            rz = math.sqrt(max(0.0, 1.0 - z * z))
            d = np.array([rz * math.cos(ph), rz * math.sin(ph), z])
            delta = 0.1 + 0.5 * u[8]
            V, F, _ = geodesic_sphere(n)
            V = radial_displace(V, amp, seed=seed, stream=st + (1 << 40))
            if newell_min_sq(V, F) < 1e-9:
                continue
            G, Fb = _box_grid(m)
            Vb = ((G / m) - 0.5) * sc
            Vb = Vb @ R.T + d * (1.0 + sc / 2.0 - delta)
            pairs.append(((V, F, None), (np.ascontiguousarray(Vb), Fb.copy(), None)))
            break
        else:
            raise RuntimeError(f"pair {k}: no valid draw")
    return pairs


def config3(cells: int = 1024, scale: float = 1.0, seed: int = 4, cut_quads: int = 64):
    """Degenerate-heavy sectioning.  Source: heightfield over a cells x cells
    grid on [0, scale]^2 (two triangles per cell, normals up), heights =
    6-octave fBm (value noise, lacunarity 2, gain 0.5, seed) rescaled to
    [0, 0.2] and rounded to multiples of 2^-12.  Cut: a planar cut_quads x
    cut_quads quad grid, triangulated, at z = 410/4096 (the representable
    height nearest 0.1, which many heightfield vertices have exactly) over
    [-0.005, 1.005]^2 * scale.  scale = cells makes every source cell 1 x 1,
    which keeps the source's face normals above the reference's 1e-9
    threshold (at scale 1 they are about cells^-4)."""
    i = np.arange(cells + 1, dtype=np.float64)
    X, Y = np.meshgrid(i / cells, i / cells, indexing="ij")
    h = value_noise_fbm(X.ravel(), Y.ravel(), 6, seed)
    h = (h - h.min()) / (h.max() - h.min()) * 0.2
    h = np.round(h * 4096.0) / 4096.0
    V = np.stack([X.ravel() * scale, Y.ravel() * scale, h], axis=1)
    a_i, a_j = np.meshgrid(np.arange(cells), np.arange(cells), indexing="ij")
    a = (a_i * (cells + 1) + a_j).ravel()
    b = a + (cells + 1)
    F = np.stack([np.stack([a, b, b + 1], 1), np.stack([a, b + 1, a + 1], 1)], 1).reshape(-1)
    G, Fc, _ = grid_patch(cut_quads, 1.01 * scale)
    Vc = G + np.array([0.5 * scale, 0.5 * scale, 410.0 / 4096.0])
    return (V, F.astype(np.uint32), None), (np.ascontiguousarray(Vc), Fc, None)


def config5_source(n: int = 91, amp: float = 0.02, seed: int = 6):
    """Served stream's resident source: Class-I geodesic sphere of frequency n
    (n = 91: 165,620 triangles), radial noise U(-amp, amp)."""
    V, F, _ = geodesic_sphere(n)
    return radial_displace(V, amp, seed=seed, stream=7), F, None


def crack_sheet(k: int, seed: int = 6, tri_range=(2_000, 50_000), side: float = 3.0, amp: float = 0.03):
    """Request k of the served stream: an open m x m quad patch (two triangles
    per quad), m = round(sqrt(T / 2)) with T ~ logU(tri_range), edge `side`,
    normal uniform on the sphere, offset along the normal U(-0.6, 0.6), and
    heights amp * fBm (4 octaves, centred) along the normal."""
    u = u01(seed, 50_000 + k, 8)
    T = math.exp(math.log(tri_range[0]) + u[0] * math.log(tri_range[1] / tri_range[0]))
    m = max(1, int(round(math.sqrt(T / 2.0))))
    z = 2.0 * u[1] - 1.0
    ph = 2.0 * math.pi * u[2]
    rz = math.sqrt(max(0.0, 1.0 - z * z))
    nrm = np.array([rz * math.cos(ph), rz * math.sin(ph), z])
    a = np.array([1.0, 0.0, 0.0]) if abs(nrm[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(nrm, a)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(nrm, e1)
    rot = 2.0 * math.pi * u[3]
    e1, e2 = math.cos(rot) * e1 + math.sin(rot) * e2, -math.sin(rot) * e1 + math.cos(rot) * e2
    off = -0.6 + 1.2 * u[4]
    G, F, _ = grid_patch(m, side)
    h = value_noise_fbm(G[:, 0] / side + 0.5, G[:, 1] / side + 0.5, 4, seed * 1000 + k)
    h = amp * (h - 0.9375) / 0.9375
    V = off * nrm[None, :] + G[:, 0:1] * e1[None, :] + G[:, 1:2] * e2[None, :] + h[:, None] * nrm[None, :]
    return np.ascontiguousarray(V), F, None


def read_off(path: str):
    """Read an OFF file; returns (V, F, S) with S None when all faces are triangles."""
    with open(path, "r") as fh:
        toks = []
        for line in fh:
            line = line.split("#", 1)[0].strip()
            if line:
                toks.extend(line.split())
    assert toks[0] == "OFF", path
    nv, nf = int(toks[1]), int(toks[2])
    p = 4
    V = np.asarray(toks[p:p + 3 * nv], dtype=np.float64).reshape(nv, 3)
    p += 3 * nv
    F, S = [], []
    for _ in range(nf):
        k = int(toks[p])
        F.extend(int(t) for t in toks[p + 1:p + 1 + k])
        S.append(k)
        p += 1 + k
    S = np.asarray(S, dtype=np.uint32)
    F = np.asarray(F, dtype=np.uint32)
    return V, F, (None if np.all(S == 3) else S)


def read_obj(path: str):
    """Read the vertices and faces of an OBJ file (``v x y z``, ``f a b c ...``
    with 1-based or negative indices, ``a/t/n`` forms allowed; other records
    ignored); returns (V, F, S) like read_off."""
    V, F, S = [], [], []
    with open(path, "r") as fh:
        for line in fh:
            t = line.split("#", 1)[0].split()
            if not t:
                continue
            if t[0] == "v":
                V.append([float(t[1]), float(t[2]), float(t[3])])
            elif t[0] == "f":
                ids = []
                for w in t[1:]:
                    k = int(w.split("/", 1)[0])
                    ids.append(k - 1 if k > 0 else len(V) + k)
                F.extend(ids)
                S.append(len(ids))
    V = np.asarray(V, dtype=np.float64).reshape(-1, 3)
    S = np.asarray(S, dtype=np.uint32)
    F = np.asarray(F, dtype=np.uint32)
    return V, F, (None if np.all(S == 3) else S)


def _faces(F, S):
    F = np.asarray(F)
    if S is None:
        return F.reshape(-1, 3).tolist()
    out, at = [], 0
    for k in np.asarray(S):
        out.append(F[at:at + int(k)].tolist())
        at += int(k)
    return out


def write_off(path: str, V, F, S=None):
    """Write (V, F, S) as OFF the way hmesh.cpp:1268-1340 lays it out: the
    header, counts with 0 edges, one vertex a line, then each face's size and
    vertex ids.  Coordinates are written with 17 significant digits, so
    read_off gives back the same doubles."""
    V = np.asarray(V, dtype=np.float64).reshape(-1, 3)
    faces = _faces(F, S)
    with open(path, "w") as fh:
        fh.write("OFF\n%d %d 0\n" % (V.shape[0], len(faces)))
        for x, y, z in V:
            fh.write("%.17g %.17g %.17g\n" % (x, y, z))
        for f in faces:
            fh.write("%d %s\n" % (len(f), " ".join(str(i) for i in f)))


def write_obj(path: str, V, F, S=None):
    """Write (V, F, S) as OBJ (1-based face indices, 17 significant digits)."""
    V = np.asarray(V, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as fh:
        for x, y, z in V:
            fh.write("v %.17g %.17g %.17g\n" % (x, y, z))
        for f in _faces(F, S):
            fh.write("f %s\n" % " ".join(str(i + 1) for i in f))


def newell_min_sq(V: np.ndarray, F: np.ndarray, S=None) -> float:
    """Smallest squared Newell-normal length over the faces (plain float64)."""
    if S is None:
        T = V[F.reshape(-1, 3).astype(np.int64)]
        n = np.zeros((T.shape[0], 3))
        for i in range(3):
            c, d = T[:, i], T[:, (i + 1) % 3]
            n[:, 0] += (c[:, 1] - d[:, 1]) * (c[:, 2] + d[:, 2])
            n[:, 1] += (c[:, 2] - d[:, 2]) * (c[:, 0] + d[:, 0])
            n[:, 2] += (c[:, 0] - d[:, 0]) * (c[:, 1] + d[:, 1])
        return float(np.min(np.sum(n * n, axis=1)))
    raise NotImplementedError

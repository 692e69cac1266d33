"""Exact-rational error of intersection points (a verification helper for the
tests and the benchmark; nothing on the dispatch path calls it).

Every binary64 number is an integer times a power of two, so scaling all
coordinates of a dispatch by one power of two 2^S makes them integers, and
the exact intersection of a segment (q, r) with the plane through a triangle
(a, b, c) follows from integer determinants:

    o(x) = det[a - x; b - x; c - x]  (exact orient3d value, scaled by 2^3S)
    p* = q + o(q) / (o(q) - o(r)) * (r - q)

The reference computes its points from the face plane's unit normal and d
(math.cpp:249-287), so they carry rounding error; this measures it.
"""
from __future__ import annotations

import numpy as np


def _scaled_ints(a):
    """Python-int object array X and a shift S with a == X / 2**S exactly."""
    a = np.asarray(a, np.float64)
    if not np.all(np.isfinite(a)):
        raise ValueError("non-finite coordinate")
    m, e = np.frexp(a)
    mi = (m * 2.0 ** 53).astype(np.int64)  # exact: a = mi * 2^(e - 53)
    ex = e.astype(np.int64) - 53
    nz = mi != 0
    S = max(0, int(-ex[nz].min())) if np.any(nz) else 0
    sh = np.where(nz, ex + S, 0)
    vals = [int(x) << int(k) if x else 0 for x, k in zip(mi.ravel().tolist(), sh.ravel().tolist())]
    out = np.empty(a.size, dtype=object)
    out[:] = vals
    return out.reshape(a.shape), S


def _det3(a, b, c, d):
    A, B, C = a - d, b - d, c - d
    return (A[:, 0] * (B[:, 1] * C[:, 2] - B[:, 2] * C[:, 1]) - A[:, 1] * (B[:, 0] * C[:, 2] - B[:, 2] * C[:, 0]) +
            A[:, 2] * (B[:, 0] * C[:, 1] - B[:, 1] * C[:, 0]))


def registry_point_errors(ps_vertices, face_triangles, edge_v0, edge_v1, reg_edge, reg_face, reg_point):
    """Per registry entry: the largest relative coordinate error
    |p_i - p*_i| / max(|p*_i|, 1) of the point against the exact intersection
    of its edge's line with the plane of its (triangular) face, and whether
    the edge's endpoints lie strictly on opposite sides of that plane.

    ps_vertices: (V, 3) shifted polygon-soup coordinates; face_triangles:
    (F, 3) vertex ids of each face; edge_v0 / edge_v1: edge endpoints (q, r);
    reg_edge / reg_face / reg_point: the registry."""
    V = np.asarray(ps_vertices, np.float64).reshape(-1, 3)
    P = np.asarray(reg_point, np.float64).reshape(-1, 3)
    n = P.shape[0]
    if n == 0:
        return np.zeros(0), np.zeros(0, bool)
    X, S = _scaled_ints(np.vstack([V, P]))
    XV, XP = X[:V.shape[0]], X[V.shape[0]:]
    e = np.asarray(reg_edge, np.int64)
    g = np.asarray(reg_face, np.int64)
    tri = np.asarray(face_triangles, np.int64)[g]
    q = XV[np.asarray(edge_v0, np.int64)[e]]
    r = XV[np.asarray(edge_v1, np.int64)[e]]
    a, b, c = XV[tri[:, 0]], XV[tri[:, 1]], XV[tri[:, 2]]
    oq, orr = _det3(a, b, c, q), _det3(a, b, c, r)
    straddle = np.array([(x > 0) != (y > 0) and x != 0 and y != 0 for x, y in zip(oq.tolist(), orr.tolist())])
    den = oq - orr
    den[~straddle] = 1  # no exact point: the entry is reported by straddle
    num = q * den[:, None] + oq[:, None] * (r - q)  # p* * 2^S * den
    err = np.abs(XP * den[:, None] - num)
    unit = np.abs(den) * (1 << S)
    scale = np.maximum(np.abs(num), unit[:, None])
    rel = np.array([[float(x / y) for x, y in zip(er, sc)] for er, sc in zip(err.tolist(), scale.tolist())])
    return rel.max(axis=1), straddle


def triangle_soup_faces(src_faces, cut_faces, num_src_vertices):
    """(F, 3) polygon-soup vertex ids of two triangle meshes: source faces,
    then cut faces with their vertices offset by the source vertex count."""
    s = np.asarray(src_faces, np.int64).reshape(-1, 3)
    c = np.asarray(cut_faces, np.int64).reshape(-1, 3) + int(num_src_vertices)
    return np.vstack([s, c])

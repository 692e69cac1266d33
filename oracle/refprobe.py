"""ctypes wrapper over oracle/_ref/libref_probe.so (test infrastructure only).

The library is the reference MCUT compiled from its own sources plus a probe
(oracle/ref_probe.cpp).  Only tests and the benchmark's CPU-baseline leg load it.
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libref_probe.so")
_lib = None

MC_DISPATCH_VERTEX_ARRAY_DOUBLE = 1 << 1
MC_DISPATCH_REQUIRE_THROUGH_CUTS = 1 << 2
MC_DISPATCH_INCLUDE_VERTEX_MAP = 1 << 3
MC_DISPATCH_FILTER_SEAM_SRCMESH = 1 << 13
MC_DISPATCH_FILTER_SEAM_CUTMESH = 1 << 14
MC_DISPATCH_FILTER_ALL = sum(1 << b for b in range(5, 15))
MC_DISPATCH_ENFORCE_GENERAL_POSITION = 1 << 15
MC_DISPATCH_ENFORCE_GENERAL_POSITION_ABSOLUTE = 1 << 16

CC_FRAGMENT, CC_PATCH, CC_SEAM, CC_INPUT = 1, 4, 8, 16

_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint32)


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        mesh = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]
        L.ref_dispatch.argtypes = [C.c_uint32, C.c_int] + mesh + mesh
        L.ref_dispatch.restype = C.c_int
        L.ref_dispatch_ms.restype = C.c_double
        L.ref_dispatch_debug.argtypes = [C.c_uint32] + mesh + mesh
        L.ref_dispatch_debug.restype = C.c_int
        L.ref_kernel_message.restype = C.c_char_p
        L.ref_cc_count.restype = C.c_uint32
        L.ref_cc_sizes.argtypes = [C.c_uint32, C.POINTER(C.c_uint64)]
        L.ref_cc_sizes.restype = C.c_uint32
        L.ref_cc_copy.argtypes = [C.c_uint32] + [C.c_void_p] * 6
        L.ref_broad_phase.argtypes = [C.c_int, C.c_double] + mesh + mesh + [C.c_void_p]
        L.ref_broad_phase.restype = C.c_int64
        L.ref_broad_copy.argtypes = [C.c_void_p] * 6
        L.ref_plane.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.ref_plane.restype = C.c_int
        L.ref_segment_type.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        L.ref_segment_type.restype = C.c_int
        L.ref_segment_point.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_segment_point.restype = C.c_int
        L.ref_point_in_polygon.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        L.ref_point_in_polygon.restype = C.c_int
        L.ref_orient3d.argtypes = [C.c_void_p] * 4
        L.ref_orient3d.restype = C.c_double
        L.ref_orient2d.argtypes = [C.c_void_p] * 3
        L.ref_orient2d.restype = C.c_double
        L.ref_best_triple.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.ref_best_triple.restype = C.c_int
        L.ref_ps_edges.argtypes = mesh + mesh
        L.ref_ps_edges.restype = C.c_int64
        L.ref_ps_edges_copy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _mesh_args(m):
    V, F, S = m
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.uint32)
    S = None if S is None else np.ascontiguousarray(S, dtype=np.uint32)
    nf = (F.size // 3) if S is None else S.size
    return (V, F, S), [_ptr(V), _ptr(F), _ptr(S), V.shape[0], nf]


def dispatch(src, cut, flags, helpers=0, workdir=None):
    """mcDispatch through the reference's public API.

    The reference writes debug meshes into the current directory on every
    dispatch (kernel_input.verbose is hard-wired on, preproc.cpp:2371), so the
    call runs inside a scratch directory.  Returns (status, ms, ccs) where ccs
    is a list of dicts with type, vertices, faces, face_sizes, seam, vmap,
    perturbation.
    """
    L = lib()
    keep_s, sa = _mesh_args(src)
    keep_c, ca = _mesh_args(cut)
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        os.chdir(d)
        try:
            st = L.ref_dispatch(flags, helpers, *sa, *ca)
        finally:
            os.chdir(cwd)
    ms = L.ref_dispatch_ms()
    ccs = []
    for i in range(L.ref_cc_count()):
        sz = (C.c_uint64 * 5)()
        typ = L.ref_cc_sizes(i, sz)
        v = np.zeros(sz[0], np.float64)
        f = np.zeros(sz[1], np.uint32)
        fs = np.zeros(sz[2], np.uint32)
        se = np.zeros(sz[3], np.uint32)
        vm = np.zeros(sz[4], np.uint32)
        pv = np.zeros(3, np.float64)
        L.ref_cc_copy(i, _ptr(v), _ptr(f), _ptr(fs), _ptr(se), _ptr(vm), _ptr(pv))
        ccs.append(dict(type=typ, vertices=v.reshape(-1, 3), faces=f, face_sizes=fs, seam=se, vmap=vm,
                        perturbation=pv))
    return st, ms, ccs


KERNEL_STATUS = ("INVALID_SRC_MESH", "INVALID_CUT_MESH", "INVALID_MESH_INTERSECTION", "GENERAL_POSITION_VIOLATION")


def dispatch_status(src, cut, flags, workdir=None):
    """mcDispatch in a debug context: (McResult, kernel status name).  The
    reference maps several kernel statuses onto one McResult
    (preproc.cpp:580-600); the status is read from its reports to the debug
    callback: "<status> : <reason>" before it throws (preproc.cpp:2947-2955),
    or a general-position violation once the perturbation attempts run out
    (preproc.cpp:2596-2612).  "" when the dispatch succeeded."""
    L = lib()
    keep_s, sa = _mesh_args(src)
    keep_c, ca = _mesh_args(cut)
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        os.chdir(d)
        try:
            st = L.ref_dispatch_debug(flags, *sa, *ca)
        finally:
            os.chdir(cwd)
    lines = L.ref_kernel_message().decode().splitlines()
    if st == 0:
        return st, ""
    for line in lines:
        head = line.split(" : ")[0]
        if head in KERNEL_STATUS:
            return st, head
    if "general position assumption violated!" in lines:
        return st, "GENERAL_POSITION_VIOLATION"
    return st, ""


def seam_points(ccs):
    """Set of intersection-point coordinates (as byte-exact tuples) over all CCs."""
    pts = set()
    for cc in ccs:
        if cc["type"] == CC_INPUT:
            continue
        for s in cc["seam"]:
            pts.add(tuple(cc["vertices"][s].tolist()))
    return pts


def broad_phase(src, cut, helpers=0, gp_constant=1e-4):
    L = lib()
    keep_s, sa = _mesh_args(src)
    keep_c, ca = _mesh_args(cut)
    tm = np.zeros(3, np.float64)
    n = L.ref_broad_phase(helpers, gp_constant, *sa, *ca, _ptr(tm))
    if n < 0:
        raise RuntimeError("reference broad phase failed")
    com = np.zeros(3)
    t = np.zeros(3)
    eps = np.zeros(1)
    pairs = np.zeros(2 * n, np.uint32)
    sa_ = np.zeros((sa[4], 6))
    ca_ = np.zeros((ca[4], 6))
    L.ref_broad_copy(_ptr(com), _ptr(t), _ptr(eps), _ptr(pairs), _ptr(sa_), _ptr(ca_))
    return dict(com=com, t=t, eps=float(eps[0]), pairs=pairs.reshape(-1, 2), src_face_aabb=sa_,
                cut_face_aabb=ca_, ms=tm)


def plane(verts):
    v = np.ascontiguousarray(verts, np.float64)
    n = np.zeros(3)
    d = np.zeros(1)
    k = lib().ref_plane(_ptr(v), v.shape[0], _ptr(n), _ptr(d))
    return n, float(d[0]), k


def segment_type(q, r, verts, normal, k):
    q, r, v, n = (np.ascontiguousarray(a, np.float64) for a in (q, r, verts, normal))
    return chr(lib().ref_segment_type(_ptr(q), _ptr(r), _ptr(v), v.shape[0], _ptr(n), k))


def segment_point(normal, d, q, r):
    n, q, r = (np.ascontiguousarray(a, np.float64) for a in (normal, q, r))
    p = np.zeros(3)
    c = lib().ref_segment_point(_ptr(n), d, _ptr(q), _ptr(r), _ptr(p))
    return chr(c), p


def point_in_polygon(p, verts, normal, k):
    p, v, n = (np.ascontiguousarray(a, np.float64) for a in (p, verts, normal))
    return chr(lib().ref_point_in_polygon(_ptr(p), _ptr(v), v.shape[0], _ptr(n), k))


def orient3d(a, b, c, d):
    a, b, c, d = (np.ascontiguousarray(x, np.float64) for x in (a, b, c, d))
    return lib().ref_orient3d(_ptr(a), _ptr(b), _ptr(c), _ptr(d))


def orient2d(a, b, c):
    a, b, c = (np.ascontiguousarray(x, np.float64) for x in (a, b, c))
    return lib().ref_orient2d(_ptr(a), _ptr(b), _ptr(c))


def best_triple(verts, normal, k):
    v, n = (np.ascontiguousarray(a, np.float64) for a in (verts, normal))
    ijk = np.zeros(3, np.int32)
    ok = lib().ref_best_triple(_ptr(v), v.shape[0], _ptr(n), k, _ptr(ijk))
    return ok, tuple(ijk.tolist())


def ps_edges(src, cut):
    """Polygon-soup edge table built with the reference's hmesh_t: (v0, v1, f0, f1) per edge."""
    L = lib()
    keep_s, sa = _mesh_args(src)
    keep_c, ca = _mesh_args(cut)
    n = L.ref_ps_edges(*sa, *ca)
    out = np.zeros((n, 4), np.int32)
    L.ref_ps_edges_copy(_ptr(out))
    return out

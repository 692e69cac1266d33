"""ctypes wrapper over the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
benchmark's CPU-baseline leg as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

SUCCESS, INVALID_SRC_MESH, INVALID_CUT_MESH, INVALID_MESH_INTERSECTION, GENERAL_POSITION_VIOLATION = 0, -1, -2, -3, -4
UNSUPPORTED = -100


class _Mesh(C.Structure):
    _fields_ = [("v", C.c_void_p), ("f", C.c_void_p), ("sz", C.c_void_p), ("nv", C.c_uint32), ("nf", C.c_uint32)]


class _Opt(C.Structure):
    _fields_ = [("gp_constant", C.c_double), ("eps_absolute", C.c_int), ("enforce_gp", C.c_int),
                ("enforcement_count", C.c_int), ("require_through_cuts", C.c_int), ("has_perturbation", C.c_int),
                ("perturbation", C.c_double * 3)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "oracle"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.om_frontend.argtypes = [C.POINTER(_Mesh), C.POINTER(_Mesh), C.POINTER(_Opt)]
        L.om_frontend.restype = C.c_void_p
        L.om_free.argtypes = [C.c_void_p]
        for fn in ("om_status", "om_stage", "om_partial"):
            getattr(L, fn).argtypes = [C.c_void_p]
            getattr(L, fn).restype = C.c_int
        L.om_time_ms.argtypes = [C.c_void_p]
        L.om_time_ms.restype = C.c_double
        L.om_shift.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.om_count.argtypes = [C.c_void_p, C.c_int]
        L.om_count.restype = C.c_int64
        L.om_copy_pairs.argtypes = [C.c_void_p, C.c_void_p]
        L.om_copy_tests.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.om_copy_registry.argtypes = [C.c_void_p] * 4
        L.om_copy_edges.argtypes = [C.c_void_p] * 5
        L.om_copy_planes.argtypes = [C.c_void_p] * 6
        L.om_copy_ps_vertices.argtypes = [C.c_void_p] * 2
        L.om_orient2d.argtypes = [C.c_void_p] * 3
        L.om_orient2d.restype = C.c_double
        L.om_orient3d_sign.argtypes = [C.c_void_p] * 4
        L.om_orient3d_sign.restype = C.c_int
        L.om_plane.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.om_plane.restype = C.c_int
        L.om_point_in_polygon.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        L.om_point_in_polygon.restype = C.c_char
        L.om_segment_point.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        L.om_best_triple.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.om_cutpath.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_int64, C.POINTER(C.c_int)]
        L.om_cutpath.restype = C.c_int64
        L.om_best_triple.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _mesh(m, keep):
    V, F, S = m
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.uint32)
    S = None if S is None else np.ascontiguousarray(S, dtype=np.uint32)
    keep += [V, F, S]
    nf = F.size // 3 if S is None else S.size
    return _Mesh(_ptr(V), _ptr(F), _ptr(S), V.shape[0], nf)


class Result(dict):
    """Front-end outputs of the oracle as numpy arrays."""


def frontend(src, cut, gp_constant=1e-4, enforce_gp=False, eps_absolute=False, enforcement_count=0,
             require_through_cuts=False, perturbation=None, detail=True, cutpath=None):
    """The oracle's front end.  cutpath: None (skip), True (the cut-path
    edges of the registry in its own order) or an order array (order[k] =
    registry index of point k): R["cutpath"] holds them (om_cutpath)."""
    L = lib()
    keep = []
    ms, mc = _mesh(src, keep), _mesh(cut, keep)
    o = _Opt(gp_constant, int(eps_absolute), int(enforce_gp), enforcement_count, int(require_through_cuts),
             0 if perturbation is None else 1, (C.c_double * 3)(*(perturbation if perturbation is not None else (0, 0, 0))))
    h = L.om_frontend(C.byref(ms), C.byref(mc), C.byref(o))
    try:
        R = Result()
        R["status"] = L.om_status(h)
        R["stage"] = L.om_stage(h)
        R["partial"] = L.om_partial(h)
        R["ms"] = L.om_time_ms(h)
        com, t, eps = np.zeros(3), np.zeros(3), np.zeros(1)
        L.om_shift(h, _ptr(com), _ptr(t), _ptr(eps))
        R.update(com=com, t=t, eps=float(eps[0]))
        n = {k: L.om_count(h, i) for i, k in enumerate(["pairs", "tests", "registry", "edges", "cand"])}
        R["counts"] = n
        pairs = np.zeros((max(n["pairs"], 0), 2), np.uint32)
        L.om_copy_pairs(h, _ptr(pairs))
        R["pairs"] = pairs
        re = np.zeros(n["registry"], np.uint32)
        rf = np.zeros(n["registry"], np.uint32)
        rp = np.zeros((n["registry"], 3), np.float64)
        L.om_copy_registry(h, _ptr(re), _ptr(rf), _ptr(rp))
        R.update(reg_edge=re, reg_face=rf, reg_point=rp)
        if detail:
            T = n["tests"]
            te, tf = np.zeros(T, np.uint32), np.zeros(T, np.uint32)
            qs, rs = np.zeros(T, np.int8), np.zeros(T, np.int8)
            ty, pp = np.zeros(T, np.uint8), np.zeros(T, np.uint8)
            if T:
                L.om_copy_tests(h, _ptr(te), _ptr(tf), _ptr(qs), _ptr(rs), _ptr(ty), _ptr(pp))
            R.update(test_edge=te, test_face=tf, qsign=qs, rsign=rs, type=ty, pip=pp)
            E = n["edges"]
            v0, v1 = np.zeros(E, np.uint32), np.zeros(E, np.uint32)
            f0, f1 = np.zeros(E, np.int32), np.zeros(E, np.int32)
            if E:
                L.om_copy_edges(h, _ptr(v0), _ptr(v1), _ptr(f0), _ptr(f1))
            R.update(edge_v0=v0, edge_v1=v1, edge_f0=f0, edge_f1=f1)
            NC = n["cand"]
            cf = np.zeros(NC, np.uint32)
            cn = np.zeros((NC, 3))
            cd = np.zeros(NC)
            ck = np.zeros(NC, np.int32)
            ct = np.zeros((NC, 3), np.int32)
            if NC:
                L.om_copy_planes(h, _ptr(cf), _ptr(cn), _ptr(cd), _ptr(ck), _ptr(ct))
            R.update(cand_face=cf, plane_n=cn, plane_d=cd, plane_k=ck, plane_triple=ct)
            nv = src[0].shape[0] + cut[0].shape[0]
            psv = np.zeros((nv, 3))
            L.om_copy_ps_vertices(h, _ptr(psv))
            R["ps_vertices"] = psv
        if cutpath is not None:
            nr = n["registry"]
            order = None
            if cutpath is not True:
                order = np.ascontiguousarray(cutpath, np.uint32)
                assert order.size == nr
            cap = 2 * nr + 1
            cs, ct, cg = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
            cf, ck = np.zeros((cap, 2), np.uint32), np.zeros(cap, np.uint8)
            sg = C.c_int(0)
            ne = L.om_cutpath(h, None if order is None else _ptr(order), _ptr(cs), _ptr(ct), _ptr(cf), _ptr(cg),
                              _ptr(ck), cap, C.byref(sg))
            R["cutpath"] = dict(src=cs[:ne], tgt=ct[:ne], faces=cf[:ne], groups=cg[:ne], kinds=ck[:ne],
                                short_group=sg.value)
        return R
    finally:
        L.om_free(h)


def orient2d(a, b, c):
    a, b, c = (np.ascontiguousarray(x, np.float64) for x in (a, b, c))
    return lib().om_orient2d(_ptr(a), _ptr(b), _ptr(c))


def orient3d_sign(a, b, c, d):
    a, b, c, d = (np.ascontiguousarray(x, np.float64) for x in (a, b, c, d))
    return lib().om_orient3d_sign(_ptr(a), _ptr(b), _ptr(c), _ptr(d))


def plane(verts):
    v = np.ascontiguousarray(verts, np.float64)
    n = np.zeros(3)
    d = np.zeros(1)
    k = lib().om_plane(_ptr(v), v.shape[0], _ptr(n), _ptr(d))
    return n, float(d[0]), k


def point_in_polygon(p, verts, normal, k):
    p, v, n = (np.ascontiguousarray(a, np.float64) for a in (p, verts, normal))
    return lib().om_point_in_polygon(_ptr(p), _ptr(v), v.shape[0], _ptr(n), k).decode()


def segment_point(normal, d, q, r):
    n, q, r = (np.ascontiguousarray(a, np.float64) for a in (normal, q, r))
    p = np.zeros(3)
    lib().om_segment_point(_ptr(n), d, _ptr(q), _ptr(r), _ptr(p))
    return p

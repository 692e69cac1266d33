"""Python binding of libmcgpu (include/mcgpu.h) through its C ABI.

``Context.dispatch(src, cut, flags)`` runs MCUT's intersection front end on
the GPU for one source/cut mesh pair given as ``(V, F, S)`` arrays
(``S`` = face sizes or ``None`` for triangles), the same inputs mcDispatch
takes (include/mcut/mcut.h:1008-1020).  Results stay on the device; the
``Result`` object copies them out on request.

The library must be built (``make -C mcut_amd/csrc``); loading fails loudly
when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmcgpu.so")

# MC_DISPATCH_* bits the front end reads (mcut.h:379-441)
REQUIRE_THROUGH_CUTS = 1 << 2
ENFORCE_GENERAL_POSITION = 1 << 15
ENFORCE_GENERAL_POSITION_ABSOLUTE = 1 << 16
CHECK_COMPONENTS = 1 << 30  # libmcgpu's own: refuse meshes of more than one connected component

STATUS = {0: "SUCCESS", -1: "INVALID_SRC_MESH", -2: "INVALID_CUT_MESH", -3: "INVALID_MESH_INTERSECTION",
          -4: "GENERAL_POSITION_VIOLATION"}
RESULT = {0: "success", -1: "invalid value", -2: "HIP runtime error", -3: "out of device memory",
          -4: "unsupported input", -5: "no dispatch result"}


class McgpuError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {RESULT.get(code, code)} ({code})")
        self.code = code


class Mesh(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("face_indices", C.c_void_p), ("face_sizes", C.c_void_p),
                ("num_vertices", C.c_uint32), ("num_faces", C.c_uint32)]


class DispatchDesc(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("gp_constant", C.c_double), ("enforcement_count", C.c_uint32),
                ("input_on_device", C.c_uint32), ("keep_tests", C.c_uint32), ("reserved", C.c_uint32)]


class FrontendInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("gp_violation", C.c_uint32), ("num_candidate_pairs", C.c_uint64),
                ("num_candidate_faces", C.c_uint64), ("num_edges", C.c_uint64), ("num_tests", C.c_uint64),
                ("num_escalated", C.c_uint64), ("num_registry", C.c_uint64), ("num_raw_pairs", C.c_uint64),
                ("num_decided_f32", C.c_uint64), ("num_decided_f64", C.c_uint64),
                ("com", C.c_double * 3),
                ("translation", C.c_double * 3), ("eps", C.c_double), ("device_ms", C.c_float),
                ("partial_cut", C.c_uint32)]


class SoupTopology(C.Structure):
    _fields_ = [("corner_edges", C.c_void_p * 2), ("edge_v0", C.c_void_p), ("edge_v1", C.c_void_p),
                ("edge_f0", C.c_void_p), ("edge_f1", C.c_void_p), ("num_edges", C.c_uint32), ("reserved", C.c_uint32)]


class Batch(C.Structure):
    _fields_ = [("num_pairs", C.c_uint32), ("src", Mesh), ("cut", Mesh), ("src_vertex_offsets", C.c_void_p),
                ("src_face_offsets", C.c_void_p), ("cut_vertex_offsets", C.c_void_p), ("cut_face_offsets", C.c_void_p)]


class ServerLimits(C.Structure):
    _fields_ = [("max_cut_vertices", C.c_uint32), ("max_cut_faces", C.c_uint32), ("max_cut_corners", C.c_uint32),
                ("max_slots", C.c_uint32), ("hbm_budget", C.c_uint64)]


class RequestResult(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("start_s", C.c_double), ("done_s", C.c_double), ("status", C.c_int32),
                ("partial_cut", C.c_uint32), ("num_registry", C.c_uint64), ("num_candidate_pairs", C.c_uint64)]


_lib = None


class CutpathInfo(C.Structure):
    """mcgpu_cutpath_info"""
    _fields_ = [("num_edges", C.c_uint64), ("num_groups", C.c_uint64), ("num_candidates", C.c_uint64),
                ("short_group", C.c_uint32), ("unsupported", C.c_uint32), ("device_ms", C.c_float),
                ("reserved", C.c_float)]


class Tuning(C.Structure):
    """mcgpu_tuning: the traversal's launch shape and the CU split."""
    _fields_ = [("wave_leaves", C.c_uint32), ("tree_stride", C.c_uint32), ("queues", C.c_uint32),
                ("chunk", C.c_uint32), ("broad_share_pct", C.c_uint32), ("candidates", C.c_uint32),
                ("traversal_ms", C.c_float), ("default_traversal_ms", C.c_float), ("dispatch_ms", C.c_float),
                ("default_dispatch_ms", C.c_float)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise McgpuError(-2, f"libmcgpu is not built ({LIB_PATH}); run make -C mcut_amd/csrc")
        L = C.CDLL(LIB_PATH)
        vp, u64 = C.c_void_p, C.c_uint64
        L.mcgpu_create_context.argtypes = [C.c_int, C.POINTER(vp)]
        L.mcgpu_release_context.argtypes = [vp]
        L.mcgpu_result_string.argtypes = [C.c_int]
        L.mcgpu_result_string.restype = C.c_char_p
        L.mcgpu_dispatch_frontend.argtypes = [vp, C.POINTER(Mesh), C.POINTER(Mesh), C.POINTER(DispatchDesc),
                                              C.POINTER(FrontendInfo)]
        L.mcgpu_redispatch_perturbed.argtypes = [vp, vp, C.c_uint32, C.POINTER(FrontendInfo)]
        L.mcgpu_dispatch_narrow.argtypes = [vp, C.POINTER(Mesh), C.POINTER(Mesh), vp, vp, vp, u64,
                                            C.POINTER(SoupTopology), C.POINTER(DispatchDesc), C.POINTER(FrontendInfo)]
        L.mcgpu_get_registry.argtypes = [vp, vp, vp, vp, u64]
        L.mcgpu_get_candidate_pairs.argtypes = [vp, vp, u64]
        L.mcgpu_get_edges.argtypes = [vp, vp, vp, vp, vp, u64]
        L.mcgpu_get_tests.argtypes = [vp, vp, vp, vp, vp, vp, vp, u64]
        L.mcgpu_get_face_planes.argtypes = [vp, vp, vp, vp, vp, vp, u64]
        L.mcgpu_get_ps_vertices.argtypes = [vp, vp, u64]
        L.mcgpu_registry_device.argtypes = [vp, vp, vp, vp, vp]
        L.mcgpu_broad_phase.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, C.c_uint32, C.POINTER(u64)]
        L.mcgpu_broad_phase_noalloc.argtypes = [C.c_int, vp, C.c_uint32, vp, C.c_uint32, vp, u64, C.POINTER(u64), vp]
        L.mcgpu_dispatch_batch.argtypes = [vp, C.POINTER(Batch), C.POINTER(DispatchDesc), C.POINTER(FrontendInfo)]
        L.mcgpu_get_batch_status.argtypes = [vp, vp, vp, u64]
        L.mcgpu_get_batch_candidate_pairs.argtypes = [vp, vp, u64]
        L.mcgpu_get_batch_partial_cuts.argtypes = [vp, vp, u64]
        L.mcgpu_server_create.argtypes = [C.c_int, C.POINTER(Mesh), C.POINTER(DispatchDesc), C.POINTER(ServerLimits),
                                          C.POINTER(vp)]
        L.mcgpu_server_release.argtypes = [vp]
        L.mcgpu_server_info.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(u64), C.POINTER(C.c_uint32)]
        L.mcgpu_server_run.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, vp]
        L.mcgpu_server_get_registry.argtypes = [vp, C.c_uint32, vp, vp, vp, u64]
        L.mcgpu_dispatch_enforced.argtypes = [vp, C.POINTER(Mesh), C.POINTER(Mesh), C.POINTER(DispatchDesc), C.c_uint32,
                                              C.POINTER(FrontendInfo), vp, C.POINTER(C.c_uint32)]
        L.mcgpu_predicate_stage.argtypes = [vp, vp, vp, vp, u64, C.POINTER(C.c_float)]
        L.mcgpu_stage_times.argtypes = [vp, vp, C.c_int]
        L.mcgpu_set_stage_timing.argtypes = [vp, C.c_int]
        L.mcgpu_stage_name.argtypes = [C.c_int]
        L.mcgpu_autotune.argtypes = [vp, C.POINTER(Tuning)]
        L.mcgpu_cutpath_edges.argtypes = [vp, vp, C.POINTER(CutpathInfo)]
        L.mcgpu_dispatch_frontend_async.argtypes = [vp, C.POINTER(Mesh), C.POINTER(Mesh), C.POINTER(DispatchDesc), vp]
        L.mcgpu_dispatch_query.argtypes = [vp]
        L.mcgpu_dispatch_wait.argtypes = [vp, C.POINTER(FrontendInfo), vp]
        L.mcgpu_get_cutpath_edges.argtypes = [vp, vp, vp, vp, vp, vp, u64]
        L.mcgpu_set_graphs.argtypes = [vp, C.c_int]
        L.mcgpu_set_batch_split.argtypes = [vp, C.c_uint32]
        L.mcgpu_graph_stats.argtypes = [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(C.c_uint32)]
        L.mcgpu_set_autotune.argtypes = [vp, C.c_int]
        L.mcgpu_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.mcgpu_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.mcgpu_stage_name.restype = C.c_char_p
        for fn in ("mcgpu_create_context", "mcgpu_release_context", "mcgpu_dispatch_frontend",
                   "mcgpu_redispatch_perturbed", "mcgpu_get_registry", "mcgpu_get_candidate_pairs", "mcgpu_get_edges",
                   "mcgpu_get_tests", "mcgpu_get_face_planes", "mcgpu_get_ps_vertices", "mcgpu_registry_device",
                   "mcgpu_stage_times", "mcgpu_broad_phase", "mcgpu_dispatch_batch", "mcgpu_get_batch_status",
                   "mcgpu_get_batch_candidate_pairs", "mcgpu_get_batch_partial_cuts", "mcgpu_server_create", "mcgpu_server_release",
                   "mcgpu_server_info", "mcgpu_server_run", "mcgpu_server_get_registry", "mcgpu_predicate_stage",
                   "mcgpu_dispatch_enforced", "mcgpu_set_stage_timing", "mcgpu_broad_phase_noalloc",
                   "mcgpu_dispatch_narrow"):
            getattr(L, fn).restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _check(code, what):
    if code != 0:
        raise McgpuError(code, what)


def _host_mesh(m, keep):
    V, F, S = m
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.uint32)
    S = None if S is None else np.ascontiguousarray(S, dtype=np.uint32)
    keep += [V, F, S]
    nf = F.size // 3 if S is None else S.size
    return Mesh(_ptr(V), _ptr(F), _ptr(S), V.shape[0], nf)


def device_mesh(V_ptr, F_ptr, S_ptr, nv, nf):
    """Mesh descriptor over device pointers (e.g. torch tensors' data_ptr())."""
    return Mesh(V_ptr, F_ptr, S_ptr, nv, nf)


class Result:
    """Front-end outputs of one dispatch; arrays are copied from the device lazily."""

    def __init__(self, ctx, info: FrontendInfo, nv: int):
        self._ctx = ctx
        self.info = info
        self.status = info.status
        self.status_name = STATUS.get(info.status, str(info.status))
        self.counts = dict(pairs=info.num_candidate_pairs, cand=info.num_candidate_faces, edges=info.num_edges,
                           tests=info.num_tests, escalated=info.num_escalated, registry=info.num_registry,
                           raw_pairs=info.num_raw_pairs, decided_f32=info.num_decided_f32,
                           decided_f64=info.num_decided_f64)
        self.com = np.array(info.com[:])
        self.t = np.array(info.translation[:])
        self.eps = info.eps
        self.device_ms = info.device_ms
        self.partial_cut = int(info.partial_cut)
        self._nv = nv

    def registry(self):
        n = self.info.num_registry
        e, g, p = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros((n, 3), np.float64)
        _check(lib().mcgpu_get_registry(self._ctx.h, _ptr(e), _ptr(g), _ptr(p), n), "mcgpu_get_registry")
        return e, g, p

    def pairs(self):
        n = self.info.num_candidate_pairs
        out = np.zeros((n, 2), np.uint32)
        _check(lib().mcgpu_get_candidate_pairs(self._ctx.h, _ptr(out), n), "mcgpu_get_candidate_pairs")
        return out

    def edges(self):
        n = self.info.num_edges
        v0, v1 = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        f0, f1 = np.zeros(n, np.int32), np.zeros(n, np.int32)
        _check(lib().mcgpu_get_edges(self._ctx.h, _ptr(v0), _ptr(v1), _ptr(f0), _ptr(f1), n), "mcgpu_get_edges")
        return v0, v1, f0, f1

    def tests(self):
        n = self.info.num_tests
        e, g = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        qs, rs = np.zeros(n, np.int8), np.zeros(n, np.int8)
        ty, pp = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        _check(lib().mcgpu_get_tests(self._ctx.h, _ptr(e), _ptr(g), _ptr(qs), _ptr(rs), _ptr(ty), _ptr(pp), n),
               "mcgpu_get_tests")
        return dict(edge=e, face=g, qsign=qs, rsign=rs, type=ty, pip=pp)

    def planes(self):
        n = self.info.num_candidate_faces
        f = np.zeros(n, np.uint32)
        nrm = np.zeros((n, 3))
        d = np.zeros(n)
        k = np.zeros(n, np.int32)
        tri = np.zeros((n, 3), np.int32)
        _check(lib().mcgpu_get_face_planes(self._ctx.h, _ptr(f), _ptr(nrm), _ptr(d), _ptr(k), _ptr(tri), n),
               "mcgpu_get_face_planes")
        return dict(face=f, normal=nrm, d=d, k=k, triple=tri)

    def ps_vertices(self):
        out = np.zeros((self._nv, 3))
        _check(lib().mcgpu_get_ps_vertices(self._ctx.h, _ptr(out), self._nv), "mcgpu_get_ps_vertices")
        return out

    def predicate_stage(self, signs=False):
        """Both orient3d signs of every test through the staged filter, alone:
        counts per stage and per sign, the kernel time, and (signs=True) the
        signs in tests() order."""
        cnt = np.zeros(6, np.uint64)
        n = self.info.num_tests
        qs = np.zeros(n, np.int8) if signs else None
        rs = np.zeros(n, np.int8) if signs else None
        ms = C.c_float(0)
        _check(lib().mcgpu_predicate_stage(self._ctx.h, _ptr(cnt), _ptr(qs), _ptr(rs), n, C.byref(ms)),
               "mcgpu_predicate_stage")
        out = dict(f32=int(cnt[0]), f64=int(cnt[1]), exact=int(cnt[2]), negative=int(cnt[3]), zero=int(cnt[4]),
                   positive=int(cnt[5]), device_ms=ms.value)
        if signs:
            out["qsign"], out["rsign"] = qs, rs
        return out

    def cutpath(self, order=None):
        """Cut-path edges of this dispatch (mcgpu_cutpath_edges): order = None
        (registry order) or order[k] = registry index of point k.  Returns the
        edges (src, tgt point ranks, faces (n, 2), groups, kinds) and the info."""
        info = CutpathInfo()
        o = None if order is None else np.ascontiguousarray(order, np.uint32)
        _check(lib().mcgpu_cutpath_edges(self._ctx.h, None if o is None else _ptr(o), C.byref(info)),
               "mcgpu_cutpath_edges")
        n = info.num_edges
        src, tgt, grp = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        faces, kinds = np.zeros((n, 2), np.uint32), np.zeros(n, np.uint8)
        _check(lib().mcgpu_get_cutpath_edges(self._ctx.h, _ptr(src), _ptr(tgt), _ptr(faces), _ptr(grp), _ptr(kinds), n),
               "mcgpu_get_cutpath_edges")
        return dict(src=src, tgt=tgt, faces=faces, groups=grp, kinds=kinds, short_group=info.short_group,
                    unsupported=info.unsupported, num_groups=info.num_groups, num_candidates=info.num_candidates,
                    device_ms=info.device_ms)

    def stage_ms(self):
        """Device time of each pipeline stage of this dispatch, in ms (HIP events);
        empty unless the context had stage timing on (Context.set_stage_timing)."""
        ms = np.zeros(16, np.float32)
        n = lib().mcgpu_stage_times(self._ctx.h, _ptr(ms), 16)
        return {name: float(ms[i]) for i, name in enumerate(stage_names()[:n])}


_stage_names = None


# limits of mcgpu_broad_phase_noalloc (include/mcgpu.h)
NOALLOC_MAX_FACES = 65536
NOALLOC_MAX_PAIRS = 1 << 20


def broad_phase_noalloc(device, src_boxes_ptr, num_src, cut_boxes_ptr, num_cut, pairs_ptr, capacity, stream=None):
    """mcgpu_broad_phase_noalloc over device buffers: face boxes (6 doubles a
    face) in, packed (source face << 32 | cut face) pairs out, unsorted; no
    device allocation.  Returns (result code, pair count)."""
    n = C.c_uint64(0)
    r = lib().mcgpu_broad_phase_noalloc(device, src_boxes_ptr, num_src, cut_boxes_ptr, num_cut, pairs_ptr, capacity,
                                        C.byref(n), stream)
    return r, n.value


def stage_names():
    global _stage_names
    if _stage_names is None:
        names = []
        while True:
            nm = lib().mcgpu_stage_name(len(names)).decode()
            if not nm:
                break
            names.append(nm)
        _stage_names = names
    return _stage_names


def pack_batch(pairs):
    """Concatenates (src, cut) mesh pairs into the arrays of an mcgpu_batch:
    every source mesh in pair order, every cut mesh in pair order, face
    indices local to their mesh, and the vertex / face offsets per pair."""
    out = {}
    for side, k in (("src", 0), ("cut", 1)):
        meshes = [pr[k] for pr in pairs]
        V = np.ascontiguousarray(np.concatenate([np.asarray(m[0], np.float64).reshape(-1, 3) for m in meshes]))
        F = np.ascontiguousarray(np.concatenate([np.asarray(m[1], np.uint32).ravel() for m in meshes]))
        nf = [(np.asarray(m[1]).size // 3 if m[2] is None else len(m[2])) for m in meshes]
        S = None
        if any(m[2] is not None for m in meshes):
            S = np.ascontiguousarray(np.concatenate([
                np.full(n, 3, np.uint32) if m[2] is None else np.asarray(m[2], np.uint32) for m, n in zip(meshes, nf)]))
        out[side] = (V, F, S)
        out[side + "_vo"] = np.concatenate([[0], np.cumsum([np.asarray(m[0]).reshape(-1, 3).shape[0]
                                                            for m in meshes])]).astype(np.uint32)
        out[side + "_fo"] = np.concatenate([[0], np.cumsum(nf)]).astype(np.uint32)
    out["num_pairs"] = len(pairs)
    return out


class BatchResult:
    """Per-pair outputs of one batched dispatch."""

    def __init__(self, ctx, info: FrontendInfo, num_pairs: int):
        self._ctx = ctx
        self.info = info
        self.num_pairs = num_pairs
        self.counts = dict(pairs=info.num_candidate_pairs, cand=info.num_candidate_faces, edges=info.num_edges,
                           tests=info.num_tests, escalated=info.num_escalated, registry=info.num_registry,
                           raw_pairs=info.num_raw_pairs, decided_f32=info.num_decided_f32,
                           decided_f64=info.num_decided_f64)
        self.device_ms = info.device_ms
        self.status = np.zeros(num_pairs, np.int32)
        self.reg_offsets = np.zeros(num_pairs + 1, np.uint64)
        _check(lib().mcgpu_get_batch_status(ctx.h, _ptr(self.status), _ptr(self.reg_offsets), num_pairs + 1),
               "mcgpu_get_batch_status")
        self.partial_cut = np.zeros(num_pairs, np.uint32)
        _check(lib().mcgpu_get_batch_partial_cuts(ctx.h, _ptr(self.partial_cut), num_pairs),
               "mcgpu_get_batch_partial_cuts")
        self._reg = None

    def registry(self):
        """The whole batch registry (pair-local ids, pairs in order)."""
        if self._reg is None:
            n = self.info.num_registry
            e, g, p = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros((n, 3), np.float64)
            _check(lib().mcgpu_get_registry(self._ctx.h, _ptr(e), _ptr(g), _ptr(p), n), "mcgpu_get_registry")
            self._reg = (e, g, p)
        return self._reg

    def pair_registry(self, i):
        e, g, p = self.registry()
        a, b = int(self.reg_offsets[i]), int(self.reg_offsets[i + 1])
        return e[a:b], g[a:b], p[a:b]

    def candidate_pairs(self):
        """(pair, source face, cut face) triples, pair-local ids, sorted."""
        n = self.info.num_candidate_pairs
        out = np.zeros((n, 3), np.uint32)
        if n:
            _check(lib().mcgpu_get_batch_candidate_pairs(self._ctx.h, _ptr(out), n), "mcgpu_get_batch_candidate_pairs")
        return out

    def stage_ms(self):
        return Result.stage_ms(self)


class Context:
    """One libmcgpu context: a HIP stream and its device buffers on one GPU."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().mcgpu_create_context(device, C.byref(h)), "mcgpu_create_context")
        self.h = h

    def close(self):
        if self.h:
            lib().mcgpu_release_context(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self):
        """Tunes the traversal shape for the last dispatch's workload class
        (re-running it); returns the report as a dict."""
        t = Tuning()
        _check(lib().mcgpu_autotune(self.h, C.byref(t)), "mcgpu_autotune")
        return t.as_dict()

    def set_graphs(self, on):
        """Record each dispatch shape once as HIP graphs and replay them."""
        _check(lib().mcgpu_set_graphs(self.h, 1 if on else 0), "mcgpu_set_graphs")

    def graph_stats(self):
        r, p, n = C.c_uint64(0), C.c_uint64(0), C.c_uint32(0)
        _check(lib().mcgpu_graph_stats(self.h, C.byref(r), C.byref(p), C.byref(n)), "mcgpu_graph_stats")
        return dict(recordings=r.value, replays=p.value, graph_launches=n.value)

    def set_batch_split(self, broad_share_pct):
        """CU split of the batch streams; only before the context's first batch."""
        _check(lib().mcgpu_set_batch_split(self.h, broad_share_pct), "mcgpu_set_batch_split")

    def set_autotune(self, on):
        _check(lib().mcgpu_set_autotune(self.h, 1 if on else 0), "mcgpu_set_autotune")

    def set_tuning(self, wave_leaves=None, tree_stride=3, queues=0, chunk=1):
        """Forces a traversal shape (wave_leaves=None: back to tuned/default)."""
        if wave_leaves is None:
            _check(lib().mcgpu_set_tuning(self.h, None), "mcgpu_set_tuning")
            return
        t = Tuning(wave_leaves=wave_leaves, tree_stride=tree_stride, queues=queues, chunk=chunk)
        _check(lib().mcgpu_set_tuning(self.h, C.byref(t)), "mcgpu_set_tuning")

    def tuning(self):
        """The shape and split the last dispatch ran with."""
        t = Tuning()
        _check(lib().mcgpu_get_tuning(self.h, C.byref(t)), "mcgpu_get_tuning")
        return t.as_dict()

    def set_stage_timing(self, on):
        """Record per-stage events in later dispatches (Result.stage_ms); off by
        default.  on = "kernel": only the traversal kernel's timestamps."""
        mode = 2 if on == "kernel" else (1 if on else 0)
        _check(lib().mcgpu_set_stage_timing(self.h, mode), "mcgpu_set_stage_timing")

    def dispatch(self, src, cut, flags=0, gp_constant=1e-4, enforcement_count=0, keep_tests=False):
        keep = []
        ms, mc = _host_mesh(src, keep), _host_mesh(cut, keep)
        d = DispatchDesc(flags, gp_constant, enforcement_count, 0, 1 if keep_tests else 0, 0)
        info = FrontendInfo()
        _check(lib().mcgpu_dispatch_frontend(self.h, C.byref(ms), C.byref(mc), C.byref(d), C.byref(info)),
               "mcgpu_dispatch_frontend")
        return Result(self, info, ms.num_vertices + mc.num_vertices)

    def dispatch_enforced(self, src, cut, flags=1 << 15, gp_constant=1e-4, attempts=4):
        """mcDispatch's general-position loop on the GPU front end; returns
        (Result, perturbation vector of the last rerun, number of perturbations)."""
        keep = []
        ms, mc = _host_mesh(src, keep), _host_mesh(cut, keep)
        d = DispatchDesc(flags, gp_constant, 0, 0, 0, 0)
        info = FrontendInfo()
        p = np.zeros(3)
        k = C.c_uint32(0)
        _check(lib().mcgpu_dispatch_enforced(self.h, C.byref(ms), C.byref(mc), C.byref(d), attempts, C.byref(info),
                                             _ptr(p), C.byref(k)), "mcgpu_dispatch_enforced")
        return Result(self, info, ms.num_vertices + mc.num_vertices), p, k.value

    def dispatch_async(self, src, cut, flags=0, gp_constant=1e-4, stream=None):
        """Enqueues a dispatch of host meshes and returns at once; wait() ends it."""
        keep = []
        ms, mc = _host_mesh(src, keep), _host_mesh(cut, keep)
        d = DispatchDesc(flags, gp_constant, 0, 0, 0, 0)
        _check(lib().mcgpu_dispatch_frontend_async(self.h, C.byref(ms), C.byref(mc), C.byref(d), stream),
               "mcgpu_dispatch_frontend_async")
        self._pending_nv = ms.num_vertices + mc.num_vertices
        self._pending_keep = keep  # the host arrays stay alive until the wait

    def query(self):
        return bool(lib().mcgpu_dispatch_query(self.h))

    def wait(self, stream=None):
        info = FrontendInfo()
        try:
            _check(lib().mcgpu_dispatch_wait(self.h, C.byref(info), stream), "mcgpu_dispatch_wait")
        finally:
            self._pending_keep = None
        return Result(self, info, self._pending_nv)

    def dispatch_device(self, src: Mesh, cut: Mesh, flags=0, gp_constant=1e-4, enforcement_count=0, keep_tests=False):
        d = DispatchDesc(flags, gp_constant, enforcement_count, 1, 1 if keep_tests else 0, 0)
        info = FrontendInfo()
        _check(lib().mcgpu_dispatch_frontend(self.h, C.byref(src), C.byref(cut), C.byref(d), C.byref(info)),
               "mcgpu_dispatch_frontend")
        return Result(self, info, src.num_vertices + cut.num_vertices)

    def dispatch_narrow(self, src, cut, src_boxes, cut_boxes, pairs, flags=0, enforcement_count=0, topology=None):
        """dispatch()'s front-end block alone (kernel.cpp:1779-3229) on the
        inputs preproc hands it: meshes with the soup's own (shifted)
        coordinates, face AABBs (F x 6, cut faces enlarged) and candidate
        (source face, cut face) pairs; topology (optional) = (corner edges of
        the source mesh, of the cut mesh, edge v0, v1, f0, f1)."""
        keep = []
        ms, mc = _host_mesh(src, keep), _host_mesh(cut, keep)
        a = np.ascontiguousarray(src_boxes, dtype=np.float64).reshape(-1, 6)
        b = np.ascontiguousarray(cut_boxes, dtype=np.float64).reshape(-1, 6)
        p = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        t = None
        if topology is not None:
            arrs = [np.ascontiguousarray(x, dt) for x, dt in zip(topology, (np.uint32, np.uint32, np.uint32, np.uint32,
                                                                           np.int32, np.int32))]
            keep += arrs
            t = SoupTopology((C.c_void_p * 2)(_ptr(arrs[0]), _ptr(arrs[1])), _ptr(arrs[2]), _ptr(arrs[3]),
                             _ptr(arrs[4]), _ptr(arrs[5]), arrs[2].size, 0)
        d = DispatchDesc(flags, 1e-4, enforcement_count, 0, 0, 0)
        info = FrontendInfo()
        _check(lib().mcgpu_dispatch_narrow(self.h, C.byref(ms), C.byref(mc), _ptr(a), _ptr(b), _ptr(p), p.shape[0],
                                           None if t is None else C.byref(t), C.byref(d), C.byref(info)),
               "mcgpu_dispatch_narrow")
        return Result(self, info, ms.num_vertices + mc.num_vertices)

    def broad_phase(self, src_boxes, cut_boxes):
        """Candidate (source face, cut face) pairs of two sets of face AABBs
        (F x 6 float64: min xyz, max xyz), sorted; the seam of build_oibvh +
        intersectOIBVHs (bvh.cpp:219, 638)."""
        a = np.ascontiguousarray(src_boxes, dtype=np.float64).reshape(-1, 6)
        b = np.ascontiguousarray(cut_boxes, dtype=np.float64).reshape(-1, 6)
        n = C.c_uint64(0)
        _check(lib().mcgpu_broad_phase(self.h, _ptr(a), a.shape[0], _ptr(b), b.shape[0], 0, C.byref(n)),
               "mcgpu_broad_phase")
        out = np.zeros((n.value, 2), np.uint32)
        if n.value:
            _check(lib().mcgpu_get_candidate_pairs(self.h, _ptr(out), n.value), "mcgpu_get_candidate_pairs")
        return out

    def dispatch_batch(self, pairs=None, packed=None, flags=0, gp_constant=1e-4):
        """Batched front end over a list of (src, cut) pairs (or pack_batch's output)."""
        P = packed if packed is not None else pack_batch(pairs)
        keep = []
        ms, mc = _host_mesh(P["src"], keep), _host_mesh(P["cut"], keep)
        offs = [np.ascontiguousarray(P[k], np.uint32) for k in ("src_vo", "src_fo", "cut_vo", "cut_fo")]
        b = Batch(P["num_pairs"], ms, mc, *[_ptr(o) for o in offs])
        return self._dispatch_batch(b, DispatchDesc(flags, gp_constant, 0, 0, 0, 0))

    def dispatch_batch_device(self, num_pairs, src: Mesh, cut: Mesh, offsets, flags=0, gp_constant=1e-4):
        """Batched front end over device-resident meshes; offsets = host uint32
        arrays (src vertex, src face, cut vertex, cut face), num_pairs + 1 each."""
        offs = [np.ascontiguousarray(o, np.uint32) for o in offsets]
        b = Batch(num_pairs, src, cut, *[_ptr(o) for o in offs])
        return self._dispatch_batch(b, DispatchDesc(flags, gp_constant, 0, 1, 0, 0))

    def _dispatch_batch(self, b, d):
        info = FrontendInfo()
        _check(lib().mcgpu_dispatch_batch(self.h, C.byref(b), C.byref(d), C.byref(info)), "mcgpu_dispatch_batch")
        return BatchResult(self, info, b.num_pairs)

    def redispatch_perturbed(self, perturbation, enforcement_count, nv):
        p = np.ascontiguousarray(perturbation, dtype=np.float64)
        info = FrontendInfo()
        _check(lib().mcgpu_redispatch_perturbed(self.h, _ptr(p), enforcement_count, C.byref(info)),
               "mcgpu_redispatch_perturbed")
        return Result(self, info, nv)


class Server:
    """mcgpu_server: one resident source mesh, cut requests run in overlapping
    slots as they arrive (C ABI in include/mcgpu.h)."""

    def __init__(self, src, max_cut_vertices, max_cut_faces, slots=4, hbm_budget=8 << 30, device=0, flags=0,
                 gp_constant=1e-4, max_cut_corners=0):
        self._keep = []
        ms = _host_mesh(src, self._keep)
        d = DispatchDesc(flags, gp_constant, 0, 0, 0, 0)
        lim = ServerLimits(max_cut_vertices, max_cut_faces, max_cut_corners, slots, hbm_budget)
        h = C.c_void_p()
        _check(lib().mcgpu_server_create(device, C.byref(ms), C.byref(d), C.byref(lim), C.byref(h)),
               "mcgpu_server_create")
        self.h = h
        n, b, e = C.c_uint32(), C.c_uint64(), C.c_uint32()
        lib().mcgpu_server_info(self.h, C.byref(n), C.byref(b), C.byref(e))
        self.slots, self.slot_bytes, self.source_edges = n.value, b.value, e.value

    def run(self, cuts, arrival_s, keep_registries=False):
        """Runs the schedule; returns a structured array of per-request results."""
        keep = []
        meshes = (Mesh * len(cuts))(*[_host_mesh(c, keep) for c in cuts])
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        out = (RequestResult * len(cuts))()
        _check(lib().mcgpu_server_run(self.h, meshes, _ptr(arr), len(cuts), 1 if keep_registries else 0, out),
               "mcgpu_server_run")
        dt = np.dtype([("arrival_s", "f8"), ("start_s", "f8"), ("done_s", "f8"), ("status", "i4"),
                       ("partial_cut", "u4"), ("num_registry", "u8"), ("num_candidate_pairs", "u8")])
        return np.frombuffer(bytes(out), dtype=dt).copy()

    def registry(self, i, n):
        e, g, p = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros((n, 3), np.float64)
        _check(lib().mcgpu_server_get_registry(self.h, i, _ptr(e), _ptr(g), _ptr(p), n), "mcgpu_server_get_registry")
        return e, g, p

    def close(self):
        if self.h:
            lib().mcgpu_server_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""ctypes driver of MCUT's public C API (include/mcut/mcut.h), for any build of
the library: the seam build (mcut_amd/seam/_build/libmcut_seam.so, MCUT with
its broad phase on the GPU) or MCUT compiled unchanged."""
import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SEAM_LIB = os.path.join(HERE, "_build", "libmcut_seam.so")
GPU_LIB = os.path.join(HERE, "_build", "libmcut_gpu.so")      # front end and cut-path edges on libmcgpu
PROBE_LIB = os.path.join(HERE, "_build", "libmcut_probe.so")  # MCUT unchanged, cut-path edges recorded

# mcut.h:202-208, 367-441, 213-224, 292-306
MC_NO_ERROR = 0
MC_DISPATCH_VERTEX_ARRAY_DOUBLE = 1 << 1
MC_DISPATCH_REQUIRE_THROUGH_CUTS = 1 << 2
MC_DISPATCH_FILTER_ALL = sum(1 << b for b in range(5, 15))
MC_DISPATCH_ENFORCE_GENERAL_POSITION = 1 << 15
MC_DISPATCH_INCLUDE_INTERSECTION_TYPE = 1 << 17
MC_CONTEXT_DISPATCH_INTERSECTION_TYPE = 1 << 13
MC_CONNECTED_COMPONENT_TYPE_ALL = 0xFFFFFFFF
MC_CONNECTED_COMPONENT_TYPE_INPUT = 1 << 4
MC_CONNECTED_COMPONENT_DATA_VERTEX_DOUBLE = 1 << 2
MC_CONNECTED_COMPONENT_DATA_FACE = 1 << 5
MC_CONNECTED_COMPONENT_DATA_FACE_SIZE = 1 << 6
MC_CONNECTED_COMPONENT_DATA_TYPE = 1 << 9
MC_CONNECTED_COMPONENT_DATA_SEAM_VERTEX = 1 << 13


class McutError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed with McResult {code}")
        self.code = code


class Mcut:
    """One loaded MCUT library."""

    def __init__(self, path=SEAM_LIB):
        if not os.path.exists(path):
            raise McutError(-1, f"loading {path} (not built)")
        L = C.CDLL(path)
        vp, u32 = C.c_void_p, C.c_uint32
        L.mcCreateContextWithHelpers.argtypes = [C.POINTER(vp), u32, u32]
        L.mcCreateContext.argtypes = [C.POINTER(vp), u32]
        L.mcDispatch.argtypes = [vp, u32, vp, vp, vp, u32, u32, vp, vp, vp, u32, u32]
        L.mcGetConnectedComponents.argtypes = [vp, u32, u32, vp, C.POINTER(u32)]
        L.mcGetConnectedComponentData.argtypes = [vp, vp, u32, C.c_size_t, vp, C.POINTER(C.c_size_t)]
        L.mcReleaseConnectedComponents.argtypes = [vp, u32, vp]
        L.mcReleaseContext.argtypes = [vp]
        L.mcGetInfo.argtypes = [vp, u32, C.c_size_t, vp, C.POINTER(C.c_size_t)]
        for f in ("mcCreateContextWithHelpers", "mcCreateContext", "mcDispatch", "mcGetConnectedComponents",
                  "mcGetConnectedComponentData", "mcReleaseConnectedComponents", "mcReleaseContext", "mcGetInfo"):
            getattr(L, f).restype = C.c_int
        self.L = L
        # McDispatchIntersectionType of the last dispatch that asked for it
        # (MC_DISPATCH_INCLUDE_INTERSECTION_TYPE), else None
        self.intersection_type = None

    def cutpath_record(self):
        """The cut-path edges of this library's last dispatch that made them
        (builds with the cut-path recorder: libmcut_gpu.so, libmcut_probe.so):
        the registry in the dispatch's point order (edge, face), the edges as
        (source, target) point ranks in creation order, and every face's edge
        list as (face, edge) rows, faces ascending."""
        L = self.L
        if not hasattr(L, "mcut_probe_cutpath_sizes"):
            raise McutError(-1, "this build records no cut-path edges")
        u64 = C.c_uint64
        L.mcut_probe_cutpath_sizes.argtypes = [C.POINTER(u64)] * 3
        L.mcut_probe_cutpath_sizes.restype = u64
        L.mcut_probe_cutpath_copy.argtypes = [C.c_void_p] * 5
        npt, ne, nfe = u64(0), u64(0), u64(0)
        count = L.mcut_probe_cutpath_sizes(C.byref(npt), C.byref(ne), C.byref(nfe))
        re, rf = np.zeros(npt.value, np.uint32), np.zeros(npt.value, np.uint32)
        src, tgt = np.zeros(ne.value, np.uint32), np.zeros(ne.value, np.uint32)
        fe = np.zeros((nfe.value, 2), np.uint32)
        L.mcut_probe_cutpath_copy(re.ctypes.data, rf.ctypes.data, src.ctypes.data, tgt.ctypes.data, fe.ctypes.data)
        return dict(dispatches=count, reg_edge=re, reg_face=rf, src=src, tgt=tgt, face_edges=fe)

    def _ok(self, code, what):
        if code != MC_NO_ERROR:
            raise McutError(code, what)

    def time_dispatch(self, src, cut, flags=MC_DISPATCH_FILTER_ALL, helpers=0, repeats=3):
        """Wall-clock ms of mcDispatch alone, `repeats` times in one context
        (connected components released between calls, outside the timing).
        The library's banner and any debug text go to stderr, not stdout."""
        import time
        L = self.L
        out_fd = os.dup(1)
        os.dup2(2, 1)
        cwd = os.getcwd()
        ctx = C.c_void_p()
        times = []
        try:
            if helpers:
                self._ok(L.mcCreateContextWithHelpers(C.byref(ctx), 0, helpers), "mcCreateContextWithHelpers")
            else:
                self._ok(L.mcCreateContext(C.byref(ctx), 0), "mcCreateContext")
            keep = []

            def arrs(m):
                V, F, S = m
                v = np.ascontiguousarray(V, np.float64)
                f = np.ascontiguousarray(F, np.uint32)
                s = None if S is None else np.ascontiguousarray(S, np.uint32)
                keep.extend([v, f, s])
                nf = f.size // 3 if s is None else s.size
                return v.ctypes.data, f.ctypes.data, (None if s is None else s.ctypes.data), v.shape[0], nf
            sv, sf, ss, snv, snf = arrs(src)
            cv, cf, cs, cnv, cnf = arrs(cut)
            with tempfile.TemporaryDirectory() as tmp:
                os.chdir(tmp)
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    self._ok(L.mcDispatch(ctx, flags | MC_DISPATCH_VERTEX_ARRAY_DOUBLE, sv, sf, ss, snv, snf, cv, cf,
                                          cs, cnv, cnf), "mcDispatch")
                    times.append((time.perf_counter() - t0) * 1e3)
                    self._ok(L.mcReleaseConnectedComponents(ctx, 0, None), "mcReleaseConnectedComponents")
                os.chdir(cwd)
        finally:
            os.chdir(cwd)
            if ctx.value:
                L.mcReleaseContext(ctx)
            C.CDLL(None).fflush(None)  # C-level buffered text leaves while fd 1 is still stderr
            os.dup2(out_fd, 1)
            os.close(out_fd)
        return times

    def dispatch(self, src, cut, flags=MC_DISPATCH_FILTER_ALL, helpers=0):
        """mcDispatch on (V, F, S) meshes (S None = triangles) in a fresh context;
        returns the connected components as dicts of numpy arrays."""
        L = self.L
        ctx = C.c_void_p()
        if helpers:
            self._ok(L.mcCreateContextWithHelpers(C.byref(ctx), 0, helpers), "mcCreateContextWithHelpers")
        else:
            self._ok(L.mcCreateContext(C.byref(ctx), 0), "mcCreateContext")
        cwd = os.getcwd()
        try:
            keep = []

            def arrs(m):
                V, F, S = m
                v = np.ascontiguousarray(V, np.float64)
                f = np.ascontiguousarray(F, np.uint32)
                s = None if S is None else np.ascontiguousarray(S, np.uint32)
                keep.extend([v, f, s])
                nf = f.size // 3 if s is None else s.size
                return v.ctypes.data, f.ctypes.data, (None if s is None else s.ctypes.data), v.shape[0], nf
            sv, sf, ss, snv, snf = arrs(src)
            cv, cf, cs, cnv, cnf = arrs(cut)
            with tempfile.TemporaryDirectory() as tmp:
                os.chdir(tmp)  # MCUT may write debug files into the working directory
                self._ok(L.mcDispatch(ctx, flags | MC_DISPATCH_VERTEX_ARRAY_DOUBLE, sv, sf, ss, snv, snf, cv, cf, cs,
                                      cnv, cnf), "mcDispatch")
                os.chdir(cwd)
            self.intersection_type = None
            if flags & MC_DISPATCH_INCLUDE_INTERSECTION_TYPE:
                t = C.c_uint32(0)
                self._ok(L.mcGetInfo(ctx, MC_CONTEXT_DISPATCH_INTERSECTION_TYPE, C.sizeof(t), C.byref(t), None),
                         "mcGetInfo")
                self.intersection_type = t.value
            n = C.c_uint32(0)
            self._ok(L.mcGetConnectedComponents(ctx, MC_CONNECTED_COMPONENT_TYPE_ALL, 0, None, C.byref(n)),
                     "mcGetConnectedComponents")
            ccs = (C.c_void_p * max(n.value, 1))()
            # the count pointer is for the size query only (NULL with an array)
            self._ok(L.mcGetConnectedComponents(ctx, MC_CONNECTED_COMPONENT_TYPE_ALL, n.value, ccs, None),
                     "mcGetConnectedComponents")
            out = []
            for i in range(n.value):
                cc = ccs[i]

                def query(flag, dtype):
                    nb = C.c_size_t(0)
                    self._ok(L.mcGetConnectedComponentData(ctx, cc, flag, 0, None, C.byref(nb)),
                             "mcGetConnectedComponentData")
                    a = np.zeros(nb.value // np.dtype(dtype).itemsize, dtype)
                    if nb.value:
                        self._ok(L.mcGetConnectedComponentData(ctx, cc, flag, nb.value, a.ctypes.data, None),
                                 "mcGetConnectedComponentData")
                    return a
                typ = int(query(MC_CONNECTED_COMPONENT_DATA_TYPE, np.int32)[0])
                out.append(dict(type=typ,
                                vertices=query(MC_CONNECTED_COMPONENT_DATA_VERTEX_DOUBLE, np.float64).reshape(-1, 3),
                                faces=query(MC_CONNECTED_COMPONENT_DATA_FACE, np.uint32),
                                face_sizes=query(MC_CONNECTED_COMPONENT_DATA_FACE_SIZE, np.uint32),
                                # input components have no seam (frontend.cpp refuses the query)
                                seam=(np.zeros(0, np.uint32) if typ == MC_CONNECTED_COMPONENT_TYPE_INPUT
                                      else query(MC_CONNECTED_COMPONENT_DATA_SEAM_VERTEX, np.uint32))))
            L.mcReleaseConnectedComponents(ctx, 0, None)
            return out
        finally:
            os.chdir(cwd)
            L.mcReleaseContext(ctx)


def write_components(ccs, directory, fmt="off", prefix="cc"):
    """Writes connected components (Mcut.dispatch's output: the arrays of
    mcGetConnectedComponentData's VERTEX_DOUBLE / FACE / FACE_SIZE queries,
    frontend.cpp:2006-2400) as OFF or OBJ files, one per component; returns
    the paths."""
    from mcut_amd import meshgen as mg
    write = {"off": mg.write_off, "obj": mg.write_obj}[fmt]
    paths = []
    for i, cc in enumerate(ccs):
        path = os.path.join(directory, f"{prefix}{i}.{fmt}")
        sizes = cc["face_sizes"]
        write(path, cc["vertices"], cc["faces"], None if np.all(sizes == 3) else sizes)
        paths.append(path)
    return paths


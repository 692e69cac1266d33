"""Times the reference MCUT's own front end, from its own stage timers.

TEST AND MEASUREMENT INFRASTRUCTURE ONLY (the benchmark's CPU-baseline leg).
oracle/_ref/libmcut_ref_prof.so is MCUT compiled from its sources with its
compile-time switches PROFILING_BUILD and MCUT_WITH_API_EVENT_LOGGING
(oracle/Makefile ref_prof): every TIMESTACK_PUSH/POP pair and SCOPED_TIMER
then writes `[MCUT][PROF:<thread>]: "<name>" (<ms>ms)` to stderr when it
ends (include/mcut/internal/timer.h:50-99).  A dispatch is run through MCUT's
public API (mcCreateContext[WithHelpers] + mcDispatch) with stderr captured,
and the front-end time is the sum of the timers of the stages the GPU front
end replaces, between preproc's entry (preproc.cpp:2304) and the end of the
narrow phase (kernel.cpp:3229):

  client_input_arrays_to_hmesh  preproc.cpp:57-468 (shifted hmesh, both meshes)
  build_oibvh                   bvh.cpp:219-636 (both meshes)
  intersectOIBVHs               bvh.cpp:638-783
  Create ps / Create m0         kernel.cpp:1593-1772
  Prepare edge-to-face pairs    kernel.cpp:1779-1983
  Build edge bounding boxes     kernel.cpp:1989-2069
  Cull redundant edge-face pairs kernel.cpp:2074-2177
  Compute intersecting face properties kernel.cpp:2184-2356
  Calculate intersection points (edge-to-face) kernel.cpp:2415-3231

The timers leave out what lies between them: the coordinate sums
(calculate_vertex_parameters, two sequential passes), the single-component
checks (check_input_mesh) and MCUT's debug mesh dumps, which its build writes
on every dispatch (kernel_input.verbose is hard-wired on, preproc.cpp:2371)
before "Create ps" and between "Create ps" and "Create m0".  Each timer
truncates to whole milliseconds.  The sum is therefore a lower bound of the
reference's front-end time.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import statistics
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PROF_LIB = os.path.join(_HERE, "_ref", "libmcut_ref_prof.so")

FRONT_END_STAGES = ("client_input_arrays_to_hmesh", "build_oibvh", "intersectOIBVHs", "Create ps", "Create m0",
                    "Prepare edge-to-face pairs", "Build edge bounding boxes", "Cull redundant edge-face pairs",
                    "Compute intersecting face properties", "Calculate intersection points (edge-to-face)")
_LINE = re.compile(r'\[MCUT\]\[PROF:[^\]]*\]: "([^"]+)" \((\d+)ms\)')

MC_DISPATCH_VERTEX_ARRAY_DOUBLE = 1 << 1
MC_DISPATCH_FILTER_ALL = sum(1 << b for b in range(5, 15))


def available() -> bool:
    return os.path.exists(PROF_LIB)


def parse_stage_ms(text):
    """{stage name: summed ms} of the timer lines in `text`."""
    out = {}
    for name, ms in _LINE.findall(text):
        out[name] = out.get(name, 0) + int(ms)
    return out


def usable_cores():
    """CPUs this process may run on: its affinity set, capped by a cgroup
    CPU quota when one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, -(-int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class _Lib:
    def __init__(self):
        L = C.CDLL(PROF_LIB)
        vp, u32 = C.c_void_p, C.c_uint32
        L.mcCreateContextWithHelpers.argtypes = [C.POINTER(vp), u32, u32]
        L.mcCreateContext.argtypes = [C.POINTER(vp), u32]
        L.mcDispatch.argtypes = [vp, u32, vp, vp, vp, u32, u32, vp, vp, vp, u32, u32]
        L.mcReleaseConnectedComponents.argtypes = [vp, u32, vp]
        L.mcReleaseContext.argtypes = [vp]
        for f in ("mcCreateContextWithHelpers", "mcCreateContext", "mcDispatch", "mcReleaseConnectedComponents",
                  "mcReleaseContext"):
            getattr(L, f).restype = C.c_int
        self.L = L


def time_frontend(src, cut, helpers=0, runs=5, warmup=1, flags=MC_DISPATCH_FILTER_ALL, workdir=None):
    """The reference's front end on (src, cut) in one context with `helpers`
    helper threads: `warmup` untimed dispatches, then `runs` timed ones.
    Returns a dict with the per-run front-end ms (sum of the stage timers),
    median / min / max, the per-stage ms of the median run and the dispatch's
    McResult (MCUT may fail after its front end, in its cutting stage)."""
    lib = _Lib().L
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
    ctx = C.c_void_p()
    r = (lib.mcCreateContextWithHelpers(C.byref(ctx), 0, helpers) if helpers
         else lib.mcCreateContext(C.byref(ctx), 0))
    if r != 0:
        raise RuntimeError(f"mcCreateContext failed ({r})")
    totals, stages, results = [], [], []
    cwd = os.getcwd()
    saved = [os.dup(1), os.dup(2)]
    libc = C.CDLL(None)
    try:
        with tempfile.TemporaryDirectory(dir=workdir) as tmp:
            os.chdir(tmp)  # MCUT writes its debug meshes into the working directory
            for it in range(warmup + runs):
                log = os.path.join(tmp, "prof.txt")
                fd = os.open(log, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
                libc.fflush(None)
                os.dup2(fd, 2)
                os.dup2(fd, 1)  # the licence banner goes to stdout
                os.close(fd)
                try:
                    res = lib.mcDispatch(ctx, flags | MC_DISPATCH_VERTEX_ARRAY_DOUBLE, sv, sf, ss, snv, snf, cv, cf,
                                         cs, cnv, cnf)
                    if res == 0:
                        lib.mcReleaseConnectedComponents(ctx, 0, None)
                finally:
                    libc.fflush(None)
                    os.dup2(saved[0], 1)
                    os.dup2(saved[1], 2)
                st = parse_stage_ms(open(log).read())
                for f in os.listdir(tmp):  # the dumps of this dispatch
                    if f.endswith(".off"):
                        os.remove(os.path.join(tmp, f))
                if it < warmup:
                    continue
                missing = [s for s in FRONT_END_STAGES if s not in st]
                if missing:
                    raise RuntimeError(f"reference stage timers missing: {missing}")
                totals.append(sum(st[s] for s in FRONT_END_STAGES))
                stages.append({s: st[s] for s in FRONT_END_STAGES})
                results.append(res)
            os.chdir(cwd)
    finally:
        os.chdir(cwd)
        for fd in saved:
            os.close(fd)
        lib.mcReleaseContext(ctx)
    med = statistics.median(totals)
    k = min(range(len(totals)), key=lambda i: abs(totals[i] - med))
    return {"runs_ms": totals, "median_ms": med, "min_ms": min(totals), "max_ms": max(totals),
            "stage_ms": stages[k], "dispatch_result": results[k], "helpers": helpers}

"""The C ABI of libmcgpu without a GPU: every function include/mcgpu.h declares
is exported, the Python binding's structures have the C layout, and the entry
points that need no device behave as documented."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mcgpu.h")


def declared_functions():
    text = open(HEADER).read()
    return re.findall(r"MCGPU_API\s+[\w\s\*]+?\b(mcgpu_\w+)\s*\(", text)


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("mcgpu_create_context", "mcgpu_release_context", "mcgpu_dispatch_frontend",
              "mcgpu_redispatch_perturbed", "mcgpu_get_registry", "mcgpu_registry_device"):
        assert n in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_function():
    from mcut_amd import frontend
    L = frontend.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", frontend.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\b(mcgpu_\w+)\b", out))
    assert exported == set(declared_functions()), exported ^ set(declared_functions())


def _c_layout(tmp_path):
    """sizeof/offsetof of the ABI structures, from the header compiled by gcc."""
    fields = {
        "mcgpu_mesh": ["vertices", "face_indices", "face_sizes", "num_vertices", "num_faces"],
        "mcgpu_dispatch_desc": ["flags", "gp_constant", "enforcement_count", "input_on_device", "keep_tests",
                                "reserved"],
        "mcgpu_frontend_info": ["status", "gp_violation", "num_candidate_pairs", "num_candidate_faces", "num_edges",
                                "num_tests", "num_escalated", "num_registry", "num_raw_pairs", "num_decided_f32",
                                "num_decided_f64", "com",
                                "translation", "eps", "device_ms", "partial_cut"],
        "mcgpu_batch": ["num_pairs", "src", "cut", "src_vertex_offsets", "src_face_offsets", "cut_vertex_offsets",
                        "cut_face_offsets"],
        "mcgpu_server_limits": ["max_cut_vertices", "max_cut_faces", "max_cut_corners", "max_slots", "hbm_budget"],
        "mcgpu_request_result": ["arrival_s", "start_s", "done_s", "status", "partial_cut", "num_registry",
                                 "num_candidate_pairs"],
        "mcgpu_cutpath_info": ["num_edges", "num_groups", "num_candidates", "short_group", "unsupported",
                               "device_ms", "reserved"],
        "mcgpu_tuning": ["wave_leaves", "tree_stride", "queues", "chunk", "broad_share_pct", "candidates",
                         "traversal_ms", "default_traversal_ms", "dispatch_ms", "default_dispatch_ms"],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mcgpu.h"', "int main(void) {"]
    for s, fs in fields.items():
        lines.append(f'printf("{s} sizeof %zu\\n", sizeof({s}));')
        for f in fs:
            lines.append(f'printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    exe = tmp_path / "layout"
    src.write_text("\n".join(lines))
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    res = {}
    for line in out.splitlines():
        a, b, c = line.split()
        res[(a, b)] = int(c)
    return fields, res


def test_struct_layouts_match_the_header(tmp_path):
    from mcut_amd import frontend
    py = {"mcgpu_mesh": frontend.Mesh, "mcgpu_dispatch_desc": frontend.DispatchDesc,
          "mcgpu_frontend_info": frontend.FrontendInfo, "mcgpu_batch": frontend.Batch,
          "mcgpu_server_limits": frontend.ServerLimits, "mcgpu_request_result": frontend.RequestResult,
          "mcgpu_tuning": frontend.Tuning, "mcgpu_cutpath_info": frontend.CutpathInfo}
    fields, res = _c_layout(tmp_path)
    for s, fs in fields.items():
        T = py[s]
        assert C.sizeof(T) == res[(s, "sizeof")], s
        for f in fs:
            assert getattr(T, f).offset == res[(s, f)], (s, f)


def test_result_strings_and_stage_names():
    from mcut_amd import frontend
    L = frontend.lib()
    L.mcgpu_result_string.restype = C.c_char_p
    L.mcgpu_result_string.argtypes = [C.c_int]
    for code in (0, -1, -2, -3, -4, -5):
        s = L.mcgpu_result_string(code).decode()
        assert s and s != "unknown"
    assert L.mcgpu_result_string(7).decode() == "unknown"
    names = [L.mcgpu_stage_name(i).decode() for i in range(13)]
    assert names[0] == "coordinate_sums" and names[11] == "registry" and names[12] == "traversal_kernel" and all(names)
    assert L.mcgpu_stage_name(13) == b"" and L.mcgpu_stage_name(-1) == b""


def test_null_arguments_rejected_without_a_device():
    from mcut_amd import frontend
    L = frontend.lib()
    assert L.mcgpu_create_context(0, None) == -1
    assert L.mcgpu_release_context(None) == -1
    assert L.mcgpu_dispatch_frontend(None, None, None, None, None) == -1
    assert L.mcgpu_redispatch_perturbed(None, None, 0, None) == -1
    assert L.mcgpu_get_registry(None, None, None, None, 0) == -1
    assert L.mcgpu_stage_times(None, None, 0) == 0
    assert L.mcgpu_autotune(None, None) == -1
    assert L.mcgpu_set_autotune(None, 1) == -1
    assert L.mcgpu_set_tuning(None, None) == -1
    assert L.mcgpu_get_tuning(None, None) == -1
    assert L.mcgpu_set_graphs(None, 1) == -1
    assert L.mcgpu_set_batch_split(None, 50) == -1
    assert L.mcgpu_cutpath_edges(None, None, None) == -1
    assert L.mcgpu_dispatch_frontend_async(None, None, None, None, None) == -1
    assert L.mcgpu_dispatch_wait(None, None, None) == -1
    assert L.mcgpu_dispatch_query(None) == 1
    assert L.mcgpu_get_cutpath_edges(None, None, None, None, None, None, 0) == -1
    assert L.mcgpu_graph_stats(None, None, None, None) == -1


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is present")
def test_context_without_a_device_is_a_device_error():
    from mcut_amd import frontend
    L = frontend.lib()
    h = C.c_void_p()
    assert L.mcgpu_create_context(0, C.byref(h)) == -2
    assert not h.value

"""The triplet ranking's sort (math.cpp:373-381 sorts with std::sort) as the
GPU runs it (mcut_amd/csrc/tripsort.hpp, compiled for the host) and as the
oracle restates it, against libstdc++'s std::sort itself on inputs that reach
the introsort's every path, its heap-sort fallback included (McIlroy's
adversary); and the oracle's triplet choice for polygons of 17-40 and 256-330
vertices against the reference's determine_three_noncollinear_vertices
(golden vectors written by tests/golden/make_golden.py)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import meshes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC) or shutil.which("gcc") is None, reason="hipcc and gcc are needed")
def test_sort_replay_equals_libstdcxx_std_sort(tmp_path):
    obj = tmp_path / "oracle.o"
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-c", os.path.join(ROOT, "oracle", "frontend_oracle.c"),
                    "-o", str(obj)], check=True)
    chk = tmp_path / "tripsort_check.o"
    subprocess.run([HIPCC, "-std=c++17", "-O1", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "mcut_amd", "csrc"),
                    "-c", os.path.join(ROOT, "tests", "support", "tripsort_check.cpp"), "-o", str(chk)],
                   check=True, capture_output=True)
    exe = tmp_path / "tripsort_check"
    subprocess.run([HIPCC, str(chk), str(obj), "-lm", "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    f = out.stdout.split()
    assert out.returncode == 0 and f[0] == "cases" and f[3] == "0", out.stdout
    assert int(f[1]) > 150 and int(f[5]) > 50  # the heap-sort fallback was exercised
    assert f[6] == "first_mismatches" and f[7] == "0" and f[8] == "pair_mismatches" and f[9] == "0", out.stdout


@pytest.mark.parametrize("fixture,nmin,nmax", [("ref_big_polygons.npz", 17, 40), ("ref_huge_polygons.npz", 256, 300)])
def test_big_polygon_triplets_against_reference(oracle_lib, fixture, nmin, nmax):
    """Polygons of 17-40 vertices, and of 256-330 (past 8-bit vertex indices,
    up to 6 M triplets each)."""
    from oracle import oracle as om
    G = np.load(os.path.join(meshes.GOLDEN, fixture))
    L = oracle_lib.lib()
    n_all = G["n"]
    assert n_all.min() >= nmin and n_all.max() >= nmax
    for i, n in enumerate(n_all):
        P = np.ascontiguousarray(G["poly"][i][:n])
        nrm, d, k = om.plane(P)
        assert np.array_equal(nrm, G["normal"][i]) and k == G["k"][i], i
        ijk = np.zeros(3, np.int32)
        ok = L.om_best_triple(P.ctypes.data, int(n), np.ascontiguousarray(nrm).ctypes.data, int(k), ijk.ctypes.data)
        assert ok == G["ok"][i] and tuple(ijk) == tuple(G["triple"][i]), (i, int(n), tuple(ijk), tuple(G["triple"][i]))

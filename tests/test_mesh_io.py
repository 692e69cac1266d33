"""Mesh files (the output-format row: OFF/OBJ ingest for fixtures, and the
connected components MCUT's VERTEX_DOUBLE / FACE / FACE_SIZE queries return,
written out): OFF laid out as hmesh.cpp:1268-1340 writes it, OBJ with 1-based,
negative and v/t/n indices; every round trip gives back the same doubles and
faces."""
import os

import numpy as np
import pytest

import meshes
from mcut_amd import meshgen as mg


def same(a, b):
    (V0, F0, S0), (V1, F1, S1) = a, b
    assert np.array_equal(np.asarray(V0).view(np.uint64), np.asarray(V1).view(np.uint64))
    assert np.array_equal(F0, F1)
    assert (S0 is None and S1 is None) or np.array_equal(S0, S1)


def random_polygons(seed=3):
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((50, 3)) * 10.0 ** rng.integers(-20, 20, size=(50, 1))
    S = rng.integers(3, 9, size=30).astype(np.uint32)
    F = rng.integers(0, 50, size=int(S.sum())).astype(np.uint32)
    return V, F, S


@pytest.mark.parametrize("fmt", ["off", "obj"])
def test_round_trip_is_exact(tmp_path, fmt):
    write, read = {"off": (mg.write_off, mg.read_off), "obj": (mg.write_obj, mg.read_obj)}[fmt]
    for m in (random_polygons(), meshes.bench_pair(2)[0], mg.icosphere(2)):
        path = str(tmp_path / f"m.{fmt}")
        write(path, *m)
        same(m, read(path))


def test_obj_index_forms(tmp_path):
    path = tmp_path / "forms.obj"
    path.write_text("# a quad and a triangle\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
                    "f 1/1/1 2/1/1 3//1 4\nf -4 -2 -1\ng ignored\n")
    V, F, S = mg.read_obj(str(path))
    assert V.shape == (4, 3) and np.array_equal(S, [4, 3]) and np.array_equal(F, [0, 1, 2, 3, 0, 2, 3])


def test_reference_meshes_read_the_same_as_off_and_obj(tmp_path):
    for i in (0, 2, 14, 33):
        for m in meshes.bench_pair(i):
            path = str(tmp_path / "m.obj")
            mg.write_obj(path, *m)
            same(m, mg.read_obj(path))


def test_connected_components_written_and_read_back(tmp_path):
    """MCUT's components (unchanged build) as OFF and OBJ files: the files
    read back to the arrays the queries returned."""
    from mcut_amd.seam import mcut_api
    ref_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                           "libmcut_ref.so")
    if not os.path.exists(ref_lib):
        pytest.skip("the reference build is missing")
    ccs = mcut_api.Mcut(ref_lib).dispatch(*meshes.config1_rotated())
    assert len(ccs) > 2
    for fmt, read in (("off", mg.read_off), ("obj", mg.read_obj)):
        paths = mcut_api.write_components(ccs, str(tmp_path), fmt)
        for cc, p in zip(ccs, paths):
            sz = cc["face_sizes"]
            same((cc["vertices"], cc["faces"], None if np.all(sz == 3) else sz), read(p))


@pytest.mark.gpu
def test_registry_exported_without_a_copy():
    """mcgpu_registry_device hands out the device arrays the getters copy
    from: read through them, they equal mcgpu_get_registry's output."""
    import ctypes as C
    from mcut_amd import frontend
    from test_zz_no_alloc import Hip
    hip = Hip()
    ctx = frontend.Context(0)
    try:
        for name, pair in (("config1_rotated", meshes.config1_rotated()), ("bench014", meshes.bench_pair(14))):
            print(f"mesh_io: registry export of {name}", flush=True)
            R = ctx.dispatch(*pair)
            e, g, p = R.registry()
            pe, pg, pp, n = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_uint64()
            rc = frontend.lib().mcgpu_registry_device(ctx.h, C.byref(pe), C.byref(pg), C.byref(pp), C.byref(n))
            assert rc == 0 and n.value == len(e) > 0
            de = hip.download(pe.value, np.zeros_like(e))
            dg = hip.download(pg.value, np.zeros_like(g))
            dp = hip.download(pp.value, np.zeros_like(p))
            assert np.array_equal(de, e) and np.array_equal(dg, g)
            assert np.array_equal(dp.view(np.uint64), p.view(np.uint64))
    finally:
        ctx.close()
